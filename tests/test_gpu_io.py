"""Output and restart through the C ABI against the oracle
(src/Field/Field.hxx:403-522, src/Mesh/Mesh.cpp:224-365): the library's
EnSight files are byte for byte the oracle's on periodic, walled and
stretched grids; reading a file back gives the oracle's float32 values; the
FP64 checkpoint restores fields bit for bit; a field overwritten right
after its write call is written as it was at the call."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
TWO_PI = 2.0 * math.pi


@pytest.fixture(scope="module")
def nv():
    import torch
    import novaclone_amd as nv
    assert torch.cuda.is_available(), "GPU tests need a device"
    nv.lib()
    return nv


def dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


CASES = [
    dict(N=(12, 10, 8), L=(TWO_PI, 2.0, 3.0), wall=False, hyp=None),
    dict(N=(9, 14, 11), L=(1.0, math.pi, 2.0), wall=True, hyp=None),
    dict(N=(8, 6, 16), L=(1.0, 1.0, 2.0), wall=False, hyp=(0.02, 0.03)),
    dict(N=(70, 6, 130), L=(1.0, 1.0, 2.0), wall=True, hyp=None),  # several 64-node tiles
]


def meshes(nv, orc, case, origin=(0.0, 0.0, 0.0)):
    bc = ("periodic", "wall" if case["wall"] else "periodic", "periodic")
    G = nv.Mesh(case["N"], case["L"], bc=bc, origin=origin, hyp=case["hyp"])
    O = orc.OMesh(case["N"], case["L"], origin=origin, wall_y=case["wall"], hyp=case["hyp"])
    return G, O


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c['N']}-wall{int(c['wall'])}-hyp{int(bool(c['hyp']))}")
def test_ensight_files_match_oracle_bytes(nv, orc, tmp_path, case):
    print("case", case, flush=True)
    G, O = meshes(nv, orc, case, origin=(0.5, -0.25, 0.125))
    rng = np.random.default_rng(1)
    U = rng.standard_normal((3, *O.stored)) * np.array([1.0, 1e-4, 3e3])[:, None, None, None]
    U[0, 0, 0, :3] = (1e-40, -3e-44, 1e39)
    if case["wall"]:
        U[1, :, 0, :] = U[1, :, -1, :] = 0.0
    p = rng.standard_normal((1, *O.stored))
    t = nv.Time(0.1, 1.0, 2)
    mine, ref = tmp_path / "lib", tmp_path / "orc"
    mine.mkdir()
    ref.mkdir()
    io = nv.IO(G)
    Ud, pd = dev(U), dev(p)
    for n in range(5):
        t.advance()
        io.write_mesh(t, mine)
        io.write_field(t, mine, "U", Ud)
        io.write_field(t, mine, "p", pd)
    io.wait()
    Uo, po = O.from_compact(U), O.from_compact(p)
    for fn in range(1, t.fileNum() + 1):
        with np.errstate(over="ignore"):
            O.field_write(Uo, ref, "U", fn)
            O.field_write(po, ref, "p", fn)
    O.mesh_write(ref, t.steps())
    names = sorted(x.name for x in ref.iterdir())
    assert sorted(x.name for x in mine.iterdir()) == names
    assert len(names) == 3 + 2 * 3
    for nm in names:
        assert (mine / nm).read_bytes() == (ref / nm).read_bytes(), nm
    io.close()


@pytest.mark.parametrize("case", CASES[:2], ids=["periodic", "wall"])
def test_read_field_matches_oracle(nv, orc, tmp_path, case):
    G, O = meshes(nv, orc, case)
    rng = np.random.default_rng(2)
    U = rng.standard_normal((3, *O.stored))
    if case["wall"]:  # the wall-normal velocity is odd: zero on the walls
        U[1, :, 0, :] = U[1, :, -1, :] = 0.0
    Uo = O.from_compact(U)
    O.field_write(Uo, tmp_path, "U", 3)
    io = nv.IO(G)
    Ud = dev(np.zeros_like(U))
    io.read_field(tmp_path, "U", 3, Ud)
    back = O.field(3)
    O.field_read(back, tmp_path, "U", 3)
    assert np.array_equal(host(Ud), O.interior(back))
    assert np.array_equal(host(Ud), U.astype(np.float32).astype(np.float64))
    with pytest.raises(nv.NovaError):
        io.read_field(tmp_path, "U", 4, Ud)  # no such file
    io.close()


def test_checkpoint_round_trip_is_bitwise(nv, tmp_path):
    import torch
    G = nv.Mesh((24, 16, 20), (TWO_PI, 2.0, 3.0), bc=("periodic", "wall", "periodic"))
    gen = torch.Generator(device="cuda").manual_seed(4)
    U = torch.randn((3, *G.shape), generator=gen, dtype=torch.float64, device="cuda")
    p = torch.randn((1, *G.shape), generator=gen, dtype=torch.float64, device="cuda")
    io = nv.IO(G)
    io.checkpoint_write(tmp_path / "c.ckpt", [U, p], time=1.25, step=42)
    io.wait()
    U2, p2 = torch.zeros_like(U), torch.zeros_like(p)
    t, st = io.checkpoint_read(tmp_path / "c.ckpt", [U2, p2])
    assert (t, st) == (1.25, 42)
    assert torch.equal(U2, U) and torch.equal(p2, p)
    with pytest.raises(nv.NovaError):  # component counts differ
        io.checkpoint_read(tmp_path / "c.ckpt", [p2, U2])
    H = nv.Mesh((24, 16, 21), (TWO_PI, 2.0, 3.0), bc=("periodic", "wall", "periodic"))
    ioh = nv.IO(H)
    with pytest.raises(nv.NovaError):  # another grid
        ioh.checkpoint_read(tmp_path / "c.ckpt", [H.field(3), H.field(1)])
    ioh.close()
    io.close()


def test_write_sees_the_field_as_it_was_at_the_call(nv, orc, tmp_path):
    import torch
    N = (32, 24, 40)
    G = nv.Mesh(N, (TWO_PI,) * 3)
    O = orc.OMesh(N, (TWO_PI,) * 3)
    gen = torch.Generator(device="cuda").manual_seed(5)
    U = torch.randn((3, *G.shape), generator=gen, dtype=torch.float64, device="cuda")
    U0 = U.cpu().numpy()
    t = nv.Time(0.1, 1.0, 1)
    t.advance()
    io = nv.IO(G)
    for _ in range(4):  # more writes than staging slots, each followed by an overwrite
        io.write_field(t, tmp_path, "U", U)
        U.mul_(-2.0)
        t.advance()
    io.wait()
    for q in range(4):
        O.field_write(O.from_compact(U0 * (-2.0) ** q), tmp_path, "R", q + 1)
        assert (tmp_path / f"U.0000.{q + 1:05d}.dat").read_bytes() == (tmp_path / f"R.0000.{q + 1:05d}.dat").read_bytes()
    io.close()


def test_io_rejects_bad_arguments(nv, tmp_path):
    import torch
    G = nv.Mesh((16, 16, 16), (TWO_PI,) * 3)
    io = nv.IO(G)
    t = nv.Time(0.1, 1.0, 1)
    t.advance()
    F = torch.zeros((2, 16, 16, 16), dtype=torch.float64, device="cuda")
    with pytest.raises(nv.NovaError):
        io.write_field(t, tmp_path, "F", F)  # 2 components
    io.write_field(t, tmp_path / "absent_dir", "U", G.field(3))
    with pytest.raises(nv.NovaError):  # the worker's error comes back at wait
        io.wait()
    io.wait()  # reported once
    io.close()
