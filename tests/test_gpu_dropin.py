"""The C++ drop-in API (novaclone_amd/host/nova/nova.hpp) driving the
reference's TG_Poisson time loop (frozen-stage RK4 + one projection) must
reproduce the oracle's kinetic-energy history."""
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent


def test_tg_poisson_app_matches_oracle(orc, tmp_path):
    exe = ROOT / "novaclone_amd" / "host" / "bin" / "tg_poisson"
    assert exe.exists(), "build first (__graft_entry__.build)"
    N, steps, dt, nu = 32, 5, 2e-3, 1.0 / 1600.0
    out = tmp_path / "dat.dat"
    print("running", exe, flush=True)
    r = subprocess.run([str(exe), str(N), str(steps), repr(dt), repr(nu), str(out)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.loadtxt(out)
    assert got.shape == (steps, 3)
    O = orc.OMesh((N, N, N), (2 * math.pi,) * 3)
    x = np.arange(N) * (2 * math.pi / N)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    U = O.from_compact(np.stack([np.sin(X) * np.cos(Y) * np.cos(Z), -np.cos(X) * np.sin(Y) * np.cos(Z), 0 * X]))
    P = orc.OPoisson(O)
    p = O.field(1)
    ke = []
    for _ in range(steps):
        P.ns_step(U, p, nu, dt, orc.RK_LITERAL4)
        ke.append(O.ke(U))
    assert np.allclose(got[:, 0], dt * np.arange(1, steps + 1), rtol=1e-12)
    assert np.max(np.abs(got[:, 1] - ke) / np.array(ke)) < 1e-12


@pytest.mark.parametrize("N", [16, 20])
def test_field_value_semantics(tmp_path, N):
    """The drop-in's Field assignment keeps the reference's value semantics
    (src/Field/Field.h:160-367: every operator returns a new temporary,
    assignment copies) while an unshared expression result hands over its
    buffer: a named FieldPtr keeps its value after being assigned from, an
    expression that reads its own target (W = W + W) gives its value, and host
    writes through the page-locked mirror reach the device."""
    exe = ROOT / "novaclone_amd" / "host" / "bin" / "field_semantics"
    assert exe.exists(), "build first (__graft_entry__.build)"
    print("running", exe, N, flush=True)
    r = subprocess.run([str(exe), str(N)], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    assert "field_semantics: ok" in r.stdout


def test_tg_poisson_app_writes_ensight_output(orc, tmp_path):
    # app/TG_Piosson/TG.cpp:78-80: mesh, U and p after every step
    exe = ROOT / "novaclone_amd" / "host" / "bin" / "tg_poisson"
    N, steps, dt, nu = 16, 3, 2e-3, 1.0 / 1600.0
    outdir = tmp_path / "data"
    outdir.mkdir()
    r = subprocess.run([str(exe), str(N), str(steps), repr(dt), repr(nu), str(tmp_path / "ke.dat"), str(outdir)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    names = sorted(x.name for x in outdir.iterdir())
    want = sorted(["geom.p0000.geo", "fd.p0000.case", "restart.p.0000.case"] +
                  [f"{f}.0000.{q:05d}.dat" for f in ("U", "p") for q in range(1, steps + 1)])
    assert names == want
    # the clock records the time before each step: 0, dt, 2 dt
    assert (outdir / "restart.p.0000.case").read_text().split() == ["3", "0", "0.002", "0.004"]
    O = orc.OMesh((N, N, N), (2 * math.pi,) * 3)
    x = np.arange(N) * (2 * math.pi / N)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    U = O.from_compact(np.stack([np.sin(X) * np.cos(Y) * np.cos(Z), -np.cos(X) * np.sin(Y) * np.cos(Z), 0 * X]))
    P = orc.OPoisson(O)
    p = O.field(1)
    ref = tmp_path / "ref"
    ref.mkdir()
    E = N + 1
    for q in range(1, steps + 1):
        P.ns_step(U, p, nu, dt, orc.RK_LITERAL4)
        O.field_write(U, ref, "U", q)
        got = np.fromfile(outdir / f"U.0000.{q:05d}.dat", np.float32, offset=244).reshape(3, E, E, E)
        exp = np.fromfile(ref / f"U.0000.{q:05d}.dat", np.float32, offset=244).reshape(3, E, E, E)
        assert (outdir / f"U.0000.{q:05d}.dat").read_bytes()[:244] == (ref / f"U.0000.{q:05d}.dat").read_bytes()[:244]
        # the states agree to 1e-12; their float32 roundings to one float32 ulp
        assert np.max(np.abs(got - exp)) <= 2.0 ** -23 * np.max(np.abs(exp))


def _ot_ic(O):
    # app/OT_Poisson/OT.cpp:73-88
    x = np.arange(O.stored[0]) * (2 * math.pi / O.N[0])
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    u = np.stack([-2.0 * np.sin(Y), 2.0 * np.sin(X), 0 * X])
    b = np.stack([(-2.0 * np.sin(2.0 * Y)) + np.sin(Z), (2.0 * np.sin(X)) + np.sin(Z), np.sin(X) + np.sin(Y)])
    return u, 0.8 * b


def test_ot_poisson_app_matches_oracle(orc, tmp_path):
    """app/OT_Poisson as written (UEqn.H + BEqn.H with B&B, the RK4a loops,
    the cleaning and J = curl(B)) through the drop-in API: its Ek, Em, Jav,
    Jmax, omega, epsilon, Umax history against the oracle's restatement of the
    same step (mode 0) and OT.cpp's diagnostics, 1e-10 relative."""
    exe = ROOT / "novaclone_amd" / "host" / "bin" / "ot_poisson"
    N, steps, dt, mu, eta = 16, 4, 1e-3, 0.057, 0.057
    out = tmp_path / "dat.dat"
    print("running", exe, flush=True)
    r = subprocess.run([str(exe), str(N), str(steps), repr(dt), repr(mu), repr(eta), str(out)], capture_output=True,
                       text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    assert "CFL x:" in r.stdout
    got = np.loadtxt(out)
    assert got.shape == (steps, 10)
    O = orc.OMesh((N, N, N), (2 * math.pi,) * 3)
    u, b = _ot_ic(O)
    U, B = O.from_compact(u), O.from_compact(b)
    P = orc.OPoisson(O)
    p, pB = O.field(1), O.field(1)
    for s in range(steps):
        P.mhd_step(U, B, p, pB, mu, eta, dt, 0)
        d = O.mhd_diag(U, B, mu, eta)
        # columns: t Ek Em Jav Jmax omega epsilon -dEk/dt dEm/dt Umax
        exp = [d[0], d[1], d[2], d[3], d[4], d[5]]
        assert np.max(np.abs(got[s, 1:7] - exp) / np.abs(exp)) <= 1e-10, (s, got[s], d)
        assert abs(got[s, 9] - d[6]) <= 1e-10 * d[6]


def test_tg_wc_app_matches_oracle(orc, tmp_path):
    """app/TG as written (rhoEqn.H with -ex::div(rho*U), pEqn.H's
    50.0*(rho*rho), UEqn.H with rhoUOld/rho + c*(drhoU/rho)) through the
    drop-in API: its kinetic-energy history against the oracle's
    weakly-compressible step, 1e-10 relative."""
    exe = ROOT / "novaclone_amd" / "host" / "bin" / "tg_wc"
    N, steps, dt, mu = 16, 4, 1e-3, 1.0 / 1600.0
    out = tmp_path / "dat.dat"
    r = subprocess.run([str(exe), str(N), str(steps), repr(dt), repr(mu), str(out)], capture_output=True,
                       text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    got = np.loadtxt(out)
    assert got.shape == (steps, 3)
    O = orc.OMesh((N, N, N), (2 * math.pi,) * 3)
    pi = 3.1415926536  # TG.cpp:25
    x = -pi + np.arange(N) * (2 * math.pi / N)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    U = O.from_compact(np.stack([np.sin(X) * np.cos(Y) * np.cos(Z), -np.cos(X) * np.sin(Y) * np.cos(Z), 0 * X]))
    pc = 50 + 1.0 / 16.0 * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2)
    p = O.from_compact(pc[None])
    rho = O.from_compact(np.power(pc / 50.0, 0.5)[None])
    for s in range(steps):
        O.wc_step(U, rho, p, mu, 50.0, dt)
        ek = O.ke(U)
        assert abs(got[s, 1] - ek) <= 1e-10 * ek, (s, got[s, 1], ek)


def test_bc_files_are_checked_against_the_device_rule(tmp_path):
    # Field::setBC (Field.hxx:519-547) reads <zone>/<name>BC; an entry the
    # device cannot apply is an error, not a silent change of the problem
    exe = ROOT / "novaclone_amd" / "host" / "bin" / "ot_poisson"
    lines = [f"{d} periodic 0 {c}" for c in range(3) for d in ("north", "south", "east", "west", "top", "bottom")]
    (tmp_path / "UBC").write_text("\n".join(lines) + "\n")
    (tmp_path / "BBC").write_text("\n".join(lines) + "\n")
    r = subprocess.run([str(exe), "8", "1", "1e-3", "0.05", "0.05", "d.dat"], capture_output=True, text=True,
                       timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    (tmp_path / "BBC").write_text("north neumann 0 0\n")
    r = subprocess.run([str(exe), "8", "1", "1e-3", "0.05", "0.05", "d.dat"], capture_output=True, text=True,
                       timeout=120, cwd=tmp_path)
    assert r.returncode == 1 and "BBC" in r.stderr and "periodic" in r.stderr, r.stderr


def test_mhdjet_app_hdf5_dumps_match_oracle(orc, tmp_path):
    """app/mhdjet as written (the OT_Poisson step with the jet's state) through
    the drop-in API, with its HDF5 dumps at every write step: the files carry
    the reference's names, and their /3DArray contents equal the oracle's
    state (U, p, B, J = literal curl B), started from the driver's own initial
    state (its random draws).  Bar: 1e-12 of the field's norm per dumped
    component (the components U_y, U_z, B_y, B_z are 1e-4 .. 1e-8 of their
    field, so their own relative differences are the field's rounding
    amplified); for J = curl B, whose size is that of the small part of B,
    the scale is |J| + |B| / h (the curl of B's rounding); for p, which the
    correction applies as dt grad p, |p| + |U| h / dt (a difference of p on
    that scale moves U by 1e-12 of |U|)."""
    import novaclone_amd as nv
    exe = ROOT / "novaclone_amd" / "host" / "bin" / "mhdjet"
    N, steps, dt = (16, 32, 16), 2, 0.02
    mu = eta = 0.000285714
    ic = tmp_path / "ic"
    ic.mkdir()
    r = subprocess.run([str(exe), *map(str, N), str(steps), repr(dt), "1", str(ic)], capture_output=True, text=True,
                       timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    O = orc.OMesh(N, (13.96, 20.0, 13.96))
    u0 = np.stack([nv.hdf5_read_host(ic / f"ic_U{c}.h5") for c in "xyz"])
    b0 = np.stack([nv.hdf5_read_host(ic / f"ic_B{c}.h5") for c in "xyz"])
    assert u0.shape == (3, *N) and np.all(b0[0] == 0.015) and np.all(b0[1:] == 0.0)
    U, B = O.from_compact(u0), O.from_compact(b0)
    P = orc.OPoisson(O)
    p, pB = O.field(1), O.field(1)
    h5 = tmp_path / "h5data"
    for s in range(steps):
        P.mhd_step(U, B, p, pB, mu, eta, dt, 0)
        J = O.curl_literal(B)
        t = f"{s * dt:.6f}.h5"
        Ui, Bi, Ji, pi = O.interior(U), O.interior(B), O.interior(J), O.interior(p)
        hmin = min(O.L[a] / N[a] for a in range(3))
        scale = {"U": np.linalg.norm(Ui), "B": np.linalg.norm(Bi), "p": np.linalg.norm(pi) + np.linalg.norm(Ui) * hmin / dt,
                 "J": np.linalg.norm(Ji) + np.linalg.norm(Bi) / hmin}
        ref = {"Ux": Ui[0], "Uy_": Ui[1], "Uz_": Ui[2], "p_": pi[0], "Bx_": Bi[0], "By_": Bi[1], "Bz_": Bi[2],
               "Jx_": Ji[0], "Jy_": Ji[1], "Jz_": Ji[2]}
        for name, exp in ref.items():
            got = nv.hdf5_read_host(h5 / f"{name}{t}")
            e = np.linalg.norm(got - exp) / scale[name[0]]
            print(name, t, e, flush=True)
            assert got.shape == N and e <= 1e-12, (name, t, e)


def test_shear_layer_app_matches_oracle(orc, tmp_path):
    """app/shearLayer as written, from a zone holding the example's mesh and
    BC files (examples/shearLayer), through the drop-in API: Field::setBC
    installs the reference's wall fills, poisson takes the reference's FFT
    solve, UEqn.H runs operator by operator; U and p after three steps
    against the oracle's restatement, 1e-11 relative L2."""
    from test_gpu_shear_literal import BBC, DT, L, MU, N, PBC, UBC, initial_state
    exe = ROOT / "novaclone_amd" / "host" / "bin" / "shear_layer"
    (tmp_path / "mesh").write_text(f"origin: 0 0 0\nN: {N[0]} {N[1]} {N[2]} \nL: {L[0]} {L[1]} {L[2]}\n")
    (tmp_path / "UBC").write_text(UBC)
    (tmp_path / "pBC").write_text(PBC)
    (tmp_path / "BBC").write_text(BBC)
    steps = 3
    out = tmp_path / "state.bin"
    print("running", exe, flush=True)
    r = subprocess.run([str(exe), str(tmp_path), str(steps), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "CFL x:" in r.stdout and "Velocity divergence:" in r.stdout
    s = (N[0], N[1] + 1, N[2])
    got = np.fromfile(out, dtype=np.float64).reshape(4, *s)
    O = orc.OMesh(N, L, wall_y=True)
    P = orc.OPoisson(orc.OMesh(N, L))
    ubc, pbc, bbc = orc.bc_entries(UBC, 3), orc.bc_entries(PBC, 1), orc.bc_entries(BBC, 3)
    U, B, p = initial_state(O)
    O.bc_apply(U, ubc)
    O.bc_apply(p, pbc)
    O.bc_apply(B, bbc)
    for _ in range(steps):
        orc.shear_layer_step(O, P, U, B, p, MU, DT, ubc, pbc)
    u, q = O.interior(U), O.interior(p)
    eu = np.linalg.norm(got[:3] - u) / np.linalg.norm(u)
    ep = np.linalg.norm(got[3] - q[0]) / np.linalg.norm(q[0])
    print("shear_layer app: rel L2 U", eu, "p", ep, flush=True)
    assert eu <= 1e-11 and ep <= 1e-11
