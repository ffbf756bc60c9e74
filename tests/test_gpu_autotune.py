"""Launch-shape autotuner and variant equivalence (no reference counterpart:
every launch shape must give the same step bit for bit, and tuning must not
change results or break a recorded graph)."""
import math

import pytest

pytestmark = pytest.mark.gpu
TWO_PI = 2.0 * math.pi


@pytest.fixture(scope="module")
def nv():
    import torch
    import novaclone_amd as nv
    assert torch.cuda.is_available()
    nv.lib()
    return nv


def tg(nv, G):
    import torch
    s = G.shape
    x = [torch.arange(s[a], dtype=torch.float64, device="cuda") * (TWO_PI / s[a]) for a in range(3)]
    X, Y, Z = torch.meshgrid(*x, indexing="ij")
    return torch.stack([torch.sin(X) * torch.cos(Y) * torch.cos(Z), -torch.cos(X) * torch.sin(Y) * torch.cos(Z),
                        0 * X]).contiguous()


def rel(a, b):
    return ((a - b).norm() / b.norm()).item()


@pytest.mark.parametrize("wall", [False, True])
def test_every_stage_variant_gives_the_same_step(nv, wall):
    import torch
    N = (24, 12, 32)
    G = nv.Mesh(N, (TWO_PI, 3.0 if wall else TWO_PI, TWO_PI), bc=("periodic", "wall" if wall else "periodic",
                                                                    "periodic"))
    P = nv.Poisson(G)
    S = nv.Solver(P, nv.MODEL_NS, nv.RK_SSP3, nu=1e-2)
    gen = torch.Generator(device="cuda").manual_seed(11)
    U0 = torch.randn((3, *G.shape), generator=gen, dtype=torch.float64, device="cuda")
    if wall:
        U0[1, :, 0, :] = 0.0
        U0[1, :, -1, :] = 0.0
    P.project(U0, 1.0)
    ref = None
    for v in (3, 0, 1, 2, 4, 5, 6, 7, 8, 9, 10, 11, 12):
        for c in (0, 5):
            G.tune(v, c, 1)
            U = U0.clone()
            S.step(1e-3, 2, U, G.field(1))
            if ref is None:
                ref = U
            else:
                print("variant", v, "chunk", c, "rel", rel(U, ref), flush=True)
                # multiply-adds are fused per source expression, so every
                # launch shape rounds identically
                assert torch.equal(U, ref)


def test_first_stage_slot_gives_the_same_step(nv):
    # the first SSP stage takes its launch shape from a slot of its own
    import torch
    G = nv.Mesh((24, 12, 32), (TWO_PI,) * 3)
    P = nv.Poisson(G)
    S = nv.Solver(P, nv.MODEL_NS, nv.RK_SSP3, nu=1e-2)
    gen = torch.Generator(device="cuda").manual_seed(9)
    U0 = torch.randn((3, *G.shape), generator=gen, dtype=torch.float64, device="cuda")
    P.project(U0, 1.0)
    G.tune(8, 0, 2)
    ref = U0.clone()
    S.step(1e-3, 2, ref, G.field(1))
    for v, c in ((11, 16), (1, 0), (5, 128), (0, 8)):
        G.tune_kernel(nv.KERNEL_NS_FIRST_STAGE, v, c)
        U = U0.clone()
        S.step(1e-3, 2, U, G.field(1))
        assert torch.equal(U, ref), (v, c)
    with pytest.raises(nv.NovaError):
        G.tune_kernel(5, 3, 0)


@pytest.mark.parametrize("wall", [False, True])
def test_two_field_stage_variants_give_the_same_results(nv, wall):
    # MHD, tangent-linear and adjoint stages under every launch shape they
    # take (a variant a model does not have runs its default shape)
    import torch
    N = (24, 12, 32)
    G = nv.Mesh(N, (TWO_PI, 3.0 if wall else TWO_PI, TWO_PI), bc=("periodic", "wall" if wall else "periodic",
                                                                    "periodic"))
    P = nv.Poisson(G)
    gen = torch.Generator(device="cuda").manual_seed(5)
    U0 = torch.randn((3, *G.shape), generator=gen, dtype=torch.float64, device="cuda")
    B0 = torch.randn((3, *G.shape), generator=gen, dtype=torch.float64, device="cuda")
    if wall:
        U0[1, :, 0, :] = U0[1, :, -1, :] = 0.0
        B0[1, :, 0, :] = B0[1, :, -1, :] = 0.0
    P.project(U0, 1.0)
    P.project(B0, 1.0)
    S = nv.Solver(P, nv.MODEL_MHD, nv.RK_SSP3, nu=1e-2, eta=2e-2)
    A = nv.Adjoint(nv.Solver(P, nv.MODEL_NS, nv.RK_SSP3, nu=1e-2), 2, 4)
    ref = None
    for v, c in ((3, 0), (0, 0), (2, 0), (4, 0), (5, 0), (5, 5), (6, 0), (6, 7), (11, 0), (12, 0), (13, 0),
                 (14, 0)):
        G.tune(v, c, 1)
        U, B = U0.clone(), B0.clone()
        S.step(1e-3, 2, U, G.field(1), B, G.field(1))
        A.forward(1e-3, 3, U0)
        lam = A.backward(B0)
        d = A.tlm(B0)
        out = (U, B, lam, d)
        if ref is None:
            ref = tuple(t.clone() for t in out)
        else:
            print("variant", v, "chunk", c, [rel(o, r) for o, r in zip(out, ref)], flush=True)
            assert all(torch.equal(o, r) for o, r in zip(out, ref))


def test_solver_autotune_keeps_results_and_rerecords_graph(nv):
    import torch
    G = nv.Mesh((32, 32, 32), (TWO_PI,) * 3)
    P = nv.Poisson(G)
    U0 = tg(nv, G)
    Se = nv.Solver(P, nv.MODEL_NS, nv.RK_SSP3, nu=1e-2)
    Sg = nv.Solver(P, nv.MODEL_NS, nv.RK_SSP3, nu=1e-2, graph=True)
    Ue, Ug = U0.clone(), U0.clone()
    p = G.field(1)
    Se.step(2e-3, 3, Ue, p)
    Sg.step(2e-3, 3, Ug, p)
    rep = Sg.autotune(U0, reps=2)
    print(rep, flush=True)
    assert "ns_stage chosen variant" in rep and "projection chosen variant" in rep
    assert rep.count("ns_stage variant") == 54
    assert "ns_first_stage chosen variant" in rep and rep.count("ns_first_stage variant") == 66
    Se.step(2e-3, 3, Ue, p)
    Sg.step(2e-3, 3, Ug, p)  # the graph is recorded again with the tuned shapes
    torch.cuda.synchronize()
    assert rel(Ug, Ue) <= 1e-14


def test_mhd_autotune(nv):
    import torch
    G = nv.Mesh((24, 24, 24), (TWO_PI,) * 3)
    P = nv.Poisson(G)
    U = tg(nv, G)
    B = U.flip(0).contiguous()
    S = nv.Solver(P, nv.MODEL_MHD, nv.RK_SSP3, nu=1e-2, eta=1e-2)
    U1, B1 = U.clone(), B.clone()
    S.step(1e-3, 2, U1, G.field(1), B1, G.field(1))
    rep = S.autotune(U, B, reps=1)
    assert "mhd_stage chosen variant" in rep
    U2, B2 = U.clone(), B.clone()
    S.step(1e-3, 2, U2, G.field(1), B2, G.field(1))
    assert rel(U2, U1) <= 1e-14 and rel(B2, B1) <= 1e-14


def test_adjoint_autotune_keeps_gradient(nv):
    import torch
    G = nv.Mesh((24, 24, 24), (TWO_PI,) * 3)
    P = nv.Poisson(G)
    U0 = tg(nv, G)
    A = nv.Adjoint(nv.Solver(P, nv.MODEL_NS, nv.RK_SSP3, nu=1e-2), 2, 6)
    UT = A.forward(2e-3, 5, U0)
    _, g = nv.enstrophy(G, UT, grad=True)
    l1 = A.backward(g)
    d1 = A.tlm(g)
    rep = A.autotune(reps=1)
    assert "adj_stage chosen variant" in rep and "tlm_stage chosen variant" in rep
    l2 = A.backward(g)
    d2 = A.tlm(g)
    assert rel(l2, l1) <= 1e-14 and rel(d2, d1) <= 1e-14


@pytest.mark.parametrize("wall", [False, True])
def test_projection_variants_agree(nv, wall):
    """The LDS-staged divergence/correction (variant 2) evaluates the same
    expressions as the row kernels (variant 1): identical bits; the fused
    divergence (variant 0) agrees to rounding."""
    import torch
    N = (24, 12, 32)
    G = nv.Mesh(N, (TWO_PI, 3.0 if wall else TWO_PI, TWO_PI), bc=("periodic", "wall" if wall else "periodic",
                                                                    "periodic"))
    P = nv.Poisson(G)
    gen = torch.Generator(device="cuda").manual_seed(21)
    U0 = torch.randn((3, *G.shape), generator=gen, dtype=torch.float64, device="cuda")
    if wall:
        U0[1, :, 0, :] = 0.0
        U0[1, :, -1, :] = 0.0
    out = {}
    for pv in (1, 2, 0):
        G.tune(3, 0, pv)
        U = U0.clone()
        p = G.field(1)
        P.project(U, 0.7, p)
        out[pv] = (U, p)
    assert torch.equal(out[2][0], out[1][0]) and torch.equal(out[2][1], out[1][1])
    assert rel(out[0][0], out[1][0]) <= 1e-14
