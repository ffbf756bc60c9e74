"""Batched ensemble of uneven instances (no reference counterpart; each
instance follows app/TG_Piosson/RKLoop.H or app/OT_Poisson/{UEqn,BEqn}.H
with the SSP-RK3 composition).  Pinned by construction: every instance must
equal its standalone solver run bit for bit."""
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


def initial(nv, N, model, seed):
    """TG velocity (NS) or Orszag-Tang u, B (MHD) plus 1e-3 noise, projected."""
    import torch
    G = nv.Mesh(N, (TWO_PI,) * 3)
    P = nv.Poisson(G)
    s = G.shape
    x = [torch.arange(s[a], dtype=torch.float64, device="cuda") * (TWO_PI / N[a]) for a in range(3)]
    X, Y, Z = torch.meshgrid(*x, indexing="ij")
    gen = torch.Generator(device="cuda").manual_seed(seed)
    noise = lambda: 1e-3 * torch.randn((3, *s), generator=gen, dtype=torch.float64, device="cuda")
    if model == nv.MODEL_NS:
        U = torch.stack([torch.sin(X) * torch.cos(Y) * torch.cos(Z), -torch.cos(X) * torch.sin(Y) * torch.cos(Z),
                         0 * X]) + noise()
        B = None
    else:
        U = torch.stack([-torch.sin(Y), torch.sin(X), 0 * X]) + noise()
        B = torch.stack([-torch.sin(Y), torch.sin(2 * X), 0 * X]) + noise()
        B = B.contiguous()
        P.project(B, 1.0)
    U = U.contiguous()
    P.project(U, 1.0)
    return G, P, U, B


INSTANCES = [((16, 16, 16), 0, 1 / 400), ((24, 24, 24), 1, 1 / 900), ((16, 16, 16), 1, 1 / 800),
             ((32, 32, 32), 0, 1 / 1200), ((24, 24, 24), 0, 1 / 600), ((16, 16, 16), 0, 1 / 1600)]


@pytest.mark.parametrize("graph", [True, False])
def test_each_instance_equals_its_standalone_run(nv, graph):
    import torch
    dt, steps = 2e-3, 3
    E = nv.Ensemble([(N, (TWO_PI,) * 3, model, nu, nu) for N, model, nu in INSTANCES], graph=graph)
    info = E.info()
    assert info["instances"] == len(INSTANCES) and info["classes"] == 3
    assert info["slots"] == 4 + 2 * 2
    runs = []
    for q, (N, model, nu) in enumerate(INSTANCES):
        G, P, U, B = initial(nv, N, model, 40 + q)
        E.load(q, U, B)
        runs.append((G, P, U.clone(), None if B is None else B.clone(), model, nu))
    E.step(dt, steps)
    for q, (G, P, U, B, model, nu) in enumerate(runs):
        S = nv.Solver(P, model, nv.RK_SSP3, nu=nu, eta=nu)
        p, pB = G.field(1), G.field(1)
        S.step(dt, steps, U, p, B, pB if B is not None else None)
        Ue, Be, pe, pBe = E.store(q)
        torch.cuda.synchronize()
        print("instance", q, G.N, model, "max|dU|", (Ue - U).abs().max().item(), flush=True)
        assert torch.equal(Ue, U)
        assert torch.equal(pe, p)
        if B is not None:
            assert torch.equal(Be, B)
            assert torch.equal(pBe, pB)
    E.close()


def test_ensemble_rejects_bad_descriptions(nv):
    with pytest.raises(nv.NovaError):
        nv.Ensemble([((16,) * 3, (TWO_PI,) * 3, 0, 1e-3, 0.0)], m=5)
    with pytest.raises(nv.NovaError):
        nv.Ensemble([((16,) * 3, (TWO_PI,) * 3, 0, 1e-3, 0.0), ((20,) * 3, (TWO_PI,) * 3, 0, 1e-3, 0.0)])


def test_benchmark_size_classes_equal_their_standalone_runs(nv):
    """The benchmark's radix-3 size classes (48^3 and 96^3, FFT lengths with a
    factor 3) and a 64^3 class, NS and MHD members, two steps: every member
    bit for bit its standalone run under graph replay."""
    import torch
    dt, steps = 2e-3, 2
    inst = [((48, 48, 48), 0, 1 / 1600), ((48, 48, 48), 1, 1 / 1000), ((96, 96, 96), 0, 1 / 1600),
            ((96, 96, 96), 1, 1 / 1000), ((64, 64, 64), 0, 1 / 1600)]
    E = nv.Ensemble([(N, (TWO_PI,) * 3, model, nu, nu) for N, model, nu in inst], graph=True)
    runs = []
    for q, (N, model, nu) in enumerate(inst):
        G, P, U, B = initial(nv, N, model, 70 + q)
        E.load(q, U, B)
        runs.append((G, P, U.clone(), None if B is None else B.clone(), model, nu))
    E.step(dt, steps)
    for q, (G, P, U, B, model, nu) in enumerate(runs):
        S = nv.Solver(P, model, nv.RK_SSP3, nu=nu, eta=nu)
        p, pB = G.field(1), G.field(1)
        S.step(dt, steps, U, p, B, pB if B is not None else None)
        Ue, Be, pe, pBe = E.store(q)
        torch.cuda.synchronize()
        print("instance", q, G.N, model, "max|dU|", (Ue - U).abs().max().item(), flush=True)
        assert torch.equal(Ue, U) and torch.equal(pe, p)
        if B is not None:
            assert torch.equal(Be, B) and torch.equal(pBe, pB)
    E.close()
