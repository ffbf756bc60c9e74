"""Parity of the compact sixth-order path (gfx950, through the C ABI) with
the oracle's compact scheme (no reference counterpart; the oracle is pinned
by closed forms in test_oracle_compact.py).  Bar: relative L2 <= 1e-12."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TWO_PI = 2.0 * math.pi
RTOL = 1e-12


def rel_l2(a, b):
    nb = np.linalg.norm(np.ravel(b))
    return np.linalg.norm(np.ravel(a - b)) / (nb if nb > 0 else 1.0)


@pytest.fixture(scope="module")
def nv():
    import torch
    import novaclone_amd as nv
    assert torch.cuda.is_available()
    nv.lib()
    return nv


def dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


SHAPES = [((24, 16, 32), (TWO_PI, 2.0, 3.0)), ((48, 64, 96), (TWO_PI,) * 3), ((128, 8, 256), (4.0, 1.0, 8.0))]


@pytest.mark.parametrize("N,L", SHAPES)
def test_compact_derivatives_match_oracle(nv, orc, N, L):
    rng = np.random.default_rng(sum(N))
    G = nv.Mesh(N, L)
    O = orc.OMesh(N, L)
    c = rng.standard_normal((3, *O.stored))
    f = O.from_compact(c, -1)
    F = dev(c)
    for d in range(3):
        for order in (1, 2):
            ref = O.interior(O.compact_deriv(f, d, order))
            got = host(nv.compact_deriv(G, F, d, order))
            r = rel_l2(got, ref)
            print(N, "dir", d, "order", order, "rel", r, flush=True)
            assert r <= 1e-13


def tg_noise(O, seed, amp=1e-2):
    rng = np.random.default_rng(seed)
    x = [np.arange(O.stored[a]) * (O.L[a] / O.N[a]) for a in range(3)]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    c = np.stack([np.sin(X) * np.cos(Y) * np.cos(Z), -np.cos(X) * np.sin(Y) * np.cos(Z), 0 * X])
    return c + amp * rng.standard_normal(c.shape)


@pytest.mark.parametrize("N,L", SHAPES[:2])
def test_compact_projection_matches_oracle(nv, orc, N, L):
    G = nv.Mesh(N, L)
    O = orc.OMesh(N, L)
    P = nv.Poisson(G, scheme=nv.SCHEME_COMPACT6)
    OP = orc.OPoisson(O, compact=True)
    c = tg_noise(O, 3, amp=0.3)
    F = O.from_compact(c)
    p = O.field(1)
    OP.project(F, 0.7, p)
    U = dev(c)
    pg = G.field(1)
    P.project(U, 0.7, pg)
    assert rel_l2(host(U), O.interior(F)) <= RTOL
    assert rel_l2(host(pg), O.interior(p)) <= RTOL
    assert O.compact_maxabs_div(O.from_compact(host(U))) < 1e-11


@pytest.mark.parametrize("N,L", SHAPES[:2])
def test_compact_ns_step_matches_oracle(nv, orc, N, L):
    import torch
    nu, dt = 1e-2, 2e-3
    G = nv.Mesh(N, L)
    O = orc.OMesh(N, L)
    P = nv.Poisson(G, scheme=nv.SCHEME_COMPACT6)
    OP = orc.OPoisson(O, compact=True)
    c = tg_noise(O, 5)
    U0 = O.from_compact(c)
    OP.project(U0, 1.0)
    c = O.interior(U0)
    S = nv.Solver(P, nv.MODEL_NS, nv.RK_SSP3, nu=nu)
    Sg = nv.Solver(P, nv.MODEL_NS, nv.RK_SSP3, nu=nu, graph=True)
    U = dev(c)
    Ug = dev(c)
    p = G.field(1)
    pg = G.field(1)
    Uo = U0.copy()
    po = O.field(1)
    for step in range(3):
        OP.compact_ns_step(Uo, po, nu, dt)
        S.step(dt, 1, U, p)
        r = rel_l2(host(U), O.interior(Uo))
        print(N, "step", step + 1, "rel", r, flush=True)
        assert r <= RTOL
        assert rel_l2(host(p), O.interior(po)) <= 1e-11
    Sg.step(dt, 3, Ug, pg)
    torch.cuda.synchronize()
    assert torch.equal(Ug, U)


def test_compact_scheme_rejections(nv):
    Gw = nv.Mesh((16, 16, 16), (TWO_PI,) * 3, bc=("periodic", "wall", "periodic"))
    with pytest.raises(nv.NovaError):
        nv.Poisson(Gw, scheme=nv.SCHEME_COMPACT6)
    G = nv.Mesh((16, 16, 16), (TWO_PI,) * 3)
    S = nv.Solver(nv.Poisson(G), nv.MODEL_NS, nv.RK_SSP3, nu=1e-2)
    # the solver's Poisson solver is explicit: the scheme change is refused
    assert nv.lib().nova_solver_set_scheme(S._h, nv.SCHEME_COMPACT6) != 0
    with pytest.raises(nv.NovaError):  # compact stages are SSP-RK3 only
        nv.Solver(nv.Poisson(G, scheme=nv.SCHEME_COMPACT6), nv.MODEL_NS, nv.RK_LITERAL4, nu=1e-2)


@pytest.mark.parametrize("N", [(24, 16, 32), (256, 12, 16), (12, 256, 16), (16, 12, 256)],
                         ids=lambda n: "x".join(map(str, n)))
def test_compact_mhd_step_matches_oracle(nv, orc, N):
    """Orszag-Tang type fields with noise; SSP-RK3 with both fields projected
    at every stage, compact operators throughout.  A 256-point direction
    runs that pass with four points per lane (the 256^3 shape)."""
    L = (TWO_PI,) * 3
    nu, eta, dt = 1e-2, 2e-2, 2e-3
    G = nv.Mesh(N, L)
    O = orc.OMesh(N, L)
    P = nv.Poisson(G, scheme=nv.SCHEME_COMPACT6)
    OP = orc.OPoisson(O, compact=True)
    rng = np.random.default_rng(3)
    x = [np.arange(O.stored[a]) * (L[a] / N[a]) for a in range(3)]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    Uc = np.stack([-np.sin(Y), np.sin(X), 0 * X]) + 1e-2 * rng.standard_normal((3, *O.stored))
    Bc = np.stack([-np.sin(Y), np.sin(2 * X), 0 * X]) + 1e-2 * rng.standard_normal((3, *O.stored))
    Uo, Bo = O.from_compact(Uc), O.from_compact(Bc)
    OP.project(Uo, 1.0)
    OP.project(Bo, 1.0)
    U, B = dev(O.interior(Uo)), dev(O.interior(Bo))
    S = nv.Solver(P, nv.MODEL_MHD, nv.RK_SSP3, nu=nu, eta=eta)
    p, pB = G.field(1), G.field(1)
    po, pBo = O.field(1), O.field(1)
    for step in range(2):
        OP.compact_mhd_step(Uo, Bo, po, pBo, nu, eta, dt)
        S.step(dt, 1, U, p, B, pB)
        ru, rb = rel_l2(host(U), O.interior(Uo)), rel_l2(host(B), O.interior(Bo))
        print("compact MHD step", step + 1, "rel U", ru, "rel B", rb, flush=True)
        assert ru <= RTOL and rb <= RTOL
    assert O.compact_maxabs_div(O.from_compact(host(B))) < 1e-11
