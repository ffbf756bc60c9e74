"""Discrete adjoint of the SSP-RK3 + projection step (no reference
counterpart; pinned by mathematics): the dot-product identity between the
tangent-linear model and the adjoint, central finite differences of the
enstrophy objective, and the torch.autograd harness.  Tolerances are the
north star's 1e-8 for FD agreement and roundoff-level for the dot product."""
import math

import numpy as np
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


def flow(nv, N, L, wall=False, seed=0, amp=1e-2, m=6):
    import torch
    bc = ("periodic", "wall" if wall else "periodic", "periodic")
    G = nv.Mesh(N, L, m=m, bc=bc)
    P = nv.Poisson(G)
    gen = torch.Generator(device="cuda").manual_seed(seed)
    s = G.shape
    x = [torch.arange(s[a], dtype=torch.float64, device="cuda") * (L[a] / N[a]) for a in range(3)]
    X, Y, Z = torch.meshgrid(*x, indexing="ij")
    if wall:
        U = torch.stack([torch.tanh(2 * (Y - L[1] / 2)), 0 * X, 0 * X])
    else:
        U = torch.stack([torch.sin(X) * torch.cos(Y) * torch.cos(Z), -torch.cos(X) * torch.sin(Y) * torch.cos(Z),
                         0 * X])
    U = U + amp * torch.randn(U.shape, generator=gen, dtype=torch.float64, device="cuda")
    if wall:
        U[1, :, 0, :] = 0.0
        U[1, :, -1, :] = 0.0
    U = U.contiguous()
    P.project(U, 1.0)
    return G, P, U


def rand_like(U, seed, wall=False):
    import torch
    gen = torch.Generator(device="cuda").manual_seed(seed)
    r = torch.randn(U.shape, generator=gen, dtype=torch.float64, device="cuda")
    if wall:
        r[1, :, 0, :] = 0.0
        r[1, :, -1, :] = 0.0
    return r.contiguous()


@pytest.mark.parametrize("wall,nsteps,every", [(False, 1, 1), (False, 7, 3), (True, 5, 2)])
def test_dot_product_tlm_adjoint(nv, wall, nsteps, every):
    N = (24, 16, 32) if not wall else (16, 16, 24)
    L = (TWO_PI,) * 3 if not wall else (4.0, 3.0, 4.0)
    G, P, U0 = flow(nv, N, L, wall=wall, seed=1)
    S = nv.Solver(P, nv.MODEL_NS, nv.RK_SSP3, nu=2e-2)
    A = nv.Adjoint(S, every, 16)
    A.forward(0.01, nsteps, U0)
    d = rand_like(U0, 2, wall)
    m = rand_like(U0, 3, wall)
    Ld = A.tlm(d)
    LTm = A.backward(m)
    lhs = nv.weighted_dot(G, Ld, m)
    rhs = nv.weighted_dot(G, d, LTm)
    print("dot", wall, nsteps, lhs, rhs, flush=True)
    assert abs(lhs - rhs) <= 1e-11 * max(abs(lhs), abs(rhs))


@pytest.mark.parametrize("m,wall", [(2, False), (4, False), (8, False), (2, True), (4, True), (8, True)])
def test_dot_product_tlm_adjoint_other_stencil_widths(nv, m, wall):
    """<L d, m> = <d, L^T m> over three steps (checkpoints every 2) with the
    adjoint stage kernels of the other stencil widths."""
    N = (16, 12, 24)
    L = (TWO_PI,) * 3 if not wall else (4.0, 3.0, 4.0)
    G, P, U0 = flow(nv, N, L, wall=wall, seed=5, m=m)
    S = nv.Solver(P, nv.MODEL_NS, nv.RK_SSP3, nu=2e-2)
    A = nv.Adjoint(S, 2, 8)
    A.forward(0.01, 3, U0)
    d = rand_like(U0, 6, wall)
    mm = rand_like(U0, 7, wall)
    lhs = nv.weighted_dot(G, A.tlm(d), mm)
    rhs = nv.weighted_dot(G, d, A.backward(mm))
    print("dot m", m, "wall", wall, lhs, rhs, flush=True)
    assert abs(lhs - rhs) <= 1e-11 * max(abs(lhs), abs(rhs))


@pytest.mark.parametrize("wall", [False, True])
def test_enstrophy_gradient_matches_central_differences(nv, wall):
    """16 seeded directions.  The 4th-order central difference (eps 1e-4,
    truncation ~eps^4, roundoff ~u J / eps) must agree to 1e-8; the
    2nd-order one at eps = 1e-6 must agree to 1e-8 plus its roundoff bound
    4 u J / eps (u = 2^-53), which is of the same size for J ~ 1e3."""
    import torch
    if wall:
        N, L = (16, 16, 16), (4.0, 3.0, 4.0)
    else:
        N, L = (32, 32, 32), (TWO_PI,) * 3
    G, P, U0 = flow(nv, N, L, wall=wall, seed=5, amp=5e-2)
    S = nv.Solver(P, nv.MODEL_NS, nv.RK_SSP3, nu=1e-2)
    A = nv.Adjoint(S, 4, 16)
    dt, n = 0.01, 10
    U = A.forward(dt, n, U0)
    J, gT = nv.enstrophy(G, U, grad=True)
    g0 = A.backward(gT)
    Jf = lambda X: nv.enstrophy(G, A.forward(dt, n, X.contiguous()))
    u = 2.0 ** -53
    for seed in range(16):
        d = rand_like(U0, 100 + seed, wall)
        d = d / d.abs().max()
        ad = nv.weighted_dot(G, g0, d)
        e4 = 1e-4
        fd4 = (-Jf(U0 + 2 * e4 * d) + 8 * Jf(U0 + e4 * d) - 8 * Jf(U0 - e4 * d) + Jf(U0 - 2 * e4 * d)) / (12 * e4)
        e2 = 1e-6
        fd2 = (Jf(U0 + e2 * d) - Jf(U0 - e2 * d)) / (2 * e2)
        print(f"fd wall={wall} seed={seed} ad={ad:.12e} fd4 rel={abs(fd4 - ad) / abs(ad):.2e} "
              f"fd2 rel={abs(fd2 - ad) / abs(ad):.2e} J={J:.3e}", flush=True)
        assert abs(fd4 - ad) <= 1e-8 * abs(ad)
        assert abs(fd2 - ad) <= 1e-8 * abs(ad) + 4 * u * J / e2


def test_autograd_harness(nv):
    import torch
    N, L = (16, 16, 16), (TWO_PI,) * 3
    G, P, U0 = flow(nv, N, L, seed=7)
    S = nv.Solver(P, nv.MODEL_NS, nv.RK_SSP3, nu=1e-2)
    A = nv.Adjoint(S, 2, 8)
    X = U0.clone().requires_grad_(True)
    J = nv.EnstrophyAfterSteps.apply(X, A, 0.01, 5)
    J.backward()
    U = A.forward(0.01, 5, U0)
    J2, gT = nv.enstrophy(G, U, grad=True)
    g0 = A.backward(gT)
    assert abs(J.item() - J2) <= 1e-14 * abs(J2)
    assert torch.equal(X.grad, g0)


def shear_layer(nv, N, L, seed=5, kmax=4):
    """u = tanh(2y), y in [-L_y/2, L_y/2], plus a seeded random-phase
    perturbation of amplitude 1e-3 on the modes kx, kz <= kmax localised in
    y, projected: the adjoint benchmark's initial state on a smaller grid."""
    import torch
    G = nv.Mesh(N, L, bc=("periodic", "wall", "periodic"))
    P = nv.Poisson(G)
    gen = torch.Generator(device="cpu").manual_seed(seed)
    sx, sy, sz = G.shape
    x = torch.arange(sx, dtype=torch.float64, device="cuda") * (L[0] / N[0])
    y = torch.arange(sy, dtype=torch.float64, device="cuda") * (L[1] / N[1]) - L[1] / 2
    z = torch.arange(sz, dtype=torch.float64, device="cuda") * (L[2] / N[2])
    U = torch.zeros((3, sx, sy, sz), dtype=torch.float64, device="cuda")
    U[0] += torch.tanh(2 * y)[None, :, None]
    env = torch.exp(-y * y)[None, :, None]
    for c in range(3):
        for kx in range(kmax + 1):
            for kz in range(kmax + 1):
                if kx or kz:
                    ph = float(torch.rand(1, generator=gen)) * TWO_PI
                    U[c] += 1e-3 * torch.cos(TWO_PI * kx * x[:, None, None] / L[0]
                                             + TWO_PI * kz * z[None, None, :] / L[2] + ph) * env
    U[1, :, 0, :] = 0.0
    U[1, :, -1, :] = 0.0
    U = U.contiguous()
    P.project(U, 1.0)
    return G, P, U


def test_shear_layer_gradient_midsize(nv):
    """The adjoint benchmark's setting on a 64 x 96 x 64 grid (the same
    spacings, free-slip y walls, Re = 1000, dt = 0.02): 50 SSP-RK3 steps with
    checkpoints every 5, so the reverse sweep recomputes nine segments and
    takes the last one from the forward sweep.  Dot product of the
    tangent-linear model and the adjoint; the enstrophy gradient against
    16-direction central differences (as the small-grid test); a second
    reverse sweep (which recomputes the last segment too) equals the first
    bit for bit."""
    import torch
    N, L = (64, 96, 64), (TWO_PI, 3.0, TWO_PI)
    G, P, U0 = shear_layer(nv, N, L)
    A = nv.Adjoint(nv.Solver(P, nv.MODEL_NS, nv.RK_SSP3, nu=1e-3), 5, 50)
    dt, n = 0.02, 50
    U = A.forward(dt, n, U0)
    d = rand_like(U0, 21, True)
    m = rand_like(U0, 22, True)
    Ld = A.tlm(d)
    LTm = A.backward(m)
    lhs = nv.weighted_dot(G, Ld, m)
    rhs = nv.weighted_dot(G, d, LTm)
    print("dot 64x96x64 50 steps", lhs, rhs, abs(lhs - rhs) / abs(lhs), flush=True)
    assert abs(lhs - rhs) <= 1e-10 * max(abs(lhs), abs(rhs))
    U = A.forward(dt, n, U0)
    J, gT = nv.enstrophy(G, U, grad=True)
    g0 = A.backward(gT)
    g1 = A.backward(gT)
    assert torch.equal(g0, g1)
    Jf = lambda X: nv.enstrophy(G, A.forward(dt, n, X.contiguous()))
    u = 2.0 ** -53
    for seed in range(16):
        dd = rand_like(U0, 300 + seed, True)
        dd = dd / dd.abs().max()
        ad = nv.weighted_dot(G, g0, dd)
        e4 = 1e-4
        fd4 = (-Jf(U0 + 2 * e4 * dd) + 8 * Jf(U0 + e4 * dd) - 8 * Jf(U0 - e4 * dd) + Jf(U0 - 2 * e4 * dd)) / (12 * e4)
        e2 = 1e-6
        fd2 = (Jf(U0 + e2 * dd) - Jf(U0 - e2 * dd)) / (2 * e2)
        print(f"fd 64x96x64 seed={seed} ad={ad:.12e} fd4 rel={abs(fd4 - ad) / abs(ad):.2e} "
              f"fd2 rel={abs(fd2 - ad) / abs(ad):.2e} J={J:.3e}", flush=True)
        assert abs(fd4 - ad) <= 1e-8 * abs(ad)
        assert abs(fd2 - ad) <= 1e-8 * abs(ad) + 4 * u * J / e2


def test_reserved_pool_fields_give_the_same_results(nv):
    """Step-start states kept from the forward sweep (the pool of 3C fields
    from nova_adjoint_create plus those of nova_adjoint_reserve) replace
    recomputed segments bit for bit, in the reverse sweep and the
    tangent-linear model, for every pool size: resident steps starting at a
    segment boundary, inside a segment, every step resident, and the default
    pool; a partial last segment (20 steps, checkpoints every 3)."""
    import torch
    G, P, U0 = shear_layer(nv, (32, 48, 32), (TWO_PI / 2, 1.5, TWO_PI / 2), kmax=2)
    A = nv.Adjoint(nv.Solver(P, nv.MODEL_NS, nv.RK_SSP3, nu=1e-3), 3, 20)
    dt, n = 0.01, 20
    U = A.forward(dt, n, U0)
    _, g = nv.enstrophy(G, U, grad=True)
    d = rand_like(U0, 31, True)
    l0 = A.backward(g)
    t0 = A.tlm(d)
    field = 3 * G.np * 8  # one pool field
    for want, expect in ((2 * field, 2), (7 * field + 5, 7), (1 << 50, 11), (0, 0)):
        got = A.reserve(want)
        print("pool fields", want, got, flush=True)
        assert got == expect
        assert torch.equal(A.forward(dt, n, U0), U)
        assert torch.equal(A.backward(g), l0)
        assert torch.equal(A.backward(g), l0)  # after a reverse sweep that recomputed
        assert torch.equal(A.tlm(d), t0)
        assert torch.equal(A.backward(g), l0)
        assert torch.equal(A.tlm(d), t0)
    A.close()


def test_adjoint_cost_ratio(nv):
    """Adjoint sweep (segment recompute + reverse stages) against the plain
    forward solve of the same steps, at 64^3 (timing only, reported)."""
    import time
    import torch
    N, L = (64, 64, 64), (TWO_PI,) * 3
    G, P, U0 = flow(nv, N, L, seed=9)
    S = nv.Solver(P, nv.MODEL_NS, nv.RK_SSP3, nu=1e-2)
    A = nv.Adjoint(S, 5, 20)
    U = U0.clone()
    p = G.field(1)
    S.step(0.01, 2, U, p)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    S.step(0.01, 20, U, p)
    torch.cuda.synchronize()
    tf = time.perf_counter() - t0
    UT = A.forward(0.01, 20, U0)
    _, g = nv.enstrophy(G, UT, grad=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    A.backward(g)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    print(f"adjoint/forward time ratio at 64^3: {tb / tf:.2f}", flush=True)
    assert tb / tf < 10.0


def compact_flow(nv, N, seed=0, amp=1e-2):
    import torch
    G = nv.Mesh(N, (TWO_PI,) * 3)
    P = nv.Poisson(G, scheme=nv.SCHEME_COMPACT6)
    gen = torch.Generator(device="cuda").manual_seed(seed)
    s = G.shape
    x = [torch.arange(s[a], dtype=torch.float64, device="cuda") * (TWO_PI / N[a]) for a in range(3)]
    X, Y, Z = torch.meshgrid(*x, indexing="ij")
    U = torch.stack([torch.sin(X) * torch.cos(Y) * torch.cos(Z), -torch.cos(X) * torch.sin(Y) * torch.cos(Z), 0 * X])
    U = (U + amp * torch.randn(U.shape, generator=gen, dtype=torch.float64, device="cuda")).contiguous()
    P.project(U, 1.0)
    return G, P, U


@pytest.mark.parametrize("nsteps,every", [(1, 1), (5, 2)])
def test_compact_dot_product_tlm_adjoint(nv, nsteps, every):
    """Compact scheme: the transposed compact stencils and tridiagonal solves
    (D^T = -D, D2^T = D2 on periodic lines) make the adjoint exact."""
    G, P, U0 = compact_flow(nv, (24, 16, 32), seed=11)
    S = nv.Solver(P, nv.MODEL_NS, nv.RK_SSP3, nu=2e-2)
    A = nv.Adjoint(S, every, 8)
    A.forward(0.01, nsteps, U0)
    d = rand_like(U0, 12)
    m = rand_like(U0, 13)
    lhs = nv.weighted_dot(G, A.tlm(d), m)
    rhs = nv.weighted_dot(G, d, A.backward(m))
    print("compact dot", nsteps, lhs, rhs, flush=True)
    assert abs(lhs - rhs) <= 1e-11 * max(abs(lhs), abs(rhs))


def test_compact_enstrophy_gradient_matches_central_differences(nv):
    G, P, U0 = compact_flow(nv, (32, 32, 32), seed=15, amp=5e-2)
    S = nv.Solver(P, nv.MODEL_NS, nv.RK_SSP3, nu=1e-2)
    A = nv.Adjoint(S, 4, 16)
    dt, n = 0.01, 10
    U = A.forward(dt, n, U0)
    J, gT = nv.enstrophy(G, U, grad=True)
    g0 = A.backward(gT)
    Jf = lambda X: nv.enstrophy(G, A.forward(dt, n, X.contiguous()))
    u = 2.0 ** -53
    for seed in range(8):
        d = rand_like(U0, 200 + seed)
        d = d / d.abs().max()
        ad = nv.weighted_dot(G, g0, d)
        e4 = 1e-4
        fd4 = (-Jf(U0 + 2 * e4 * d) + 8 * Jf(U0 + e4 * d) - 8 * Jf(U0 - e4 * d) + Jf(U0 - 2 * e4 * d)) / (12 * e4)
        print(f"compact fd seed={seed} J={J:.3e} ad={ad:.12e} fd4 rel={abs(fd4 - ad) / abs(ad):.2e}", flush=True)
        # 1e-8 relative, plus the roundoff of the four J evaluations,
        # (1 + 8 + 8 + 1) / 12 u J / eps, for directions with a small derivative
        assert abs(fd4 - ad) <= 1e-8 * abs(ad) + 2 * u * J / e4
