"""Pins the oracle (CPU restatement of the reference hot path).

The reference ships no tests, fixtures or golden vectors for this path, so
every function of the oracle is pinned here by what the mathematics fixes:
the exact discrete eigenvalues of the central stencils on Fourier modes
(modified wavenumbers), structural identities between operators, numpy's
independent FFT, Fourier-space brute force for the Poisson inverse, the
stability polynomials of the RK compositions on a linear mode, and the
reference's BC formulas on hand-sized cases.  Each check is chosen so that
a dropped term, a wrong sign, index or spacing fails it.
"""
import math

import numpy as np
import pytest

TWO_PI = 2.0 * math.pi


def kprime(kappa, h, m=6):
    """First-derivative modified wavenumber of the order-m central stencil."""
    import pyoracle
    c = pyoracle.coef(m)
    return sum(c[l] / (l + 1) * math.sin((l + 1) * kappa * h) / h for l in range(m // 2))


def k2second(kappa, h, m=6):
    """Second-derivative modified wavenumber (squared) of the stencil."""
    import pyoracle
    c = pyoracle.coef(m)
    return sum(2.0 * c[l] * (1.0 - math.cos((l + 1) * kappa * h)) / ((l + 1) * h) ** 2 for l in range(m // 2))


def coords(M, axis):
    """Node coordinates of the padded index along one axis."""
    n = M.n[axis]
    d = M.L[axis] / M.N[axis]
    return (np.arange(n) - M.h) * d


def grid(M):
    x, y, z = (coords(M, a) for a in range(3))
    return np.meshgrid(x, y, z, indexing="ij")


def interior_slices(M, endpoint=True):
    h = M.h
    e = 1 if endpoint else 0
    return tuple(slice(h, h + M.N[a] + e) for a in range(3))


def test_coefficients_consistent(orc):
    # sum c_l = 1 makes both the first and the second derivative exact to
    # leading order; the known central weights for order 6
    for m in (2, 4, 6, 8):
        c = orc.coef(m)
        assert abs(sum(c) - 1.0) < 1e-15
    c6 = orc.coef(6)
    assert np.allclose(c6[:3], [1.5, -0.6, 0.1], rtol=0, atol=1e-16)
    # order-6 consistency: sum c_l (l+1)^2 = 0 (the h^2 term vanishes)
    assert abs(sum(c6[l] * (l + 1) ** 2 for l in range(3))) < 1e-14


@pytest.mark.parametrize("m", [2, 4, 6, 8])
def test_grad_div_lap_are_exact_on_fourier_modes(orc, m):
    N = (12, 16, 24)
    L = (TWO_PI, 2 * TWO_PI, TWO_PI)
    M = orc.OMesh(N, L, m=m)
    X, Y, Z = grid(M)
    a, b, c = 2.0, 1.5, 5.0
    hx, hy, hz = (L[i] / N[i] for i in range(3))
    ka, kb, kc = kprime(a, hx, m), kprime(b, hy, m), kprime(c, hz, m)
    p = (np.sin(a * X) * np.cos(b * Y) * np.cos(c * Z))[..., None].copy()
    g = M.grad(p)
    S = interior_slices(M)
    ex = np.stack([ka * np.cos(a * X) * np.cos(b * Y) * np.cos(c * Z),
                   -kb * np.sin(a * X) * np.sin(b * Y) * np.cos(c * Z),
                   -kc * np.sin(a * X) * np.cos(b * Y) * np.sin(c * Z)], -1)
    assert np.max(np.abs(g[S] - ex[S])) < 1e-12 * max(ka, kb, kc)
    # ghosts of operator results stay zero (pool temporaries)
    assert np.all(g[: M.h] == 0) and np.all(g[:, :, -M.h:] == 0)
    # laplacian eigenvalue, per component with different modes
    nu = 0.37
    U = np.stack([np.sin(a * X) * np.cos(b * Y) * np.cos(c * Z),
                  np.cos(2 * X) * np.sin(1.0 * Y) * np.cos(3 * Z),
                  np.cos(X) * np.cos(0.5 * Y) * np.sin(Z)], -1)
    lap = M.laplacian(nu, U)
    lam = [k2second(a, hx, m) + k2second(b, hy, m) + k2second(c, hz, m),
           k2second(2, hx, m) + k2second(1.0, hy, m) + k2second(3, hz, m),
           k2second(1, hx, m) + k2second(0.5, hy, m) + k2second(1, hz, m)]
    for q in range(3):
        assert np.max(np.abs(lap[S][..., q] + nu * lam[q] * U[S][..., q])) < 1e-11 * nu * max(lam)
    # divergence of the same vector field
    d = M.div(U)
    exd = (ka * np.cos(a * X) * np.cos(b * Y) * np.cos(c * Z)
           + kprime(1.0, hy, m) * np.cos(2 * X) * np.cos(1.0 * Y) * np.cos(3 * Z)
           + kprime(1.0, hz, m) * np.cos(X) * np.cos(0.5 * Y) * np.cos(Z))
    assert np.max(np.abs(d[S][..., 0] - exd[S])) < 1e-12 * 10


def test_div_adv_is_mean_of_conservative_and_expanded_forms(orc):
    """div(U,phi) = 0.5 (div(U phi) + U.grad(phi) + phi div(U)), built from
    the separately pinned div and grad (div.hxx:120-165 is exactly this)."""
    rng = np.random.default_rng(11)
    M = orc.OMesh((10, 12, 14), (1.0, 1.3, 0.7))
    U = M.from_compact(rng.standard_normal((3, *M.stored)))
    for nc in (1, 3):
        phi = M.from_compact(rng.standard_normal((nc, *M.stored)))
        got = M.div_adv(U, phi)
        divU = M.div(U)[..., 0]
        S = interior_slices(M)
        for q in range(nc):
            Up = U * phi[..., q:q + 1]
            gq = M.grad(np.ascontiguousarray(phi[..., q:q + 1]))
            ref = 0.5 * (M.div(Up)[..., 0] + (U * gq).sum(-1) + phi[..., q] * divU)
            assert np.max(np.abs(got[S][..., q] - ref[S])) < 1e-11 * np.max(np.abs(ref[S]))


def test_div_adv_skew_symmetric_for_discretely_solenoidal_u(orc):
    # the TG velocity is divergence-free for the discrete operator (equal
    # modified wavenumbers in x and y), so sum phi * div(U, phi) = 0
    M = orc.OMesh((16, 16, 16), (TWO_PI,) * 3)
    X, Y, Z = grid(M)
    U = np.stack([np.sin(X) * np.cos(Y) * np.cos(Z), -np.cos(X) * np.sin(Y) * np.cos(Z), 0 * X], -1)
    assert M.maxabs_div(U) < 1e-14
    rng = np.random.default_rng(3)
    phi = M.from_compact(rng.standard_normal((1, *M.stored)))
    r = M.div_adv(U, phi)
    S = interior_slices(M, endpoint=False)
    s = np.sum(phi[S] * r[S])
    assert abs(s) < 1e-11 * np.sum(np.abs(phi[S] * r[S]))


def test_literal_vector_gradient_on_linear_field(orc):
    # grad.cpp:84-94 uses f(i+l) - f(i-1-l): on f = x it returns
    # sum c_l (2l+1)/(l+1) = 1.5 - 0.9 + 1/6 for order 6, not 1
    M = orc.OMesh((16, 16, 16), (1.0, 1.0, 1.0))
    X, Y, Z = grid(M)
    U = np.stack([X, 2.0 * Y, -Z], -1)
    g = M.grad_vec_literal(np.ascontiguousarray(U))
    core = tuple(slice(2 * M.h, M.n[a] - 2 * M.h) for a in range(3))
    val = 1.5 - 0.9 + 0.1 * 5.0 / 3.0
    assert np.allclose(g[core][..., 0], val, atol=1e-13)
    assert np.allclose(g[core][..., 4], 2 * val, atol=1e-13)
    assert np.allclose(g[core][..., 8], -val, atol=1e-13)
    assert np.allclose(g[core][..., [1, 2, 3, 5, 6, 7]], 0.0, atol=1e-13)
    # curl = antisymmetric part: for U = (-y, x, 0) the literal curl_z = 2*val
    U2 = np.stack([-Y, X, 0 * X], -1)
    w = M.curl_literal(np.ascontiguousarray(U2))
    assert np.allclose(w[core][..., 2], 2 * val, atol=1e-13)
    wc = M.curl_central(np.ascontiguousarray(U2))
    assert np.allclose(wc[core][..., 2], 2.0, atol=1e-12)


def test_periodic_fill_images_bitwise(orc):
    rng = np.random.default_rng(5)
    M = orc.OMesh((8, 10, 12), (1.0, 1.0, 1.0))
    f = M.field(3)
    h = M.h
    f[h:h + 9, h:h + 11, h:h + 13] = rng.standard_normal((9, 11, 13, 3))
    M.fill(f)
    for ax in range(3):
        N = M.N[ax]
        t = np.moveaxis(f, ax, 0)
        assert np.array_equal(t[0:h], t[N:N + h])              # lower ghosts
        assert np.array_equal(t[N + h:N + 2 * h + 1], t[h:2 * h + 1])  # endpoint + upper ghosts


def test_wall_fills_and_reference_bc_formulas(orc):
    rng = np.random.default_rng(6)
    M = orc.OMesh((8, 8, 8), (1.0, 1.0, 1.0))
    h, n = M.h, M.n
    f = rng.standard_normal((*n, 3))
    g = f.copy()
    # dirichletBC, south face, component 0 (dirichletBC.hxx: south)
    M.bc_apply(g, [(orc.SOUTH, orc.BC_DIRICHLET, 0, 0.3)])
    assert np.all(g[:, h, :, 0] == 0.3)
    for l in range(h):
        assert np.array_equal(g[:, h - 1 - l, :, 0], 2.0 * (l + 1) * 0.3 - g[:, h + 1 + l, :, 0])
    assert np.array_equal(g[..., 1:], f[..., 1:])
    # neumannBC, top face: wall node from its neighbour + val*dz, ghosts
    # mirror - 2(l+1) val dz (neumannBC.hxx: top)
    g = f.copy()
    M.bc_apply(g, [(orc.TOP, orc.BC_NEUMANN, 2, 0.5)])
    w = n[2] - h - 1
    dz = 1.0 / 8
    assert np.array_equal(g[:, :, w, 2], f[:, :, w - 1, 2] + 0.5 * dz)
    for l in range(h):
        assert np.array_equal(g[:, :, w + 1 + l, 2], g[:, :, w - 1 - l, 2] - 2.0 * (l + 1) * 0.5 * dz)
    # pure reflections used by the wall case
    Mw = orc.OMesh((8, 8, 8), (1.0, 1.0, 1.0), wall_y=True)
    f = rng.standard_normal((*Mw.n, 3))
    Mw.fill(f, odd_comp=1)
    for l in range(h):
        assert np.array_equal(f[:, h - 1 - l, :, 0], f[:, h + 1 + l, :, 0])
        assert np.array_equal(f[:, h - 1 - l, :, 1], -f[:, h + 1 + l, :, 1])
        top = h + 8
        assert np.array_equal(f[:, top + 1 + l, :, 2], f[:, top - 1 - l, :, 2])
    assert np.all(f[:, h, :, 1] == 0) and np.all(f[:, h + 8, :, 1] == 0)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 6, 8, 12, 16, 48, 96, 256, 384, 5, 7])
def test_fft1d_matches_numpy(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.allclose(orc.fft1d(x, -1), np.fft.fft(x), rtol=0, atol=1e-12 * max(1, n))
    assert np.allclose(orc.fft1d(x, +1), np.fft.ifft(x) * n, rtol=0, atol=1e-12 * max(1, n))


def test_fft3d_matches_numpy(orc):
    rng = np.random.default_rng(1)
    x = rng.standard_normal((6, 8, 12)) + 1j * rng.standard_normal((6, 8, 12))
    assert np.allclose(orc.fft3d(x, -1), np.fft.fftn(x), atol=1e-11)


def test_poisson_single_mode_closed_form(orc):
    N, L = (16, 12, 24), (TWO_PI, 1.5, 3.0)
    M = orc.OMesh(N, L)
    P = orc.OPoisson(M)
    X, Y, Z = grid(M)
    a, b, c = 3.0, 2 * TWO_PI / 1.5, TWO_PI / 3.0 * 5
    rhs = (np.cos(a * X) * np.cos(b * Y) * np.cos(c * Z))[..., None].copy()
    p = P.solve(rhs)
    hx, hy, hz = (L[i] / N[i] for i in range(3))
    k2 = kprime(a, hx) ** 2 + kprime(b, hy) ** 2 + kprime(c, hz) ** 2
    S = interior_slices(M, endpoint=False)
    assert np.max(np.abs(p[S] + rhs[S] / k2)) < 1e-13 / k2


def test_poisson_inverts_div_grad_outside_null_modes(orc):
    """lap^-1 in the solve is the inverse of div(grad(.)) (k2 = sum k'^2);
    modes with k' = 0 in every direction (mean and Nyquist combinations)
    are removed (poisson.cpp:237-244)."""
    rng = np.random.default_rng(2)
    N = (12, 8, 6)
    M = orc.OMesh(N, (1.0, 2.0, 0.5))
    P = orc.OPoisson(M)
    r = rng.standard_normal((1, *N))
    rhs = M.from_compact(r)
    p = P.solve(rhs)
    M.fill(p, -1)
    gp = M.grad(p)
    M.fill(gp)
    lp = M.div(gp)
    got = M.interior(lp)[0]
    R = np.fft.fftn(r[0])
    for i in (0, N[0] // 2):
        for j in (0, N[1] // 2):
            for k in (0, N[2] // 2):
                R[i, j, k] = 0.0
    want = np.real(np.fft.ifftn(R))
    assert np.max(np.abs(got - want)) < 1e-11


def test_poisson_even_extension_for_walls(orc):
    # cos modes in y are even about both walls: p = -rhs / k2 with the
    # modified wavenumber of the y stencil
    N, L = (12, 10, 8), (TWO_PI, 2.0, TWO_PI)
    M = orc.OMesh(N, L, wall_y=True)
    P = orc.OPoisson(M)
    X, Y, Z = grid(M)
    a, b, c = 2.0, 3 * math.pi / 2.0, 1.0
    rhs = (np.cos(a * X) * np.cos(b * Y) * np.sin(c * Z))[..., None].copy()
    p = P.solve(rhs)
    k2 = kprime(a, L[0] / N[0]) ** 2 + kprime(b, L[1] / N[1]) ** 2 + kprime(c, L[2] / N[2]) ** 2
    h = M.h
    S = (slice(h, h + N[0]), slice(h, h + N[1] + 1), slice(h, h + N[2]))
    assert np.max(np.abs(p[S] + rhs[S] / k2)) < 1e-12 / k2


def test_projection_removes_divergence(orc):
    rng = np.random.default_rng(4)
    for wall in (False, True):
        M = orc.OMesh((16, 12, 8), (1.0, 1.0, 1.0), wall_y=wall)
        P = orc.OPoisson(M)
        U = M.from_compact(rng.standard_normal((3, *M.stored)))
        if wall:
            h = M.h
            U[:, h, :, 1] = 0.0
            U[:, h + 12, :, 1] = 0.0
            M.fill(U)
        assert M.maxabs_div(U) > 1.0
        P.project(U, 0.01)
        assert M.maxabs_div(U) < 1e-11


def tg_field(M, amp=1.0):
    X, Y, Z = grid(M)
    U = np.stack([np.sin(X) * np.cos(Y) * np.cos(Z), -np.cos(X) * np.sin(Y) * np.cos(Z), 0 * X], -1) * amp
    return np.ascontiguousarray(U)


def test_tg_initial_energy_is_one_eighth(orc):
    for n in (4, 8, 32):
        M = orc.OMesh((n, n, n), (TWO_PI,) * 3)
        assert abs(M.ke(tg_field(M)) - 0.125) < 1e-13  # serial sum of n^3 terms


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_rk_compositions_stability_polynomials(orc, mode):
    """On a small-amplitude TG mode the advection is negligible and the
    projection is the identity, so one step multiplies U by R(z),
    z = -nu dt k''^2_total."""
    n = 16
    M = orc.OMesh((n, n, n), (TWO_PI,) * 3)
    P = orc.OPoisson(M)
    nu, dt = 0.2, 0.1
    U = tg_field(M, 1e-10)
    p = M.field(1)
    U0 = U.copy()
    P.ns_step(U, p, nu, dt, mode)
    z = -nu * dt * 3 * k2second(1.0, TWO_PI / n)
    R = {0: 1 + z,
         1: 1 + z + z * z / 2 + z ** 3 / 6,
         2: 1 + (z / 6) * (1 + (z / 3) * (1 + (z / 3) * (1 + z / 6)))}[mode]
    S = interior_slices(M, endpoint=False)
    ratio = U[S][..., :2] / np.where(np.abs(U0[S][..., :2]) > 5e-11, U0[S][..., :2], np.nan)
    assert np.nanmax(np.abs(ratio - R)) < 1e-9


def test_mhd_step_keeps_both_fields_solenoidal(orc):
    rng = np.random.default_rng(9)
    M = orc.OMesh((12, 12, 12), (TWO_PI,) * 3)
    P = orc.OPoisson(M)
    X, Y, Z = grid(M)
    U = np.ascontiguousarray(np.stack([-np.sin(Y), np.sin(X), 0 * X], -1))
    B = np.ascontiguousarray(np.stack([-np.sin(Y), np.sin(2 * X), 0 * X], -1))
    B += M.from_compact(1e-3 * rng.standard_normal((3, *M.stored)))
    P.project(B, 1.0)
    p, pB = M.field(1), M.field(1)
    for mode in (0, 1):
        U1, B1 = U.copy(), B.copy()
        P.mhd_step(U1, B1, p, pB, 1e-3, 1e-3, 2e-3, mode)
        assert M.maxabs_div(U1) < 1e-11 and M.maxabs_div(B1) < 1e-11
        # the step changes the fields by O(dt)
        assert 1e-6 < np.max(np.abs(U1 - U)) < 0.1


def test_tdma_matches_dense_solve(orc):
    rng = np.random.default_rng(7)
    n = 17
    a, c = rng.standard_normal(n), rng.standard_normal(n)
    b = 4.0 + np.abs(rng.standard_normal(n))
    d = rng.standard_normal(n)
    A = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    assert np.allclose(orc.tdma(a, b, c, d), np.linalg.solve(A, d), atol=1e-13)


def test_mesh_metric_uniform_and_stretched(orc):
    M = orc.OMesh((8, 8, 16), (1.0, 1.0, 2.0))
    dz, d2 = M.metric()
    h = M.h
    assert np.allclose(dz[h:-h], 1.0, atol=1e-13)
    assert np.max(np.abs(d2[h:-h])) < 1e-8
    Ms = orc.OMesh((8, 8, 16), (1.0, 1.0, 2.0), hyp=(0.02, 0.03))  # (N-1) sqrt(s1 s2) < 1: the b fixed point is non-zero
    from pyoracle import lib
    z0 = lib().orc_mesh_z(Ms._h, h)
    z1 = lib().orc_mesh_z(Ms._h, h + 16)
    assert abs(z0) < 1e-14 and abs(z1 - 2.0) < 1e-14
    zs = [lib().orc_mesh_z(Ms._h, k) for k in range(h, h + 17)]
    assert np.all(np.diff(zs) > 0)


def test_stretched_mesh_end_slopes_and_symmetry(orc):
    """Interior pins of the stretched z (src/Mesh/Mesh.cpp:370-404), from what
    its definition implies rather than from the formula: with xi = k/N, b the
    fixed point of b = asinh(b / (N sqrt(s1 s2))) and A = sqrt(s1/s2), the map
    has dz/dk = L s2 at the bottom wall and L s1 at the top (the end slopes
    2b/sinh(b) of U times 1/(2A) and A/2, with sinh(b)/b = 1/(N sqrt(s1 s2)));
    with s1 = s2 it is symmetric, z(k) + z(N - k) = L.  A wrong fixed-point
    argument, an inverted A or a sign in the tanh argument fails these."""
    from pyoracle import lib
    N, L = 256, 2.0
    # one-sided sixth-order first derivative at the end node
    c = np.array([-49 / 20, 6, -15 / 2, 20 / 3, -15 / 4, 6 / 5, -1 / 6])
    for s1, s2 in [(0.002, 0.003), (0.003, 0.0015)]:
        M = orc.OMesh((8, 8, N), (1.0, 1.0, L), hyp=(s1, s2))
        h = M.h
        z = np.array([lib().orc_mesh_z(M._h, h + k) for k in range(N + 1)])
        assert abs(z[0]) < 1e-14 and abs(z[-1] - L) < 1e-14
        d0, dN = c @ z[:7], -(c @ z[::-1][:7])
        assert abs(d0 / (L * s2) - 1) < 1e-9, (s1, s2, d0)
        assert abs(dN / (L * s1) - 1) < 1e-9, (s1, s2, dN)
    M = orc.OMesh((8, 8, N), (1.0, 1.0, L), hyp=(0.002, 0.002))
    z = np.array([lib().orc_mesh_z(M._h, M.h + k) for k in range(N + 1)])
    assert np.max(np.abs(z + z[::-1] - L)) < 1e-13


@pytest.mark.parametrize("wall_y,hyp", [(False, None), (True, None), (False, (0.02, 0.03))])
def test_fused_cpu_baseline_step_equals_the_faithful_step(orc, wall_y, hyp):
    # the fused variant times the same arithmetic: it must agree with the
    # operator-at-a-time step to rounding
    rng = np.random.default_rng(5)
    M = orc.OMesh((12, 16, 12), (TWO_PI, 3.0 if wall_y else TWO_PI, TWO_PI), wall_y=wall_y, hyp=hyp)
    P = orc.OPoisson(M)
    c = rng.standard_normal((3, *M.stored))
    if wall_y:
        c[1, :, 0, :] = c[1, :, -1, :] = 0.0
    U = M.from_compact(c)
    P.project(U, 1.0)
    V = U.copy()
    p, q = M.field(1), M.field(1)
    P.ns_step(U, p, 0.05, 0.01, orc.RK_SSP3)
    P.ns_step_fused(V, q, 0.05, 0.01)
    S = interior_slices(M)
    assert np.max(np.abs(U[S] - V[S])) < 1e-13 * np.max(np.abs(U[S]))
    assert np.max(np.abs(p[S] - q[S])) < 1e-12 * np.max(np.abs(p[S]))
