"""Pins the oracle's compact 6th-order scheme (no reference counterpart:
the reference has only explicit stencils).  Lele (1992) tridiagonal schemes:
first derivative alpha = 1/3, a = 14/9, b = 1/9; second derivative
alpha = 2/11, a = 12/11, b = 3/11.  Pinned by a dense linear solve, by sixth
order convergence to analytic derivatives of non-trigonometric functions,
by the operators' symmetries, and by closed forms on Fourier modes and on
the Taylor-Green vortex."""
import math

import numpy as np
import pytest

TWO_PI = 2.0 * math.pi


def coords(M, axis):
    n = M.n[axis]
    d = M.L[axis] / M.N[axis]
    return (np.arange(n) - M.h) * d


def grid(M):
    x, y, z = (coords(M, a) for a in range(3))
    return np.meshgrid(x, y, z, indexing="ij")


def inner(M):
    h = M.h
    return tuple(slice(h, h + M.N[a]) for a in range(3))


@pytest.mark.parametrize("n,a,b,c", [(5, 1 / 3, 1.0, 1 / 3), (17, 2 / 11, 1.0, 2 / 11), (64, 0.2, 1.1, 0.3)])
def test_cyclic_tridiag_matches_dense_solve(orc, n, a, b, c):
    rng = np.random.default_rng(n)
    d = rng.standard_normal(n)
    A = np.diag(np.full(n, b)) + np.diag(np.full(n - 1, a), -1) + np.diag(np.full(n - 1, c), 1)
    A[0, n - 1] = a
    A[n - 1, 0] = c
    x = orc.cyclic_tridiag(a, b, c, d)
    assert np.max(np.abs(A @ x - d)) < 1e-14
    assert np.allclose(x, np.linalg.solve(A, d), atol=1e-13)


@pytest.mark.parametrize("order", [1, 2])
def test_sixth_order_convergence_on_non_trigonometric_function(orc, order):
    """f = exp(sin x): f' = cos x f, f'' = (cos^2 x - sin x) f.  The error
    of a sixth-order scheme falls by ~2^6 per halving of h."""
    errs = []
    for n in (16, 32):
        M = orc.OMesh((n, 8, 8), (TWO_PI,) * 3)
        X, _, _ = grid(M)
        f = np.ascontiguousarray(np.exp(np.sin(X))[..., None])
        out = M.compact_deriv(f, 0, order)
        exact = np.cos(X) * np.exp(np.sin(X)) if order == 1 else (np.cos(X) ** 2 - np.sin(X)) * np.exp(np.sin(X))
        S = inner(M)
        errs.append(np.max(np.abs(out[S][..., 0] - exact[S])))
    rate = math.log2(errs[0] / errs[1])
    print("order", order, "errors", errs, "rate", rate)
    assert 5.6 < rate < 6.6
    assert errs[1] < 2e-5


def test_modified_wavenumbers_on_fourier_modes(orc):
    """D sin(kx) = k' cos(kx), D2 sin(kx) = -k''^2 sin(kx) with the closed
    forms of the tridiagonal schemes; the direction, spacing and component
    handling are exercised on all three axes."""
    N = (16, 24, 32)
    L = (TWO_PI, 3.0, 5.0)
    M = orc.OMesh(N, L)
    G = grid(M)
    for ax in range(3):
        h = L[ax] / N[ax]
        for mode in (1, 3, N[ax] // 2 - 1):
            k = TWO_PI * mode / L[ax]
            kh = k * h
            k1 = (14 / 9 * math.sin(kh) + 1 / 18 * math.sin(2 * kh)) / (1 + 2 / 3 * math.cos(kh)) / h
            k2 = (24 / 11 * (1 - math.cos(kh)) + 3 / 22 * (1 - math.cos(2 * kh))) / (1 + 4 / 11 * math.cos(kh)) / h ** 2
            assert abs(orc.compact_kmod(kh, 1) / h - k1) < 1e-12 * k1
            assert abs(orc.compact_kmod(kh, 2) / h ** 2 - k2) < 1e-12 * k2
            f = np.ascontiguousarray(np.stack([np.sin(k * G[ax]), 2 * np.cos(k * G[ax]), 0 * G[ax]], -1))
            d1 = M.compact_deriv(f, ax, 1)
            d2 = M.compact_deriv(f, ax, 2)
            S = inner(M)
            assert np.max(np.abs(d1[S][..., 0] - k1 * np.cos(k * G[ax])[S])) < 1e-11 * k1
            assert np.max(np.abs(d1[S][..., 1] + 2 * k1 * np.sin(k * G[ax])[S])) < 1e-11 * k1
            assert np.max(np.abs(d2[S][..., 0] + k2 * np.sin(k * G[ax])[S])) < 1e-11 * k2
            assert np.max(np.abs(d1[S][..., 2])) == 0.0


def test_compact_first_derivative_is_skew_and_second_symmetric(orc):
    rng = np.random.default_rng(3)
    M = orc.OMesh((12, 8, 16), (TWO_PI, 2.0, 3.0))
    f = M.from_compact(rng.standard_normal((1, *M.stored)))
    g = M.from_compact(rng.standard_normal((1, *M.stored)))
    S = inner(M)
    for ax in range(3):
        a = np.sum(g[S] * M.compact_deriv(f, ax, 1)[S])
        b = np.sum(f[S] * M.compact_deriv(g, ax, 1)[S])
        assert abs(a + b) < 1e-11 * (abs(a) + 1.0)
        a2 = np.sum(g[S] * M.compact_deriv(f, ax, 2)[S])
        b2 = np.sum(f[S] * M.compact_deriv(g, ax, 2)[S])
        assert abs(a2 - b2) < 1e-11 * (abs(a2) + 1.0)
        assert np.sum(f[S] * M.compact_deriv(f, ax, 2)[S]) < 0.0


def test_compact_rhs_of_a_shear_mode_and_of_taylor_green(orc):
    """u = (sin y, 0, 0): advection vanishes identically and the viscous term
    is -nu k''^2 u.  Taylor-Green: the inviscid right-hand side converges at
    sixth order to -(u.grad)u = (-sin 2x cos^2 z / 2, -sin 2y cos^2 z / 2, 0)."""
    nu = 0.3
    M = orc.OMesh((16, 16, 16), (TWO_PI,) * 3)
    X, Y, Z = grid(M)
    U = np.ascontiguousarray(np.stack([np.sin(Y), 0 * X, 0 * X], -1))
    R = M.compact_ns_rhs(U, nu)
    h = TWO_PI / 16
    k2 = orc.compact_kmod(h, 2) / h ** 2
    S = inner(M)
    assert np.max(np.abs(R[S][..., 0] + nu * k2 * np.sin(Y)[S])) < 1e-13
    assert np.max(np.abs(R[S][..., 1:])) < 1e-15
    errs = []
    for n in (16, 32):
        M = orc.OMesh((n, n, n), (TWO_PI,) * 3)
        X, Y, Z = grid(M)
        U = np.ascontiguousarray(np.stack([np.sin(X) * np.cos(Y) * np.cos(Z), -np.cos(X) * np.sin(Y) * np.cos(Z),
                                           0 * X], -1))
        R = M.compact_ns_rhs(U, 0.0)
        gp = np.stack([-0.5 * np.sin(2 * X) * np.cos(Z) ** 2, -0.5 * np.sin(2 * Y) * np.cos(Z) ** 2, 0 * X], -1)
        errs.append(np.max(np.abs(R[inner(M)] - gp[inner(M)])))
    rate = math.log2(errs[0] / errs[1])
    print("TG rhs errors", errs, "rate", rate)
    assert 5.5 < rate < 6.8


def test_compact_projection_and_linear_step(orc):
    rng = np.random.default_rng(5)
    n = 16
    M = orc.OMesh((n, n, n), (TWO_PI,) * 3)
    P = orc.OPoisson(M, compact=True)
    U = M.from_compact(rng.standard_normal((3, *M.stored)))
    P.project(U, 0.3)
    assert M.compact_maxabs_div(U) < 1e-12 * np.max(np.abs(U)) * n
    V = U.copy()
    P.project(V, 1.0)  # idempotent on a compact-solenoidal field
    assert np.max(np.abs(V - U)) < 1e-13
    # small-amplitude TG: one SSP-RK3 step multiplies U by the stability
    # polynomial of z = -nu dt k''^2 (compact second-derivative symbol)
    X, Y, Z = grid(M)
    amp = 1e-10
    U = np.ascontiguousarray(amp * np.stack([np.sin(X) * np.cos(Y) * np.cos(Z), -np.cos(X) * np.sin(Y) * np.cos(Z),
                                             0 * X], -1))
    U0 = U.copy()
    nu, dt = 0.2, 0.1
    p = M.field(1)
    P.compact_ns_step(U, p, nu, dt)
    h = TWO_PI / n
    z = -nu * dt * 3 * orc.compact_kmod(h, 2) / h ** 2
    Rz = 1 + z + z * z / 2 + z ** 3 / 6
    S = inner(M)
    ratio = U[S][..., :2] / np.where(np.abs(U0[S][..., :2]) > 5e-11, U0[S][..., :2], np.nan)
    assert np.nanmax(np.abs(ratio - Rz)) < 1e-9


def test_compact_mhd_rhs_closed_forms(orc):
    """Lorentz force of B = (sin y, 0, 0) on U = 0: -D(|B|^2 / 2) = (0, -k'(2) sin 2y / 4, 0);
    induction of U = (0, sin x, 0) in B = (1, 0, 0): B.grad U = (0, k'(1) cos x, 0);
    with B = 0 the velocity right-hand side is the NS one."""
    n = 16
    h = TWO_PI / n
    M = orc.OMesh((n, n, n), (TWO_PI,) * 3)
    X, Y, Z = grid(M)
    S = inner(M)
    zero = np.zeros((*M.n, 3))
    nu, eta = 0.3, 0.2
    k1 = orc.compact_kmod(h, 1) / h
    k2 = orc.compact_kmod(2 * h, 1) / h
    kk = orc.compact_kmod(h, 2) / h ** 2
    B = np.ascontiguousarray(np.stack([np.sin(Y), 0 * X, 0 * X], -1))
    dU, dB = M.compact_mhd_rhs(zero, B, nu, eta)
    assert np.max(np.abs(dU[S][..., 1] + 0.25 * k2 * np.sin(2 * Y)[S])) < 1e-13
    assert np.max(np.abs(dU[S][..., [0, 2]])) < 1e-14
    assert np.max(np.abs(dB[S][..., 0] + eta * kk * np.sin(Y)[S])) < 1e-13
    U = np.ascontiguousarray(np.stack([0 * X, np.sin(X), 0 * X], -1))
    B = np.ascontiguousarray(np.stack([1.0 + 0 * X, 0 * X, 0 * X], -1))
    dU, dB = M.compact_mhd_rhs(U, B, nu, 0.0)
    assert np.max(np.abs(dB[S][..., 1] - k1 * np.cos(X)[S])) < 1e-13
    assert np.max(np.abs(dB[S][..., [0, 2]])) < 1e-14
    assert np.max(np.abs(dU[S][..., 1] + nu * kk * np.sin(X)[S])) < 1e-13
    rng = np.random.default_rng(2)
    U = M.from_compact(rng.standard_normal((3, *M.stored)))
    dU, dB = M.compact_mhd_rhs(U, np.ascontiguousarray(zero), nu, eta)
    assert np.max(np.abs(dU - M.compact_ns_rhs(U, nu))[S]) < 1e-12
    assert np.max(np.abs(dB[S])) == 0.0


def test_compact_mhd_linear_step(orc):
    """Small-amplitude solenoidal modes: one step multiplies U and B by the
    SSP-RK3 polynomial of their own z = -(nu|eta) dt k''^2."""
    n = 16
    h = TWO_PI / n
    M = orc.OMesh((n, n, n), (TWO_PI,) * 3)
    P = orc.OPoisson(M, compact=True)
    X, Y, Z = grid(M)
    amp = 1e-10
    U = np.ascontiguousarray(amp * np.stack([np.sin(Y), 0 * X, 0 * X], -1))
    B = np.ascontiguousarray(amp * np.stack([0 * X, 0 * X, np.sin(X)], -1))
    U0, B0 = U.copy(), B.copy()
    nu, eta, dt = 0.2, 0.1, 0.1
    P.compact_mhd_step(U, B, M.field(1), M.field(1), nu, eta, dt)
    kk = orc.compact_kmod(h, 2) / h ** 2
    R = lambda z: 1 + z + z * z / 2 + z ** 3 / 6
    S = inner(M)
    m = np.abs(U0[S][..., 0]) > 5e-11
    assert np.max(np.abs(U[S][..., 0][m] / U0[S][..., 0][m] - R(-nu * dt * kk))) < 1e-9
    m = np.abs(B0[S][..., 2]) > 5e-11
    assert np.max(np.abs(B[S][..., 2][m] / B0[S][..., 2][m] - R(-eta * dt * kk))) < 1e-9
