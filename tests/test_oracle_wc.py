"""Pins the oracle's weakly-compressible step (app/TG/TG.cpp:71-76 with
rhoEqn.H, pEqn.H, UEqn.H).  The reference has no tests for this path; the
step is pinned by what the mathematics fixes: a uniform state is a fixed
point, the periodic divergence form conserves mass, a density wave carried
by a uniform velocity is multiplied by the classic RK4 polynomial of the
advection operator's modified wavenumber, a shear wave at uniform density
by the RK4 polynomial of the diffusion operator's modified wavenumber over
rho, and from rest the velocity after one short step is -dt grad(p)/rho to
second order in dt."""
import math

import numpy as np
import pytest

TWO_PI = 2.0 * math.pi
COEF6 = (1.5, -0.6, 0.1)


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


def rk4(z):
    return 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24


def k1(k, h):
    """modified wavenumber of the sixth-order central first derivative"""
    return sum(c * math.sin((l + 1) * k * h) / ((l + 1) * h) for l, c in enumerate(COEF6))


def k2(k, h):
    """modified wavenumber (squared) of the sixth-order central second derivative"""
    return sum(2 * c * (1 - math.cos((l + 1) * k * h)) / ((l + 1) * h) ** 2 for l, c in enumerate(COEF6))


def filled(M, arr, nc):
    f = M.field(nc)
    f[...] = arr.reshape(*M.n, nc)
    M.fill(f, 1 if nc == 3 else -1)
    return f


def test_uniform_state_is_a_fixed_point(orc):
    M = orc.OMesh((12, 8, 16), (TWO_PI, 2.0, 3.0))
    U, rho, p = M.field(3), M.field(1), M.field(1)
    U[...] = (0.3, -0.2, 0.1)
    rho[...] = 1.7
    U0, r0 = U.copy(), rho.copy()
    M.wc_step(U, rho, p, 0.05, 50.0, 0.01)
    assert np.array_equal(U, U0) and np.array_equal(rho, r0)
    assert np.all(p[..., 0][inner(M)] == 50.0 * (1.7 * 1.7))


def test_mass_is_conserved(orc):
    M = orc.OMesh((16, 12, 8), (TWO_PI,) * 3)
    X, Y, Z = grid(M)
    U = filled(M, np.stack([np.sin(X) * np.cos(2 * Y), np.cos(Z) + 0.2 * np.sin(Y), 0.3 * np.sin(X + Z)], -1), 3)
    rho = filled(M, 1.0 + 0.1 * np.cos(X) * np.sin(Y + Z), 1)
    p = M.field(1)
    m0 = rho[inner(M)].sum()
    for _ in range(3):
        M.wc_step(U, rho, p, 1e-3, 50.0, 0.01)
    assert abs(rho[inner(M)].sum() - m0) < 1e-13 * m0
    assert np.abs(rho[inner(M)] - 1.0).max() > 0.05  # the field did move


@pytest.mark.parametrize("k,u0,dt", [(3, 0.7, 0.05), (5, -1.3, 0.02)])
def test_density_wave_follows_rk4_of_the_advection_symbol(orc, k, u0, dt):
    # rhoEqn runs first, with U frozen: d rho/dt = -u0 d rho/dx, so every
    # Fourier mode is multiplied by R(-i u0 k' dt)
    N = (32, 8, 8)
    M = orc.OMesh(N, (TWO_PI,) * 3)
    X, Y, Z = grid(M)
    eps = 1e-3
    rho = filled(M, 1.0 + eps * np.cos(k * X), 1)
    U = M.field(3)
    U[..., 0] = u0
    p = M.field(1)
    M.wc_step(U, rho, p, 0.0, 50.0, dt)
    h = TWO_PI / N[0]
    g = rk4(-1j * u0 * k1(k, h) * dt)
    want = 1.0 + eps * np.real(g * np.exp(1j * k * X))
    err = np.abs(rho[..., 0][inner(M)] - want[inner(M)]).max()
    assert err < 1e-15 * 10, err
    # a mistake in a stage weight or offset moves the amplitude by O(z^2)
    assert abs(abs(g) - 1) > 1e-9


@pytest.mark.parametrize("rho0", [1.0, 2.0])
def test_shear_wave_decays_with_rk4_of_the_diffusion_symbol(orc, rho0):
    # uniform rho, p = pc rho^2 uniform: U = (0, 0, w(x)) has no advection
    # and d(rho w)/dt = mu w'', so w is multiplied by R(-(mu/rho) k''^2 dt)
    N = (32, 8, 8)
    k, mu, dt, eps = 4, 0.05, 0.04, 1e-2
    M = orc.OMesh(N, (TWO_PI,) * 3)
    X, Y, Z = grid(M)
    U = M.field(3)
    U[..., 2] = eps * np.sin(k * X)
    rho = M.field(1)
    rho[...] = rho0
    p = M.field(1)
    M.wc_step(U, rho, p, mu, 50.0, dt)
    h = TWO_PI / N[0]
    g = rk4(-(mu / rho0) * k2(k, h) * dt)
    want = g * eps * np.sin(k * X)
    assert np.abs(U[inner(M)][..., 2] - want[inner(M)]).max() < 1e-15 * 10
    assert np.abs(U[inner(M)][..., :2]).max() == 0.0
    assert np.all(rho[..., 0][inner(M)] == rho0)


def test_flow_from_rest_is_driven_by_minus_grad_p_over_rho(orc):
    N = (24, 16, 8)
    M = orc.OMesh(N, (TWO_PI,) * 3)
    X, Y, Z = grid(M)
    rho = filled(M, 1.0 + 0.01 * np.cos(2 * X) * np.cos(Y), 1)
    pc = 50.0
    errs = []
    for dt in (2e-4, 1e-4):
        U, r, p = M.field(3), rho.copy(), M.field(1)
        M.wc_step(U, r, p, 0.0, pc, dt)
        assert np.array_equal(r, rho)  # no flow: rhoEqn leaves rho unchanged
        pf = M.field(1)
        pf[...] = pc * (rho * rho)
        G = M.grad(pf)
        want = -dt * G / rho
        s = inner(M)
        errs.append(np.abs(U[s] - want[s]).max() / np.abs(want[s]).max())
    assert errs[0] < 1e-6
    # the remainder is second order in dt (the advection of the O(dt) flow)
    assert 2.5 < errs[0] / errs[1] < 6.0 or errs[1] < 1e-13


def test_lorentz_force_of_a_frozen_field(orc):
    # app/OT/UEqn.H:22-23 with B = (b(x), 0, 0) frozen, rho = 1, p uniform,
    # from rest: -div(2 D B, B) + grad(D B.B) = -2 D b D1(b) on the x
    # component (D1 the sixth-order first derivative), so for
    # b = eps sin(kx) the flow after a short step is -dt D eps^2 k' sin(2kx)
    N = (32, 8, 8)
    M = orc.OMesh(N, (TWO_PI,) * 3)
    X, Y, Z = grid(M)
    k, eps, dbu = 3, 0.2, 0.5
    B = M.field(3)
    B[..., 0] = eps * np.sin(k * X)
    h = TWO_PI / N[0]
    errs = []
    for dt in (2e-4, 1e-4):
        U, rho, p = M.field(3), M.field(1), M.field(1)
        rho[...] = 1.0
        M.wc_step(U, rho, p, 0.0, 50.0, dt, B=B, dbu=dbu)
        want = -dt * dbu * eps * eps * k1(k, h) * np.sin(2 * k * X)
        s = inner(M)
        assert np.abs(U[s][..., 1:]).max() == 0.0
        errs.append(np.abs(U[s][..., 0] - want[s]).max() / np.abs(want[s]).max())
    assert errs[0] < 1e-5
    assert 2.5 < errs[0] / errs[1] < 6.0 or errs[1] < 1e-13


# app/shearLayerWC (shearLayer.cpp:197-200): UEqn with buoyancy, rhoEqn,
# pEqn, TEqn


@pytest.mark.parametrize("k,u0,dt", [(3, 0.7, 0.05), (4, -1.1, 0.03)])
def test_temperature_wave_follows_rk4_of_the_advection_symbol(orc, k, u0, dt):
    # uniform rho and U = (u0, 0, 0), g = 0: UEqn and rhoEqn leave the
    # uniform state alone and d(rho T)/dt = -rho u0 dT/dx
    N = (32, 8, 8)
    M = orc.OMesh(N, (TWO_PI,) * 3)
    X, Y, Z = grid(M)
    eps = 1e-2
    U, rho, p = M.field(3), M.field(1), M.field(1)
    U[..., 0] = u0
    rho[...] = 1.3
    p[...] = 7.0
    T = filled(M, 0.5 + eps * np.cos(k * X), 1)
    M.wcs_step(U, rho, p, T, 0.0, 0.0, (0.0, 0.0, 0.0), 2.0, dt)
    h = TWO_PI / N[0]
    g = rk4(-1j * u0 * k1(k, h) * dt)
    want = 0.5 + eps * np.real(g * np.exp(1j * k * X))
    err = np.abs(T[..., 0][inner(M)] - want[inner(M)]).max()
    assert err < 1e-14, err
    assert np.all(U[..., 0][inner(M)] == u0) and np.all(rho[..., 0][inner(M)] == 1.3)
    assert np.all(p[..., 0][inner(M)] == (2.0 * 1.3) * 1.3)


@pytest.mark.parametrize("rho0", [1.0, 3.0])
def test_temperature_diffuses_with_rk4_of_the_diffusion_symbol(orc, rho0):
    N = (32, 8, 8)
    k, ka, dt, eps = 5, 0.02, 0.05, 1e-2
    M = orc.OMesh(N, (TWO_PI,) * 3)
    X, Y, Z = grid(M)
    U, rho, p = M.field(3), M.field(1), M.field(1)
    rho[...] = rho0
    T = filled(M, eps * np.sin(k * X), 1)
    M.wcs_step(U, rho, p, T, 0.0, ka, (0.0, 0.0, 0.0), 1.0, dt)
    h = TWO_PI / N[0]
    want = rk4(-(ka / rho0) * k2(k, h) * dt) * eps * np.sin(k * X)
    assert np.abs(T[..., 0][inner(M)] - want[inner(M)]).max() < 1e-16 * 10
    assert np.abs(U[inner(M)]).max() == 0.0


def test_buoyancy_drives_the_flow_along_minus_g(orc):
    # from rest, rho = 1, p uniform, T = T(x): -(rho T) g is the only force,
    # no term moves x-only fields along z, so after one step U_z = dt g0 T
    N = (24, 8, 8)
    M = orc.OMesh(N, (TWO_PI,) * 3)
    X, Y, Z = grid(M)
    g0, dt = 9.81, 0.01
    U, rho, p = M.field(3), M.field(1), M.field(1)
    rho[...] = 1.0
    p[...] = 4.0
    T0 = 0.3 + 0.1 * np.cos(2 * X)
    T = filled(M, T0, 1)
    M.wcs_step(U, rho, p, T, 0.0, 0.0, (0.0, 0.0, -g0), 4.0, dt)
    s = inner(M)
    want = dt * g0 * T0
    assert np.abs(U[s][..., 2] - want[s]).max() < 1e-15
    assert np.abs(U[s][..., :2]).max() == 0.0
    assert np.array_equal(T[..., 0][s], T0[s]) and np.all(rho[..., 0][s] == 1.0)


# app/shearLayerWC/hydrostaticP.H and the initial density (shearLayer.cpp:139-151)


@pytest.mark.parametrize("wall", [False, True])
def test_hydrostatic_pressure_is_the_column_sum_from_the_top(orc, wall):
    # T = a + b z: node k of a column gets sum_{k'=k}^{N-1} (a + b k' dz) dz g rho0
    # = ((N - k) a + b dz (N(N-1)/2 - k(k-1)/2)) dz g rho0; the node k = N
    # (the top node) and the columns of the last i node get nothing
    N = (12, 10, 16)
    M = orc.OMesh(N, (2.0, 3.0, 1.6), wall_y=wall)
    h = M.h
    dz = 1.6 / N[2]
    a, b, g2, rho0, pc = 0.3, -0.7, -2.5, 999.57, 8.7e-7
    kk = np.arange(M.n[2]) - h
    T = M.field(1)
    T[..., 0] = (a + b * kk * dz)[None, None, :]
    p0 = M.field(1)
    p0[...] = 0.8 + 0.01 * np.arange(M.n[1])[None, :, None, None]
    p, rho = p0.copy(), M.field(1)
    M.wc_hydrostatic(T, p, rho, (7.0, 11.0, g2), rho0, pc)
    k = np.arange(N[2] + 1)
    s = np.where(k < N[2], (N[2] - k) * a + b * dz * (N[2] * (N[2] - 1) / 2 - k * (k - 1) / 2), 0.0)
    want = s * dz * g2 * rho0
    dp = (p - p0)[..., 0]
    I, J, K = slice(h, h + N[0]), slice(h, h + N[1] + 1), slice(h, h + N[2] + 1)
    assert np.abs(dp[I, J, K] - want[None, None, :]).max() < 1e-12 * np.abs(want).max()
    assert np.all(dp[h + N[0], J, K] == 0.0)          # i = N (periodic image) excluded
    assert np.all(dp[:h] == 0.0) and np.all(dp[:, :, :h] == 0.0)  # ghosts untouched
    # the density follows p: rho^2 pc = p on every interior node
    R = rho[..., 0][h:h + N[0] + 1, J, K]
    P = p[..., 0][h:h + N[0] + 1, J, K]
    assert np.abs(R * R * pc / P - 1.0).max() < 1e-15
    assert np.all(rho[..., 0][:h] == 0.0)
