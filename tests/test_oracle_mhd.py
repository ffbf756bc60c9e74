"""Pins of the oracle's MHD step, viscous dissipation, MHD diagnostics and
stretched z mesh (CPU).

What fixes each expected value:
* the explicit MHD right-hand sides (app/OT_Poisson/UEqn.H:16-23,
  BEqn.H:15-23) on trigonometric fields, through the exact discrete
  eigenvalue D e^{ikx} = i k' e^{ikx} of the central stencil (k' the
  modified wavenumber);
* both MHD step compositions on an Alfven-type mode U = (0, a sin x, 0),
  B = (1, b cos x, 0), on which every nonlinear term is zero or a pure
  gradient that the projection removes, so that one step is a 2 x 2 linear
  map in closed form: for the reference's composition (UEqn.H: frozen-stage
  U, projection; BEqn.H: classic RK4 of B with the new U frozen; cleaning)
  and for SSP-RK3 with both fields projected every stage;
* the viscous dissipation nu <|grad u|^2> of the Taylor-Green field;
* the MHD diagnostics of app/OT_Poisson/OT.cpp:134-176 on single-mode
  fields, the literal curl through the sum-to-product form of its
  half-shifted stencil (src/ex/grad/grad.cpp:84-94);
* Mesh::z (src/Mesh/Mesh.cpp:373-407), Vinokur's two-sided tanh stretching:
  its inverse map and the end slopes its design conditions fix.

tests/test_oracle_planted.py shows that a planted mistake in any of these
functions fails one of the tests below.
"""
import math

import numpy as np
import pytest
from scipy.optimize import brentq

TWO_PI = 2.0 * math.pi


def kp(kappa, h, m=6):
    import pyoracle
    c = pyoracle.coef(m)
    return sum(c[l] / (l + 1) * math.sin((l + 1) * kappa * h) / h for l in range(m // 2))


def kpp(kappa, h, m=6):
    import pyoracle
    c = pyoracle.coef(m)
    return sum(2.0 * c[l] * (1.0 - math.cos((l + 1) * kappa * h)) / ((l + 1) * h) ** 2 for l in range(m // 2))


def grid(M):
    cs = [(np.arange(M.n[a]) - M.h) * (M.L[a] / M.N[a]) for a in range(3)]
    return np.meshgrid(*cs, indexing="ij")


def inner(M):
    h = M.h
    return tuple(slice(h, h + M.N[a]) for a in range(3))


def vec(*c):
    return np.ascontiguousarray(np.stack(c, -1))


N0 = (16, 12, 8)
L0 = (TWO_PI, TWO_PI, TWO_PI)


def test_mhd_rhs_lorentz_and_induction(orc):
    M = orc.OMesh(N0, L0)
    X, Y, Z = grid(M)
    hx, hy = L0[0] / N0[0], L0[1] / N0[1]
    S = inner(M)
    nu, eta = 0.3, 0.2
    zero = 0 * X
    # Lorentz force on U = 0 of B = (sin y, sin 2x, 0):
    #   div(B, B) = (k'_y(1) sin 2x cos y, k'_x(2) cos 2x sin y, 0)
    #   grad(|B|^2 / 2) = (k'_x(4) sin 4x / 4, k'_y(2) sin 2y / 4, 0)
    B = vec(np.sin(Y), np.sin(2 * X), zero)
    dU, dB = M.mhd_rhs(vec(zero, zero, zero), B, nu, eta)
    ex = [kp(1, hy) * np.sin(2 * X) * np.cos(Y) - 0.25 * kp(4, hx) * np.sin(4 * X),
          kp(2, hx) * np.cos(2 * X) * np.sin(Y) - 0.25 * kp(2, hy) * np.sin(2 * Y), zero]
    for c in range(3):
        assert np.max(np.abs(dU[S][..., c] - ex[c][S])) < 1e-13
    # with U = 0 the induction vanishes and B only diffuses
    exb = [-eta * kpp(1, hy) * np.sin(Y), -eta * kpp(2, hx) * np.sin(2 * X), zero]
    for c in range(3):
        assert np.max(np.abs(dB[S][..., c] - exb[c][S])) < 1e-13
    # induction of U = (0, sin x, 0) in B = (1, 0, 0): -div(U, B) + div(B, U) = (0, k'_x cos x, 0);
    # the uniform B exerts no force, U (a shear in x) does not advect itself
    U = vec(zero, np.sin(X), zero)
    B = vec(1.0 + zero, zero, zero)
    dU, dB = M.mhd_rhs(U, B, nu, eta)
    for c, e in enumerate([zero, kp(1, hx) * np.cos(X), zero]):
        assert np.max(np.abs(dB[S][..., c] - e[S])) < 1e-13
    for c, e in enumerate([zero, -nu * kpp(1, hx) * np.sin(X), zero]):
        assert np.max(np.abs(dU[S][..., c] - e[S])) < 1e-13


def test_mhd_rhs_without_field_is_the_ns_rhs(orc):
    rng = np.random.default_rng(4)
    M = orc.OMesh(N0, L0)
    S = inner(M)
    U = M.from_compact(rng.standard_normal((3, *M.stored)))
    dU, dB = M.mhd_rhs(U, M.field(3), 0.3, 0.2)
    ref = -M.div_adv(U, U) + M.laplacian(0.3, U)
    assert np.max(np.abs(dU[S] - ref[S])) < 1e-13 * np.max(np.abs(ref))
    assert np.max(np.abs(dB[S])) == 0.0


def alfven(M, a, b):
    X, Y, Z = grid(M)
    zero = 0 * X
    return vec(zero, a * np.sin(X), zero), vec(1.0 + zero, b * np.cos(X), zero)


def amplitudes(M, U, B):
    """(a, b) of U_y = a sin x, B_y = b cos x, checking the other parts."""
    X, Y, Z = grid(M)
    S = inner(M)
    s, c = np.sin(X)[S], np.cos(X)[S]
    a = np.sum(U[S][..., 1] * s) / np.sum(s * s)
    b = np.sum(B[S][..., 1] * c) / np.sum(c * c)
    assert np.max(np.abs(U[S][..., 1] - a * s)) < 1e-13
    assert np.max(np.abs(B[S][..., 1] - b * c)) < 1e-13
    assert np.max(np.abs(U[S][..., [0, 2]])) < 1e-13
    assert np.max(np.abs(B[S][..., 0] - 1.0)) < 1e-13 and np.max(np.abs(B[S][..., 2])) < 1e-13
    return a, b


def test_mhd_reference_composition_closed_form(orc):
    """UEqn.H: U* = U + (sum of RK4 weights) dt R_U(U, B), projected;
    BEqn.H: classic RK4 of B' = k' u1 + lambda B with the projected u1
    frozen, lambda = -eta k''^2:  b1 = R4(z) b0 + dt phi(z) k' u1."""
    M = orc.OMesh(N0, L0)
    P = orc.OPoisson(M)
    a0, b0, nu, eta, dt = 1.0, 0.5, 0.2, 0.1, 0.1
    U, B = alfven(M, a0, b0)
    p, pB = M.field(1), M.field(1)
    P.mhd_step(U, B, p, pB, nu, eta, dt, 0)
    h = L0[0] / N0[0]
    k1, kk = kp(1, h), kpp(1, h)
    u1 = a0 + dt * (-nu * kk * a0 - k1 * b0)
    z = -eta * kk * dt
    R4 = 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24
    phi = 1 + z / 2 + z * z / 6 + z ** 3 / 24
    b1 = R4 * b0 + dt * phi * k1 * u1
    a, b = amplitudes(M, U, B)
    assert abs(a - u1) < 1e-13 and abs(b - b1) < 1e-13


def test_mhd_ssp3_composition_closed_form(orc):
    """SSP-RK3 with both fields projected every stage: (a, b) advances by
    I + Z + Z^2/2 + Z^3/6, Z = dt [[-nu k''^2, -k'], [k', -eta k''^2]]."""
    M = orc.OMesh(N0, L0)
    P = orc.OPoisson(M)
    a0, b0, nu, eta, dt = 1.0, 0.5, 0.2, 0.1, 0.1
    U, B = alfven(M, a0, b0)
    P.mhd_step(U, B, M.field(1), M.field(1), nu, eta, dt, 1)
    h = L0[0] / N0[0]
    k1, kk = kp(1, h), kpp(1, h)
    Z = dt * np.array([[-nu * kk, -k1], [k1, -eta * kk]])
    R = np.eye(2) + Z + Z @ Z / 2 + Z @ Z @ Z / 6
    a1, b1 = R @ np.array([a0, b0])
    a, b = amplitudes(M, U, B)
    assert abs(a - a1) < 1e-13 and abs(b - b1) < 1e-13


def test_dissipation_taylor_green_closed_form(orc):
    """nu <|grad u|^2> of u = sin x cos y cos z, v = -cos x sin y cos z on an
    uneven grid: each of the six non-zero derivatives has mean square
    k'_d^2 / 8, so the mean is (k'_x^2 + k'_y^2 + k'_z^2) / 4."""
    N, L = (12, 16, 20), (TWO_PI, 2 * TWO_PI, TWO_PI)
    M = orc.OMesh(N, L)
    X, Y, Z = grid(M)
    U = vec(np.sin(X) * np.cos(Y) * np.cos(Z), -np.cos(X) * np.sin(Y) * np.cos(Z), 0 * X)
    nu = 0.3
    ks = [kp(1, L[a] / N[a]) for a in range(3)]
    exp = nu * sum(k * k for k in ks) / 4
    assert abs(M.dissipation(U, nu) - exp) < 1e-12 * exp  # a sum over 3840 nodes


def lit_symbol(kappa, h, m=6):
    """Half-shifted literal derivative: sin(kappa x) -> S cos(kappa (x - h/2))."""
    import pyoracle
    c = pyoracle.coef(m)
    return sum(2.0 * c[l] * math.sin((2 * l + 1) * kappa * h / 2) / ((l + 1) * h) for l in range(m // 2))


def test_mhd_diag_closed_forms(orc):
    """U = (a sin y, 0, 0), B = (0, 0, b sin x): Ek = a^2/4, Em = b^2/4,
    |curl U| = |a| S_y |cos(y - h_y/2)|, |curl B| = |b| S_x |cos(x - h_x/2)|
    (the literal curl), epsilon = -mu omega^2 - eta Jav^2, Umax = |a|."""
    N, L = (12, 16, 8), (TWO_PI, TWO_PI, TWO_PI)
    M = orc.OMesh(N, L)
    X, Y, Z = grid(M)
    zero = 0 * X
    a, b, mu, eta = 0.7, -1.3, 0.05, 0.02
    U = vec(a * np.sin(Y), zero, zero)
    B = vec(zero, zero, b * np.sin(X))
    out = M.mhd_diag(U, B, mu, eta)
    hx, hy = L[0] / N[0], L[1] / N[1]
    xs, ys = np.arange(N[0]) * hx, np.arange(N[1]) * hy
    om = abs(a) * lit_symbol(1, hy) * np.mean(np.abs(np.cos(ys - hy / 2)))
    jav = abs(b) * lit_symbol(1, hx) * np.mean(np.abs(np.cos(xs - hx / 2)))
    jmax = abs(b) * lit_symbol(1, hx) * np.max(np.abs(np.cos(xs - hx / 2)))
    exp = [a * a / 4, b * b / 4, jav, jmax, om, -mu * om * om - eta * jav * jav, abs(a)]
    for q, e in enumerate(exp):
        assert abs(out[q] - e) < 1e-13 * max(1.0, abs(e)), (q, out[q], e)
    # the literal curl's symbol differs from the central stencil's (the
    # pin would not tell them apart otherwise)
    assert abs(lit_symbol(1, hx) - kp(1, hx)) > 1e-3


@pytest.mark.parametrize("s1,s2", [(0.003, 0.002), (0.0008, 0.004), (0.0025, 0.0025)])
def test_stretched_z_is_vinokur_two_sided(orc, s1, s2):
    """Mesh::z with hypz (s1, s2) on N_z cells: zeta = z/L_z = u/(2A + (1-A)u),
    u = 1 + tanh(b (xi - 1/2))/tanh(b/2), xi = k/N_z.  Checked through the
    inverse map (zeta -> u -> xi must return k/N_z, with A and b recovered
    from the end slopes) and the end slopes themselves, which the design
    conditions sinh(b)/b = 1/(N_z sqrt(s1 s2)), A = sqrt(s1/s2) make
    dzeta/dxi = N_z s2 at z = 0 and N_z s1 at z = L_z."""
    Nz, Lz = 256, 2.0  # the stretching needs N_z sqrt(s1 s2) < 1
    M = orc.OMesh((8, 8, Nz), (1.0, 1.0, Lz), hyp=(s1, s2))
    h = M.h
    zeta = np.array([M.z(h + k) for k in range(Nz + 1)]) / Lz
    assert abs(zeta[0]) < 1e-15 and abs(zeta[-1] - 1.0) < 1e-15
    # end slopes from one-sided fourth-order differences in xi
    d0 = Nz * (-25 * zeta[0] + 48 * zeta[1] - 36 * zeta[2] + 16 * zeta[3] - 3 * zeta[4]) / 12
    d1 = Nz * (25 * zeta[-1] - 48 * zeta[-2] + 36 * zeta[-3] - 16 * zeta[-4] + 3 * zeta[-5]) / 12
    assert abs(d0 - Nz * s2) < 1e-5 * Nz * s2
    assert abs(d1 - Nz * s1) < 1e-5 * Nz * s1
    # inverse map with A and b from the slopes
    A = math.sqrt(d1 / d0)
    Bv = 1.0 / math.sqrt(d0 * d1)
    b = brentq(lambda x: math.sinh(x) / x - Bv, 1e-9, 50.0, xtol=1e-15) if Bv > 1 + 1e-12 else 0.0
    xi = np.arange(Nz + 1) / Nz
    u = 2 * A * zeta / (1 - (1 - A) * zeta)
    if b > 0:
        back = 0.5 + np.arctanh(np.clip((u - 1) * math.tanh(b / 2), -1 + 1e-16, 1 - 1e-16)) / b
        assert np.max(np.abs(back - xi)) < 1e-5
    # symmetric stretching (s1 = s2): z(N - k) = L - z(k)
    if s1 == s2:
        assert np.max(np.abs(zeta[::-1] - (1 - zeta))) < 1e-14


def test_cfl_as_written(orc):
    """tools::CFL: cfl_x is the maximum of U_x dt/dx from 0; cfl_y and cfl_z
    keep max(cfl_x, U_d dt/h_d) of the last node of the loop (the padded
    index n - m/2 - 1 in every direction: the image of node 0)."""
    N, L = (8, 12, 16), (1.0, 2.0, 3.0)
    M = orc.OMesh(N, L)
    hx, hy, hz = (L[a] / N[a] for a in range(3))
    dt = 0.01
    rng = np.random.default_rng(3)
    U = M.from_compact(rng.standard_normal((3, *M.stored)))
    U[..., 1] *= 5.0
    h = M.h
    last = (M.n[0] - h - 1, M.n[1] - h - 1, M.n[2] - h - 1)
    S = inner(M)
    cx = max(0.0, np.max(U[S][..., 0]) * dt / hx)
    out = M.cfl(U, dt)
    assert out[0] == pytest.approx(cx, rel=1e-15)
    assert out[1] == pytest.approx(max(cx, U[last][1] * dt / hy), rel=1e-15)
    assert out[2] == pytest.approx(max(cx, U[last][2] * dt / hz), rel=1e-15)
    for d, hd in enumerate((hx, hy, hz)):
        assert out[3 + d] == pytest.approx(np.max(np.abs(U[S][..., d])) * dt / hd, rel=1e-15)
    # the last node is the periodic image of node 0
    assert np.array_equal(U[last], U[h, h, h])
    # a uniform field: cfl_y = max(cfl_x, v dt/dy)
    V = np.zeros_like(U)
    V[..., 0], V[..., 1], V[..., 2] = -1.0, 2.0, 0.5
    out = M.cfl(V, dt)
    assert out[0] == 0.0 and out[1] == 2.0 * dt / hy and out[2] == 0.5 * dt / hz
