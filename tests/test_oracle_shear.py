"""Pins of the oracle's restatement of app/shearLayer's step
(pyoracle.shear_layer_step, CPU).

* A shear wave U = (0, a sin kx, 0) with B = 0 on the walled mesh: it is
  y-independent, so the Neumann fills of u and v leave it alone, its
  advection vanishes and its divergence is zero (p = 0 after the solve);
  each stage of UEqn.H then sets U_rk = UOld + w_rk dt L U_{rk-1} with the
  stage weights (1/6, 1/3, 1/3, 1/6) of shearLayer.cpp:145, L U = -mu k''^2 U
  (k'' the sixth-order second-derivative symbol), so one step multiplies a
  by 1 + z/6 (1 + z/3 (1 + z/3 (1 + z/6))), z = -mu k''^2 dt.
* The fills as the reference leaves them after a step of the example's
  double shear layer: u, v at each wall node equal to the node next to it
  (neumann 0), w zero there (dirichlet 0), p copied likewise, the periodic
  image nodes equal to node 0.
* The solve takes nodes 0 .. N-1 of y as one period: a divergence that is a
  single y mode of period N dy gives the closed-form potential there.
"""
import math

import numpy as np

N = (16, 24, 8)
L = (1.0, 1.5, 0.5)
UBC = "\n".join([f"{d} {t} 0 {c}" for c, t in ((0, "neumann"), (1, "neumann"), (2, "dirichlet"))
                 for d in ("north", "south")] +
                [f"{d} periodic 0 {c}" for c in range(3) for d in ("east", "west", "top", "bottom")])
PBC = "north neumann 0 0\nsouth neumann 0 0\neast periodic 0 0\nwest periodic 0 0\ntop periodic 0 0\n" \
      "bottom periodic 0 0\n"


def kpp(kappa, h, m=6):
    import pyoracle
    c = pyoracle.coef(m)
    return sum(2.0 * c[l] * (1.0 - math.cos((l + 1) * kappa * h)) / ((l + 1) * h) ** 2 for l in range(m // 2))


def kp(kappa, h, m=6):
    import pyoracle
    c = pyoracle.coef(m)
    return sum(c[l] / (l + 1) * math.sin((l + 1) * kappa * h) / h for l in range(m // 2))


def setup(orc):
    O = orc.OMesh(N, L, wall_y=True)
    P = orc.OPoisson(orc.OMesh(N, L))
    return O, P, orc.bc_entries(UBC, 3), orc.bc_entries(PBC, 1)


def nodes(O):
    h = O.h
    cs = [(np.arange(O.n[a]) - h) * (L[a] / N[a]) for a in range(3)]
    return np.meshgrid(*cs, indexing="ij")


def test_shear_wave_follows_the_stage_polynomial(orc):
    O, P, ubc, pbc = setup(orc)
    X, Y, Z = nodes(O)
    kappa = 2 * math.pi * 2 / L[0]
    a, mu, dt = 0.7, 0.01, 0.05
    U = np.ascontiguousarray(np.stack([0 * X, a * np.sin(kappa * X), 0 * X], -1))
    O.bc_apply(U, ubc)
    B = O.field(3)
    p = O.field(1) + 0.87
    orc.shear_layer_step(O, P, U, B, p, mu, dt, ubc, pbc)
    z = -mu * kpp(kappa, L[0] / N[0]) * dt
    amp = 1 + z / 6 * (1 + z / 3 * (1 + z / 3 * (1 + z / 6)))
    S = tuple(slice(O.h, O.h + N[a] + (1 if a == 1 else 0)) for a in range(3))
    assert np.max(np.abs(U[S][..., 1] - amp * a * np.sin(kappa * X[S]))) < 1e-14
    assert np.max(np.abs(U[S][..., [0, 2]])) < 1e-14
    assert np.max(np.abs(p[S])) < 1e-14
    # a planted stage-weight mistake (the classic RK4 sum) is told apart
    classic = 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24
    assert abs(classic - amp) * a > 1e-6


def test_fills_after_a_step(orc):
    from test_gpu_shear_literal import initial_state, N as NE, L as LE, UBC as UE, PBC as PE, BBC as BE
    O = orc.OMesh(NE, LE, wall_y=True)
    P = orc.OPoisson(orc.OMesh(NE, LE))
    ubc, pbc, bbc = orc.bc_entries(UE, 3), orc.bc_entries(PE, 1), orc.bc_entries(BE, 3)
    U, B, p = initial_state(O)
    O.bc_apply(U, ubc)
    O.bc_apply(p, pbc)
    O.bc_apply(B, bbc)
    orc.shear_layer_step(O, P, U, B, p, 1.0 / 28000.0, 1e-3, ubc, pbc)
    h, n1 = O.h, NE[1]
    w0, wN = h, h + n1
    for c in (0, 1):
        assert np.array_equal(U[:, w0, :, c], U[:, w0 + 1, :, c])
        assert np.array_equal(U[:, wN, :, c], U[:, wN - 1, :, c])
    assert not np.any(U[:, w0, :, 2]) and not np.any(U[:, wN, :, 2])
    assert np.array_equal(p[:, w0, :, 0], p[:, w0 + 1, :, 0]) and np.array_equal(p[:, wN, :, 0], p[:, wN - 1, :, 0])
    # x and z periodic: node N is node 0's image
    assert np.array_equal(U[h + NE[0]], U[h]) and np.array_equal(U[:, :, h + NE[2]], U[:, :, h])
    # the periodic solve made U discretely solenoidal wherever the stencils
    # of div(grad p) (6 nodes) and div U (3 nodes) stay clear of the walls,
    # up to the modes the solve drops (k'^2 = 0: the mean and the Nyquist
    # modes), which the walls' non-periodic divergence feeds
    d = O.div(U)
    S = (slice(h, h + NE[0]), slice(h + 7, h + n1 - 6), slice(h, h + NE[2]))
    Sw = (slice(h, h + NE[0]), slice(h, h + 4), slice(h, h + NE[2]))
    assert np.max(np.abs(d[S])) < 1e-4 * np.max(np.abs(d[Sw]))


def test_solve_takes_the_first_n_nodes_as_one_period(orc):
    """rhs = cos(2 pi q y / (N dy)) on nodes 0 .. N-1 gives
    p = -rhs / k'_y(q)^2 there, whatever node N holds."""
    O, P, ubc, pbc = setup(orc)
    X, Y, Z = nodes(O)
    q = 3
    ky = 2 * math.pi * q / L[1]
    rhs = np.ascontiguousarray(np.cos(ky * Y)[..., None])
    rhs[:, O.h + N[1], :, 0] = 123.0  # the wall node is not part of the solve
    p = O.field(1) + 5.0
    P.solve(rhs, p)
    S = tuple(slice(O.h, O.h + N[a]) for a in range(3))
    k2 = kp(ky, L[1] / N[1]) ** 2
    assert np.max(np.abs(p[S][..., 0] + np.cos(ky * Y[S]) / k2)) < 1e-12
    assert np.all(p[:, O.h + N[1], :, 0] == 5.0)
