"""Parity of the weakly-compressible step (app/TG/TG.cpp:71-76 with
rhoEqn.H, pEqn.H, UEqn.H; app/OT with its frozen magnetic field;
app/shearLayerWC with the temperature equation and buoyancy) through
the C ABI with the oracle: one step per-field relative L2 <= 1e-12,
kinetic-energy history over 20 steps <= 1e-10 relative, mass conserved to
rounding; periodic and y-walled grids, stencil orders m = 4, 6, 8."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TWO_PI = 2.0 * math.pi
RTOL = 1e-12


def rel_l2(a, b):
    nb = np.linalg.norm(np.ravel(b))
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / (nb if nb > 0 else 1.0)


@pytest.fixture(scope="module")
def nv():
    import torch
    import novaclone_amd as nv
    assert torch.cuda.is_available(), "GPU tests need a device"
    nv.lib()
    return nv


def dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def tg_wc(O, seed):
    """The WC Taylor-Green state of app/TG/TG.cpp:39-55 on the stored nodes
    (origin 0, so x = 2 pi i / N), plus a seeded 1e-3 perturbation of U."""
    s0, s1, s2 = O.stored
    x = [np.arange(O.stored[a]) * (O.L[a] / O.N[a]) for a in range(3)]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    U = np.stack([np.sin(X) * np.cos(Y) * np.cos(Z), -np.cos(X) * np.sin(Y) * np.cos(Z), 0 * X])
    U += 1e-3 * np.random.default_rng(seed).standard_normal(U.shape)
    if O.wall_y:
        U[1, :, 0, :] = 0.0
        U[1, :, -1, :] = 0.0
    p = 50 + 1.0 / 16.0 * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2)
    rho = np.sqrt(p / 50.0)
    return U, rho[None], p[None]


CASES = [
    dict(N=(32, 24, 16), L=(TWO_PI,) * 3, m=6, wall_y=False),
    dict(N=(16, 16, 24), L=(TWO_PI,) * 3, m=4, wall_y=False),
    dict(N=(16, 12, 16), L=(TWO_PI,) * 3, m=8, wall_y=False),
    dict(N=(24, 16, 16), L=(TWO_PI, math.pi, TWO_PI), m=6, wall_y=True),
    # walls at every other stencil width: m = 2 with two k tiles (70 = 64 + 6)
    # and j rows that do not fill the last tile, m = 4 with several i chunks,
    # m = 8 with no plane of loads in flight in the momentum stage
    dict(N=(20, 14, 70), L=(TWO_PI, math.pi, TWO_PI), m=2, wall_y=True),
    dict(N=(40, 10, 20), L=(TWO_PI, math.pi, TWO_PI), m=4, wall_y=True),
    dict(N=(16, 12, 16), L=(TWO_PI, math.pi, TWO_PI), m=8, wall_y=True),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c['N']}-m{c['m']}-wall{int(c['wall_y'])}")
def test_wc_step_matches_oracle(nv, orc, case):
    print("case", case, flush=True)
    bc = ("periodic", "wall" if case["wall_y"] else "periodic", "periodic")
    G = nv.Mesh(case["N"], case["L"], m=case["m"], bc=bc)
    O = orc.OMesh(case["N"], case["L"], m=case["m"], wall_y=case["wall_y"])
    U, rho, p = tg_wc(O, 7)
    mu, pc, dt = 1.0 / 1600.0, 50.0, 0.01
    W = nv.WC(G, mu, pc)
    Ud, rd, pd = dev(U), dev(rho), dev(np.zeros_like(p))
    W.step(dt, 1, Ud, rd, pd)
    Uo, ro, po = O.from_compact(U), O.from_compact(rho), O.field(1)
    O.wc_step(Uo, ro, po, mu, pc, dt)
    eu = rel_l2(host(Ud), O.interior(Uo))
    er = rel_l2(host(rd), O.interior(ro))
    ep = rel_l2(host(pd), O.interior(po))
    print("rel L2 U", eu, "rho", er, "p", ep, flush=True)
    assert eu < RTOL and er < RTOL and ep < RTOL
    # the step moved the fields well beyond the tolerance
    assert rel_l2(host(rd), rho) > 1e-6 and rel_l2(host(Ud), U) > 1e-6


def test_wc_history_and_mass(nv, orc):
    N, L = (24, 24, 24), (TWO_PI,) * 3
    G = nv.Mesh(N, L)
    O = orc.OMesh(N, L)
    U, rho, p = tg_wc(O, 3)
    mu, pc, dt = 1.0 / 1600.0, 50.0, 0.01
    W = nv.WC(G, mu, pc)
    Ud, rd, pd = dev(U), dev(rho), dev(np.zeros_like(p))
    Uo, ro, po = O.from_compact(U), O.from_compact(rho), O.field(1)
    m0 = rho.sum()
    ke_g, ke_o = [], []
    for _ in range(20):
        W.step(dt, 1, Ud, rd, pd)
        O.wc_step(Uo, ro, po, mu, pc, dt)
        ke_g.append(nv.ke(G, Ud))
        ke_o.append(O.ke(Uo))
    ke_g, ke_o = np.array(ke_g), np.array(ke_o)
    err = np.abs(ke_g - ke_o).max() / np.abs(ke_o).max()
    print("KE history rel", err, "mass", host(rd).sum() / m0 - 1.0, flush=True)
    assert err < 1e-10
    assert abs(host(rd).sum() - m0) < 1e-12 * m0
    assert rel_l2(host(Ud), O.interior(Uo)) < 1e-10


@pytest.mark.parametrize("wall", [False, True])
def test_wc_ot_with_frozen_field_matches_oracle(nv, orc, wall):
    # app/OT (OT.cpp:50-76): U = (-2 sin y, 2 sin x, 0), B = 0.8 (-2 sin 2y +
    # sin z, 2 sin x + sin z, sin x + sin y), rho = 1, mu = DB = 0.1256,
    # DBU = 1/2; BEqn.H leaves B as it is, UEqn.H adds its Lorentz force
    N, L = (24, 16, 32), (TWO_PI, math.pi if wall else TWO_PI, TWO_PI)
    bc = ("periodic", "wall" if wall else "periodic", "periodic")
    G = nv.Mesh(N, L, bc=bc)
    O = orc.OMesh(N, L, wall_y=wall)
    x = [np.arange(O.stored[a]) * (O.L[a] / O.N[a]) for a in range(3)]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    U = np.stack([-2 * np.sin(Y), 2 * np.sin(X), 0 * X])
    B = 0.8 * np.stack([-2 * np.sin(2 * Y) + np.sin(Z), 2 * np.sin(X) + np.sin(Z), np.sin(X) + np.sin(Y)])
    if wall:
        U[1, :, 0, :] = U[1, :, -1, :] = 0.0
        B[1, :, 0, :] = B[1, :, -1, :] = 0.0
    rho = np.ones((1, *O.stored))
    mu, pc, dbu, dt = 0.1256, 50.0, 0.5, 1e-3
    W = nv.WC(G, mu, pc)
    Bd = dev(B)
    W.set_magnetic(Bd, dbu)
    Ud, rd, pd = dev(U), dev(rho), dev(np.zeros_like(rho))
    W.step(dt, 3, Ud, rd, pd)
    Uo, ro, po, Bo = O.from_compact(U), O.from_compact(rho), O.field(1), O.from_compact(B)
    for _ in range(3):
        O.wc_step(Uo, ro, po, mu, pc, dt, B=Bo, dbu=dbu)
    eu = rel_l2(host(Ud), O.interior(Uo))
    er = rel_l2(host(rd), O.interior(ro))
    ep = rel_l2(host(pd), O.interior(po))
    print("wall", wall, "rel L2 U", eu, "rho", er, "p", ep, flush=True)
    assert eu < RTOL and er < RTOL and ep < RTOL
    assert np.array_equal(host(Bd), B)
    # the Lorentz force moved the flow: without B the step differs
    W.set_magnetic(None, 0.0)
    U2, r2, p2 = dev(U), dev(rho), dev(np.zeros_like(rho))
    W.step(dt, 3, U2, r2, p2)
    assert rel_l2(host(U2), host(Ud)) > 1e-6


@pytest.mark.parametrize("wall", [False, True])
def test_wc_shear_layer_with_temperature_matches_oracle(nv, orc, wall):
    # app/shearLayerWC (shearLayer.cpp:31-45, 114-150, 197-200) with the shear
    # along y (walls in y: the oracle's walled direction): u = u0/2 tanh(2y/h0)
    # plus a seeded perturbation, T = t0/2 tanh(2y/h0) plus noise, rho0
    # uniform, p = c^2/(2 rho0) rho^2, buoyancy along -y
    mach, re0, ri0, pr, rho0, c = 0.2, 1967.0, 0.08, 7.0, 999.57, 0.0417
    u0, h0, t0 = c * mach, 0.2, 0.1
    mu = rho0 * u0 * h0 / re0
    ka = mu / pr
    g = (0.0, -ri0 * u0 * u0 / t0 / h0, 0.0)
    pc = c * c / 2.0 / rho0
    N, L = (24, 16, 16), (14 * h0, 15 * h0, 1.7 * h0)
    bc = ("periodic", "wall" if wall else "periodic", "periodic")
    G = nv.Mesh(N, L, bc=bc)
    O = orc.OMesh(N, L, wall_y=wall)
    x = [np.arange(O.stored[a]) * (O.L[a] / O.N[a]) for a in range(3)]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    y = Y - 7.5 * h0
    rng = np.random.default_rng(13)
    env = 1.0 / np.cosh(2 * y / h0) ** 2
    U = np.stack([u0 / 2 * np.tanh(2 * y / h0), 0 * X, 0 * X]) + 0.005 * u0 / 2 * rng.uniform(-1, 1, (3, *O.stored)) * env
    if wall:
        U[1, :, 0, :] = U[1, :, -1, :] = 0.0
    T = (t0 / 2 * np.tanh(2 * y / h0) + 0.005 * t0 / 2 * rng.uniform(-1, 1, O.stored) * env)[None]
    rho = np.full((1, *O.stored), rho0)
    p = pc * rho * rho
    dt = 0.5
    W = nv.WC(G, mu, pc)
    W.set_temperature(ka, g)
    Ud, rd, pd, Td = dev(U), dev(rho), dev(p), dev(T)
    W.step_t(dt, 3, Ud, rd, pd, Td)
    Uo, ro, po, To = O.from_compact(U), O.from_compact(rho), O.from_compact(p), O.from_compact(T)
    for _ in range(3):
        O.wcs_step(Uo, ro, po, To, mu, ka, g, pc, dt)
    errs = [rel_l2(host(a), O.interior(b)) for a, b in ((Ud, Uo), (rd, ro), (pd, po), (Td, To))]
    print("wall", wall, "rel L2 U, rho, p, T", errs, flush=True)
    assert max(errs) < RTOL
    assert rel_l2(host(Td), T) > 1e-8 and rel_l2(host(rd), rho) > 1e-12


@pytest.mark.parametrize("wall", [False, True])
def test_wc_hydrostatic_initial_state_matches_oracle(nv, orc, wall):
    # app/shearLayerWC/hydrostaticP.H and shearLayer.cpp:141-151: p from
    # pEqn.H at rho = rho0, plus the column sum of T dz g_z rho0 from the top;
    # then rho = (p/pc)^(1/2).  The sum runs in the reference's order, so the
    # pressure agrees to rounding of the last add at most.
    rho0, c, h0, t0 = 999.57, 0.0417, 0.2, 0.1
    pc = c * c / 2.0 / rho0
    g = (0.0, 0.0, -0.08 * (0.2 * c) ** 2 / t0 / h0)
    N, L = (16, 24, 20), (14 * h0, 15 * h0, 1.7 * h0)
    bc = ("periodic", "wall" if wall else "periodic", "periodic")
    G = nv.Mesh(N, L, bc=bc)
    O = orc.OMesh(N, L, wall_y=wall)
    x = [np.arange(O.stored[a]) * (O.L[a] / O.N[a]) for a in range(3)]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    rng = np.random.default_rng(17)
    T = (t0 / 2 * np.tanh(2 * (Z - 0.85 * h0) / h0) + 0.005 * rng.uniform(-1, 1, O.stored))[None]
    p = np.full((1, *O.stored), (pc * rho0) * rho0)
    W = nv.WC(G, 1e-3, pc)
    W.set_temperature(1e-4, g)
    Td, pd, rd = dev(T), dev(p), dev(np.zeros_like(p))
    W.hydrostatic(Td, rho0, pd, rd)
    To, po, ro = O.from_compact(T), O.from_compact(p), O.field(1)
    O.wc_hydrostatic(To, po, ro, g, rho0, pc)
    ep, er = rel_l2(host(pd), O.interior(po)), rel_l2(host(rd), O.interior(ro))
    print("wall", wall, "rel L2 p", ep, "rho", er, flush=True)
    assert ep < 1e-15 and er < 1e-15
    # the column sums moved p well beyond rounding, and rho follows
    assert rel_l2(host(pd), p) > 1e-6 and rel_l2(host(rd), np.full_like(p, rho0)) > 1e-6


def test_wc_rejects_bad_arguments(nv):
    G = nv.Mesh((16, 16, 16), (TWO_PI,) * 3)
    with pytest.raises(nv.NovaError):
        nv.WC(G, -1.0)
    W = nv.WC(G, 0.01)
    import torch
    U = torch.zeros((3, 16, 16, 16), dtype=torch.float64, device="cuda")
    r = torch.ones((1, 16, 16, 16), dtype=torch.float64, device="cuda")
    with pytest.raises(nv.NovaError):
        W.step(0.01, 1, U, r, None)
    with pytest.raises(nv.NovaError):  # the temperature equation is not set up
        W.step_t(0.01, 1, U, r, r.clone(), r.clone())
    with pytest.raises(nv.NovaError):  # needs the temperature equation's g
        W.hydrostatic(r, 1.0, r.clone(), r.clone())
    with pytest.raises(nv.NovaError):
        W.set_temperature(-1.0, (0.0, 0.0, -1.0))
