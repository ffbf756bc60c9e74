"""app/shearLayer as the reference runs it, on the GPU, against the oracle's
restatement (pyoracle.shear_layer_step): the example case of
examples/shearLayer (mesh 64 x 128 x 8 on 1 x 2 x 0.125, y walled), its BC
files (UBC: u and v neumann 0, w dirichlet 0 on north and south; pBC:
neumann 0; BBC: periodic everywhere), the initial state of
shearLayer.cpp:86-136 and UEqn.H's per-stage projection with the frozen
magnetic field's force.  The device takes the reference's fills
(nova_ctx_set_wall_bc: odd or even reflections per component, wall nodes
overwritten after every assignment) and the reference's FFT solve
(y periodic over its first N nodes, nova_poisson_set_wall_solve).  Bar:
1e-12 relative L2 in U and p after one step, 1e-11 after five."""
import math

import numpy as np
import pytest

from test_gpu_parity import dev, host, rel_l2

pytestmark = pytest.mark.gpu

N = (64, 128, 8)          # examples/shearLayer/mesh
L = (1.0, 2.0, 0.125)
UBC = """north neumann 0 0
south neumann 0 0
east  periodic 0 0
west  periodic 0 0
top periodic 0 0
bottom periodic 0 0
north neumann 0 1
south neumann 0 1
east  periodic 0 1
west  periodic 0 1
top periodic 0 1
bottom periodic 0 1
north dirichlet 0 2
south dirichlet 0 2
east  periodic 0 2
west  periodic 0 2
top periodic 0 2
bottom periodic 0 2
"""
PBC = """east periodic 0 0
west periodic 0 0
bottom periodic 0 0
north neumann 0 0
top periodic 0 0
south neumann 0 0
"""
BBC = "\n".join(f"{d} periodic 0 {c}" for c in range(3)
                for d in ("north", "south", "east", "west", "top", "bottom")) + "\n"
MU = 1.0 / 28000.0        # shearLayer.cpp:65
DT = 0.001                # Time(0.001, ...), shearLayer.cpp:24


@pytest.fixture(scope="module")
def nv():
    import torch
    import novaclone_amd as nv
    assert torch.cuda.is_available(), "GPU tests need a device"
    nv.lib()
    return nv


def initial_state(O):
    """shearLayer.cpp:86-136 on the nodes 0 .. N of every direction (padded
    AoS); U, p and B then get their BC files' fills."""
    u0, delta, cn = 1.0, 1.0 / 28.0, 0.001
    h = O.h
    dx, dy = L[0] / N[0], L[1] / N[1]
    x = np.arange(N[0] + 1) * dx
    y = np.arange(N[1] + 1) * dy
    X, Y = np.meshgrid(x, y, indexing="ij")
    psi = u0 * np.exp(-(((Y - 0.5) * (Y - 0.5)) / (delta * delta))) * (np.cos(8.0 * math.pi * X)
                                                                       + np.cos(20.0 * math.pi * X))
    psi2 = u0 * np.exp(-(((Y - 1.5) * (Y - 1.5)) / (delta * delta))) * (np.cos(8.0 * math.pi * X)
                                                                        + np.cos(20.0 * math.pi * X))
    low = Y < 1.0
    ux = np.where(low, u0 * np.tanh(((4.0 * Y) - 2.0) / delta), -u0 * np.tanh(((4.0 * Y) - 6.0) / delta))
    bx = ux.copy()
    ux = ux + np.where(low, cn * dy * psi, cn * dy * psi2)
    uy = np.where(low, -cn * dx * psi, -cn * dx * psi2)
    U, B, p = O.field(3), O.field(3), O.field(1)
    sl = (slice(h, h + N[0] + 1), slice(h, h + N[1] + 1))
    for k in range(N[2] + 1):
        U[sl[0], sl[1], h + k, 0] = ux
        U[sl[0], sl[1], h + k, 1] = uy
        B[sl[0], sl[1], h + k, 0] = bx
    p[...] = 0.87
    return U, B, p


def setup(nv, orc):
    import torch
    O = orc.OMesh(N, L, wall_y=True)
    P = orc.OPoisson(orc.OMesh(N, L))
    ubc, pbc, bbc = orc.bc_entries(UBC, 3), orc.bc_entries(PBC, 1), orc.bc_entries(BBC, 3)
    U, B, p = initial_state(O)
    O.bc_apply(U, ubc)
    O.bc_apply(p, pbc)
    O.bc_apply(B, bbc)
    G = nv.Mesh(N, L, bc=("periodic", "wall", "periodic"))
    G.set_wall_bc(1, [nv.WALL_NEUMANN0, nv.WALL_NEUMANN0, nv.WALL_DIRICHLET0])
    GP = nv.Poisson(G)
    GP.set_wall_solve(nv.WALL_SOLVE_REFERENCE)
    S = nv.Solver(GP, nv.MODEL_NS, nv.RK_SHEAR4, nu=MU)
    # the frozen field's force div(B, B) - grad(0.5 (B & B)) once, on the
    # periodic grid B's BC file describes; its node N_y is node 0's image
    Gp = nv.Mesh(N, L)
    Bp = dev(O.interior(B)[:, :, :N[1], :])
    BB = nv.field_op(Gp, nv.OP_DOT, Bp, Bp)
    Fp = nv.ex.div(Gp, Bp, Bp) - nv.ex.grad(Gp, (0.5 * BB).contiguous())
    F = torch.cat([Fp, Fp[:, :, :1, :]], dim=2).contiguous()
    S.set_force(F)
    Ud, pd = dev(O.interior(U)), dev(O.interior(p))
    return O, P, G, S, (U, B, p, ubc, pbc), (Ud, pd, F)


def test_reference_shear_layer_steps_match_oracle(nv, orc):
    O, P, G, S, (U, B, p, ubc, pbc), (Ud, pd, F) = setup(nv, orc)
    for step in range(5):
        orc.shear_layer_step(O, P, U, B, p, MU, DT, ubc, pbc)
        S.step(DT, 1, Ud, pd)
        eu = rel_l2(host(Ud), O.interior(U))
        ep = rel_l2(host(pd), O.interior(p))
        print("shearLayer step", step + 1, "rel L2 U", eu, "p", ep, flush=True)
        assert eu <= (1e-12 if step == 0 else 1e-11) and ep <= (1e-12 if step == 0 else 1e-11)
    # the fills as the reference leaves them: u, v copied from the node next
    # to each wall, w zero there, p copied likewise
    u = host(Ud)
    assert np.array_equal(u[:2, :, 0, :], u[:2, :, 1, :]) and np.array_equal(u[:2, :, -1, :], u[:2, :, -2, :])
    assert not np.any(u[2, :, 0, :]) and not np.any(u[2, :, -1, :])
    q = host(pd)
    assert np.array_equal(q[0, :, 0, :], q[0, :, 1, :]) and np.array_equal(q[0, :, -1, :], q[0, :, -2, :])


def test_reference_fills_change_the_problem(nv, orc):
    # the same state under the device's default rule (v odd, u and w even,
    # wall nodes kept) and the even-extension solve is a different step
    O, P, G, S, (U, B, p, ubc, pbc), (Ud, pd, F) = setup(nv, orc)
    G2 = nv.Mesh(N, L, bc=("periodic", "wall", "periodic"))
    S2 = nv.Solver(nv.Poisson(G2), nv.MODEL_NS, nv.RK_SHEAR4, nu=MU)
    S2.set_force(F)
    U2, p2 = Ud.clone(), pd.clone()
    S.step(DT, 1, Ud, pd)
    S2.step(DT, 1, U2, p2)
    assert rel_l2(host(U2), host(Ud)) > 1e-8


def test_reference_shear_layer_graph_replay_and_refusals(nv, orc):
    import torch
    O, P, G, S, _, (Ud, pd, F) = setup(nv, orc)
    Sg = nv.Solver(S.poisson, nv.MODEL_NS, nv.RK_SHEAR4, nu=MU, graph=True)
    Sg.set_force(F)
    Ue, pe = Ud.clone(), pd.clone()
    S.step(DT, 3, Ue, pe)
    Sg.step(DT, 3, Ud, pd)
    torch.cuda.synchronize()
    assert torch.equal(Ue, Ud) and torch.equal(pe, pd)
    # the models and steps that keep the reflection rule refuse these fills
    with pytest.raises(nv.NovaError):
        nv.Adjoint(nv.Solver(S.poisson, nv.MODEL_NS, nv.RK_SSP3, nu=MU), 2, 4)
    with pytest.raises(nv.NovaError):
        M = nv.Solver(S.poisson, nv.MODEL_MHD, nv.RK_SSP3, nu=MU, eta=MU)
        M.step(DT, 1, Ue, pe, Ud, pd)
    with pytest.raises(nv.NovaError):
        G.set_wall_bc(0, [nv.WALL_NEUMANN0] * 3)  # x is periodic
