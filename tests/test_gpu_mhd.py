"""The reference's own MHD step on the GPU (app/OT_Poisson/UEqn.H:14-40,
BEqn.H:14-45; app/mhdjet alike) against the oracle's restatement (mode 0):
frozen-stage RK4 of U with the Lorentz terms and its projection, then a
classic RK4 of B with the new U frozen and the s = 1 cleaning.  Bar: one
step <= 1e-12 relative L2 per field; the OT diagnostics of OT.cpp:134-176
over 100 steps within 1e-10 relative."""
import math

import numpy as np
import pytest

from test_gpu_parity import RTOL, dev, host, pair, rel_l2, wall_clean

pytestmark = pytest.mark.gpu
TWO_PI = 2.0 * math.pi


@pytest.fixture(scope="module")
def nv():
    import torch
    import novaclone_amd as nv
    assert torch.cuda.is_available(), "GPU tests need a device"
    nv.lib()
    return nv


def ot_fields(O, seed, amp=1e-3):
    """u = (-sin y, sin x, 0), B = (-sin y, sin 2x, 0) plus seeded noise,
    both projected once (solenoidal on both sides)."""
    x = [np.arange(O.stored[a]) * O.L[a] / O.N[a] for a in range(3)]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    rng = np.random.default_rng(seed)
    cu = np.stack([-np.sin(Y), np.sin(X), 0 * X]) + amp * rng.standard_normal((3, *O.stored))
    cb = np.stack([-np.sin(Y), np.sin(2 * X), 0 * X]) + amp * rng.standard_normal((3, *O.stored))
    cu, cb = wall_clean(O, cu), wall_clean(O, cb)
    import pyoracle
    OP = pyoracle.OPoisson(O)
    Uo, Bo = O.from_compact(cu), O.from_compact(cb)
    OP.project(Uo, 1.0)
    OP.project(Bo, 1.0)
    return OP, Uo, Bo


CASES = [
    dict(N=(24, 24, 24), L=(TWO_PI,) * 3, m=6, wall_y=False),
    dict(N=(16, 24, 12), L=(TWO_PI, 1.5 * TWO_PI, TWO_PI), m=4, wall_y=False),
    dict(N=(16, 12, 16), L=(TWO_PI, 3.0, TWO_PI), m=6, wall_y=True),
    dict(N=(12, 16, 16), L=(TWO_PI,) * 3, m=8, wall_y=False),
    dict(N=(16, 16, 24), L=(TWO_PI,) * 3, m=2, wall_y=False),
    dict(N=(16, 12, 16), L=(TWO_PI, 3.0, TWO_PI), m=8, wall_y=True),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c['N']}-m{c['m']}-wall{int(c['wall_y'])}")
def test_reference_mhd_step_matches_oracle(nv, orc, case):
    G, O = pair(nv, orc, case["N"], case["L"], m=case["m"], wall_y=case["wall_y"])
    OP, Uo, Bo = ot_fields(O, 3)
    U, B, p, pB = dev(O.interior(Uo)), dev(O.interior(Bo)), G.field(1), G.field(1)
    po, pBo = O.field(1), O.field(1)
    nu, eta, dt = 1e-3, 2e-3, 2e-3
    for _ in range(2):
        OP.mhd_step(Uo, Bo, po, pBo, nu, eta, dt, 0)
    S = nv.Solver(nv.Poisson(G), nv.MODEL_MHD, nv.RK_LITERAL4, nu=nu, eta=eta)
    S.step(dt, 2, U, p, B, pB)
    for got, ref in ((U, Uo), (B, Bo), (p, po)):
        e = rel_l2(host(got), O.interior(ref))
        print(case, e, flush=True)
        assert e <= RTOL
    # the cleaning potential pB is the potential of the divergence the four
    # B stages leave behind, about 2e-7 of B in size, so its own relative
    # difference is roundoff of B amplified; it is compared on the scale of
    # the field it corrects
    dpb = np.linalg.norm(host(pB) - O.interior(pBo))
    scale = np.linalg.norm(O.interior(pBo)) + np.linalg.norm(O.interior(Bo))
    print(case, "pB", dpb / scale, flush=True)
    assert dpb <= RTOL * scale
    assert nv.maxabs_div(G, U) < 1e-11 and nv.maxabs_div(G, B) < 1e-11


def test_reference_mhd_step_is_not_the_ssp_step(nv, orc):
    # the two compositions differ at O(dt^2): the test above cannot pass by
    # running the SSP-RK3 step
    G, O = pair(nv, orc, (16, 16, 16), (TWO_PI,) * 3)
    OP, Uo, Bo = ot_fields(O, 5)
    outs = []
    for mode in (nv.RK_LITERAL4, nv.RK_SSP3):
        U, B = dev(O.interior(Uo)), dev(O.interior(Bo))
        nv.Solver(nv.Poisson(G), nv.MODEL_MHD, mode, nu=1e-3, eta=1e-3).step(0.02, 1, U, G.field(1), B, G.field(1))
        outs.append((host(U), host(B)))
    assert rel_l2(outs[0][1], outs[1][1]) > 1e-7


def test_reference_mhd_step_same_under_every_launch_shape_and_graph(nv):
    import torch
    G = nv.Mesh((24, 12, 32), (TWO_PI, 3.0, TWO_PI), bc=("periodic", "wall", "periodic"))
    P = nv.Poisson(G)
    gen = torch.Generator(device="cuda").manual_seed(7)
    U0 = torch.randn((3, *G.shape), generator=gen, dtype=torch.float64, device="cuda")
    B0 = torch.randn((3, *G.shape), generator=gen, dtype=torch.float64, device="cuda")
    U0[1, :, 0, :] = U0[1, :, -1, :] = 0.0
    B0[1, :, 0, :] = B0[1, :, -1, :] = 0.0
    P.project(U0, 1.0)
    P.project(B0, 1.0)
    ref = None
    for v, c in ((3, 0), (0, 5), (2, 0), (4, 8), (11, 0), (12, 5), (1, 0)):
        G.tune(v, c, 2)
        for graph in (False, True):
            S = nv.Solver(P, nv.MODEL_MHD, nv.RK_LITERAL4, nu=1e-2, eta=2e-2, graph=graph)
            U, B = U0.clone(), B0.clone()
            S.step(1e-3, 2, U, G.field(1), B, G.field(1))
            if ref is None:
                ref = (U, B)
            else:
                assert torch.equal(U, ref[0]) and torch.equal(B, ref[1]), (v, c, graph)


def test_reference_mhd_diagnostics_100_steps(nv, orc):
    """Ek, Em, Jav, Jmax, omega, epsilon, Umax of OT.cpp:134-176 every step
    for 100 steps of the reference composition, 32^3, nu = eta = 1e-3."""
    N = (32, 32, 32)
    G, O = pair(nv, orc, N, (TWO_PI,) * 3)
    OP, Uo, Bo = ot_fields(O, 3)
    U, B, p, pB = dev(O.interior(Uo)), dev(O.interior(Bo)), G.field(1), G.field(1)
    po, pBo = O.field(1), O.field(1)
    nu = eta = 1e-3
    dt = 2e-3
    S = nv.Solver(nv.Poisson(G), nv.MODEL_MHD, nv.RK_LITERAL4, nu=nu, eta=eta, graph=True)
    hg, ho = [], []
    for step in range(100):
        OP.mhd_step(Uo, Bo, po, pBo, nu, eta, dt, 0)
        S.step(dt, 1, U, p, B, pB)
        hg.append(nv.mhd_diag(G, U, B, nu, eta)[:7])
        ho.append(O.mhd_diag(Uo, Bo, nu, eta)[:7])
        if step % 25 == 24:
            print("step", step + 1, hg[-1], ho[-1], flush=True)
    hg, ho = np.array(hg), np.array(ho)
    err = np.max(np.abs(hg - ho) / np.abs(ho), axis=0)
    print("max relative difference per diagnostic", err, flush=True)
    assert np.all(err <= 1e-10)
    assert rel_l2(host(B), O.interior(Bo)) <= 1e-10
