"""The timed configuration itself: Taylor-Green 256^3, FP64, SSP-RK3 + one
FFT projection per step (app/TG_Piosson/RKLoop.H:1-54 with the SSP-RK3
tableau), under the launch shapes the benchmark times, against the oracle
and against the default shapes; plus the multi-chunk streaming paths of
those shapes on grids small enough for the oracle, and one compact-scheme
step at 256^3."""
import math

import numpy as np
import pytest

from test_gpu_parity import RTOL, dev, host, pair, rel_l2

pytestmark = pytest.mark.gpu
TWO_PI = 2.0 * math.pi
NU = 1.0 / 1600.0
DT = 2e-3
# the shapes of the profiled benchmark run: NS stages variant 8 (64 x 8
# tiles) and the first stage variant 11 (two staged planes), both streaming
# 128-plane chunks, projection variant 2
BENCH_SHAPES = ((8, 128), (11, 128), 2)


@pytest.fixture(scope="module")
def nv():
    import torch
    import novaclone_amd as nv
    assert torch.cuda.is_available(), "GPU tests need a device"
    nv.lib()
    return nv


def tg(O, seed=None, amp=0.0):
    x = [np.arange(O.stored[a]) * O.L[a] / O.N[a] for a in range(3)]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    c = np.stack([np.sin(X) * np.cos(Y) * np.cos(Z), -np.cos(X) * np.sin(Y) * np.cos(Z), 0 * X])
    if seed is not None:
        c = c + amp * np.random.default_rng(seed).standard_normal(c.shape)
    return np.ascontiguousarray(c)


def set_shapes(nv, G, shapes):
    (v, c), (vf, cf), pv = shapes
    G.tune(v, c, pv)
    G.tune_kernel(nv.KERNEL_NS_FIRST_STAGE, vf, cf)


def gpu_step(nv, G, c, shapes, nsteps=1, autotune=False):
    import torch
    U, p = dev(c), G.field(1)
    S = nv.Solver(nv.Poisson(G), nv.MODEL_NS, nv.RK_SSP3, nu=NU, graph=True)
    if shapes is not None:
        set_shapes(nv, G, shapes)
    if autotune:
        print(S.autotune(U, reps=2).splitlines()[-3:], flush=True)
    S.step(DT, nsteps, U, p)
    torch.cuda.synchronize()
    return U, p


def test_headline_step_at_256_matches_oracle_and_default_shapes(nv, orc):
    import torch
    N = (256, 256, 256)
    G, O = pair(nv, orc, N, (TWO_PI,) * 3)
    c = tg(O)
    print("256^3 TG, one SSP-RK3 step, bench shapes", BENCH_SHAPES, flush=True)
    U, p = gpu_step(nv, G, c, BENCH_SHAPES)
    Ud, pd = gpu_step(nv, G, c, ((3, 0), (3, 0), 2))
    assert torch.equal(U, Ud) and torch.equal(p, pd)
    del Ud, pd
    Uo, po = O.from_compact(c), O.field(1)
    orc.OPoisson(O).ns_step(Uo, po, NU, DT, orc.RK_SSP3)
    eu = rel_l2(host(U), O.interior(Uo))
    ep = rel_l2(host(p), O.interior(po))
    print("rel L2 U", eu, "p", ep, flush=True)
    assert eu <= RTOL and ep <= RTOL


def test_headline_autotuned_shapes_give_the_default_step(nv):
    # whatever the autotuner picks on this box is what the benchmark times
    import torch
    N = (256, 256, 256)
    G = nv.Mesh(N, (TWO_PI,) * 3)

    class _O:
        stored, L = N, (TWO_PI,) * 3
    _O.N = N
    c = tg(_O)
    U, p = gpu_step(nv, G, c, None, nsteps=2, autotune=True)
    Ud, pd = gpu_step(nv, G, c, ((3, 0), (3, 0), 2), nsteps=2)
    assert torch.equal(U, Ud) and torch.equal(p, pd)


@pytest.mark.parametrize("N,shapes", [
    ((192, 16, 192), BENCH_SHAPES),            # 192 planes in a 128-plane and a 64-plane chunk, 3 k tiles
    ((48, 24, 96), ((8, 16), (11, 32), 2)),    # 16-plane chunks; the first stage a 32 and a 16; a partial k tile
    ((48, 24, 96), ((7, 16), (12, 8), 1)),     # 32 x 16 tiles, row kernels in the projection
    ((48, 24, 192), ((9, 32), (10, 32), 0)),   # ring staging, divergence fused into the first FFT pass
])
def test_multi_chunk_streaming_matches_oracle(nv, orc, N, shapes):
    import torch
    L = (TWO_PI, TWO_PI, TWO_PI)
    G, O = pair(nv, orc, N, L)
    c = tg(O, seed=2, amp=1e-2)
    Uo = O.from_compact(c)
    OP = orc.OPoisson(O)
    OP.project(Uo, 1.0)
    c = O.interior(Uo)
    U, p = gpu_step(nv, G, c, shapes, nsteps=2)
    Ud, pd = gpu_step(nv, G, c, ((3, 0), (3, 0), 2), nsteps=2)
    assert torch.equal(U, Ud)
    po = O.field(1)
    for _ in range(2):
        OP.ns_step(Uo, po, NU, DT, orc.RK_SSP3)
    e = rel_l2(host(U), O.interior(Uo))
    print(N, shapes, "rel L2", e, flush=True)
    assert e <= RTOL


def test_compact_step_at_256_matches_oracle(nv, orc):
    N = (256, 256, 256)
    G, O = pair(nv, orc, N, (TWO_PI,) * 3)
    c = tg(O)
    U, p = dev(c), G.field(1)
    S = nv.Solver(nv.Poisson(G, scheme=nv.SCHEME_COMPACT6), nv.MODEL_NS, nv.RK_SSP3, nu=NU)
    S.step(DT, 1, U, p)
    Uo, po = O.from_compact(c), O.field(1)
    orc.OPoisson(O, compact=True).compact_ns_step(Uo, po, NU, DT)
    e = rel_l2(host(U), O.interior(Uo))
    print("compact 256^3 rel L2", e, flush=True)
    assert e <= RTOL


def test_ot_mhd_step_at_256_matches_oracle(nv, orc):
    """The OT-type MHD configuration at its full size: 256^3 periodic,
    u = (-sin y, sin x, 0), B = (-sin y, sin 2x, 0) plus 1e-3 Gaussian noise
    (seed 3), both projected once, then one SSP-RK3 step with both fields
    projected every stage (Re = Rm = 1000, dt = 1e-3) under the autotuned
    launch shapes, against the oracle; and the OT diagnostics of the result."""
    import torch
    N = (256, 256, 256)
    G, O = pair(nv, orc, N, (TWO_PI,) * 3)
    x = [np.arange(O.stored[a]) * TWO_PI / N[a] for a in range(3)]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    rng = np.random.default_rng(3)
    cu = np.stack([-np.sin(Y), np.sin(X), 0 * X]) + 1e-3 * rng.standard_normal((3, *O.stored))
    cb = np.stack([-np.sin(Y), np.sin(2 * X), 0 * X]) + 1e-3 * rng.standard_normal((3, *O.stored))
    del X, Y, Z
    OP = orc.OPoisson(O)
    Uo, Bo = O.from_compact(cu), O.from_compact(cb)
    OP.project(Uo, 1.0)
    OP.project(Bo, 1.0)
    cu, cb = O.interior(Uo), O.interior(Bo)
    U, B, p, pB = dev(cu), dev(cb), G.field(1), G.field(1)
    S = nv.Solver(nv.Poisson(G), nv.MODEL_MHD, nv.RK_SSP3, nu=1e-3, eta=1e-3, graph=True)
    print(S.autotune(U, B, reps=2).splitlines()[-2:], flush=True)
    S.step(1e-3, 1, U, p, B, pB)
    torch.cuda.synchronize()
    po, pBo = O.field(1), O.field(1)
    OP.mhd_step(Uo, Bo, po, pBo, 1e-3, 1e-3, 1e-3, 1)
    eu, eb = rel_l2(host(U), O.interior(Uo)), rel_l2(host(B), O.interior(Bo))
    print("OT MHD 256^3 one step rel L2", eu, eb, flush=True)
    assert eu <= RTOL and eb <= RTOL
    d = nv.mhd_diag(G, U, B, 1e-3, 1e-3)
    do = O.mhd_diag(Uo, Bo, 1e-3, 1e-3)
    assert np.allclose(d[:7], do[:7], rtol=1e-12, atol=1e-14)


def test_wc_step_at_256_matches_oracle(nv, orc):
    """The weakly-compressible TG configuration of app/TG at its full size:
    256^3 periodic, the state of TG.cpp:39-55 (p = 50 + (cos 2x + cos 2y)
    (cos 2z + 2)/16, rho = (p/50)^(1/2)) with a 1e-3 seeded perturbation of
    U, mu = 1/1600, p_c = 50, dt = 1e-3, one classic-RK4 step of rho and
    rho U against the oracle."""
    N = (256, 256, 256)
    G, O = pair(nv, orc, N, (TWO_PI,) * 3)
    x = [np.arange(O.stored[a]) * TWO_PI / N[a] for a in range(3)]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    U = np.stack([np.sin(X) * np.cos(Y) * np.cos(Z), -np.cos(X) * np.sin(Y) * np.cos(Z), 0 * X])
    U += 1e-3 * np.random.default_rng(11).standard_normal(U.shape)
    p = 50 + 1.0 / 16.0 * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2)
    del X, Y, Z
    rho = np.sqrt(p / 50.0)[None]
    mu, pc, dt = 1.0 / 1600.0, 50.0, 1e-3
    W = nv.WC(G, mu, pc)
    Ud, rd, pd = dev(U), dev(rho), G.field(1)
    W.step(dt, 1, Ud, rd, pd)
    Uo, ro, po = O.from_compact(U), O.from_compact(rho), O.field(1)
    del U, rho, p
    O.wc_step(Uo, ro, po, mu, pc, dt)
    eu, er, ep = (rel_l2(host(Ud), O.interior(Uo)), rel_l2(host(rd), O.interior(ro)),
                  rel_l2(host(pd), O.interior(po)))
    print("WC 256^3 one step rel L2 U", eu, "rho", er, "p", ep, flush=True)
    assert eu <= RTOL and er <= RTOL and ep <= RTOL
