#!/usr/bin/env python3
"""Throughput of the NOVA hot path on one MI355X.

Workload: incompressible Taylor-Green vortex, 256^3 periodic on [0, 2pi)^3,
FP64, Re = 1600, explicit 6th-order stencils, SSP-RK3 + one FFT pressure
projection per step, dt = 2e-3 (the reference's TG_Poisson step,
app/TG_Piosson/RKLoop.H, with the RK3 composition).  Inputs are resident in
HBM before the timed region.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
METRIC = "TG 256^3 FP64 Mpoint-steps/s, % HBM roofline; adjoint/forward time ratio"
HBM_PEAK_GBS = 8000.0
DT = 2e-3
NU = 1.0 / 1600.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def alg_bytes_per_point(name: str, n: int) -> float:
    """Algorithmic HBM bytes per grid point for one launch of each kernel
    (unique FP64 values read once + written once)."""
    half = (n // 2 + 1) / n * 16.0  # half-spectrum complex value per real point
    return {
        "ns_stage": (48.0 + 72.0 + 72.0) / 3.0,  # stage 1 reads U only
        "fft_div_r2c_z": 24.0 + half,
        "div_rows": 24.0 + 8.0,
        "fft_r2c_z": 8.0 + half,
        "fft_y": 2 * half,
        "fft_x_solve": 2 * half,
        "fft_c2r_z": half + 8.0,
        "grad_correct": 8.0 + 24.0 + 24.0,
        "copy": 48.0,
    }.get(name, float("nan"))


def tg_field(nv, G, n):
    import torch
    x = torch.arange(n, dtype=torch.float64, device=G.device) * (2 * math.pi / n)
    X, Y, Z = torch.meshgrid(x, x, x, indexing="ij")
    U = torch.stack([torch.sin(X) * torch.cos(Y) * torch.cos(Z),
                     -torch.cos(X) * torch.sin(Y) * torch.cos(Z),
                     torch.zeros_like(X)]).contiguous()
    return U


def _median_step_s(step, warm: int, reps: int) -> float:
    for _ in range(warm):
        step()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        step()
        ts.append(time.perf_counter() - t0)
    return sorted(ts)[len(ts) // 2]


def cpu_baseline(n: int, n1: int = 64, warm: int = 2, reps: int = 3) -> dict:
    """The oracle's TG step (SSP-RK3 + FFT projection) timed on the host:
    (i) faithful (the reference's operator-at-a-time passes with zeroed
    temporaries, l-outer loops and a full complex FFT) and (ii) fused (one
    OpenMP pass per stage, same arithmetic), each on all OpenMP threads at
    n^3 and on one core at n1^3; steady_clock-like wall time per step,
    median of `reps` steps after `warm` warm-up steps.  `value` is (i) on all
    threads (the denominator of the GPU/CPU ratio)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import numpy as np
    import pyoracle
    L = (2 * math.pi,) * 3
    cores = pyoracle.num_threads()

    def setup(m):
        O = pyoracle.OMesh((m, m, m), L)
        P = pyoracle.OPoisson(O)
        x = np.arange(m) * (2 * math.pi / m)
        X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
        c = np.stack([np.sin(X) * np.cos(Y) * np.cos(Z), -np.cos(X) * np.sin(Y) * np.cos(Z), 0 * X])
        del X, Y, Z
        return O, P, O.from_compact(c)

    out = {}
    t_all = time.perf_counter()
    for label, m, threads in (("all_threads", n, cores), ("one_core", n1, 1)):
        pyoracle.set_threads(threads)
        O, P, U = setup(m)
        p, W1, W2 = O.field(1), O.field(3), O.field(3)
        faithful = _median_step_s(lambda: P.ns_step(U, p, NU, DT, pyoracle.RK_SSP3, W1), warm, reps)
        fused = _median_step_s(lambda: P.ns_step_fused(U, p, NU, DT, W1, W2), warm, reps)
        out[label] = {"N": m, "threads": threads, "faithful_s_per_step": round(faithful, 4),
                      "fused_s_per_step": round(fused, 4),
                      "faithful_Mpoint_steps_per_s": round(m ** 3 / faithful / 1e6, 3),
                      "fused_Mpoint_steps_per_s": round(m ** 3 / fused / 1e6, 3)}
        del O, P, U, p, W1, W2
    pyoracle.set_threads(cores)
    a = out["all_threads"]
    return {"value": a["faithful_Mpoint_steps_per_s"], "unit": "Mpoint-steps/s", "cores": cores, "kind": "port",
            "sample": f"oracle TG SSP-RK3 + FFT projection: faithful and fused variants at {n}^3 on {cores} "
                      f"OpenMP threads and at {n1}^3 on 1 core, median of {reps} steps after {warm} warm-up "
                      f"steps each ({time.perf_counter() - t_all:.1f} s in all); value = faithful, all threads",
            "variants": out}


def shear_layer(nv, G, seed=5):
    """u = tanh(2y) on y in [-6, 6] plus a seeded random-phase perturbation of
    amplitude 1e-3 on the modes kx, kz <= 8 (localised in y), projected."""
    import torch
    gen = torch.Generator(device="cpu").manual_seed(seed)
    sx, sy, sz = G.shape
    Lx, Ly, Lz = G.L
    x = torch.arange(sx, dtype=torch.float64, device=G.device) * (Lx / G.N[0])
    y = torch.arange(sy, dtype=torch.float64, device=G.device) * (Ly / G.N[1]) - Ly / 2
    z = torch.arange(sz, dtype=torch.float64, device=G.device) * (Lz / G.N[2])
    U = torch.zeros((3, sx, sy, sz), dtype=torch.float64, device=G.device)
    U[0] += torch.tanh(2 * y)[None, :, None]
    env = torch.exp(-y * y)[None, :, None]
    for c in range(3):
        for kx in range(9):
            for kz in range(9):
                if kx == 0 and kz == 0:
                    continue
                ph = float(torch.rand(1, generator=gen)) * 2 * math.pi
                U[c] += 1e-3 * torch.cos(2 * math.pi * kx * x[:, None, None] / Lx
                                         + 2 * math.pi * kz * z[None, None, :] / Lz + ph) * env
    U[1, :, 0, :] = 0.0
    U[1, :, -1, :] = 0.0
    return U.contiguous()


def adjoint_ratio(nv, steps: int, every: int) -> dict:
    """Time of the checkpointed adjoint sweep over the time of the forward
    solve, shear layer 256x384x256 (x, z periodic, free-slip y walls)."""
    import torch
    N, L = (256, 384, 256), (8 * math.pi, 12.0, 8 * math.pi)
    G = nv.Mesh(N, L, bc=("periodic", "wall", "periodic"))
    P = nv.Poisson(G)
    S = nv.Solver(P, nv.MODEL_NS, nv.RK_SSP3, nu=1e-3, graph=True)
    U0 = shear_layer(nv, G)
    P.project(U0, 1.0)
    A = nv.Adjoint(nv.Solver(P, nv.MODEL_NS, nv.RK_SSP3, nu=1e-3), every, steps)
    # HBM beyond the checkpoints holds the step-start states of the forward
    # sweep's last steps (one vector field each), whose segments the reverse
    # sweep then does not recompute; 8 GiB stay free for the rest of the run
    # (it allocates about six more vector fields, 3.6 GB)
    free, _ = torch.cuda.mem_get_info()
    extra = A.reserve(max(0, free - (8 << 30)))
    resident = min(steps, 3 * every + extra)
    dt = 0.02
    tuned = chosen(S.autotune(U0, reps=3))
    U, p = U0.clone(), G.field(1)
    S.step(dt, 2, U, p)
    torch.cuda.synchronize()
    U = U0.clone()
    t0 = time.perf_counter()
    S.step(dt, steps, U, p)
    torch.cuda.synchronize()
    tf = time.perf_counter() - t0
    UT = A.forward(dt, steps, U0)
    tuned += chosen(A.autotune(reps=3))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    UT = A.forward(dt, steps, U0)  # the checkpointing forward sweep every gradient needs
    torch.cuda.synchronize()
    tcf = time.perf_counter() - t0
    J, g = nv.enstrophy(G, UT, grad=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g0 = A.backward(g)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    gn = nv.weighted_dot(G, g0, g0) ** 0.5
    A.close()
    # the gradient against central differences (eps = 1e-6) on 16 seeded
    # directions at this size (untimed): J(U0 + eps d) from the replayed
    # forward solve, directions with max |d| = 1 and the wall-normal
    # component zero on the walls; the bound allows the difference quotient
    # its roundoff, 4 u J / eps (u = 2^-53)
    # (and the fourth-order quotient at eps4 = 1e-4, whose roundoff, about
    # u J / eps4, leaves the 1e-8 agreement measurable)
    eps, eps4, u = 1e-6, 1e-4, 2.0 ** -53
    fd = {"directions": 16, "eps": eps, "rel_err": [], "within_bound": True,
          "bound": "|fd - ad| <= 1e-8 |ad| + 4 u J / eps", "eps4": eps4, "rel_err4": []}

    def J_at(d, e):
        U = (U0 + e * d).contiguous()
        S.step(dt, steps, U, p)
        return nv.enstrophy(G, U)

    for seed in range(16):
        gen = torch.Generator(device="cuda").manual_seed(500 + seed)
        d = torch.randn(U0.shape, generator=gen, dtype=torch.float64, device="cuda")
        d[1, :, 0, :] = 0.0
        d[1, :, -1, :] = 0.0
        d = (d / d.abs().max()).contiguous()
        fdq = (J_at(d, eps) - J_at(d, -eps)) / (2 * eps)
        fd4 = (-J_at(d, 2 * eps4) + 8 * J_at(d, eps4) - 8 * J_at(d, -eps4) + J_at(d, -2 * eps4)) / (12 * eps4)
        ad = nv.weighted_dot(G, g0, d)
        fd["rel_err"].append(float(f"{abs(fdq - ad) / abs(ad):.3e}"))
        fd["rel_err4"].append(float(f"{abs(fd4 - ad) / abs(ad):.3e}"))
        if abs(fdq - ad) > 1e-8 * abs(ad) + 4 * u * J / eps:
            fd["within_bound"] = False
        if seed % 4 == 3:
            log(f"bench: adjoint FD check, {seed + 1} of 16 directions")
    fd["max_rel_err"] = max(fd["rel_err"])
    fd["max_rel_err4"] = max(fd["rel_err4"])
    log(f"bench: adjoint gradient vs central differences, max rel err {fd['max_rel_err']} (eps {eps}, within "
        f"bound {fd['within_bound']}), {fd['max_rel_err4']} (fourth order, eps {eps4})")
    return {"ratio": round(tb / tf, 3), "forward_s": round(tf, 3), "adjoint_s": round(tb, 3),
            "checkpoint_forward_s": round(tcf, 3), "gradient_ratio": round((tcf + tb) / tf, 3),
            "ratios": "ratio = adjoint (reverse) sweep / forward solve; gradient_ratio = (checkpointing forward "
                      "+ reverse sweep) / forward solve, the whole cost of one gradient",
            "workload": f"shear layer 256x384x256 FP64, free-slip y walls (DCT-I), Re=1000, dt=0.02, "
                        f"{steps} SSP-RK3 steps, checkpoints every {every}; gradient of final enstrophy",
            "resident_steps": resident,
            "recomputed_steps": steps - resident,
            "J": J, "grad_norm": gn, "fd_check": fd, "tuning": tuned}


def wc_bench(nv, n: int, steps: int) -> dict:
    """Weakly-compressible Taylor-Green (app/TG/TG.cpp): n^3 periodic FP64,
    mu = 1/1600, p = 50 rho^2, per step a classic RK4 of rho and one of rho U,
    one kernel per stage."""
    import torch
    G = nv.Mesh((n, n, n), (2 * math.pi,) * 3)
    x = torch.arange(n, dtype=torch.float64, device="cuda") * (2 * math.pi / n)
    X, Y, Z = torch.meshgrid(x, x, x, indexing="ij")
    U = torch.stack([torch.sin(X) * torch.cos(Y) * torch.cos(Z), -torch.cos(X) * torch.sin(Y) * torch.cos(Z),
                     0 * X]).contiguous()
    p = (50 + (torch.cos(2 * X) + torch.cos(2 * Y)) * (torch.cos(2 * Z) + 2) / 16).unsqueeze(0).contiguous()
    rho = torch.sqrt(p / 50.0).contiguous()
    del X, Y, Z
    W = nv.WC(G, 1.0 / 1600.0, 50.0)
    dt = 1e-3
    W.step(dt, 2, U, rho, p)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    W.step(dt, steps, U, rho, p)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    G.timing(True)
    W.step(dt, 2, U, rho, p)
    st = G.timing_collect()
    G.timing(False)
    # minimal-traffic model per node-step: rho stages 48 + 64 + 64 + 56 B,
    # U stages 88 + 136 + 136 + 88 B (stencil inputs once, pointwise
    # operands, outputs)
    alg = 680.0 * n ** 3
    finite = bool(torch.isfinite(U).all() and torch.isfinite(rho).all())
    # app/OT: the same step with the frozen field's Lorentz force (one more
    # pass per step, 24 B more per node in each U stage)
    B = (0.8 * torch.stack([-2 * torch.sin(2 * x)[None, :, None] + torch.sin(x)[None, None, :] + 0 * x[:, None, None],
                            2 * torch.sin(x)[:, None, None] + torch.sin(x)[None, None, :] + 0 * x[None, :, None],
                            torch.sin(x)[:, None, None] + torch.sin(x)[None, :, None] + 0 * x[None, None, :]])
         ).contiguous()
    W.set_magnetic(B, 0.5)
    W.step(dt, 2, U, rho, p)
    torch.cuda.synchronize()
    e0.record()
    W.step(dt, steps, U, rho, p)
    e1.record()
    torch.cuda.synchronize()
    ms_ot = e0.elapsed_time(e1) / steps
    alg_ot = (680.0 + 56.0 + 4 * 24.0) * n ** 3  # + the force pass (B 24 in, F 24 out) and F read per U stage
    # app/shearLayerWC order with the temperature equation and buoyancy
    W.set_magnetic(None, 0.0)
    T = (0.05 * torch.tanh(2 * (x - math.pi))[None, :, None] + 0 * x[:, None, None] + 0 * x[None, None, :]
         ).unsqueeze(0).contiguous()
    W.set_temperature(1.0 / 1600.0 / 7.0, (0.0, -0.1, 0.0))
    W.step_t(dt, 2, U, rho, p, T)
    torch.cuda.synchronize()
    e0.record()
    W.step_t(dt, steps, U, rho, p, T)
    e1.record()
    torch.cuda.synchronize()
    ms_t = e0.elapsed_time(e1) / steps
    # + four T stages: rho 8 + U 24 + T 8 stencil, T0 8, accumulator 8 in / 8 out, T 8 out; T read by
    # the U stages
    alg_t = (680.0 + 64 + 80 + 80 + 64 + 4 * 8.0) * n ** 3
    return {"workload": f"weakly-compressible TG {n}^3 periodic FP64, mu=1/1600, p=50 rho^2, classic RK4 of rho "
                        f"and of rho U (app/TG), dt={dt}, {steps} steps",
            "value": round(n ** 3 / (ms / 1e3) / 1e6, 1), "unit": "Mpoint-steps/s", "ms_per_step": round(ms, 4),
            "alg_bytes_per_step": alg, "achieved_GBps": round(alg / (ms / 1e3) / 1e9, 1),
            "frac": round(alg / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
            "kernels_us": {k: round(v[1] / v[0] * 1e3, 1) for k, v in st.items()},
            "ke": nv.ke(G, U), "finite": finite,
            "ot_frozen_B": {"ms_per_step": round(ms_ot, 4), "value": round(n ** 3 / (ms_ot / 1e3) / 1e6, 1),
                            "alg_bytes_per_step": alg_ot,
                            "frac": round(alg_ot / (ms_ot / 1e3) / 1e9 / HBM_PEAK_GBS, 4)},
            "shear_layer_T": {"ms_per_step": round(ms_t, 4), "value": round(n ** 3 / (ms_t / 1e3) / 1e6, 1),
                              "alg_bytes_per_step": alg_t,
                              "frac": round(alg_t / (ms_t / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}}


def dropin_bench(n: int, steps: int) -> dict:
    """The drop-in path: the TG_Poisson time loop written against the C++
    drop-in API (novaclone_amd/host/apps/tg_poisson.cpp, the reference app's
    per-operator expressions and its host kinetic-energy loop), run as a
    child process at n^3; the driver reports the device work per step and the
    host loop per step (of which the host mirror's download of U, into
    page-locked memory, on its own), averaged over the steps after the
    first."""
    import subprocess
    import tempfile
    exe = ROOT / "novaclone_amd" / "host" / "bin" / "tg_poisson"
    with tempfile.TemporaryDirectory(prefix="nova_dropin_", dir=os.environ.get("TMPDIR", "/tmp")) as d:
        r = subprocess.run([str(exe), str(n), str(steps), str(DT), str(NU), os.path.join(d, "ek.dat")],
                           capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        return {"error": f"tg_poisson exited {r.returncode}: {r.stderr.strip()[-300:]}"}
    f = None
    for line in r.stdout.splitlines():
        if line.startswith("timing:"):
            w = line.split()
            f = {w[k]: float(w[k + 1]) for k in range(1, len(w) - 1, 2)}
    if not f:
        return {"error": "no timing line"}
    dev, ke = f["device_ms_per_step"], f["host_ke_ms_per_step"]
    return {"workload": f"TG {n}^3 through the C++ drop-in API: the reference's TG_Poisson loop (frozen-stage RK4 "
                        f"from ex::div, ex::laplacian and Field algebra, one projection per step) and its host "
                        f"kinetic-energy loop over the host mirror, {steps} steps",
            "device_ms_per_step": round(dev, 3), "host_ke_ms_per_step": round(ke, 3),
            "mirror_download_ms_per_step": round(f.get("mirror_download_ms_per_step", float("nan")), 3),
            "ms_per_step": round(dev + ke, 3),
            "value": round(n ** 3 / ((dev + ke) * 1e-3) / 1e6, 1),
            "device_value": round(n ** 3 / (dev * 1e-3) / 1e6, 1), "unit": "Mpoint-steps/s"}


def io_bench(nv, n: int, writes: int) -> dict:
    """EnSight Gold output of the TG velocity (Field::write of a 3-component
    n^3 field): the layout change to float32, i fastest, on the GPU, then the
    copy to host and the file write behind it."""
    import shutil
    import tempfile
    import torch
    G = nv.Mesh((n, n, n), (2 * math.pi,) * 3)
    U = tg_field(nv, G, n)
    t = nv.Time(DT, 1.0, 1)
    io = nv.IO(G)
    d = tempfile.mkdtemp(prefix="nova_io_", dir=os.environ.get("TMPDIR", "/tmp"))
    try:
        t.advance()
        io.write_field(t, d, "U", U)  # sizes the staging slots
        io.wait()
        torch.cuda.synchronize()
        G.timing(True)
        t0 = time.perf_counter()
        for _ in range(writes):
            t.advance()
            io.write_field(t, d, "U", U)
        t_call = time.perf_counter() - t0
        io.wait()
        t_all = time.perf_counter() - t0
        st = G.timing_collect()
        G.timing(False)
        file_bytes = os.path.getsize(os.path.join(d, f"U.0000.{t.fileNum():05d}.dat"))
    finally:
        io.close()
        shutil.rmtree(d, ignore_errors=True)
    c, us = st["ensight_pack"]
    us = us / c * 1e3
    alg = 8.0 * 3 * n ** 3 + 4.0 * 3 * (n + 1) ** 3  # FP64 stored nodes in, float32 EnSight nodes out
    return {"workload": f"EnSight Gold write of the TG {n}^3 velocity (3 x {n + 1}^3 float32), {writes} files",
            "pack_us": round(us, 1), "pack_alg_bytes": alg,
            "pack_GBps": round(alg / (us / 1e6) / 1e9, 1), "pack_frac": round(alg / (us / 1e6) / 1e9 / HBM_PEAK_GBS, 4),
            "file_MB": round(file_bytes / 1e6, 1), "call_ms": round(t_call / writes * 1e3, 3),
            "write_MBps": round(writes * file_bytes / t_all / 1e6, 1)}


def ensemble_bench(nv, steps: int) -> dict:
    """48 uneven instances (16 x 32^3, 16 x 48^3, 12 x 64^3, 4 x 96^3), TG and
    Orszag-Tang alternating within each size, Re_i ~ U[400, 1600], 1e-3
    Gaussian noise from seed 4 + i; SSP-RK3 with one persistent launch per
    stage in a replayed step graph."""
    import torch
    sizes = [32] * 16 + [48] * 16 + [64] * 12 + [96] * 4
    gen = torch.Generator().manual_seed(4)
    inst, ics = [], []
    for i, n in enumerate(sizes):
        model = nv.MODEL_NS if i % 2 == 0 else nv.MODEL_MHD
        re = 400.0 + 1200.0 * float(torch.rand(1, generator=gen))
        inst.append(((n, n, n), (2 * math.pi,) * 3, model, 1.0 / re, 1.0 / re))
    E = nv.Ensemble(inst, graph=True)
    for i, n in enumerate(sizes):
        G = nv.Mesh((n, n, n), (2 * math.pi,) * 3)
        P = nv.Poisson(G)
        x = torch.arange(n, dtype=torch.float64, device="cuda") * (2 * math.pi / n)
        X, Y, Z = torch.meshgrid(x, x, x, indexing="ij")
        g = torch.Generator(device="cuda").manual_seed(4 + i)
        noise = lambda: 1e-3 * torch.randn((3, n, n, n), generator=g, dtype=torch.float64, device="cuda")
        if inst[i][2] == nv.MODEL_NS:
            U = (torch.stack([torch.sin(X) * torch.cos(Y) * torch.cos(Z), -torch.cos(X) * torch.sin(Y) * torch.cos(Z),
                              0 * X]) + noise()).contiguous()
            B = None
        else:
            U = (torch.stack([-torch.sin(Y), torch.sin(X), 0 * X]) + noise()).contiguous()
            B = (torch.stack([-torch.sin(Y), torch.sin(2 * X), 0 * X]) + noise()).contiguous()
            P.project(B, 1.0)
        P.project(U, 1.0)
        E.load(i, U, B)
        del G, P
    dt = 2e-3
    E.step(dt, 2)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    E.step(dt, steps)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    info = E.info()
    pts = info["nodes"]
    # minimal-traffic model: 384 B per node-step (TG), 1440 (OT, both fields
    # projected at every stage)
    alg = sum((384.0 if model == nv.MODEL_NS else 1440.0) * N[0] * N[1] * N[2] for N, _, model, _, _ in inst)
    Uf, _, _, _ = E.store(len(sizes) - 1)
    finite = bool(torch.isfinite(Uf).all())
    E.close()
    return {"workload": "48 instances: 16x32^3, 16x48^3, 12x64^3, 4x96^3, TG/OT alternating, SSP-RK3, "
                        f"dt={dt}, {steps} steps, one persistent launch per stage, replayed step graph",
            "value": round(pts * steps / (ms * steps / 1e3) / 1e6, 1), "unit": "Mpoint-steps/s",
            "ms_per_step": round(ms, 4), "nodes": pts, "alg_bytes_per_step": alg,
            "achieved_GBps": round(alg / (ms / 1e3) / 1e9, 1), "frac": round(alg / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
            "persistent_blocks": info["blocks"], "tiles_per_stage": info["tiles"], "finite": finite}


def compact_bench(nv, n: int, steps: int) -> dict:
    """TG n^3 with the compact sixth-order scheme (Lele tridiagonal first and
    second derivatives, wavefront line solves) for the right-hand side and the
    projection; SSP-RK3, replayed step graph."""
    import torch
    G = nv.Mesh((n, n, n), (2 * math.pi,) * 3)
    P = nv.Poisson(G, scheme=nv.SCHEME_COMPACT6)
    S = nv.Solver(P, nv.MODEL_NS, nv.RK_SSP3, nu=NU, graph=True)
    U = tg_field(nv, G, n)
    p = G.field(1)
    S.step(DT, 3, U, p)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    S.step(DT, steps, U, p)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    Se = nv.Solver(P, nv.MODEL_NS, nv.RK_SSP3, nu=NU)
    G.timing(True)
    Se.step(DT, 2, U, p)
    st = G.timing_collect()
    G.timing(False)
    kern = {k: {"launches": c, "avg_ms": round(t / c, 5)} for k, (c, t) in st.items()}
    # minimal traffic: stage passes i (U in, R out), j (U in, R in/out),
    # k (U, R, U0 in, out) = 192 + 216 + 216 B; projection: divergence passes
    # 16 + 24 + 24, transforms 80, gradient passes 3 x 24 = 216 B
    alg = 840.0 * n ** 3
    return {"workload": f"TG {n}^3 periodic FP64 Re=1600, compact 6th-order (Lele alpha=1/3, 2/11) derivatives and "
                        f"projection, SSP-RK3, dt={DT}, {steps} steps, replayed step graph",
            "value": round(n ** 3 / (ms / 1e3) / 1e6, 1), "unit": "Mpoint-steps/s", "ms_per_step": round(ms, 4),
            "alg_bytes_per_step": alg, "achieved_GBps": round(alg / (ms / 1e3) / 1e9, 1),
            "frac": round(alg / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4), "kernels": kern,
            "ke": nv.ke(G, U), "maxabs_div_explicit_stencil": nv.maxabs_div(G, U)}


def ot_bench(nv, n: int, steps: int, compact: bool) -> dict:
    """Orszag-Tang type MHD at n^3: u = (-sin y, sin x, 0), B = (-sin y, sin 2x, 0)
    + 1e-3 N(0,1) (seed 3), Re = Rm = 1000, SSP-RK3 with u and B projected at
    every stage, replayed step graph; explicit m=6 or compact operators."""
    import torch
    G = nv.Mesh((n, n, n), (2 * math.pi,) * 3)
    P = nv.Poisson(G, scheme=nv.SCHEME_COMPACT6 if compact else nv.SCHEME_EXPLICIT)
    S = nv.Solver(P, nv.MODEL_MHD, nv.RK_SSP3, nu=1e-3, eta=1e-3, graph=True)
    x = torch.arange(n, dtype=torch.float64, device="cuda") * (2 * math.pi / n)
    X, Y, Z = torch.meshgrid(x, x, x, indexing="ij")
    g = torch.Generator(device="cuda").manual_seed(3)
    U = (torch.stack([-torch.sin(Y), torch.sin(X), 0 * X]) +
         1e-3 * torch.randn((3, n, n, n), generator=g, dtype=torch.float64, device="cuda")).contiguous()
    B = (torch.stack([-torch.sin(Y), torch.sin(2 * X), 0 * X]) +
         1e-3 * torch.randn((3, n, n, n), generator=g, dtype=torch.float64, device="cuda")).contiguous()
    del X, Y, Z
    P.project(U, 1.0)
    P.project(B, 1.0)
    if not compact:
        S.autotune(U, B, reps=2)
    p, pB = G.field(1), G.field(1)
    S.step(1e-3, 2, U, p, B, pB)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    S.step(1e-3, steps, U, p, B, pB)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    # minimal-traffic models: explicit 3 (2 x 72) + 6 x 168 = 1440 B; compact
    # stage passes 384 + 432 + 432 and six compact projections of 216 B
    alg = (2544.0 if compact else 1440.0) * n ** 3
    return {"workload": f"OT-type MHD {n}^3 periodic FP64 Re=Rm=1000, "
                        f"{'compact 6th-order' if compact else 'explicit m=6'} operators, SSP-RK3, both fields "
                        f"projected every stage, dt=1e-3, {steps} steps, replayed step graph",
            "value": round(n ** 3 / (ms / 1e3) / 1e6, 1), "unit": "Mpoint-steps/s", "ms_per_step": round(ms, 4),
            "alg_bytes_per_step": alg, "achieved_GBps": round(alg / (ms / 1e3) / 1e9, 1),
            "frac": round(alg / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
            "finite": bool(torch.isfinite(U).all() and torch.isfinite(B).all())}


def chosen(report: str) -> list:
    """The 'chosen' lines of an autotuner report."""
    return [ln for ln in report.splitlines() if "chosen" in ln]


def _kname(full: str) -> str:
    """rocprof kernel name without return type, namespaces and arguments."""
    s = full.replace("nova::(anonymous namespace)::", "").replace("nova::", "")
    if s.startswith("void "):
        s = s[5:]
    depth = 0
    for i, ch in enumerate(s):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            return s[:i]
    return s


def _csv_rows(d: str, suffix: str) -> list:
    import csv
    rows = []
    for f in sorted(Path(d).rglob(f"*{suffix}")):
        with open(f, newline="") as fh:
            rows += list(csv.DictReader(fh))
    return rows


# algorithmic bytes per node of the TG step's kernels, by rocprof kernel name
# prefix (the seventh template argument of k_stage_stream: 2 = the first
# SSP stage, which reads no step-start operand)
def _alg_per_node(name: str, n: int) -> float | None:
    half = (n // 2 + 1) / n * 16.0
    if name.startswith("k_stage_stream<3, 0, false"):
        return 48.0 if name.split(",")[6].strip() == "2" else 72.0
    if name.startswith("k_stage_rows<3, 1"):
        return 72.0
    table = {"k_proj_stream<3, true": 32.0, "k_proj_stream<3, false": 56.0, "k_fwd_z<": 8.0 + half,
             "k_fft_y<": 2 * half, "k_fft_x<": 2 * half, "k_inv_z<": half + 8.0, "k_div_rows": 32.0,
             "k_correct_rows": 56.0}
    for k, v in table.items():
        if name.startswith(k):
            return v
    return None


def summarize(dirs: list, commit: str | None, n: int) -> int:
    """Kernel-statistics directories print a markdown table; a FETCH_SIZE
    and a WRITE_SIZE directory plus their copy calibrations (in the order
    fetch, write, fetch-calibration, write-calibration) print the per-kernel
    HBM traffic JSON that the benchmark line cites."""
    pts = n ** 3
    if len(dirs) == 1:
        rows = _csv_rows(dirs[0], "kernel_stats.csv")
        print("| kernel | calls | avg (us) | algorithmic MB / launch | algorithmic GB/s | % of 8 TB/s |")
        print("|---|---|---|---|---|---|")
        for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"]))[:24]:
            name = _kname(r["Name"])
            avg = float(r["AverageNs"]) / 1e3
            a = _alg_per_node(name, n)
            if a is None:
                print(f"| `{name[:90]}` | {r['Calls']} | {avg:.1f} | | | |")
            else:
                gbs = a * pts / (avg / 1e6) / 1e9
                print(f"| `{name[:90]}` | {r['Calls']} | {avg:.1f} | {a * pts / 1e6:.0f} | {gbs:.0f} | "
                      f"{100 * gbs / HBM_PEAK_GBS:.1f} |")
        return 0
    fetch, write, cf, cw = dirs

    def per_kernel(d, counter):
        acc = {}
        for r in _csv_rows(d, "counter_collection.csv"):
            if r.get("Counter_Name") != counter:
                continue
            name = _kname(r["Kernel_Name"])
            v = acc.setdefault(name, [0.0, 0])
            v[0] += float(r["Counter_Value"])
            v[1] += 1
        return {k: (v[0] / v[1], v[1]) for k, v in acc.items()}

    def calib(d, counter):
        vals = [v for k, (v, c) in per_kernel(d, counter).items() if "copyBuffer" in k or "copy" in k.lower()]
        return max(vals) if vals else None

    kf, kw = per_kernel(fetch, "FETCH_SIZE"), per_kernel(write, "WRITE_SIZE")
    gib_kib = float(1 << 20)  # the calibration copy moves 1 GiB = 2^20 KiB each way
    sf = gib_kib / calib(cf, "FETCH_SIZE")
    sw = gib_kib / calib(cw, "WRITE_SIZE")
    out = {"_commit": commit, "_grid": f"{n}^3",
           "_method": f"rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE (KiB per dispatch, averaged over dispatches), "
                      f"scaled by a plain 1 GiB copy's readings: fetch x {sf:.3f}, write x {sw:.3f}"}
    groups = {"ns_stage": lambda k: k.startswith("k_stage_stream<3, 0, false"),
              "div_rows": lambda k: k.startswith("k_proj_stream<3, true") or k.startswith("k_div_rows"),
              "grad_correct": lambda k: k.startswith("k_proj_stream<3, false") or k.startswith("k_correct_rows"),
              "fft_r2c_z": lambda k: k.startswith("k_fwd_z<"), "fft_y": lambda k: k.startswith("k_fft_y<"),
              "fft_x_solve": lambda k: k.startswith("k_fft_x<"), "fft_c2r_z": lambda k: k.startswith("k_inv_z<")}
    for g, sel in groups.items():
        ks = [k for k in kf if sel(k) and k in kw]
        if not ks:
            continue
        # the launches of the step dominate their group (the autotuner's trials
        # run a few calls each): weight by dispatch count
        tot = sum(kf[k][1] for k in ks)
        f = sum(kf[k][0] * kf[k][1] for k in ks) / tot * 1024 * sf
        w = sum(kw[k][0] * kw[k][1] for k in ks) / tot * 1024 * sw
        out[g] = {"kernels": {k: {"dispatches": kf[k][1], "fetch_bytes": round(kf[k][0] * 1024 * sf),
                                  "write_bytes": round(kw[k][0] * 1024 * sw)} for k in ks},
                  "bytes_per_launch": round(f + w)}
    print(json.dumps(out, indent=1))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=int(os.environ.get("WORLD_SIZE", "1")))
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=3, help="timed CPU steps per variant (median)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adjoint-steps", type=int, default=500)
    ap.add_argument("--eager", action="store_true", help="launch kernels one by one instead of graph replay")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-compact", action="store_true")
    ap.add_argument("--compact-steps", type=int, default=20)
    ap.add_argument("--no-ot", action="store_true")
    ap.add_argument("--no-wc", action="store_true")
    ap.add_argument("--no-io", action="store_true")
    ap.add_argument("--no-dropin", action="store_true")
    ap.add_argument("--dropin-steps", type=int, default=6)
    ap.add_argument("--io-writes", type=int, default=3)
    ap.add_argument("--wc-steps", type=int, default=10)
    ap.add_argument("--ot-steps", type=int, default=10)
    ap.add_argument("--ensemble-steps", type=int, default=100)
    ap.add_argument("--tune", nargs=3, type=int, metavar=("VARIANT", "CHUNK", "PROJ"),
                    help="fixed stage launch shape, i-chunk and projection variant instead of the autotuner "
                         "(reproducible profiling runs)")
    ap.add_argument("--tune-first", nargs=2, type=int, metavar=("VARIANT", "CHUNK"),
                    help="with --tune: the launch shape of the first SSP stage")
    ap.add_argument("--copy-calib", action="store_true",
                    help="only run a plain device copy of 2^30 bytes 5 times (counter calibration)")
    ap.add_argument("--summarize", nargs="+", metavar="ROCPROF_DIR",
                    help="no GPU work: summarize rocprofv3 output directories (kernel statistics, or FETCH_SIZE "
                         "and WRITE_SIZE counter runs with their copy calibrations) as markdown / JSON")
    ap.add_argument("--commit", default=None, help="with --summarize: the source commit the runs measured")
    args = ap.parse_args()
    if args.summarize:
        return summarize(args.summarize, args.commit, args.n)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU path is design-only",
              file=sys.stderr)
        return 2

    import torch
    if args.copy_calib:
        x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        log("copy-calib: 5 copies of 1073741824 bytes")
        return 0
    sys.path.insert(0, str(ROOT))
    import novaclone_amd as nv

    n = args.n
    K, W = args.steps, args.warmup
    G = nv.Mesh((n, n, n), (2 * math.pi,) * 3)
    P = nv.Poisson(G)
    S = nv.Solver(P, nv.MODEL_NS, nv.RK_SSP3, nu=NU, graph=not args.eager)
    U = tg_field(nv, G, n)
    p = G.field(1)
    if args.tune:
        G.tune(*args.tune)
        tuned = [f"fixed: ns_stage variant {args.tune[0]} chunk {args.tune[1]}, projection variant {args.tune[2]}"]
        if args.tune_first:
            G.tune_kernel(nv.KERNEL_NS_FIRST_STAGE, *args.tune_first)
            tuned.append(f"fixed: ns_first_stage variant {args.tune_first[0]} chunk {args.tune_first[1]}")
    else:
        # eight timed launches per trial: with three, launch-to-launch noise
        # decided between shapes a few microseconds apart
        tuned = chosen(S.autotune(U, reps=8))
    log(f"bench: autotuner: {tuned}")
    log(f"bench: TG {n}^3, warmup {W}, steps {K}, graph={not args.eager}")
    S.step(DT, W, U, p)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    S.step(DT, K, U, p)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = e0.elapsed_time(e1)
    ms_per_step = ms_total / K
    value = n ** 3 * K / (ms_total / 1e3) / 1e6
    log(f"bench: {ms_per_step:.3f} ms/step (events), wall {wall * 1e3 / K:.3f} ms/step, {value:.1f} Mpt-steps/s")
    ke = nv.ke(G, U)
    maxdiv = nv.maxabs_div(G, U)

    # per-kernel times with HIP events around each launch on the library stream
    S_eager = nv.Solver(P, nv.MODEL_NS, nv.RK_SSP3, nu=NU)
    G.timing(True)
    S_eager.step(DT, K, U, p)
    stats = G.timing_collect()
    G.timing(False)
    kernels = {}
    for name, (cnt, tot) in stats.items():
        avg_ms = tot / cnt
        b = alg_bytes_per_point(name, n) * n ** 3
        kernels[name] = {"launches": cnt, "avg_ms": round(avg_ms, 5), "alg_bytes": b,
                         "GBps": round(b / (avg_ms / 1e3) / 1e9, 1)}
    dom = max(stats, key=lambda k: stats[k][1])
    # HBM traffic per launch cannot be counted from inside this process (the
    # counters need rocprofv3 --pmc around it): it is read from the committed
    # counter summary, which names the command, the source commit and the
    # kernels it measured, and is labelled as such
    traffic, traffic_src = None, None
    tfile = ROOT / "profiles" / "pmc_traffic.json"
    if tfile.exists() and n == 256:
        try:
            tj = json.loads(tfile.read_text())
            entry = tj.get(dom, {})
            traffic = entry.get("bytes_per_launch")
            if traffic is not None:
                traffic_src = {"file": "profiles/pmc_traffic.json", "measured_at_commit": tj.get("_commit"),
                               "kernels": entry.get("kernels"), "method": tj.get("_method")}
        except (ValueError, AttributeError):
            traffic = None
    dk = kernels[dom]
    roofline = {"bound": "hbm", "kernel": dom, "achieved": dk["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(dk["GBps"] / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_src,
                "alg_bytes_per_launch": dk["alg_bytes"], "avg_launch_ms": dk["avg_ms"]}
    step_alg = 384.0 * n ** 3  # minimal-traffic model of one RK3 + projection step
    step_roofline = {"alg_bytes_per_step": step_alg, "achieved_GBps": round(step_alg / (ms_per_step / 1e3) / 1e9, 1),
                     "frac": round(step_alg / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}
    # the box's own ceiling: a plain device-to-device copy of 1 GiB
    # (read + write bytes per second)
    cx = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
    cy = torch.empty_like(cx)
    cy.copy_(cx)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(10):
        cy.copy_(cx)
    e1.record()
    torch.cuda.synchronize()
    copy_gbs = 2.0 * cx.numel() * 8 * 10 / (e0.elapsed_time(e1) / 1e3) / 1e9
    del cx, cy
    step_roofline["box_copy_GBps"] = round(copy_gbs, 1)
    step_roofline["frac_of_box_copy"] = round(step_alg / (ms_per_step / 1e3) / 1e9 / copy_gbs, 4)
    roofline["frac_of_box_copy"] = round(dk["GBps"] / copy_gbs, 4)
    log(f"bench: device copy {copy_gbs:.0f} GB/s; step at {100 * step_roofline['frac_of_box_copy']:.1f}% of it")

    adj = None
    if not args.no_adjoint:
        del S, S_eager, U, p
        torch.cuda.empty_cache()
        log("bench: adjoint (shear layer) ...")
        adj = adjoint_ratio(nv, args.adjoint_steps, 25)
        log(f"bench: adjoint/forward ratio {adj['ratio']}")

    cmp = None
    if not args.no_compact:
        torch.cuda.empty_cache()
        log("bench: compact scheme ...")
        cmp = compact_bench(nv, n, args.compact_steps)
        log(f"bench: compact {cmp['ms_per_step']} ms/step, {cmp['value']} Mpt-steps/s, "
            f"{cmp['achieved_GBps']} GB/s algorithmic")

    ot = None
    if not args.no_ot:
        torch.cuda.empty_cache()
        log("bench: OT MHD ...")
        ot = {"explicit": ot_bench(nv, n, args.ot_steps, False), "compact": ot_bench(nv, n, args.ot_steps, True)}
        log(f"bench: OT explicit {ot['explicit']['ms_per_step']} ms/step, compact {ot['compact']['ms_per_step']} ms/step")

    wc = None
    if not args.no_wc:
        torch.cuda.empty_cache()
        log("bench: weakly-compressible TG ...")
        wc = wc_bench(nv, n, args.wc_steps)
        log(f"bench: WC {wc['ms_per_step']} ms/step, {wc['value']} Mpt-steps/s, "
            f"{wc['achieved_GBps']} GB/s algorithmic")

    iob = None
    if not args.no_io:
        torch.cuda.empty_cache()
        log("bench: EnSight output ...")
        try:
            iob = io_bench(nv, n, args.io_writes)
            log(f"bench: EnSight pack {iob['pack_us']} us ({iob['pack_GBps']} GB/s algorithmic), "
                f"{iob['write_MBps']} MB/s to file, {iob['call_ms']} ms per call")
        except nv.NovaError as e:  # a full or absent scratch disk must not cost the line
            iob = {"error": str(e)}
            log(f"bench: EnSight output skipped: {e}")

    dropin = None
    if not args.no_dropin:
        log("bench: drop-in TG_Poisson driver ...")
        dropin = dropin_bench(n, args.dropin_steps)
        log(f"bench: drop-in {dropin}")

    ens = None
    if not args.no_ensemble:
        torch.cuda.empty_cache()
        log("bench: ensemble ...")
        ens = ensemble_bench(nv, args.ensemble_steps)
        log(f"bench: ensemble {ens['ms_per_step']} ms/step, {ens['value']} Mpt-steps/s, "
            f"{ens['achieved_GBps']} GB/s algorithmic")

    cpu = None
    if not args.no_cpu:
        log("bench: CPU baseline (oracle) ...")
        cpu = cpu_baseline(n, reps=args.cpu_steps)
        log(f"bench: CPU {cpu['value']:.3f} Mpt-steps/s on {cpu['cores']} threads")

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mpoint-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Taylor-Green initial condition)",
        "config": {"workload": f"TG {n}^3 periodic FP64 Re=1600, explicit 6th-order FD, SSP-RK3 + 1 FFT "
                               f"projection per step, dt=2e-3", "N": n, "scheme": "explicit-m6", "rk": "ssp3",
                   "graph": not args.eager, "parallelism": "single-gpu", "tuning": tuned},
        "roofline": roofline,
        "step_roofline": step_roofline,
        "kernels": kernels,
        "cpu_baseline": cpu,
        "gpu_vs_cpu": round(value / cpu["value"], 1) if cpu else None,
        "adjoint": adj,
        "ensemble": ens,
        "compact": cmp,
        "ot_mhd": ot,
        "weakly_compressible": wc,
        "ensight_output": iob,
        "dropin": dropin,
        "check": {"ke": ke, "maxabs_div": maxdiv},
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
