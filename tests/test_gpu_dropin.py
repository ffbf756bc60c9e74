"""The C++ drop-in API (novaclone_amd/host/nova/nova.hpp) driving the
reference's TG_Poisson time loop (frozen-stage RK4 + one projection) must
reproduce the oracle's kinetic-energy history."""
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent


def test_tg_poisson_app_matches_oracle(orc, tmp_path):
    exe = ROOT / "novaclone_amd" / "host" / "bin" / "tg_poisson"
    assert exe.exists(), "build first (__graft_entry__.build)"
    N, steps, dt, nu = 32, 5, 2e-3, 1.0 / 1600.0
    out = tmp_path / "dat.dat"
    print("running", exe, flush=True)
    r = subprocess.run([str(exe), str(N), str(steps), repr(dt), repr(nu), str(out)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.loadtxt(out)
    assert got.shape == (steps, 3)
    O = orc.OMesh((N, N, N), (2 * math.pi,) * 3)
    x = np.arange(N) * (2 * math.pi / N)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    U = O.from_compact(np.stack([np.sin(X) * np.cos(Y) * np.cos(Z), -np.cos(X) * np.sin(Y) * np.cos(Z), 0 * X]))
    P = orc.OPoisson(O)
    p = O.field(1)
    ke = []
    for _ in range(steps):
        P.ns_step(U, p, nu, dt, orc.RK_LITERAL4)
        ke.append(O.ke(U))
    assert np.allclose(got[:, 0], dt * np.arange(1, steps + 1), rtol=1e-12)
    assert np.max(np.abs(got[:, 1] - ke) / np.array(ke)) < 1e-12


def test_tg_poisson_app_writes_ensight_output(orc, tmp_path):
    # app/TG_Piosson/TG.cpp:78-80: mesh, U and p after every step
    exe = ROOT / "novaclone_amd" / "host" / "bin" / "tg_poisson"
    N, steps, dt, nu = 16, 3, 2e-3, 1.0 / 1600.0
    outdir = tmp_path / "data"
    outdir.mkdir()
    r = subprocess.run([str(exe), str(N), str(steps), repr(dt), repr(nu), str(tmp_path / "ke.dat"), str(outdir)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    names = sorted(x.name for x in outdir.iterdir())
    want = sorted(["geom.p0000.geo", "fd.p0000.case", "restart.p.0000.case"] +
                  [f"{f}.0000.{q:05d}.dat" for f in ("U", "p") for q in range(1, steps + 1)])
    assert names == want
    # the clock records the time before each step: 0, dt, 2 dt
    assert (outdir / "restart.p.0000.case").read_text().split() == ["3", "0", "0.002", "0.004"]
    O = orc.OMesh((N, N, N), (2 * math.pi,) * 3)
    x = np.arange(N) * (2 * math.pi / N)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    U = O.from_compact(np.stack([np.sin(X) * np.cos(Y) * np.cos(Z), -np.cos(X) * np.sin(Y) * np.cos(Z), 0 * X]))
    P = orc.OPoisson(O)
    p = O.field(1)
    ref = tmp_path / "ref"
    ref.mkdir()
    E = N + 1
    for q in range(1, steps + 1):
        P.ns_step(U, p, nu, dt, orc.RK_LITERAL4)
        O.field_write(U, ref, "U", q)
        got = np.fromfile(outdir / f"U.0000.{q:05d}.dat", np.float32, offset=244).reshape(3, E, E, E)
        exp = np.fromfile(ref / f"U.0000.{q:05d}.dat", np.float32, offset=244).reshape(3, E, E, E)
        assert (outdir / f"U.0000.{q:05d}.dat").read_bytes()[:244] == (ref / f"U.0000.{q:05d}.dat").read_bytes()[:244]
        # the states agree to 1e-12; their float32 roundings to one float32 ulp
        assert np.max(np.abs(got - exp)) <= 2.0 ** -23 * np.max(np.abs(exp))
