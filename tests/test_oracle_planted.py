"""Planted mistakes in the oracle fail its pins (CPU).

The oracle source carries switches (ORC_PLANT = n, off in every normal
build) that each put one plausible mistake into a function that the
closed-form pins of tests/test_oracle_mhd.py and tests/test_oracle.py
cover.  Each planted build is compiled here and the pins are run against it
in a child process: every planted build must fail them, and the faithful
build compiled the same way must pass.
"""
import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SRC = ROOT / "oracle" / "nova_oracle.c"

PLANTS = {
    1: "Lorentz term div(B, B) with the wrong sign",
    2: "Lorentz term div(B, B) dropped",
    3: "magnetic-pressure gradient with the wrong sign",
    4: "induction -div(U, B) + div(B, U) with the wrong sign",
    5: "reference composition: the B stages see the step-start U",
    6: "reference composition: the last row of RK4a replaced by 1/4s",
    7: "SSP-RK3 MHD: the second-stage weights swapped",
    8: "viscous dissipation without nu",
    9: "viscous dissipation without the z derivatives",
    10: "stretched z: A = sqrt(s2/s1)",
    11: "stretched z: tanh argument centred at 0.45",
    12: "MHD diagnostics: J from the central curl",
    13: "MHD diagnostics: Em without the 1/2",
    14: "MHD diagnostics: epsilon with +eta Jav^2",
}
PINS = ["tests/test_oracle_mhd.py", "tests/test_oracle.py::test_mesh_metric_uniform_and_stretched",
        "tests/test_oracle.py::test_stretched_mesh_end_slopes_and_symmetry"]


def _build(n: int, out: Path) -> Path:
    lib = out / f"libnova_oracle_plant{n}.so"
    cmd = ["gcc", "-O3", "-fPIC", "-fopenmp", "-ffp-contract=off", "-std=c11", "-w", f"-DORC_PLANT={n}",
           "-shared", "-o", str(lib), str(SRC), "-lm"]
    subprocess.run(cmd, check=True, capture_output=True)
    return lib


def _run_pins(lib: Path) -> subprocess.CompletedProcess:
    env = dict(os.environ, NOVA_ORACLE_LIB=str(lib), OMP_NUM_THREADS="2")
    return subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", *PINS],
                          cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)


@pytest.fixture(scope="module")
def results(tmp_path_factory):
    out = tmp_path_factory.mktemp("plants")
    ids = [0, *PLANTS]
    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        libs = dict(zip(ids, ex.map(lambda n: _build(n, out), ids)))
        runs = dict(zip(ids, ex.map(lambda n: _run_pins(libs[n]), ids)))
    return runs


def test_faithful_build_passes_the_pins(results):
    r = results[0]
    assert r.returncode == 0, r.stdout[-2000:]


@pytest.mark.parametrize("n", sorted(PLANTS))
def test_planted_mistake_fails_a_pin(results, n):
    r = results[n]
    assert r.returncode == 1 and " failed" in r.stdout, f"plant {n} ({PLANTS[n]}) passed the pins:\n{r.stdout[-1500:]}"
