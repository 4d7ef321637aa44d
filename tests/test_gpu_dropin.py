"""The reference solver with hgqp as its device engine, in C++.

oracle/_ref/dropin_box and oracle/_ref/solver_scale_hip are built from the
reference's own sources (``make -C oracle dropin``): solve_accelerated_box of
src/accelerated_qp.cpp compiled with SOVEREIGN_CUDA, as the reference's CUDA
build does, and tests/dropin/hip_box.cpp in the place of src/cuda_box.cu. The
reference's loop (accelerated_qp.cpp:37-59) then steps on the GPU through
HipBox : BoxIteration and checks every 25 iterations on the CPU. Each case runs
the device path and the reference's CPU path (cuda = false) of the same binary
and requires the same status, verdict, iteration count, objective and
stationarity bits, and the same primal bit for bit.

Cases: the n = 300 PSD block problem of tests/test_v2.cpp:18-24 (accepted by the
LDL^T convexity test, not diagonally dominant), the QPLIB models 8792 and 8991
read by the reference's read_qplib, the planted family of the benchmarks (C1:
n = 1e4, 8 draws per row; and n = 1e5), and the reference's scale_benchmark.cpp
driver itself (the seeded tridiagonal family, 1700 iterations) at n = 1e4 and
1e5 with backend "cuda" against backend "cpu".
"""
import gzip
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")
DROPIN = os.path.join(REF, "dropin_box")
SCALE = os.path.join(REF, "solver_scale_hip")
GOLDEN = os.path.join(ROOT, "tests", "golden")

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skipif(not (os.path.exists(DROPIN) and os.path.exists(SCALE)),
                       reason="drop-in binaries not built (make -C oracle dropin needs the reference sources)"),
]


def run(args, timeout=240):
    print("run:", " ".join(os.path.basename(a) for a in args), flush=True)
    p = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
    print(p.stdout.strip(), p.stderr.strip()[-500:], flush=True)
    return p.returncode, json.loads(p.stdout.strip().splitlines()[-1])


def assert_same(rec):
    dev, cpu = rec["device"], rec["cpu"]
    assert "error" not in rec
    for key in ("status", "verified", "iterations", "objective", "stationarity", "n"):
        assert dev[key] == cpu[key], (key, dev[key], cpu[key])
    assert dev["algorithm"] == "cuda_fista_box" and cpu["algorithm"] == "cpu_fista_box"
    assert rec["same_primal"] and rec["primal_entries_differing"] == 0


def test_device_name():
    rc, rec = run([DROPIN, "device"])
    assert rc == 0 and "gfx950" in rec["device"], rec


def test_psd300_through_the_reference_loop():
    rc, rec = run([DROPIN, "psd300"])
    assert rc == 0
    assert_same(rec)
    assert rec["device"]["status"] == "optimal" and rec["device"]["verified"]


@pytest.mark.parametrize("model", ["8792", "8991"])
def test_qplib_through_the_reference_loop(tmp_path, model):
    path = tmp_path / f"{model}.qplib"
    with gzip.open(os.path.join(GOLDEN, f"{model}.qplib.gz"), "rb") as f:
        path.write_bytes(f.read())
    rc, rec = run([DROPIN, "qplib", str(path)])
    assert rc == 0
    assert_same(rec)
    # the reference's recorded CPU run of this model (tests/golden/reference_records.json)
    recorded = json.load(open(os.path.join(GOLDEN, "reference_records.json")))["qplib"][model]
    assert rec["device"]["iterations"] == recorded["iterations"]
    assert float.fromhex(rec["device"]["objective"]) == recorded["objective"]
    assert rec["device"]["status"] == recorded["status"]


@pytest.mark.parametrize("n,draws,seed", [(10000, 8, 1), (10000, 8, 2026), (100000, 8, 7)])
def test_planted_through_the_reference_loop(n, draws, seed):
    """The BASELINE planted family (C1 is n = 1e4 with 8 draws) through the
    reference's loop at KKT 1e-6: hgqp and the reference CPU path agree bit
    for bit, and the solve ends optimal and verified."""
    rc, rec = run([DROPIN, "planted", str(n), str(draws), str(seed)])
    assert rc == 0
    assert_same(rec)
    assert rec["device"]["status"] == "optimal" and rec["device"]["verified"]


@pytest.mark.parametrize("n", [10000, 100000])
def test_reference_scale_driver(n):
    """scale_benchmark.cpp:7-27 unchanged: backend "cuda" now runs hgqp; its
    record must equal the CPU backend's in every deterministic field."""
    rc_d, dev = run([SCALE, str(n), "cuda", "1", "120"])
    rc_c, cpu = run([SCALE, str(n), "cpu", "1", "120"])
    assert rc_d == rc_c == 0 and dev["passed"] and cpu["passed"]
    assert dev["iterations"] == cpu["iterations"] == 1700
    for key in ("status", "objective", "known_objective", "relative_objective_error", "max_solution_error",
                "stationarity_residual", "hessian_nonzeros_lower"):
        assert dev[key] == cpu[key], (key, dev[key], cpu[key])
