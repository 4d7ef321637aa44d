"""The box-QP HTTP service with the real engine on the GPU.

- QPLIB 8792 through the `single` engine (the reference FP64 schedule)
  returns the reference's own recorded run bit for bit: objective, iterations,
  stationarity and the full primal vector
  (reports/cpu-validation-2026-09-17/results/8792.default.ours.json);
- the same model through the `served` engine (FP32 correction cycles, FP64
  verdict on the device) is verified, and its objective is within 1e-8 of
  the recorded FP64 one;
- planted models posted concurrently join one resident batch, and each
  record equals the record of the same model posted alone, bit for bit."""
import gzip
import json
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

fastapi = pytest.importorskip("fastapi")
from fastapi.testclient import TestClient  # noqa: E402

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KKT_TOL = 1e-6


@pytest.fixture(scope="module")
def client():
    import service
    settings = service.Settings(slots=64, pool_rows=2_000_000, pool_nnz=40_000_000, solve_time_limit=60)
    with TestClient(service.create_app(settings=settings)) as c:
        yield c


def record_8792():
    with open(os.path.join(GOLDEN, "reference_records.json")) as f:
        rec = json.load(f)["qplib"]["8792"]
    rec["primal"] = np.load(os.path.join(GOLDEN, "qplib_primal.npz"))["x8792"]
    with gzip.open(os.path.join(GOLDEN, "8792.qplib.gz"), "rt") as f:
        text = f.read()
    return text, rec


def planted_problem(hg, n, draws, seed):
    d = hg.gen_planted(n, draws, seed)
    rows, cols, vals = hg.lower_triplets(d["start"], d["index"], d["value"])
    return dict(n=n, rows=rows.tolist(), cols=cols.tolist(), values=vals.tolist(), cost=d["cost"].tolist(),
                lower=d["lower"].tolist(), upper=d["upper"].tolist())


def test_qplib_single_engine_reproduces_reference_run(client):
    text, rec = record_8792()
    print("case: 8792 single fp64", flush=True)
    r = client.post("/qp/solve", json={"model": text, "engine": "single", "tolerance": 1e-7,
                                       "iteration_limit": 200000})
    assert r.status_code == 200
    got = r.json()
    assert got["status"] == rec["status"] == "optimal" and got["algorithm"] == "hip_fista_box"
    assert got["iterations"] == rec["iterations"]
    assert got["objective"] == rec["objective"]
    assert got["stationarity_residual"] == rec["stationarity_residual"]
    assert np.array_equal(np.array(got["primal"]), rec["primal"])


def test_qplib_single_engine_fast_modes_verified(client):
    text, rec = record_8792()
    print("case: 8792 single fp32-refine restart jacobi", flush=True)
    got = client.post("/qp/solve", json={"model": text, "engine": "single", "precision": "fp32-refine",
                                         "momentum": "restart", "step": "jacobi", "tolerance": KKT_TOL}).json()
    assert got["status"] == "optimal" and got["optimality_verified"]
    assert got["algorithm"] == "hip_fista_box_fp32_refine_restart_jacobi"
    assert got["stationarity_residual"] <= KKT_TOL
    assert abs(got["objective"] - rec["objective"]) <= 1e-8 * (1 + abs(rec["objective"]))


def test_qplib_served_engine_verified(client):
    text, rec = record_8792()
    print("case: 8792 served", flush=True)
    got = client.post("/qp/solve", json={"model": text, "tolerance": KKT_TOL}).json()
    assert got["status"] == "optimal" and got["optimality_verified"]
    assert got["stationarity_residual"] <= KKT_TOL
    assert abs(got["objective"] - rec["objective"]) <= 1e-8 * (1 + abs(rec["objective"]))


def test_concurrent_requests_equal_alone(client, hg):
    problems = [planted_problem(hg, n, 8, 900 + k) for k, n in enumerate([1500, 2600, 4100, 1800, 3300, 2200])]
    print("case: service concurrency, 6 planted models", flush=True)
    alone = [client.post("/qp/solve", json={"problem": p, "tolerance": KKT_TOL}).json() for p in problems]
    with ThreadPoolExecutor(max_workers=len(problems)) as pool:
        mixed = list(pool.map(lambda p: client.post("/qp/solve", json={"problem": p, "tolerance": KKT_TOL}).json(),
                              problems))
    for a, m in zip(alone, mixed):
        assert a["status"] == m["status"] == "optimal"
        assert a["stationarity_residual"] <= KKT_TOL
        assert a["iterations"] == m["iterations"]
        assert a["objective"] == m["objective"]
        assert a["primal"] == m["primal"]


def test_linear_rows_take_the_admm_path(client, hg):
    """A QPLIB model with linear rows (type QCL) goes to hgqp_solve_qp_admm:
    min 1/2 (x1^2 + x2^2) s.t. x1 + x2 = 1 has x* = (1/2, 1/2), objective 1/4."""
    from helpers import QCL
    print("case: QCL two-variable model", flush=True)
    got = client.post("/qp/solve", json={"model": QCL, "tolerance": KKT_TOL}).json()
    assert got["status"] == "optimal" and got["optimality_verified"] and got["rows"] == 1
    assert got["algorithm"] == "hip_admm_pcg"
    assert abs(got["objective"] - 0.25) <= 1e-6  # KKT 1e-6 (the ADMM's tail is slow without the polish)
    assert np.abs(np.array(got["primal"]) - 0.5).max() <= 1e-6
