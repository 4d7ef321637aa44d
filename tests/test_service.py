"""The box-QP HTTP service: routes, limits and records (CPU).

The HTTP layer runs against a stand-in engine that records what it was given,
so every limit of the reference service (02-software/backend/tests/test_api.py:
invalid inputs :106-127, UTF-8 byte limit :130-134, body limit before parsing
and chunked upload :137-143, 429 with Retry-After and slot release :146-168,
startup failure :221-227) is exercised here without a GPU. QPLIB text is
parsed by the real reader through the C ABI (host code). The GPU tests of the
same routes with the real engine are in test_gpu_service.py."""
import gzip
import json
import math
import os
import threading
from concurrent.futures import ThreadPoolExecutor

import pytest

fastapi = pytest.importorskip("fastapi")
from fastapi.testclient import TestClient  # noqa: E402

import service  # noqa: E402
from helpers import QCL  # noqa: E402

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def qplib(name):
    with gzip.open(os.path.join(GOLDEN, f"{name}.qplib.gz"), "rt") as f:
        return f.read()


class StandIn:
    """Returns a fixed record; can hold a request until released."""

    def __init__(self, capacity=2, objective=1.5):
        self.capacity = capacity
        self.calls = []
        self.hold = None
        self.entered = threading.Event()
        self.objective = objective

    def describe(self):
        return {"name": "stand-in"}

    async def solve(self, problem, engine, tolerance, iteration_limit, time_limit, precision, momentum,
                    step="scalar"):
        import asyncio
        self.calls.append(dict(problem=problem, engine=engine, tolerance=tolerance, iteration_limit=iteration_limit,
                               time_limit=time_limit, precision=precision, momentum=momentum, step=step))
        self.entered.set()
        if self.hold is not None:
            while not self.hold.is_set():
                await asyncio.sleep(0.01)
        n = problem["n"]
        return service.result_record(n, dict(status="optimal", objective=self.objective, dual_bound=-math.inf,
                                              relative_gap=math.nan, iterations=25, optimality_verified=True,
                                              primal_residual=0.0, stationarity_residual=1e-9,
                                              complementarity_residual=0.0, primal=[0.0] * n), 0.01, "stand-in")

    async def solve_rows(self, problem, tolerance, iteration_limit, time_limit):
        self.calls.append(dict(problem=problem, engine="admm", tolerance=tolerance, iteration_limit=iteration_limit,
                               time_limit=time_limit))
        n = problem["n"]
        return service.result_record(n, dict(status="optimal", objective=self.objective, dual_bound=-math.inf,
                                              relative_gap=math.nan, iterations=25, optimality_verified=True,
                                              primal_residual=0.0, stationarity_residual=1e-9,
                                              complementarity_residual=0.0, primal=[0.0] * n), 0.01, "stand-in",
                                     rows=problem["rows"])

    def close(self):
        pass


def tiny_problem(**kw):
    p = dict(n=2, rows=[0, 1, 1], cols=[0, 0, 1], values=[2.0, 0.5, 2.0], cost=[-1.0, 1.0],
             lower=[None, -1.0], upper=[1.0, None])
    p.update(kw)
    return p


@pytest.fixture
def engine():
    return StandIn()


@pytest.fixture
def client(engine):
    with TestClient(service.create_app(engine=engine, settings=service.Settings())) as c:
        yield c


def test_discovery(client):
    assert client.get("/health").json() == {"ok": True, "engine": "hgqp"}
    caps = client.get("/capabilities").json()
    assert caps["solver"]["name"] == "stand-in" and caps["max_concurrent_requests"] == 2
    assert caps["formats"] == ["qplib", "triplets"]
    assert "/qp/solve" in client.get("/openapi.json").json()["paths"]


def test_triplets_reach_engine_and_record_is_json_safe(client, engine):
    r = client.post("/qp/solve", json={"problem": tiny_problem(), "tolerance": 1e-6})
    assert r.status_code == 200
    rec = r.json()
    json.dumps(rec, allow_nan=False)
    assert rec["engine"] == "hgqp" and rec["status"] == "optimal"
    assert rec["best_bound"] is None and rec["gap"] is None  # -inf and NaN become null
    assert rec["objective"] == 1.5 and rec["columns"] == 2 and rec["rows"] == 0
    call = engine.calls[-1]
    p = call["problem"]
    assert list(p["lower"]) == [-math.inf, -1.0] and list(p["upper"]) == [1.0, math.inf]
    assert list(p["hrow"]) == [0, 1, 1] and list(p["hval"]) == [2.0, 0.5, 2.0]
    assert call["engine"] == "served" and call["time_limit"] == 20


def test_qplib_text_is_parsed_by_the_library(client, engine):
    text = qplib("8991")
    r = client.post("/qp/solve", json={"model": text, "engine": "single", "precision": "fp32-refine",
                                       "step": "jacobi", "time_limit": 5})
    assert r.status_code == 200 and r.json()["status"] == "optimal"
    call = engine.calls[-1]
    assert call["problem"]["n"] == 14400 and call["problem"]["rows"] == 0
    assert call["engine"] == "single" and call["precision"] == "fp32-refine" and call["time_limit"] == 5
    assert call["step"] == "jacobi"


@pytest.mark.parametrize("body,code", [
    ({}, 400),
    ({"model": "  "}, 400),
    ({"model": "x", "problem": tiny_problem()}, 400),
    ({"problem": tiny_problem(cost=[1.0])}, 400),
    ({"problem": tiny_problem(rows=[0])}, 400),
    ({"problem": tiny_problem(n=0)}, 422),
    ({"problem": tiny_problem(), "tolerance": 0}, 422),
    ({"problem": tiny_problem(), "time_limit": -1}, 422),
    ({"problem": tiny_problem(), "engine": "cpu"}, 422),
    ({"problem": tiny_problem(), "precision": "fp16"}, 422),
    ({"problem": tiny_problem(), "step": "newton"}, 422),
])
def test_rejects_invalid_requests(client, engine, body, code):
    assert client.post("/qp/solve", json=body).status_code == code
    assert not engine.calls


@pytest.mark.parametrize("value", ["NaN", "Infinity", "-Infinity"])
def test_nonfinite_input_gets_a_json_safe_422(client, value):
    r = client.post("/qp/solve", content='{"problem": %s, "tolerance": %s}' % (json.dumps(tiny_problem()), value),
                    headers={"Content-Type": "application/json"})
    assert r.status_code == 422
    json.dumps(r.json(), allow_nan=False)




def test_unreadable_models_are_records_and_rows_go_to_admm(client, engine):
    bad = client.post("/qp/solve", json={"model": "8991 QCB minimize\n2\n"}).json()
    assert bad["status"] == "input_rejected" and bad["primal"] == [] and bad["objective"] is None
    assert not engine.calls
    # a continuous QP with linear rows (type QCL) takes the ADMM path
    rec = client.post("/qp/solve", json={"model": QCL, "engine": "served"}).json()
    assert rec["status"] == "optimal" and rec["rows"] == 1
    call = engine.calls[-1]
    assert call["engine"] == "admm" and call["problem"]["rows"] == 1
    assert list(call["problem"]["rstart"]) == [0, 2] and list(call["problem"]["rcol"]) == [0, 1]


def test_dimension_limits_are_records(engine):
    settings = service.Settings(max_columns=10, max_nonzeros=2)
    with TestClient(service.create_app(engine=engine, settings=settings)) as c:
        rec = c.post("/qp/solve", json={"problem": tiny_problem()}).json()
        assert rec["status"] == "input_rejected" and "limits" in rec["message"]
        big = tiny_problem(n=11, cost=[0.0] * 11, lower=[0.0] * 11, upper=[1.0] * 11, rows=[0], cols=[0],
                           values=[1.0])
        assert c.post("/qp/solve", json={"problem": big}).json()["status"] == "input_rejected"
    assert not engine.calls


def test_model_limit_counts_utf8_bytes(engine):
    with TestClient(service.create_app(engine=engine, settings=service.Settings(max_model_bytes=10))) as c:
        assert c.post("/qp/solve", json={"model": "é" * 6}).status_code == 413
        assert c.post("/qp/solve", content=json.dumps({"model": "\ud800"}),
                      headers={"Content-Type": "application/json"}).status_code == 400


def test_body_limit_before_parsing_and_chunked_upload(engine):
    with TestClient(service.create_app(engine=engine, settings=service.Settings(max_request_bytes=64))) as c:
        assert c.post("/qp/solve", content=b"x" * 65).status_code == 413
        assert c.post("/qp/solve", content=iter([b"x" * 32, b"y" * 33])).status_code == 413
        assert c.post("/qp/solve", json={}).status_code == 400  # the slot was released
        assert c.post("/not-a-route", content=b"x" * 64_001).status_code == 413
        assert c.post("/qp/solve", content=b"{}", headers={"Content-Length": "-1"}).status_code == 400


def test_busy_gets_429_keeps_health_and_releases_slots():
    engine = StandIn(capacity=1)
    engine.hold = threading.Event()
    with TestClient(service.create_app(engine=engine, settings=service.Settings())) as c:
        with ThreadPoolExecutor(max_workers=1) as pool:
            first = pool.submit(c.post, "/qp/solve", json={"problem": tiny_problem()})
            assert engine.entered.wait(5)
            try:
                busy = c.post("/qp/solve", json={"problem": tiny_problem()})
                assert busy.status_code == 429 and busy.headers["Retry-After"] == "2"
                assert c.get("/health").status_code == 200
            finally:
                engine.hold.set()
            assert first.result().status_code == 200
        assert c.post("/qp/solve", json={"problem": tiny_problem()}).status_code == 200


def test_startup_fails_without_a_device():
    import hgqp
    try:
        hgqp.device_name()
        pytest.skip("a GPU is present; the startup failure is for hosts without one")
    except hgqp.HgqpError:
        pass
    with pytest.raises(hgqp.HgqpError):
        with TestClient(service.create_app()):
            pass
