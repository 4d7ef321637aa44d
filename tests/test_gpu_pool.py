"""The instance pool (batched solves) and the serving loop against the oracle.

Every instance of a batch must meet the contract on its own (oracle check at
KKT <= 1e-6, objective within 1e-8 of the FP64 oracle solve, planted optimum
recovered), and its iterates must not depend on what else is resident."""
import threading
import time

import numpy as np
import pytest

from oracle import BoxQP

pytestmark = pytest.mark.gpu

KKT_TOL = 1e-6
OBJ_REL = 1e-8


def as_qp(hg, d):
    hr, hc, hv = hg.lower_triplets(d["start"], d["index"], d["value"])
    return BoxQP(d["n"], hr, hc, hv, d["cost"], d["lower"], d["upper"])


def check_against_oracle(hg, oracle, d, r):
    assert r["status"] == "optimal" and r["optimality_verified"], r
    q = as_qp(hg, d)
    chk = oracle.check(q, r["primal"], KKT_TOL)
    assert chk["optimality_verified"] and chk["stationarity_residual"] <= KKT_TOL
    ref = oracle.solve_box(q, tolerance=1e-9, iteration_limit=20000, time_limit=120)
    assert abs(chk["objective"] - ref["objective"]) <= OBJ_REL * (1 + abs(ref["objective"]))
    assert abs(r["objective"] - chk["objective"]) <= 1e-13 * (1 + abs(chk["objective"]))
    assert np.abs(r["primal"] - d["optimum"]).max() < 1e-4


def solve_alone(hg, d, **opts):
    with hg.Pool(max_rows=d["n"] + 1, max_nnz=len(d["value"]), max_instances=1, **opts) as p:
        assert p.admit(d, tag=0, tolerance=KKT_TOL) is not None
        (r,) = p.solve_all()
    return r


def test_batch_matches_oracle(hg, oracle):
    batch = hg.gen_batch(16, seed=7, n_min=500, n_max=4000)
    rows = sum(d["n"] + 1 for d in batch)
    nnz = sum(len(d["value"]) for d in batch)
    print(f"batch rows={rows} nnz={nnz}", flush=True)
    with hg.Pool(max_rows=rows, max_nnz=nnz, max_instances=16) as p:
        for k, d in enumerate(batch):
            assert p.admit(d, tag=k, tolerance=KKT_TOL) is not None
        res = {r["tag"]: r for r in p.solve_all()}
    assert sorted(res) == list(range(16))
    for k, d in enumerate(batch):
        check_against_oracle(hg, oracle, d, res[k])


def test_alone_equals_mix(hg):
    batch = hg.gen_batch(10, seed=11, n_min=300, n_max=3000)
    rows = sum(d["n"] + 1 for d in batch)
    nnz = sum(len(d["value"]) for d in batch)
    with hg.Pool(max_rows=rows, max_nnz=nnz, max_instances=10) as p:
        for k, d in enumerate(batch):
            p.admit(d, tag=k, tolerance=KKT_TOL)
        mix = {r["tag"]: r for r in p.solve_all()}
    for k in (0, 4, 9):
        alone = solve_alone(hg, batch[k])
        assert alone["iterations"] == mix[k]["iterations"]
        assert np.array_equal(alone["primal"], mix[k]["primal"])
        assert alone["objective"] == mix[k]["objective"]


def test_budget_refusal_and_slot_reuse(hg, oracle):
    batch = hg.gen_batch(5, seed=13, n_min=400, n_max=900)
    with hg.Pool(max_rows=2 * 1000, max_nnz=2 * 1000 * 70, max_instances=2) as p:
        assert p.admit(batch[0], tag=0) is not None
        assert p.admit(batch[1], tag=1) is not None
        assert p.admit(batch[2], tag=2) is None  # no slot: HGQP_BUSY
        done = p.solve_all()
        assert sorted(r["tag"] for r in done) == [0, 1]
        for k in (2, 3):
            assert p.admit(batch[k], tag=k) is not None
        done += p.solve_all()
    res = {r["tag"]: r for r in done}
    for k in range(4):
        check_against_oracle(hg, oracle, batch[k], res[k])
    # the row budget refuses an instance that does not fit
    big = hg.gen_planted(3000, 8, 5)
    with hg.Pool(max_rows=2000, max_nnz=10 ** 6, max_instances=4) as p:
        assert p.admit(big, tag=0) is None


def test_iteration_limit_and_autotune(hg):
    d = hg.gen_planted(3000, 12, 21, -1.5, 1.2)
    with hg.Pool(max_rows=4000, max_nnz=10 ** 6, max_instances=2) as p:
        p.admit(d, tag=0, iteration_limit=10)
        (r,) = p.solve_all()
    assert r["status"] == "iteration_limit" and not r["optimality_verified"]
    assert r["iterations"] == 10
    with hg.Pool(max_rows=4000, max_nnz=10 ** 6, max_instances=2, block_loads=0, block_rows=0) as p:
        p.admit(d, tag=0)
        loads, rows = p.autotune()
        assert loads in (4, 8, 16) and rows in (64, 256)
        (r,) = p.solve_all()
    assert r["status"] == "optimal"


def test_time_step_restores_the_pool(hg):
    """hgqp_pool_time_step reports the algorithmic bytes of a full FP32 step over
    the resident instances and leaves them as they were: the solve afterwards
    equals the solve of the same instances in a pool that was never timed."""
    batch = hg.gen_batch(6, seed=17, n_min=400, n_max=3000)
    rows = sum(d["n"] + 1 for d in batch)
    nnz = sum(len(d["value"]) for d in batch)
    expect = sum(8 * len(d["value"]) + 4 * (d["n"] + 1) + 8 * 4 * d["n"] for d in batch)  # Jacobi: 8 vectors
    out = []
    for timed in (False, True):
        with hg.Pool(max_rows=rows, max_nnz=nnz, max_instances=6) as p:
            for k, d in enumerate(batch):
                assert p.admit(d, tag=k, tolerance=KKT_TOL) is not None
            if timed:
                us, nbytes = p.time_step(4)
                assert us > 0 and nbytes == expect
            out.append({r["tag"]: r for r in p.solve_all()})
    for k in range(len(batch)):
        assert out[0][k]["status"] == out[1][k]["status"] == "optimal"
        assert out[0][k]["iterations"] == out[1][k]["iterations"]
        assert np.array_equal(out[0][k]["primal"], out[1][k]["primal"])
    with hg.Pool(max_rows=rows, max_nnz=nnz, max_instances=6) as p:
        with pytest.raises(hg.InvalidArgument):
            p.time_step(4)  # nothing resident


def test_server_results_equal_alone(hg):
    batch = hg.gen_batch(12, seed=17, n_min=300, n_max=2500)
    alone = [solve_alone(hg, d) for d in batch[:4]]
    s = hg.Server(max_rows=3 * 2600, max_nnz=3 * 2600 * 70, max_instances=3)
    try:
        def producer():
            for k, d in enumerate(batch):
                s.submit(d, tag=k, tolerance=KKT_TOL)
                time.sleep(0.002)
        th = threading.Thread(target=producer)
        th.start()
        got = {}
        t0 = time.time()
        while len(got) < len(batch) and time.time() - t0 < 60:
            for r in s.poll():
                got[r["tag"]] = r
            time.sleep(0.001)
        th.join()
    finally:
        s.stop()
    assert sorted(got) == list(range(len(batch)))
    for r in got.values():
        assert r["status"] == "optimal" and r["latency_seconds"] >= r["queued_seconds"] >= 0
    for k in range(4):
        assert np.array_equal(got[k]["primal"], alone[k]["primal"])
        assert got[k]["iterations"] == alone[k]["iterations"]


def serve(hg, submit, count, **options):
    """Runs `submit(server)` and polls until `count` results are back."""
    s = hg.Server(**options)
    try:
        submit(s)
        got = {}
        t0 = time.time()
        while len(got) < count and time.time() - t0 < 60:
            for r in s.poll():
                got[r["tag"]] = r
            time.sleep(0.001)
    finally:
        s.stop()
    assert sorted(got) == list(range(count))
    return got


def test_server_problem_submission_equals_csr_submission(hg, oracle):
    """hgqp_server_submit_problem assembles the model on the device and hands
    that CSR to the serving loop as it is (no copy back to the host): its
    results equal hgqp_server_submit's on the host CSR of the same models (the
    planted CSR is the mirrored assembly of its lower triangle) in status,
    iterations and every primal bit, and each passes the oracle's check."""
    batch = hg.gen_batch(8, seed=23, n_min=300, n_max=3000)
    opts = dict(max_rows=3 * 3100, max_nnz=3 * 3100 * 70, max_instances=3)

    def by_problem(s):
        for k, d in enumerate(batch):
            hr, hc, hv = hg.lower_triplets(d["start"], d["index"], d["value"])
            assert s.submit_problem(d["n"], hr, hc, hv, d["cost"], d["lower"], d["upper"], tag=k,
                                    tolerance=KKT_TOL) == "running"

    def by_csr(s):
        for k, d in enumerate(batch):
            s.submit(d, tag=k, tolerance=KKT_TOL)

    print("case: server submit_problem vs submit, 8 planted instances", flush=True)
    a = serve(hg, by_problem, len(batch), **opts)
    b = serve(hg, by_csr, len(batch), **opts)
    for k, d in enumerate(batch):
        assert a[k]["status"] == b[k]["status"] == "optimal"
        assert a[k]["iterations"] == b[k]["iterations"]
        assert np.array_equal(a[k]["primal"], b[k]["primal"])
        chk = oracle.check(as_qp(hg, d), a[k]["primal"], KKT_TOL)
        assert chk["optimality_verified"] and chk["stationarity_residual"] <= KKT_TOL
