"""GPU ADMM for convex QPs with linear rows (hgqp_solve_qp_admm, restating
solve_qp_admm of qp_admm.cpp:118-281 without its polish and diagonal-equality
fast path), checked by the oracle's restated check_qp_solution with rows (pinned
bit for bit to the compiled reference in test_admm_oracle.py) on the returned
point and row duals, and compared with the reference's own ADMM where
oracle/_ref is built."""
import time

import numpy as np
import pytest

from helpers import rows_qp
from oracle import BoxQP, Rows, available

pytestmark = pytest.mark.gpu

KKT_TOL = 1e-6


def solve(hg, q, rows, **kw):
    kw.setdefault("tolerance", KKT_TOL)
    return hg.solve_qp_admm(q.n, q.hrow, q.hcol, q.hval, q.cost, q.lower, q.upper, rows.start, rows.col, rows.val,
                            rows.lower, rows.upper, offset=q.offset, maximize=q.maximize, **kw)


def verified_by_oracle(oracle, q, rows, r):
    assert r["status"] == "optimal" and r["optimality_verified"], r
    assert r["algorithm"] == "hip_admm_pcg"
    chk = oracle.check_rows(q, rows, r["primal"], r["row_dual"], KKT_TOL)
    assert chk["optimality_verified"]
    # the library's verdict is the same restated check (pinned on the CPU)
    for k in ("objective", "stationarity_residual", "complementarity_residual", "primal_residual"):
        assert chk[k] == r[k], k
    return chk


@pytest.mark.parametrize("n,me,mi,seed", [(60, 6, 12, 1), (400, 40, 80, 2), (3000, 200, 600, 3),
                                          (20000, 1000, 3000, 4)])
def test_admm_solutions_verified(hg, oracle, n, me, mi, seed):
    q, rows = rows_qp(n, me, mi, seed)
    print(f"case n={n} eq={me} ineq={mi} seed={seed}", flush=True)
    t = time.perf_counter()
    r = solve(hg, q, rows, iteration_limit=20000, time_limit=120)
    print(f"gpu: {r['status']} iterations {r['iterations']} objective {r['objective']!r} "
          f"{time.perf_counter() - t:.3f}s", flush=True)
    verified_by_oracle(oracle, q, rows, r)
    if available("reference"):
        from oracle import Oracle
        ref = Oracle("reference")
        t = time.perf_counter()
        rr = ref.solve_admm(q, rows, tolerance=KKT_TOL, time_limit=120)
        print(f"reference: {rr['status']} {rr['algorithm']} iterations {rr['iterations']} objective "
              f"{rr['objective']!r} {time.perf_counter() - t:.3f}s; rel diff "
              f"{abs(rr['objective'] - r['objective']) / (1 + abs(rr['objective'])):.3e}", flush=True)
        assert rr["optimality_verified"]
        assert abs(rr["objective"] - r["objective"]) <= 1e-8 * (1 + abs(rr["objective"]))


def test_admm_maximize_offset_and_free_variables(hg, oracle):
    q, rows = rows_qp(800, 30, 90, 7, free=0.3)
    qm = BoxQP(q.n, q.hrow, q.hcol, -q.hval, -q.cost, q.lower, q.upper, offset=4.5, maximize=True)
    print("case maximize", flush=True)
    r = solve(hg, qm, rows, iteration_limit=20000, time_limit=120)
    chk = verified_by_oracle(oracle, qm, rows, r)
    rmin = solve(hg, q, rows, iteration_limit=20000, time_limit=120)
    verified_by_oracle(oracle, q, rows, rmin)
    # the maximization of -f + 4.5 is the minimization of f
    assert abs(chk["objective"] - (4.5 - rmin["objective"])) <= 1e-6 * (1 + abs(rmin["objective"]))


def test_admm_diagonal_hessian_equalities(hg, oracle):
    """The structure the reference sends to its diagonal-equality fast path
    (qp_admm.cpp:127-132); here the ADMM solves it, verified by the check."""
    rng = np.random.default_rng(11)
    n, m = 500, 60
    q = BoxQP(n, np.arange(n), np.arange(n), rng.uniform(0.5, 2.0, n), rng.normal(size=n), -np.ones(n), np.ones(n))
    x0 = rng.uniform(-0.5, 0.5, n)
    start = np.arange(0, 5 * m + 1, 5)
    col = np.concatenate([np.sort(rng.choice(n, 5, replace=False)) for _ in range(m)])
    val = rng.normal(size=5 * m)
    b = np.array([val[5 * i:5 * i + 5] @ x0[col[5 * i:5 * i + 5]] for i in range(m)])
    rows = Rows(start, col, val, b, b)
    print("case diagonal equalities", flush=True)
    r = solve(hg, q, rows, iteration_limit=20000, time_limit=120)
    verified_by_oracle(oracle, q, rows, r)


def test_admm_without_rows_matches_box_engine(hg, oracle):
    q, _ = rows_qp(2000, 0, 0, 12, free=0.0)
    rows = Rows([0], [], [], [], [])
    print("case no rows", flush=True)
    r = solve(hg, q, rows, iteration_limit=20000, time_limit=120)
    verified_by_oracle(oracle, q, rows, r)
    box = hg.solve_qp(q.n, q.hrow, q.hcol, q.hval, q.cost, q.lower, q.upper, tolerance=1e-9, iteration_limit=200000,
                      time_limit=120)
    assert box["status"] == "optimal"
    assert abs(box["objective"] - r["objective"]) <= 1e-6 * (1 + abs(box["objective"]))


def test_admm_iteration_limit(hg):
    q, rows = rows_qp(3000, 200, 600, 3)
    r = solve(hg, q, rows, iteration_limit=30, time_limit=120, tolerance=1e-12)
    assert r["status"] == "iteration_limit" and r["iterations"] == 30 and not r["optimality_verified"]


def test_admm_graphs_change_nothing(hg):
    """Graph-replayed and directly launched iterations give the same solve, bit
    for bit (each launch decides on the device whether it has work)."""
    q, rows = rows_qp(3000, 200, 600, 3)
    a = solve(hg, q, rows, iteration_limit=20000, time_limit=120, use_graph=1)
    b = solve(hg, q, rows, iteration_limit=20000, time_limit=120, use_graph=0)
    assert a["status"] == b["status"] == "optimal"
    assert a["iterations"] == b["iterations"] and a["inner_iterations"] == b["inner_iterations"]
    assert np.array_equal(a["primal"], b["primal"]) and np.array_equal(a["row_dual"], b["row_dual"])
