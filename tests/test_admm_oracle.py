"""The oracle's check for models with linear rows (orc_check_rows, restating
check_qp_solution of qp_admm.cpp:63-117 with original_violation of
mps.cpp:279-304), pinned:
  1. against the compiled reference sources (oracle/_ref), bit for bit, on
     random iterates, on the reference ADMM's own solutions and on edge cases
     (infinite row and variable bounds, zero duals, maximization, offset, a
     diagonal Hessian with its separable bound);
  2. against a hand-computed KKT point (closed form).
The GPU ADMM's results are checked by it in test_gpu_admm.py."""
import numpy as np
import pytest

from helpers import rows_qp
from oracle import BoxQP, Rows

FIELDS = ("objective", "primal_residual", "stationarity_residual", "complementarity_residual", "dual_bound",
          "relative_gap", "optimality_verified")


def same(a, b):
    for k in FIELDS:
        x, y = a[k], b[k]
        if isinstance(x, float) and np.isnan(x):
            assert np.isnan(y), k
        else:
            assert x == y, (k, x, y)


def test_closed_form_equality_row(oracle):
    # min 1/2 (x1^2 + x2^2)  s.t.  x1 + x2 = 1, -10 <= x <= 10:  x* = (1/2, 1/2), y = -1/2
    q = BoxQP(2, [0, 1], [0, 1], [1.0, 1.0], [0.0, 0.0], [-10.0, -10.0], [10.0, 10.0])
    rows = Rows([0, 2], [0, 1], [1.0, 1.0], [1.0], [1.0])
    r = oracle.check_rows(q, rows, [0.5, 0.5], [-0.5], 1e-12)
    assert r["optimality_verified"]
    assert r["objective"] == 0.25 and r["stationarity_residual"] == 0 and r["complementarity_residual"] == 0
    assert r["dual_bound"] == 0.25 and r["relative_gap"] == 0 and r["primal_residual"] == 0
    # without the dual the Lagrangian gradient is x itself: residual |0.5 - 0| / 1.5
    r = oracle.check_rows(q, rows, [0.5, 0.5], [0.0], 1e-12)
    assert not r["optimality_verified"] and r["stationarity_residual"] == pytest.approx(1 / 3, rel=1e-15)
    # an infeasible point: the row violation (1 - 0.6) / (1 + 1 + 0.6) is the primal residual
    r = oracle.check_rows(q, rows, [0.3, 0.3], [-0.3], 1e-12)
    assert r["primal_residual"] == pytest.approx(0.4 / 2.6, rel=1e-15)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_check_rows_matches_reference(oracle, reference, seed):
    q, rows = rows_qp(300, 20, 40, seed)
    rng = np.random.default_rng(seed)
    sol = reference.solve_admm(q, rows, tolerance=1e-6, time_limit=30)
    cases = [(sol["primal"], sol["row_dual"]),
             (rng.uniform(-1, 1, q.n), rng.normal(0, 1, rows.m)),
             (np.zeros(q.n), np.zeros(rows.m)),
             (sol["primal"], np.where(rng.random(rows.m) < 0.5, 0.0, sol["row_dual"]))]
    for x, y in cases:
        same(oracle.check_rows(q, rows, x, y, 1e-6), reference.check_rows(q, rows, x, y, 1e-6))


def test_check_rows_edge_cases_match_reference(oracle, reference):
    # one-sided rows (infinite far bound), free variables, maximization with an
    # offset, and a positive diagonal Hessian (separable bound with row duals)
    rng = np.random.default_rng(9)
    n, m = 40, 12
    start = np.arange(0, 4 * m + 1, 4)
    col = np.concatenate([np.sort(rng.choice(n, 4, replace=False)) for _ in range(m)])
    val = rng.normal(size=4 * m)
    lo = np.where(rng.random(m) < 0.3, -np.inf, rng.normal(size=m) - 1)
    up = np.where(rng.random(m) < 0.3, np.inf, rng.normal(size=m) + 1)
    rows = Rows(start, col, val, lo, up)
    lower = np.where(rng.random(n) < 0.2, -np.inf, -1.0)
    upper = np.where(rng.random(n) < 0.2, np.inf, 1.0)
    diag = BoxQP(n, np.arange(n), np.arange(n), rng.uniform(0.5, 2, n), rng.normal(size=n), lower, upper,
                 offset=3.5)
    full = BoxQP(n, np.r_[np.arange(n), [5, 9]], np.r_[np.arange(n), [1, 3]], np.r_[rng.uniform(2, 3, n), [0.4, -0.2]],
                 rng.normal(size=n), lower, upper, offset=-1.25, maximize=False)
    neg = BoxQP(n, np.arange(n), np.arange(n), -rng.uniform(0.5, 2, n), rng.normal(size=n), lower, upper,
                offset=2.0, maximize=True)
    for q in (diag, full, neg):
        for x, y in ((rng.uniform(-1, 1, n), rng.normal(size=m)), (rng.uniform(-1, 1, n), np.zeros(m)),
                     (np.clip(rng.normal(size=n), lower, upper), np.where(rng.random(m) < 0.5, 0, rng.normal(size=m)))):
            same(oracle.check_rows(q, rows, x, y, 1e-6), reference.check_rows(q, rows, x, y, 1e-6))


def test_check_rows_rejects_what_the_reference_rejects(oracle, reference):
    q = BoxQP(2, [0, 1], [0, 1], [1.0, 1.0], [0.0, 0.0], [-1.0, -1.0], [1.0, 1.0])
    for rows in (Rows([0, 1], [5], [1.0], [0.0], [1.0]),        # column out of range
                 Rows([0, 1], [0], [np.nan], [0.0], [1.0]),     # non-finite entry
                 Rows([0, 1], [0], [1.0], [np.inf], [np.inf])):  # lower bound +inf
        for kind in (oracle, reference):
            with pytest.raises(Exception):
                kind.check_rows(q, rows, [0.0, 0.0], [0.0], 1e-6)


def product_check(hg, q, rows, x, y, tol):
    return hg.check_qp_solution(q.n, q.hrow, q.hcol, q.hval, q.cost, q.lower, q.upper, rows.start, rows.col,
                                rows.val, rows.lower, rows.upper, x, y, tolerance=tol, offset=q.offset,
                                maximize=q.maximize)


@pytest.mark.parametrize("seed", [4, 5])
def test_product_check_equals_oracle(hg, oracle, seed):
    """hgqp_check_qp_solution (host code of the library) is the same verdict
    as the oracle's restatement, field for field, on random points and duals."""
    q, rows = rows_qp(200, 15, 30, seed)
    rng = np.random.default_rng(seed)
    for x, y in ((rng.uniform(-1, 1, q.n), rng.normal(0, 1, rows.m)), (np.zeros(q.n), np.zeros(rows.m)),
                 (rng.uniform(-1, 1, q.n), np.where(rng.random(rows.m) < 0.5, 0.0, rng.normal(size=rows.m)))):
        same(product_check(hg, q, rows, x, y, 1e-6), oracle.check_rows(q, rows, x, y, 1e-6))
    # closed form through the library
    q2 = BoxQP(2, [0, 1], [0, 1], [1.0, 1.0], [0.0, 0.0], [-10.0, -10.0], [10.0, 10.0])
    r2 = Rows([0, 2], [0, 1], [1.0, 1.0], [1.0], [1.0])
    r = product_check(hg, q2, r2, [0.5, 0.5], [-0.5], 1e-12)
    assert r["optimality_verified"] and r["objective"] == 0.25 and r["relative_gap"] == 0


def test_admm_host_outcomes_without_device(hg):
    """Model errors and early statuses of hgqp_solve_qp_admm are decided on the
    host before any device work (qp_admm.cpp:119-126)."""
    try:
        hg.device_name()
        pytest.skip("a GPU is present; this test is for hosts without one")
    except hg.HgqpError:
        pass
    base = dict(n=2, hrow=[0, 1], hcol=[0, 1], hval=[1.0, 1.0], cost=[0.0, 0.0], lower=[-1.0, -1.0],
                upper=[1.0, 1.0], rstart=[0, 2], rcol=[0, 1], rval=[1.0, 1.0], row_lower=[0.0], row_upper=[1.0])
    with pytest.raises(hg.InvalidArgument):
        hg.solve_qp_admm(**dict(base, rcol=[0, 7]))  # column out of range (mps.cpp:57)
    with pytest.raises(hg.InvalidArgument):
        hg.solve_qp_admm(**base, tolerance=0.0)  # qp_admm.cpp:119
    r = hg.solve_qp_admm(**dict(base, row_lower=[2.0], row_upper=[1.0]))
    assert r["status"] == "infeasible_bounds"
    r = hg.solve_qp_admm(**dict(base, hval=[-1.0, 1.0]))
    assert r["status"] == "unsupported_curvature"
    with pytest.raises(hg.HgqpError) as e:
        hg.solve_qp_admm(**base)
    assert e.value.code == hg.NO_DEVICE
