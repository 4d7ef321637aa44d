"""Pins the oracle (CPU restatement) before it is trusted as the checker.

Each function of oracle/boxqp_oracle.cpp is fixed three ways where possible:
  1. against the compiled reference sources (oracle/_ref), bit for bit;
  2. against results the reference recorded (tests/golden/reference_records.json);
  3. against closed forms and brute force, chosen so that a dropped term, a
     wrong sign or a transposed operand fails.
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import BoxQP, OracleError

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def records():
    with open(os.path.join(GOLDEN, "reference_records.json")) as f:
        return json.load(f)


def two_var(maximize=False):
    """test_solver.cpp:85-89: H = diag(2,2), c = (-2,-4), x in [-10,10]^2; optimum (1,2), objective -5."""
    s = -1.0 if maximize else 1.0
    return BoxQP(2, [0, 1], [0, 1], [2 * s, 2 * s], [-2 * s, -4 * s], [-10, -10], [10, 10], maximize=maximize)


def block_psd(n=300):
    """test_v2.cpp:18-24: 2x2 blocks [[1,2],[2,5]] (PSD, not diagonally dominant), c = -1, x in [0,10]."""
    hr, hc, hv = [], [], []
    for j in range(n):
        hr.append(j); hc.append(j); hv.append(5.0 if j % 2 else 1.0)
        if j % 2:
            hr.append(j); hc.append(j - 1); hv.append(2.0)
    return BoxQP(n, hr, hc, hv, -np.ones(n), np.zeros(n), 10 * np.ones(n))


def random_triplets(seed, n=40, m=120):
    rng = np.random.default_rng(seed)
    r = rng.integers(0, n, m)
    c = rng.integers(0, n, m)
    hr, hc = np.maximum(r, c), np.minimum(r, c)
    hv = rng.normal(size=m)
    hv[::7] = 0.0                      # explicit zeros are dropped
    hr = np.concatenate([hr, hr[:10]])  # duplicates are summed
    hc = np.concatenate([hc, hc[:10]])
    hv = np.concatenate([hv, -hv[:10] * 0.5])
    return BoxQP(n, hr, hc, hv, rng.normal(size=n), -np.ones(n), np.ones(n))


def dense_of(q):
    sign = -1.0 if q.maximize else 1.0
    H = np.zeros((q.n, q.n))
    for r, c, v in zip(q.hrow, q.hcol, q.hval):
        H[r, c] += sign * v
        if r != c:
            H[c, r] += sign * v
    return H


# ---- 1. the compiled reference ---------------------------------------------------

def test_build_csr_matches_reference(oracle, reference):
    for seed in range(5):
        q = random_triplets(seed)
        a, b = oracle.build_csr(q), reference.build_csr(q)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_solve_matches_reference(oracle, reference):
    from oracle import Oracle  # noqa: F401
    probs = [two_var(), two_var(True), block_psd(), random_triplets(3), oracle.gen_tridiagonal(1000)[0]]
    for q in probs:
        a = oracle.solve_box(q, tolerance=1e-7, iteration_limit=5000, time_limit=60)
        b = reference.solve_box(q, tolerance=1e-7, iteration_limit=5000, time_limit=60)
        for k in ("status", "iterations", "objective", "stationarity_residual", "dual_bound", "relative_gap",
                  "optimality_verified", "primal_residual"):
            assert a[k] == b[k] or (isinstance(a[k], float) and math.isnan(a[k]) and math.isnan(b[k])), k
        assert np.array_equal(a["primal"], b["primal"])


def test_check_matches_reference(oracle, reference):
    rng = np.random.default_rng(7)
    for q in (two_var(), two_var(True), block_psd(40), random_triplets(1)):
        for x in (rng.uniform(-1, 1, q.n), np.zeros(q.n), np.full(q.n, 20.0)):
            a, b = oracle.check(q, x, 1e-7), reference.check(q, x, 1e-7)
            for k in ("objective", "primal_residual", "stationarity_residual", "dual_bound", "relative_gap",
                      "optimality_verified"):
                assert a[k] == b[k], (k, a[k], b[k])
        bad = np.zeros(q.n)
        bad[0] = np.nan
        assert not oracle.check(q, bad)["optimality_verified"]
        assert not reference.check(q, bad)["optimality_verified"]


def test_convexity_matches_reference(oracle, reference):
    q = block_psd()
    assert oracle.check_convex(q) and reference.check_convex(q)
    q.hval[0] = -1.0  # test_v2.cpp:25
    assert not oracle.check_convex(q) and not reference.check_convex(q)
    q = BoxQP(2, [0, 1, 1], [0, 1, 0], [1.0, 1.0, 1.5], [0, 0], [-1, -1], [1, 1])  # indefinite 2x2
    assert not oracle.check_convex(q) and not reference.check_convex(q)


def test_invalid_input_matches_reference(oracle, reference):
    q = two_var()
    for kind in (oracle, reference):
        with pytest.raises(OracleError) as e:
            kind.solve_box(q, threads=0)  # test_v2.cpp:59
        assert e.value.code == 1
        bad = two_var()
        bad.cost[0] = np.nan
        with pytest.raises(OracleError):
            kind.solve_box(bad)
    inf = two_var()
    inf.lower[1], inf.upper[1] = 3.0, 2.0
    assert oracle.solve_box(inf)["status"] == reference.solve_box(inf)["status"] == "infeasible_bounds"


# ---- 2. the reference's recorded results --------------------------------------------

@pytest.mark.parametrize("n", [10000, 50000])
def test_tridiagonal_recorded_runs(oracle, n):
    rec = [r for r in records()["tridiagonal"] if r["columns"] == n and r["backend"] == "cpu" and r["threads"] == 1][0]
    q, opt, known = oracle.gen_tridiagonal(n)
    assert known == rec["known_objective"]
    r = oracle.solve_box(q, tolerance=1e-7, iteration_limit=20000, time_limit=120)
    assert r["status"] == "optimal" and r["iterations"] == rec["iterations"] == 1700
    assert r["objective"] == rec["objective"]
    assert r["stationarity_residual"] == rec["stationarity_residual"]
    assert np.abs(r["primal"] - opt).max() == rec["max_solution_error"]


def test_tridiagonal_threads_do_not_change_result(oracle):
    q, _, _ = oracle.gen_tridiagonal(20000)
    a = oracle.solve_box(q, iteration_limit=400, threads=1)
    b = oracle.solve_box(q, iteration_limit=400, threads=4)
    assert np.array_equal(a["primal"], b["primal"]) and a["objective"] == b["objective"]


# ---- 3. closed forms and brute force ---------------------------------------------

def test_two_variable_closed_form(oracle):
    r = oracle.solve_box(two_var())
    assert r["status"] == "optimal" and abs(r["objective"] + 5) < 1e-7
    assert np.allclose(r["primal"], [1, 2], atol=1e-6)
    c = oracle.check(two_var(), np.array([1.0, 2.0]))
    assert c["objective"] == -5 and c["stationarity_residual"] == 0 and c["dual_bound"] == -5
    assert c["relative_gap"] == 0 and c["optimality_verified"] and c["positive_diagonal"]
    # x = 0: gradient (-2,-4) gives stationarity 4; separable bound -5; gap 5 / 1
    c = oracle.check(two_var(), np.zeros(2))
    assert c["stationarity_residual"] == 4 and c["dual_bound"] == -5 and c["relative_gap"] == 5
    # maximization reports the original sense and the bound in minimization sense
    c = oracle.check(two_var(True), np.array([1.0, 2.0]))
    assert c["objective"] == 5 and c["dual_bound"] == -5 and c["optimality_verified"]


def test_block_psd_closed_form(oracle):
    """Per block the optimum is (1, 0) (y at its bound with gradient 1), objective -1/2."""
    r = oracle.solve_box(block_psd(), time_limit=10)
    assert r["status"] == "optimal" and r["optimality_verified"]
    assert abs(r["objective"] + 75) < 1e-6
    expect = np.tile([1.0, 0.0], 150)
    assert np.abs(r["primal"] - expect).max() < 1e-6


def test_tiny_curvature_needs_the_gap(oracle):
    """H = diag(1e-16, 1): the projected gradient at (1e8, 0) is tiny, but the
    separable bound is 0 while the value is 1/2, so the point is not verified
    (the box analogue of test_v2.cpp:53-56 and test_qp_certificate.py:44-49)."""
    q = BoxQP(2, [0, 1], [0, 1], [1e-16, 1.0], [0, 0], [-1e12, -1], [1e12, 1])
    c = oracle.check(q, np.array([1e8, 0.0]), 1e-7)
    assert c["stationarity_residual"] < 1e-7
    assert c["dual_bound"] == 0 and abs(c["objective"] - 0.5) < 1e-15
    assert c["relative_gap"] > 0.1 and not c["optimality_verified"]


def test_tangent_bound_closed_form(oracle):
    """Non-diagonal H: bound = value + sum_j g_j (target_j - x_j), target = lower if g > 0 else upper;
    an infinite target makes the bound -inf."""
    q = BoxQP(2, [0, 1, 1], [0, 1, 0], [2.0, 2.0, 0.5], [1.0, -1.0], [-1, -2], [3, 4])
    x = np.array([0.5, 0.25])
    H = dense_of(q)
    g = H @ x + q.cost
    value = 0.5 * x @ H @ x + q.cost @ x
    target = np.where(g > 0, q.lower, q.upper)
    c = oracle.check(q, x)
    assert abs(c["objective"] - value) < 1e-15
    assert abs(c["dual_bound"] - (value + g @ (target - x))) < 1e-14
    stat = np.max(np.abs(x - np.clip(x - g, q.lower, q.upper)) / (1 + np.abs(x)))
    assert abs(c["stationarity_residual"] - stat) < 1e-16
    q.upper[1] = np.inf
    if g[1] < 0:
        assert oracle.check(q, x)["dual_bound"] == -np.inf


def test_scaled_violation_closed_form(oracle):
    q = two_var()
    c = oracle.check(q, np.array([11.0, -12.0]))
    assert abs(c["primal_residual"] - max(1 / (1 + 10 + 11), 2 / (1 + 10 + 12))) < 1e-16
    assert not c["optimality_verified"]


def test_csr_and_spmv_brute_force(oracle):
    for seed in range(4):
        q = random_triplets(seed)
        start, index, value = oracle.build_csr(q)
        dense = np.zeros((q.n, q.n))
        for i in range(q.n):
            assert np.all(np.diff(index[start[i]:start[i + 1]]) > 0)  # sorted, unique
            dense[i, index[start[i]:start[i + 1]]] = value[start[i]:start[i + 1]]
        assert np.allclose(dense, dense_of(q), rtol=0, atol=1e-14)
        assert np.all(value != 0)
        x = np.random.default_rng(seed).normal(size=q.n)
        assert np.allclose(oracle.spmv(start, index, value, x), dense_of(q) @ x, rtol=1e-13, atol=1e-13)
        assert oracle.lipschitz(start, value) == pytest.approx(np.abs(dense).sum(axis=1).max(), rel=1e-14)


def test_momentum_closed_form(oracle):
    b = oracle.betas(4)
    t1 = (1 + math.sqrt(5)) / 2
    t2 = (1 + math.sqrt(1 + 4 * t1 * t1)) / 2
    assert b[0] == 0.0 and b[1] == (t1 - 1) / t2
    assert np.all(np.diff(oracle.betas(200)) > 0) and oracle.betas(200)[-1] < 1


def test_fista_step_textbook(oracle):
    """One step from w: x+ = P(w - (Hw + c)/L), w+ = x+ + beta (x+ - x) (Beck & Teboulle 2009)."""
    q = random_triplets(2)
    s, i, v = oracle.build_csr(q)
    L = oracle.lipschitz(s, v)
    H = dense_of(q)
    rng = np.random.default_rng(0)
    x = rng.uniform(-1, 1, q.n)
    w = rng.uniform(-1, 1, q.n)
    xn, wn = oracle.fista_steps(s, i, v, q.cost, q.lower, q.upper, L, np.array([0.3]), x, w)
    expect = np.clip(w - (H @ w + q.cost) / L, q.lower, q.upper)
    assert np.allclose(xn, expect, atol=1e-14)
    assert np.allclose(wn, expect + 0.3 * (expect - x), atol=1e-14)


def test_planted_optimum_is_stationary(oracle, hg):
    """The generator's x* satisfies the KKT conditions of its model and its objective."""
    from helpers import planted_qp
    q, d = planted_qp(hg, 3000, 8, 5)
    c = oracle.check(q, d["optimum"], 1e-12)
    assert c["stationarity_residual"] < 1e-14 and c["primal_residual"] == 0
    assert abs(c["objective"] - d["objective"]) <= 1e-13 * (1 + abs(d["objective"]))
    r = oracle.solve_box(q, tolerance=1e-9, iteration_limit=5000)
    assert r["status"] == "optimal" and np.abs(r["primal"] - d["optimum"]).max() < 1e-6


def test_numpy_steps_equal_oracle_in_fp64(oracle, hg):
    """fista_steps_np in FP64 with the scalar L is orc_fista_steps bit for bit
    (long rows, empty rows, a planted instance), so the numpy restatement has
    the oracle's order of operations before it is run in FP32."""
    from helpers import planted_qp, random_csr_qp
    from oracle import fista_steps_np
    cases = [random_csr_qp(400, 3, long_rows=(5, 399), density=0.02, empty_rows=(7, 8)),
             planted_qp(hg, 2000, 8, 4)[0], random_triplets(5)]
    for q in cases:
        s, i, v = oracle.build_csr(q)
        L = oracle.lipschitz(s, v)
        x0 = np.clip(np.zeros(q.n), q.lower, q.upper)
        betas = oracle.betas(12)
        xc, wc = oracle.fista_steps(s, i, v, q.cost, q.lower, q.upper, L, betas, x0, x0)
        xn, wn = fista_steps_np(s, i, v, q.cost, q.lower, q.upper, L, betas, x0, x0)
        assert xn.dtype == np.float64
        assert np.array_equal(xn, xc) and np.array_equal(wn, wc)


def test_numpy_steps_fp32_brute_force(oracle):
    """FP32 steps against scalar loops of np.float32 operations, one rounding per
    product, sum, quotient and extrapolation, rows in stored column order; with
    per-row denominators (the Jacobi metric) and with the scalar L."""
    from helpers import random_csr_qp
    from oracle import fista_steps_np, row_abs_sums
    q = random_csr_qp(60, 11, long_rows=(3,), density=0.1, empty_rows=(20,))
    s, i, v = oracle.build_csr(q)
    L = oracle.lipschitz(s, v)
    d = row_abs_sums(s, v) * (1 + 2.0 ** -20)
    d[np.diff(s) == 0] = L
    f = np.float32
    betas = oracle.betas(6)
    for den in (L, d):
        x = np.clip(np.zeros(q.n), q.lower, q.upper).astype(f)
        w = x.copy()
        dd = np.broadcast_to(np.asarray(den, dtype=np.float64), (q.n,)).astype(f)
        for b in betas:
            g = np.zeros(q.n, dtype=f)
            for r in range(q.n):
                acc = f(0)
                for k in range(s[r], s[r + 1]):
                    acc = f(acc + f(f(v[k]) * w[i[k]]))
                g[r] = acc
            nx = np.empty(q.n, dtype=f)
            nw = np.empty(q.n, dtype=f)
            for r in range(q.n):
                val = f(w[r] - f(f(g[r] + f(q.cost[r])) / dd[r]))
                m = val if val < f(q.upper[r]) else f(q.upper[r])
                nx[r] = m if f(q.lower[r]) < m else f(q.lower[r])
                nw[r] = f(nx[r] + f(f(b) * f(nx[r] - x[r])))
            x, w = nx, nw
        x0 = np.clip(np.zeros(q.n), q.lower, q.upper)
        xn, wn = fista_steps_np(s, i, v, q.cost, q.lower, q.upper, den, betas, x0, x0, dtype=np.float32)
        assert xn.dtype == np.float32
        assert np.array_equal(xn, x) and np.array_equal(wn, w)
        x64, _ = oracle.fista_steps(s, i, v, q.cost, q.lower, q.upper, L, betas, x0, x0)
        if np.isscalar(den):  # FP32 rounding only: close to the FP64 steps
            assert np.abs(xn - x64).max() < 1e-5


def test_row_abs_sums_brute_force(oracle):
    from helpers import random_csr_qp
    from oracle import row_abs_sums
    q = random_csr_qp(300, 2, long_rows=(1,), density=0.05, empty_rows=(4,))
    s, i, v = oracle.build_csr(q)
    expect = np.abs(dense_of(q)).sum(axis=1)
    got = row_abs_sums(s, v)
    assert got[4] == 0 and np.allclose(got, expect, rtol=1e-14, atol=0)
    assert got.max() == oracle.lipschitz(s, v)


def test_numpy_steps_by_column_range(oracle):
    """A row sum taken by column ranges: with one range it is the plain ordered
    sum; with ranges it equals, for every row, the range sums computed from
    masked copies of the matrix and added in range order (checked in FP32,
    where the grouping changes the rounding)."""
    from helpers import random_csr_qp
    from oracle import fista_steps_np
    q = random_csr_qp(500, 8, long_rows=(2, 499), density=0.03, empty_rows=(9, 499 - 1))
    s, i, v = oracle.build_csr(q)
    L = oracle.lipschitz(s, v)
    x0 = np.clip(np.zeros(q.n), q.lower, q.upper)
    betas = oracle.betas(5)
    for dt in (np.float64, np.float32):
        a = fista_steps_np(s, i, v, q.cost, q.lower, q.upper, L, betas, x0, x0, dtype=dt)
        b = fista_steps_np(s, i, v, q.cost, q.lower, q.upper, L, betas, x0, x0, dtype=dt, ranges=[0, q.n])
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    bounds = [0, 130, 260, 377, q.n]
    got = fista_steps_np(s, i, v, q.cost, q.lower, q.upper, L, betas[:1], x0, x0, dtype=np.float32, ranges=bounds)
    f = np.float32
    w = x0.astype(f)
    g = np.zeros(q.n, dtype=f)
    for h in range(4):
        vh = np.where((i >= bounds[h]) & (i < bounds[h + 1]), v, 0.0)  # zeros add exactly
        p = np.zeros(q.n, dtype=f)
        for r in range(q.n):
            acc = f(0)
            for k in range(s[r], s[r + 1]):
                if bounds[h] <= i[k] < bounds[h + 1]:
                    acc = f(acc + f(f(vh[k]) * w[i[k]]))
            p[r] = acc
        g = p if h == 0 else f32_add(g, p)
    expect = np.clip(w - (g + q.cost.astype(f)) / f(L), q.lower.astype(f), q.upper.astype(f))
    assert np.array_equal(got[0], expect)


def test_numpy_steps_diagonal_last(oracle):
    """The split layouts' row sum, diagonal entries out of the ranges and added
    after them: (1) with one range, on the lower triangle (each row's diagonal
    is its last stored entry) it is the plain ordered sum bit for bit, in FP32
    and FP64; (2) on the diagonal alone it is one rounded product d_i w_i;
    (3) over four ranges it equals a loop over the entries in FP32 that skips
    the diagonal and adds its product last (and differs from the sum with the
    diagonal in its range)."""
    from helpers import random_csr_qp
    from oracle import fista_steps_np
    q = random_csr_qp(400, 12, long_rows=(3, 399), density=0.04, empty_rows=(7,))
    s, i, v = oracle.build_csr(q)
    L = oracle.lipschitz(s, v)
    rng = np.random.default_rng(4)
    x0 = np.clip(rng.normal(0, 1, size=q.n), q.lower, q.upper)
    betas = oracle.betas(4)
    row_of = np.repeat(np.arange(q.n), np.diff(s))
    low = i <= row_of  # the lower triangle keeps stored order; its diagonal ends each row
    ls = np.zeros(q.n + 1, dtype=np.int64)
    np.cumsum(np.bincount(row_of[low], minlength=q.n), out=ls[1:])
    li, lv = i[low], v[low]
    assert np.all(li[ls[1:][np.diff(ls) > 0] - 1] == np.arange(q.n)[np.diff(ls) > 0])  # diagonal last
    for dt in (np.float64, np.float32):
        a = fista_steps_np(ls, li, lv, q.cost, q.lower, q.upper, L, betas, x0, x0, dtype=dt)
        b = fista_steps_np(ls, li, lv, q.cost, q.lower, q.upper, L, betas, x0, x0, dtype=dt, ranges=[0, q.n],
                           diagonal_last=True)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    f = np.float32
    w = x0.astype(f)
    on = i == row_of
    dv = np.zeros(q.n)
    dv[row_of[on]] = v[on]  # 0 on the empty row
    ds, di = np.arange(q.n + 1), np.arange(q.n)
    got = fista_steps_np(ds, di, dv, q.cost, q.lower, q.upper, L, betas[:1], x0, x0, dtype=f, ranges=[0, 200, q.n],
                         diagonal_last=True)
    expect = np.clip(w - (dv.astype(f) * w + q.cost.astype(f)) / f(L), q.lower.astype(f), q.upper.astype(f))
    assert np.array_equal(got[0], expect)
    bounds = [0, 100, 210, 333, q.n]
    got = fista_steps_np(s, i, v, q.cost, q.lower, q.upper, L, betas[:1], x0, x0, dtype=f, ranges=bounds,
                         diagonal_last=True)
    inside = fista_steps_np(s, i, v, q.cost, q.lower, q.upper, L, betas[:1], x0, x0, dtype=f, ranges=bounds)
    g = np.zeros(q.n, dtype=f)
    for r in range(q.n):
        total = f(0)
        for h in range(4):
            acc = f(0)
            for k in range(s[r], s[r + 1]):
                if bounds[h] <= i[k] < bounds[h + 1] and i[k] != r:
                    acc = f(acc + f(f(v[k]) * w[i[k]]))
            total = acc if h == 0 else f(total + acc)
        diag = [k for k in range(s[r], s[r + 1]) if i[k] == r]
        g[r] = f(total + f(f(v[diag[0]]) * w[r])) if diag else total
    expect = np.clip(w - (g + q.cost.astype(f)) / f(L), q.lower.astype(f), q.upper.astype(f))
    assert np.array_equal(got[0], expect)
    assert not np.array_equal(got[0], inside[0])


def f32_add(a, b):
    return (a.astype(np.float32) + b.astype(np.float32)).astype(np.float32)


@pytest.mark.parametrize("n,draws,seed", [(10000, 8, 1), (10000, 8, 2026), (3000, 32, 5)])
def test_planted_family_matches_reference(oracle, reference, hg, n, draws, seed):
    """The benchmarks' planted family (C1: n = 1e4, 8 draws per row), which the
    reference never ran: the oracle's solve equals the compiled reference's bit
    for bit at KKT 1e-6 (status, iterations, objective, residuals, x), and both
    recover the planted optimum."""
    from helpers import planted_qp
    q, d = planted_qp(hg, n, draws, seed)
    a = oracle.solve_box(q, tolerance=1e-6, iteration_limit=20000, time_limit=120)
    b = reference.solve_box(q, tolerance=1e-6, iteration_limit=20000, time_limit=120)
    for k in ("status", "iterations", "objective", "stationarity_residual", "dual_bound", "relative_gap",
              "optimality_verified", "primal_residual"):
        assert a[k] == b[k], (k, a[k], b[k])
    assert np.array_equal(a["primal"], b["primal"])
    assert a["status"] == "optimal" and a["optimality_verified"]
    assert np.abs(a["primal"] - d["optimum"]).max() < 1e-4
