"""FP32 steps over LDS-staged column slabs against the CSR passes.

A slab step adds each row's products in stored column order, exactly as the
ORDERED CSR pass does, so a whole FP32 solve with slabs equals the same solve
without them bit for bit: iterations, restarts, checks, every report field and
the primal. The cases cover a slab count that is not a multiple of the
three-buffer ring, a ragged last slab, rows split over several slabs, empty
rows, runs longer than a chunk (followed by the thread of their first chunk)
and tiles with more chunks than a launch has threads. Every call goes
through the C ABI of libhgqp.so.
"""
import numpy as np
import pytest

from helpers import planted_qp, tridiagonal_qp
from oracle import BoxQP

pytestmark = pytest.mark.gpu

KKT_TOL = 1e-6
SLAB_WIDTH = 12288

REPORT_KEYS = ("status", "restarts", "iterations", "fp32_iterations", "checks", "objective", "primal_residual",
               "stationarity_residual", "dual_bound", "relative_gap", "optimality_verified", "finite")


def engine_inputs(oracle, q):
    start, index, value = oracle.build_csr(q)
    return start, index, value, q.cost, q.lower, q.upper, oracle.lipschitz(start, value)


def spread_rows_qp(n, long_rows, stride, empty_rows, seed):
    """Tridiagonal diagonally dominant instance plus rows that reach every
    stride-th column (a few entries in each slab), with empty rows."""
    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []
    for i in range(1, n):
        if i in empty_rows or i - 1 in empty_rows:
            continue
        rows.append(i); cols.append(i - 1); vals.append(float(rng.normal(0, 0.3)))
    for r in long_rows:
        for j in range(0, n, stride):
            if j < r and j not in empty_rows and j != r - 1:
                rows.append(r); cols.append(j); vals.append(float(rng.normal(0, 0.3)))
            elif j > r + 1 and j not in empty_rows:
                rows.append(j); cols.append(r); vals.append(float(rng.normal(0, 0.3)))
    rows = np.array(rows, dtype=np.int64); cols = np.array(cols, dtype=np.int64); vals = np.array(vals)
    radius = np.zeros(n)
    np.add.at(radius, rows, np.abs(vals)); np.add.at(radius, cols, np.abs(vals))
    diag = np.array([i for i in range(n) if i not in empty_rows], dtype=np.int64)
    dvals = radius[diag] + rng.uniform(0.1, 1.0, size=len(diag))
    hr = np.concatenate([rows, diag]); hc = np.concatenate([cols, diag]); hv = np.concatenate([vals, dvals])
    return BoxQP(n, hr, hc, hv, rng.normal(0, 1, size=n), -rng.uniform(0.5, 2, size=n), rng.uniform(0.5, 2, size=n))


def solve(hg, data, **opt):
    s, i, v, c, lo, up, L = data
    with hg.Box(s, i, v, c, lo, up, L, **opt) as box:
        info = box.info()
        rep = box.solve(tolerance=KKT_TOL, iteration_limit=20000, time_limit=120)
        x = box.primal()
    return info, rep, x


def assert_same_solve(a, b):
    for k in REPORT_KEYS:
        va, vb = a[1][k], b[1][k]
        assert va == vb or (isinstance(va, float) and np.isnan(va) and np.isnan(vb)), (k, va, vb)
    assert np.array_equal(a[2], b[2]), np.abs(a[2] - b[2]).max()


def modes(hg):
    return {
        "refine_restart_jacobi": dict(precision=hg.PRECISION_FP32_REFINE, momentum=hg.MOMENTUM_RESTART,
                                      step_metric=hg.STEP_JACOBI),
        "polish_reference": dict(precision=hg.PRECISION_FP32_POLISH, momentum=hg.MOMENTUM_REFERENCE),
    }


def slab_case(hg, case):
    """(instance, slabs option, slab count expected)"""
    if case == "planted_400k":  # 33 slabs (not a multiple of 3), chosen by the automatic rule
        return planted_qp(hg, 400000, 8, 7)[0], 0, 33
    if case == "long_runs":  # 3 slabs; rows have up to ~20 entries in one slab, so runs span chunks
        return planted_qp(hg, 30000, 16, 5)[0], 2, 3
    if case == "one_slab_rows":  # one slab: every row is one run of up to ~65 entries
        return planted_qp(hg, 5000, 32, 6)[0], 2, 1
    if case == "planted_ragged":  # 3 slabs, the last 5424 columns wide; forced below the size rule
        return planted_qp(hg, 30000, 1, 3)[0], 2, 3
    if case == "tridiagonal":
        return tridiagonal_qp(hg, 2000)[0], 2, 1
    if case == "whole_slabs":  # n = 2 x 12288: no ragged slab
        return planted_qp(hg, 2 * SLAB_WIDTH, 2, 8)[0], 2, 2
    # rows with entries in every slab, empty rows, 4 slabs
    return spread_rows_qp(40000, (5, 20000, 39999), 3001, (10, 11, 30000), 9), 2, 4


@pytest.mark.parametrize("mode", ["refine_restart_jacobi", "polish_reference"])
@pytest.mark.parametrize("case", ["planted_400k", "planted_ragged", "tridiagonal", "spread_rows", "long_runs",
                                  "one_slab_rows", "whole_slabs"])
def test_slab_solve_equals_ordered_csr(hg, oracle, case, mode):
    q, slabs, count = slab_case(hg, case)
    data = engine_inputs(oracle, q)
    opt = dict(modes(hg)[mode], reduce=hg.REDUCE_ORDERED)
    print(f"case={case} mode={mode} n={q.n} nnz={len(data[2])} slabs={slabs}", flush=True)
    on = solve(hg, data, slabs=slabs, **opt)
    off = solve(hg, data, slabs=1, **opt)
    assert on[0]["slabs32"] == count
    assert off[0]["slabs32"] == 0
    print(on[1], flush=True)
    assert on[1]["status"] == "optimal" and on[1]["optimality_verified"]
    assert on[1]["fp32_iterations"] > 0
    assert_same_solve(on, off)


def test_tiles_wider_than_a_launch(hg, oracle):
    """Tridiagonal n = 1e6 over 82 slabs: each row block's ~3900 rows put runs of
    three entries into one tile (about 1950 chunks), more than the 512 threads of
    a launch, so threads sum several chunks of that tile in turn."""
    q = tridiagonal_qp(hg, 1000000)[0]
    data = engine_inputs(oracle, q)
    opt = dict(modes(hg)["refine_restart_jacobi"], reduce=hg.REDUCE_ORDERED)
    on = solve(hg, data, slabs=2, **opt)
    off = solve(hg, data, slabs=1, **opt)
    assert on[0]["slabs32"] == 82
    print(on[1], flush=True)
    assert on[1]["optimality_verified"]
    assert_same_solve(on, off)


def test_full_size_slabs(hg):
    """Planted n = 1e6, 16 draws: the handle picks 82 slabs for its FP32 steps by
    itself (and keeps 4 / 2 column panels for the FP64 passes and FP32 checks);
    the headline solve is verified and recovers x*, and with the ORDERED
    reduction it equals the solve without slabs bit for bit."""
    n = 1000000
    d = hg.gen_planted(n, 16, 2026)
    s, i, v, c, lo, up = d["start"], d["index"], d["value"], d["cost"], d["lower"], d["upper"]
    data = (s, i, v, c, lo, up, hg.lipschitz(s, v))
    opt = modes(hg)["refine_restart_jacobi"]
    print("case: n=1e6 headline solve with slabs", flush=True)
    on = solve(hg, data, reduce=hg.REDUCE_TREE, **opt)
    assert on[0]["slabs32"] == 82 and on[0]["column_panels"] == 4 and on[0]["column_panels32"] == 2
    print(on[1], flush=True)
    assert on[1]["status"] == "optimal" and on[1]["optimality_verified"]
    assert on[1]["stationarity_residual"] <= KKT_TOL
    assert np.abs(on[2] - d["optimum"]).max() < 1e-4
    print("case: n=1e6 ORDERED with and without slabs", flush=True)
    a = solve(hg, data, reduce=hg.REDUCE_ORDERED, **opt)
    b = solve(hg, data, reduce=hg.REDUCE_ORDERED, slabs=1, **opt)
    assert a[0]["slabs32"] == 82 and b[0]["slabs32"] == 0
    assert_same_solve(a, b)


def test_slabs_fall_back(hg, oracle):
    """Slabs need ascending columns in every row and serve FP32 steps of a
    non-diagonal Hessian only; otherwise the handle keeps the CSR passes
    (slabs32 = 0) and solves as before. A banded matrix (11 entries of a row in
    one slab) takes slabs and equals the CSR pass bit for bit."""
    q = planted_qp(hg, 30000, 1, 3)[0]
    s, i, v, c, lo, up, L = engine_inputs(oracle, q)
    fast = dict(modes(hg)["refine_restart_jacobi"], reduce=hg.REDUCE_ORDERED)
    # descending columns: a valid CSR the slabs cannot take
    i2, v2 = i.copy(), v.copy()
    for r in range(len(s) - 1):
        i2[s[r]:s[r + 1]] = i[s[r]:s[r + 1]][::-1]
        v2[s[r]:s[r + 1]] = v[s[r]:s[r + 1]][::-1]
    info, rep, _ = solve(hg, (s, i2, v2, c, lo, up, L), slabs=2, **fast)
    assert info["slabs32"] == 0 and rep["optimality_verified"]
    # FP64 handle, and the switch turned off
    assert solve(hg, (s, i, v, c, lo, up, L), slabs=2)[0]["slabs32"] == 0
    assert solve(hg, (s, i, v, c, lo, up, L), slabs=1, **fast)[0]["slabs32"] == 0
    # a banded matrix: 11 entries of almost every row in one slab
    n, hb = 20000, 5
    rng = np.random.default_rng(4)
    hr, hc, hv = [], [], []
    for k in range(1, hb + 1):
        r = np.arange(k, n)
        hr.append(r); hc.append(r - k); hv.append(rng.normal(0, 0.2, size=n - k))
    hr = np.concatenate(hr); hc = np.concatenate(hc); hv = np.concatenate(hv)
    radius = np.zeros(n)
    np.add.at(radius, hr, np.abs(hv)); np.add.at(radius, hc, np.abs(hv))
    d = np.arange(n)
    band = BoxQP(n, np.concatenate([hr, d]), np.concatenate([hc, d]), np.concatenate([hv, radius + 0.5]),
                 rng.normal(0, 1, size=n), -np.ones(n), np.ones(n))
    on = solve(hg, engine_inputs(oracle, band), slabs=2, **fast)
    off = solve(hg, engine_inputs(oracle, band), slabs=1, **fast)
    assert on[0]["slabs32"] == 2 and on[1]["optimality_verified"]
    assert_same_solve(on, off)
    # a diagonal Hessian
    dn = 5000
    ds = np.arange(dn + 1, dtype=np.int32)
    info, rep, _ = solve(hg, (ds, np.arange(dn, dtype=np.int32), np.full(dn, 2.0), np.ones(dn), -np.ones(dn),
                              np.ones(dn), 2.0), slabs=2, **fast)
    assert info["slabs32"] == 0 and rep["optimality_verified"]
