"""FP32 steps over LDS-staged column slabs against the CSR passes.

A slab step adds each row's products in stored column order, exactly as the
ORDERED CSR pass does, so a whole FP32 solve with slabs equals the same solve
without them bit for bit: iterations, restarts, checks, every report field and
the primal. The cases cover a slab count that is not a multiple of the
three-buffer ring, a ragged last slab, rows split over several slabs, empty
rows, runs longer than a chunk (followed by the thread of their first chunk)
and tiles with more chunks than a launch has threads.

The split layout (two or four workgroups per row block, each over a range of
the slabs) sums each row as its ranges' ordered sums in range order, with the
diagonal out of the ranges and its product added last; its steps are pinned
against the FP32 oracle bit for bit in test_gpu_fp32_steps.py, and its solves
here against the contract.
Every call goes through the C ABI of libhgqp.so.
"""
import numpy as np
import pytest

from helpers import planted_qp, tridiagonal_qp
from oracle import BoxQP

pytestmark = pytest.mark.gpu

KKT_TOL = 1e-6
OBJ_REL = 1e-8
SLAB_WIDTH = 12288
SPLIT_WIDTH = 10240

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
    """Planted n = 1e6, 16 draws: the handle picks slabs for its FP32 steps by
    itself (and keeps 4 / 2 column panels for the FP64 passes and FP32 checks):
    under the TREE reduction the split layout (123 slabs in four groups), whose
    solve is verified and recovers x*; under ORDERED one group of 82 slabs,
    whose solve equals the one without slabs bit for bit."""
    n = 1000000
    d = hg.gen_planted(n, 16, 2026)
    s, i, v, c, lo, up = d["start"], d["index"], d["value"], d["cost"], d["lower"], d["upper"]
    data = (s, i, v, c, lo, up, hg.lipschitz(s, v))
    opt = modes(hg)["refine_restart_jacobi"]
    print("case: n=1e6 headline solve with slabs", flush=True)
    on = solve(hg, data, reduce=hg.REDUCE_TREE, **opt)
    assert on[0]["slabs32"] == 123 and on[0]["slab_groups"] == 4
    assert on[0]["column_panels"] == 4 and on[0]["column_panels32"] == 2
    print(on[1], flush=True)
    assert on[1]["status"] == "optimal" and on[1]["optimality_verified"]
    assert on[1]["stationarity_residual"] <= KKT_TOL
    assert np.abs(on[2] - d["optimum"]).max() < 1e-4
    print("case: n=1e6 ORDERED with and without slabs", flush=True)
    a = solve(hg, data, reduce=hg.REDUCE_ORDERED, **opt)
    b = solve(hg, data, reduce=hg.REDUCE_ORDERED, slabs=1, **opt)
    assert a[0]["slabs32"] == 82 and a[0]["slab_groups"] == 1 and b[0]["slabs32"] == 0
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


def cut_band_qp(n, half_band, cuts, seed):
    """Banded diagonally dominant instance (11 entries per row in one slab:
    runs longer than a chunk) whose couplings across the columns `cuts` are
    dropped, so no row has entries on both sides of a cut."""
    rng = np.random.default_rng(seed)
    hr, hc, hv = [], [], []
    for k in range(1, half_band + 1):
        r = np.arange(k, n)
        keep = np.ones(len(r), dtype=bool)
        for cut in cuts:
            keep &= (r < cut) | (r - k >= cut)
        r = r[keep]
        hr.append(r); hc.append(r - k); hv.append(rng.normal(0, 0.2, size=len(r)))
    hr = np.concatenate(hr); hc = np.concatenate(hc); hv = np.concatenate(hv)
    radius = np.zeros(n)
    np.add.at(radius, hr, np.abs(hv)); np.add.at(radius, hc, np.abs(hv))
    d = np.arange(n)
    return BoxQP(n, np.concatenate([hr, d]), np.concatenate([hc, d]), np.concatenate([hv, radius + 0.5]),
                 rng.normal(0, 1, size=n), -np.ones(n), np.ones(n))


# slab width and first column of each range after the first, per group count,
# for n = 45000 (ranges start at slab ceil(h slabs / groups))
CUT_LAYOUTS = {2: (SPLIT_WIDTH, 5, (3 * 10240,)), 4: (8192, 6, (2 * 8192, 3 * 8192, 5 * 8192))}


@pytest.mark.parametrize("groups", [2, 4])
@pytest.mark.parametrize("mode", ["refine_restart_jacobi", "polish_reference"])
def test_split_band_meets_contract(hg, oracle, mode, groups):
    """n = 45000 split over 2 groups (5 slabs of 10240 columns, the last 4040
    wide: ranges 0-2 and 3-4) or 4 (6 slabs of 8192: ranges 0-1, 2, 3-4, 5), a
    band whose rows hold runs longer than a chunk: the split solve is verified
    on the device and by the oracle, lands on the one-group solve's objective,
    and repeats bit for bit."""
    width, count, cuts = CUT_LAYOUTS[groups]
    q = cut_band_qp(45000, 5, cuts, 12)
    data = engine_inputs(oracle, q)
    opt = dict(modes(hg)[mode], reduce=hg.REDUCE_TREE)
    print(f"case=cut_band groups={groups} mode={mode} n={q.n} nnz={len(data[2])}", flush=True)
    split = solve(hg, data, slabs=2, slab_groups=groups, **opt)
    again = solve(hg, data, slabs=2, slab_groups=groups, **opt)
    one = solve(hg, data, slabs=2, slab_groups=1, **opt)
    assert split[0]["slabs32"] == count and split[0]["slab_groups"] == groups
    assert one[0]["slabs32"] == 4 and one[0]["slab_groups"] == 1
    print(split[1], flush=True)
    assert split[1]["status"] == "optimal" and split[1]["optimality_verified"]
    assert split[1]["fp32_iterations"] > 0
    chk = oracle.check(q, split[2], KKT_TOL)
    assert chk["optimality_verified"] and chk["stationarity_residual"] <= KKT_TOL
    assert abs(split[1]["objective"] - one[1]["objective"]) <= OBJ_REL * (1 + abs(one[1]["objective"]))
    assert_same_solve(split, again)


@pytest.mark.parametrize("n,groups,expect", [(200000, 0, 4), (40960 + 777, 0, 2), (200000, 2, 2)])
def test_split_solve_meets_contract(hg, oracle, n, groups, expect):
    """Planted instances whose rows reach every range: the split solve (four
    groups or two by the automatic choice, or two asked for) is verified on the
    device and by the oracle, lands on the one-group optimum and x*, and repeats
    bit for bit (a fixed summation order)."""
    q, d = planted_qp(hg, n, 8, 21)
    data = engine_inputs(oracle, q)
    opt = dict(modes(hg)["refine_restart_jacobi"], reduce=hg.REDUCE_TREE)
    print(f"case=planted_split n={n} groups={groups}", flush=True)
    split = solve(hg, data, slabs=2, slab_groups=groups, **opt)
    again = solve(hg, data, slabs=2, slab_groups=groups, **opt)
    one = solve(hg, data, slabs=2, slab_groups=1, **opt)
    width = 8192 if expect == 4 else SPLIT_WIDTH
    assert split[0]["slab_groups"] == expect and split[0]["slabs32"] == -(-n // width)
    assert one[0]["slab_groups"] == 1
    print(split[1], flush=True)
    for r in (split, one):
        assert r[1]["status"] == "optimal" and r[1]["optimality_verified"]
        assert np.abs(r[2] - d["optimum"]).max() < 1e-4
    chk = oracle.check(q, split[2], KKT_TOL)
    assert chk["optimality_verified"] and chk["stationarity_residual"] <= KKT_TOL
    assert abs(split[1]["objective"] - one[1]["objective"]) <= OBJ_REL * (1 + abs(one[1]["objective"]))
    assert_same_solve(split, again)


def test_split_choice(hg, oracle):
    """The automatic choice under TREE: four groups once there are 8 slabs of
    8192 columns, else two once there are 4 of the two-group width, else one;
    one under ORDERED at any size. Two groups asked for on a matrix of one split
    slab leave the CSR passes in place."""
    q = planted_qp(hg, 30000, 1, 3)[0]  # 3 split slabs
    data = engine_inputs(oracle, q)
    fast = modes(hg)["refine_restart_jacobi"]
    assert solve(hg, data, slabs=2, reduce=hg.REDUCE_TREE, **fast)[0]["slab_groups"] == 1
    big = engine_inputs(oracle, planted_qp(hg, 4 * SPLIT_WIDTH, 1, 3)[0])  # 5 slabs of 8192
    assert solve(hg, big, slabs=2, reduce=hg.REDUCE_TREE, **fast)[0]["slab_groups"] == 2
    assert solve(hg, big, slabs=2, reduce=hg.REDUCE_ORDERED, **fast)[0]["slab_groups"] == 1
    bigger = engine_inputs(oracle, planted_qp(hg, 7 * 8192 + 1, 1, 3)[0])  # 8 slabs of 8192
    assert solve(hg, bigger, slabs=2, reduce=hg.REDUCE_TREE, **fast)[0]["slab_groups"] == 4
    small = engine_inputs(oracle, planted_qp(hg, 9000, 4, 3)[0])
    info, rep, _ = solve(hg, small, slabs=2, slab_groups=2, reduce=hg.REDUCE_TREE, **fast)
    assert info["slabs32"] == 0 and info["slab_groups"] == 0 and rep["optimality_verified"]
