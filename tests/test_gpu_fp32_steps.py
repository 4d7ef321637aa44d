"""FP32 steps of the device against the numpy FP32 restatement of the step.

The oracle's fista_steps_np (pinned bit for bit against the C++ oracle in
FP64, and against scalar np.float32 loops in FP32) runs the reference's step
(accelerated_qp.cpp:47-57) in FP32 arithmetic: values, vectors, denominators
and each momentum rounded to FP32, each row summed in stored column order.
The device's FP32 steps must equal it bit for bit:
  - the CSR pass (step_kernel<float>, ORDERED) and the one-group slab step,
    with the scalar 1/L and the per-row Jacobi denominators;
  - the split slab step (two or four workgroups per row block), whose row sum
    is its column ranges' ordered sums added in range order, the diagonal left
    out of the ranges and its product added last: the oracle takes the same
    ranges (slab_range_start of the layout) with diagonal_last.

The steps run through hgqp_box_solve in the FP32-polish precision with the
reference momentum (no restart), a switch threshold no residual reaches and
an iteration limit of K: K FP32 steps from clip(0), widened exactly to FP64.
Every call goes through the C ABI of libhgqp.so.
"""
import numpy as np
import pytest

from helpers import planted_qp, random_csr_qp
from oracle import BoxQP, fista_steps_np, row_abs_sums

pytestmark = pytest.mark.gpu

WIDTH = {1: 12288, 2: 10240, 4: 8192}


def range_bounds(n, groups):
    """Column boundaries of the ranges of a split layout: range h starts at
    slab ceil(h slabs / groups) (slab_range_start)."""
    w = WIDTH[groups]
    slabs = -(-n // w)
    return [min(n, ((h * slabs + groups - 1) // groups) * w) for h in range(groups)] + [n]


def denominators(start, value, L, metric):
    if metric == "scalar":
        return L
    d = row_abs_sums(start, value) * (1 + 2.0 ** -20)
    d[np.diff(start) == 0] = L
    return d


def device_steps(hg, data, steps, metric, **opt):
    s, i, v, c, lo, up, L = data
    with hg.Box(s, i, v, c, lo, up, L, precision=hg.PRECISION_FP32_POLISH, momentum=hg.MOMENTUM_REFERENCE,
                step_metric=hg.STEP_JACOBI if metric == "jacobi" else hg.STEP_SCALAR, **opt) as box:
        info = box.info()
        rep = box.solve(tolerance=1e-6, iteration_limit=steps, check_every=steps, fp32_switch=1e-300,
                        time_limit=120)
        x = box.primal()
    assert rep["status"] == "iteration_limit" and rep["fp32_iterations"] == steps, rep
    return info, x


def oracle_steps(oracle, data, steps, metric, ranges=None):
    """K FP32 steps of the oracle; with `ranges`, the split layouts' row sum
    (ranges without the diagonal, then the diagonal's product)."""
    s, i, v, c, lo, up, L = data
    x0 = np.clip(np.zeros(len(c)), lo, up)
    x32, _ = fista_steps_np(s, i, v, c, lo, up, denominators(s, v, L, metric), oracle.betas(steps), x0, x0,
                            dtype=np.float32, ranges=ranges, diagonal_last=ranges is not None)
    return x32


def widen(x32, lo, up):
    """FP64 of the FP32 iterate, exact FP64 bounds where it sits on an FP32 bound."""
    x = x32.astype(np.float64)
    x = np.where(x32 == lo.astype(np.float32), lo, x)
    return np.where(x32 == up.astype(np.float32), up, x)


def inputs(oracle, q):
    """The mirrored CSR and L as the reference assembles them (minimize: no sign change)."""
    start, index, value = oracle.build_csr(q)
    return start, index, value, q.cost, q.lower, q.upper, oracle.lipschitz(start, value)


def spread_qp(n, seed):
    """Tridiagonal rows plus rows with entries in every slab and empty rows."""
    from test_gpu_slabs import spread_rows_qp
    return spread_rows_qp(n, (5, n // 2, n - 1), 3001, (10, 11, 3 * n // 4), seed)


def one_group_case(hg, case):
    """(instance, slabs option, expected slab count)"""
    if case == "planted_runs":       # 3 slabs, runs of up to ~20 entries (longer than a chunk)
        return planted_qp(hg, 30000, 16, 5)[0], 2, 3
    if case == "one_slab":           # one slab: each row one run of up to ~65 entries
        return planted_qp(hg, 5000, 32, 6)[0], 2, 1
    if case == "spread_rows":        # rows in every slab, empty rows
        return spread_qp(40000, 9), 2, 4
    return random_csr_qp(1500, 5, long_rows=(3, 700, 1499), density=0.004, empty_rows=(10, 11)), 2, 1


@pytest.mark.parametrize("metric", ["scalar", "jacobi"])
@pytest.mark.parametrize("case", ["planted_runs", "one_slab", "spread_rows", "long_rows"])
def test_fp32_steps_bit_exact(hg, oracle, case, metric):
    """20 FP32 steps over the CSR pass and over one-group slabs equal the FP32
    oracle bit for bit."""
    q, slabs, count = one_group_case(hg, case)
    data = inputs(oracle, q)
    steps = 20
    expect = widen(oracle_steps(oracle, data, steps, metric), q.lower, q.upper)
    for opt in (dict(slabs=1), dict(slabs=slabs)):
        print(f"case={case} metric={metric} n={q.n} nnz={len(data[2])} {opt}", flush=True)
        info, x = device_steps(hg, data, steps, metric, reduce=hg.REDUCE_ORDERED, **opt)
        assert info["slabs32"] == (0 if opt["slabs"] == 1 else count)
        assert np.array_equal(x, expect), np.abs(x - expect).max()


def test_fp32_steps_automatic_slabs(hg, oracle):
    """Planted n = 4e5: the handle chooses 33 slabs (and column panels) by itself
    under ORDERED; 10 steps equal the FP32 oracle bit for bit."""
    q = planted_qp(hg, 400000, 8, 7)[0]
    data = inputs(oracle, q)
    expect = widen(oracle_steps(oracle, data, 10, "jacobi"), q.lower, q.upper)
    info, x = device_steps(hg, data, 10, "jacobi", reduce=hg.REDUCE_ORDERED)
    assert info["slabs32"] == 33 and info["slab_groups"] == 1
    assert np.array_equal(x, expect), np.abs(x - expect).max()


@pytest.mark.parametrize("groups", [2, 4])
@pytest.mark.parametrize("metric", ["scalar", "jacobi"])
def test_split_fp32_steps_bit_exact(hg, oracle, groups, metric):
    """Planted n = 2e5 (rows reach every range) over a split layout: each row is
    its ranges' ordered sums added in range order, then its diagonal product;
    12 steps equal the oracle with the same ranges bit for bit, and differ from
    the one-range sum."""
    q = planted_qp(hg, 200000, 8, 21)[0]
    data = inputs(oracle, q)
    bounds = range_bounds(q.n, groups)
    expect = widen(oracle_steps(oracle, data, 12, metric, ranges=bounds), q.lower, q.upper)
    print(f"groups={groups} metric={metric} ranges={bounds}", flush=True)
    info, x = device_steps(hg, data, 12, metric, reduce=hg.REDUCE_TREE, slabs=2, slab_groups=groups)
    assert info["slab_groups"] == groups and info["slabs32"] == -(-q.n // WIDTH[groups])
    assert np.array_equal(x, expect), np.abs(x - expect).max()
    plain = widen(oracle_steps(oracle, data, 12, metric), q.lower, q.upper)
    assert not np.array_equal(plain, expect)  # the grouping is visible in FP32


@pytest.mark.parametrize("groups", [2, 4])
def test_split_fp32_steps_long_runs(hg, oracle, groups):
    """A band of 11 entries per row (runs longer than a chunk once the diagonal
    is out: 10 entries), no row coupled across a range boundary: 10 split
    steps equal the oracle over the layout's ranges bit for bit."""
    from test_gpu_slabs import CUT_LAYOUTS, cut_band_qp
    width, count, cuts = CUT_LAYOUTS[groups]
    q = cut_band_qp(45000, 5, cuts, 12)
    data = inputs(oracle, q)
    bounds = range_bounds(q.n, groups)
    assert bounds[1:-1] == list(cuts)
    print(f"case=cut_band groups={groups} ranges={bounds}", flush=True)
    expect = widen(oracle_steps(oracle, data, 10, "jacobi", ranges=bounds), q.lower, q.upper)
    info, x = device_steps(hg, data, 10, "jacobi", reduce=hg.REDUCE_TREE, slabs=2, slab_groups=groups)
    assert info["slab_groups"] == groups and info["slabs32"] == count
    assert np.array_equal(x, expect), np.abs(x - expect).max()


def tall_block_qp(n, run, seed):
    """A banded instance tall enough that four-group row blocks reach their
    16383-row cap, with a run of `run` entries (one slab, many chunks) in the
    last row of the first blocks and in the first row after each: the local
    row 16382 carries both run marks."""
    rng = np.random.default_rng(seed)
    r = np.arange(1, n)
    hr = [r]
    hc = [r - 1]
    hv = [rng.normal(0, 0.3, size=n - 1)]
    for blk in range(1, 4):
        last = blk * 16383 - 1
        for row in (last, last + 1):
            cols = np.arange(2, 2 + run)  # columns of one slab, below the row
            hr.append(np.full(run, row)); hc.append(cols); hv.append(rng.normal(0, 0.1, size=run))
    hr = np.concatenate(hr); hc = np.concatenate(hc); hv = np.concatenate(hv)
    radius = np.zeros(n)
    np.add.at(radius, hr, np.abs(hv)); np.add.at(radius, hc, np.abs(hv))
    d = np.arange(n)
    return BoxQP(n, np.concatenate([hr, d]), np.concatenate([hc, d]), np.concatenate([hv, radius + 0.5]),
                 rng.normal(0, 1, size=n), -np.ones(n), np.ones(n))


def test_four_group_blocks_at_the_row_cap(hg, oracle):
    """Four-group row blocks of 16383 rows whose last row holds a 40-entry run
    (continuation and follow marks on local row 16382): 8 split steps equal the
    oracle over the four ranges bit for bit, so no entry of the run is taken
    for padding."""
    q = tall_block_qp(1200000, 40, 3)
    data = inputs(oracle, q)
    bounds = range_bounds(q.n, 4)
    expect = widen(oracle_steps(oracle, data, 8, "jacobi", ranges=bounds), q.lower, q.upper)
    info, x = device_steps(hg, data, 8, "jacobi", reduce=hg.REDUCE_TREE, slabs=2, slab_groups=4)
    assert info["slab_groups"] == 4
    assert np.array_equal(x, expect), np.abs(x - expect).max()
