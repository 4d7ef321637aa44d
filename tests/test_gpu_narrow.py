"""The narrow form of the step: one row per thread for a one-panel matrix
whose rows hold at most 3, 4 or 8 stored entries.

A thread loads its row's entries (slots past the row re-read its last entry),
gathers, and adds the rounded products in stored order, so the FP64 steps
under ORDERED must equal the CPU path (the oracle's fista_steps) bit for bit,
TREE steps must equal ORDERED ones bit for bit (every row is one sequential
sum), and FP32 steps must equal the FP32 oracle bit for bit. The matrices are
banded with empty rows at the start, in the middle and at the end (an empty
last row ends at nnz, the edge of the entry arrays), for each width class.
Every call goes through the C ABI of libhgqp.so.
"""
import numpy as np
import pytest

from oracle import BoxQP, fista_steps_np

pytestmark = pytest.mark.gpu


def banded_qp(n, width, seed):
    """Lower triplets of a diagonally dominant banded matrix whose longest row
    has 3 (width 3), 4 (width 4) or 7 (width 8) entries; rows 0, n // 2 and
    n - 1 hold none."""
    rng = np.random.default_rng(seed)
    empty = {0, n // 2, n - 1}
    pairs = []
    for i in range(1, n):
        offs = {3: (1,), 4: (1,), 8: (1, 2, 3)}[width]
        for o in offs:
            if i - o >= 0:
                pairs.append((i, i - o))
        if width == 4 and i % 10 == 0 and i - 5 >= 0:  # one extra entry in rows i and i - 5
            pairs.append((i, i - 5))
    pairs = [(r, c) for r, c in pairs if r not in empty and c not in empty]
    r = np.array([p[0] for p in pairs], dtype=np.int64)
    c = np.array([p[1] for p in pairs], dtype=np.int64)
    v = rng.normal(0, 0.3, size=len(r))
    radius = np.zeros(n)
    np.add.at(radius, r, np.abs(v))
    np.add.at(radius, c, np.abs(v))
    diag = np.array([i for i in range(n) if i not in empty], dtype=np.int64)
    dv = radius[diag] + rng.uniform(0.1, 1.0, size=len(diag))
    return BoxQP(n, np.concatenate([r, diag]), np.concatenate([c, diag]), np.concatenate([v, dv]),
                 rng.normal(0, 1, size=n), -rng.uniform(0.5, 2, size=n), rng.uniform(0.5, 2, size=n))


def inputs(oracle, q):
    start, index, value = oracle.build_csr(q)
    return start, index, value, q.cost, q.lower, q.upper, oracle.lipschitz(start, value)


def widen(x32, lo, up):
    x = x32.astype(np.float64)
    x = np.where(x32 == lo.astype(np.float32), lo, x)
    return np.where(x32 == up.astype(np.float32), up, x)


@pytest.mark.parametrize("width", [3, 4, 8])
def test_narrow_fp64_steps_bit_exact(hg, oracle, width):
    q = banded_qp(20000, width, 40 + width)
    s, i, v, c, lo, up, L = inputs(oracle, q)
    lengths = np.diff(s)
    assert lengths.max() <= width and lengths.max() > (width // 2) and (lengths == 0).sum() == 3
    betas = oracle.betas(40)
    x0 = np.clip(np.zeros(q.n), lo, up)
    x_cpu, _ = oracle.fista_steps(s, i, v, c, lo, up, L, betas, x0, x0)
    print(f"case=narrow width={width} n={q.n} nnz={len(v)}", flush=True)
    out = {}
    for reduce in (hg.REDUCE_ORDERED, hg.REDUCE_TREE):
        with hg.Box(s, i, v, c, lo, up, L, reduce=reduce) as box:
            info = box.info()
            assert info["narrow_rows"] == width and info["column_panels"] == 1, info
            for b in betas:
                box.step(b)
            out[reduce] = box.primal()
    assert np.array_equal(out[hg.REDUCE_ORDERED], x_cpu), np.abs(out[hg.REDUCE_ORDERED] - x_cpu).max()
    assert np.array_equal(out[hg.REDUCE_TREE], x_cpu)


@pytest.mark.parametrize("width", [3, 8])
def test_narrow_fp32_steps_bit_exact(hg, oracle, width):
    q = banded_qp(20000, width, 60 + width)
    s, i, v, c, lo, up, L = inputs(oracle, q)
    steps = 20
    x0 = np.clip(np.zeros(q.n), lo, up)
    x32, _ = fista_steps_np(s, i, v, c, lo, up, L, oracle.betas(steps), x0, x0, dtype=np.float32)
    expect = widen(x32, lo, up)
    with hg.Box(s, i, v, c, lo, up, L, precision=hg.PRECISION_FP32_POLISH, momentum=hg.MOMENTUM_REFERENCE,
                slabs=1) as box:
        info = box.info()
        rep = box.solve(tolerance=1e-6, iteration_limit=steps, check_every=steps, fp32_switch=1e-300,
                        time_limit=120)
        x = box.primal()
    assert info["narrow_rows"] == width and info["slabs32"] == 0, info
    assert rep["status"] == "iteration_limit" and rep["fp32_iterations"] == steps, rep
    assert np.array_equal(x, expect), np.abs(x - expect).max()


def test_row_blocks_keep_longer_rows(hg, oracle):
    """A matrix with a row of 11 entries keeps the staged row blocks."""
    q = banded_qp(3000, 8, 7)
    q = BoxQP(q.n, np.concatenate([q.hrow, np.full(4, 1000)]), np.concatenate([q.hcol, [1, 2, 3, 4]]),
              np.concatenate([q.hval, np.full(4, 0.01)]), q.cost, q.lower, q.upper)
    s, i, v, c, lo, up, L = inputs(oracle, q)
    assert np.diff(s).max() > 8
    with hg.Box(s, i, v, c, lo, up, L) as box:
        assert box.info()["narrow_rows"] == 0
