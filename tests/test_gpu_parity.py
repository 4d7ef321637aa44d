"""HIP engine against the oracle (CPU restatement of the reference path).

Bit-exact where the arithmetic is the reference's (ORDERED reduction, FP64,
reference momentum); within the stated tolerances otherwise. Every call goes
through the C ABI of libhgqp.so.
"""
import numpy as np
import pytest

from helpers import planted_qp, random_csr_qp, tridiagonal_qp

pytestmark = pytest.mark.gpu

# Tolerances of the box-QP contract: KKT (stationarity) <= 1e-6 and objective
# within 1e-8 relative of an FP64 CPU solve.
KKT_TOL = 1e-6
OBJ_REL = 1e-8


def engine_inputs(oracle, q):
    """The sign-normalized mirrored CSR, c and L exactly as the reference assembles them."""
    start, index, value = oracle.build_csr(q)
    sign = -1.0 if q.maximize else 1.0
    return start, index, value, sign * q.cost, q.lower, q.upper, oracle.lipschitz(start, value)


def cases(hg):
    return {
        "tridiagonal": tridiagonal_qp(hg, 2000)[0],
        "planted": planted_qp(hg, 3000, 8, 11)[0],
        "long_rows": random_csr_qp(1500, 5, long_rows=(3, 700, 1499), density=0.004, empty_rows=(10, 11)),
    }


def test_device_is_gfx950(hg):
    assert "gfx950" in hg.device_name()


@pytest.mark.parametrize("case", ["tridiagonal", "planted", "long_rows"])
def test_steps_bit_exact(hg, oracle, case):
    q = cases(hg)[case]
    s, i, v, c, lo, up, L = engine_inputs(oracle, q)
    betas = oracle.betas(40)
    x0 = np.clip(np.zeros(q.n), lo, up)
    x_cpu, _ = oracle.fista_steps(s, i, v, c, lo, up, L, betas, x0, x0)
    print(f"case={case} n={q.n} nnz={len(v)}", flush=True)
    with hg.Box(s, i, v, c, lo, up, L) as box:
        for b in betas:
            box.step(b)
        x_gpu = box.primal()
    assert np.array_equal(x_gpu, x_cpu), np.abs(x_gpu - x_cpu).max()
    with hg.Box(s, i, v, c, lo, up, L, reduce=hg.REDUCE_TREE) as box:
        for b in betas:
            box.step(b)
        x_tree = box.primal()
    assert np.abs(x_tree - x_cpu).max() <= 1e-12 * (1 + np.abs(x_cpu).max())


@pytest.mark.parametrize("case", ["tridiagonal", "planted", "long_rows"])
def test_check_matches_oracle(hg, oracle, case):
    q = cases(hg)[case]
    s, i, v, c, lo, up, L = engine_inputs(oracle, q)
    betas = oracle.betas(30)
    with hg.Box(s, i, v, c, lo, up, L) as box:
        for b in betas:
            box.step(b)
        x = box.primal()
        dev = box.check(1e-7)
    ref = oracle.check(q, x, 1e-7)
    print(case, dev, ref, flush=True)
    assert dev["optimality_verified"] == ref["optimality_verified"]
    assert abs(dev["objective"] - ref["objective"]) <= 1e-13 * (1 + abs(ref["objective"]))
    assert abs(dev["stationarity_residual"] - ref["stationarity_residual"]) <= 1e-12 * (1 + ref["stationarity_residual"])
    assert dev["primal_residual"] == ref["primal_residual"] == 0.0
    assert abs(dev["dual_bound"] - ref["dual_bound"]) <= 1e-12 * (1 + abs(ref["dual_bound"]))
    assert abs(dev["relative_gap"] - ref["relative_gap"]) <= 1e-12


@pytest.mark.parametrize("panels", [1, 4])
def test_resident_solve_bit_exact(hg, oracle, panels):
    """Device-resident loop with the reference schedule equals the CPU path bit for
    bit, with the sparse passes in one panel or split into column panels."""
    q, d = planted_qp(hg, 10000, 8, 1)
    s, i, v, c, lo, up, L = engine_inputs(oracle, q)
    ref = oracle.solve_box(q, tolerance=1e-7, iteration_limit=20000, time_limit=120)
    for graph in (1, 0):
        print(f"panels={panels} graph={graph}", flush=True)
        with hg.Box(s, i, v, c, lo, up, L, use_graph=graph, column_panels=panels) as box:
            assert box.info()["column_panels"] == panels
            rep = box.solve(tolerance=1e-7, iteration_limit=20000, time_limit=120)
            x = box.primal()
        print(rep, flush=True)
        assert rep["status"] == ref["status"] == "optimal"
        assert rep["iterations"] == ref["iterations"]
        assert np.array_equal(x, ref["primal"])
        assert abs(rep["objective"] - ref["objective"]) <= 1e-13 * (1 + abs(ref["objective"]))


def test_tridiagonal_golden(hg):
    """The reference's recorded run (kaggle-gpu-03 results.json, n=10000): 1700 iterations,
    objective -947.1248894454228, stationarity 7.093582844702874e-08, max error 5.3349e-06."""
    d = hg.gen_tridiagonal(10000)
    r = hg.solve_qp(10000, d["hrow"], d["hcol"], d["hval"], d["cost"], d["lower"], d["upper"],
                    tolerance=1e-7, iteration_limit=20000, time_limit=120)
    print(r["status"], r["iterations"], repr(r["objective"]), r["stationarity_residual"], flush=True)
    assert r["status"] == "optimal" and r["optimality_verified"]
    assert r["iterations"] == 1700
    assert abs(r["objective"] - -947.1248894454228) <= 1e-13 * 947
    assert abs(r["stationarity_residual"] - 7.093582844702874e-08) <= 1e-12
    assert abs(np.abs(r["primal"] - d["optimum"]).max() - 5.334931689526812e-06) <= 1e-12


@pytest.mark.parametrize("momentum,precision,metric,panels", [
    (1, 0, 0, 0), (1, 1, 0, 0), (0, 1, 0, 0), (1, 2, 0, 0), (0, 2, 0, 0),
    (1, 0, 1, 0), (1, 1, 1, 0), (1, 2, 1, 0), (0, 2, 1, 0),
    (1, 0, 1, 3), (1, 1, 1, 3), (1, 2, 1, 3), (1, 2, 0, 5)])
def test_fast_modes_meet_contract(hg, oracle, momentum, precision, metric, panels):
    """Restart, FP32 with FP64 polish, FP32 correction cycles anchored at FP64,
    each with the scalar 1/L step or the per-row Jacobi step, in one panel or
    column panels: KKT <= 1e-6 by the oracle's own check, objective within 1e-8
    of the FP64 oracle, x* recovered."""
    q, d = planted_qp(hg, 10000, 8, 3)
    s, i, v, c, lo, up, L = engine_inputs(oracle, q)
    ref = oracle.solve_box(q, tolerance=1e-9, iteration_limit=20000, time_limit=120)
    with hg.Box(s, i, v, c, lo, up, L, momentum=momentum, precision=precision, reduce=hg.REDUCE_TREE,
                step_metric=metric, column_panels=panels) as box:
        rep = box.solve(tolerance=KKT_TOL, iteration_limit=20000, time_limit=120)
        x = box.primal()
    print(rep, flush=True)
    assert rep["status"] == "optimal"
    chk = oracle.check(q, x, KKT_TOL)
    assert chk["optimality_verified"] and chk["stationarity_residual"] <= KKT_TOL
    assert abs(chk["objective"] - ref["objective"]) <= OBJ_REL * (1 + abs(ref["objective"]))
    assert np.abs(x - d["optimum"]).max() < 1e-4


def test_jacobi_step_is_the_row_scaled_projection(hg, oracle):
    """Jacobi steps with momentum 0 from x0 = clip(0): x+_i = clip(x_i - (Hx + c)_i / D_i)
    with D_i = (sum_k |H_ik|)(1 + 2^-20), against numpy on the engine's own CSR
    (the second step starts from the device's first iterate, so Hx is nonzero)."""
    q, d = planted_qp(hg, 3000, 8, 11)
    s, i, v, c, lo, up, L = engine_inputs(oracle, q)
    with hg.Box(s, i, v, c, lo, up, L, step_metric=hg.STEP_JACOBI) as box:
        box.step(0.0)
        x1 = box.primal()
        box.step(0.0)
        x2 = box.primal()
    n = len(c)
    rows = np.repeat(np.arange(n), np.diff(s))
    D = np.bincount(rows, weights=np.abs(v), minlength=n) * (1.0 + 2.0 ** -20)

    def step(x):
        return np.clip(x - (np.bincount(rows, weights=v * x[i], minlength=n) + c) / D, lo, up)

    x0 = np.clip(np.zeros(n), lo, up)
    for got, expect in ((x1, step(x0)), (x2, step(x1))):
        assert np.abs(got - expect).max() <= 1e-13 * (1 + np.abs(expect).max())
    assert np.abs(x2 - x1).max() > 1e-3  # the gradient moved the iterate


@pytest.mark.parametrize("panels", [2, 3, 16])
@pytest.mark.parametrize("case", ["tridiagonal", "planted", "long_rows"])
def test_column_panels_bit_exact(hg, oracle, case, panels):
    """Sparse passes split into column panels continue each row's running sum
    from the earlier panels, so FP64 ORDERED steps equal the CPU path bit for bit
    for any panel count (rows longer than a wave's stage, empty rows and rows
    with no entries in a panel included); the TREE reduction stays within
    1e-12; the check equals the one-panel check (stationarity bit for bit)."""
    q = cases(hg)[case]
    s, i, v, c, lo, up, L = engine_inputs(oracle, q)
    betas = oracle.betas(40)
    x0 = np.clip(np.zeros(q.n), lo, up)
    x_cpu, _ = oracle.fista_steps(s, i, v, c, lo, up, L, betas, x0, x0)
    print(f"case={case} panels={panels} n={q.n} nnz={len(v)}", flush=True)
    with hg.Box(s, i, v, c, lo, up, L, column_panels=panels) as box:
        assert box.info()["column_panels"] == panels
        for b in betas:
            box.step(b)
        x_gpu = box.primal()
        dev = box.check(1e-7)
    assert np.array_equal(x_gpu, x_cpu), np.abs(x_gpu - x_cpu).max()
    with hg.Box(s, i, v, c, lo, up, L) as box:
        assert box.info()["column_panels"] == 1
        for b in betas:
            box.step(b)
        one = box.check(1e-7)
    assert dev["stationarity_residual"] == one["stationarity_residual"]
    assert dev["optimality_verified"] == one["optimality_verified"]
    for k in ("objective", "dual_bound"):
        assert abs(dev[k] - one[k]) <= 1e-15 * (1 + abs(one[k]))
    with hg.Box(s, i, v, c, lo, up, L, column_panels=panels, reduce=hg.REDUCE_TREE) as box:
        for b in betas:
            box.step(b)
        x_tree = box.primal()
    assert np.abs(x_tree - x_cpu).max() <= 1e-12 * (1 + np.abs(x_cpu).max())


def test_full_size_panels_chosen_and_bit_exact(hg, oracle):
    """Planted n = 1e6, 16 draws (33M stored entries): the handle splits its
    passes into 4 FP64 and 2 FP32 column panels by itself; FP64 steps still equal
    the CPU path bit for bit; the FP32 correction-cycle solve with restart and
    the Jacobi step is verified and recovers x*."""
    n = 1000000
    d = hg.gen_planted(n, 16, 2026)
    s, i, v, c, lo, up = d["start"], d["index"], d["value"], d["cost"], d["lower"], d["upper"]
    L = hg.lipschitz(s, v)
    betas = oracle.betas(6)
    x0 = np.clip(np.zeros(n), lo, up)
    x_cpu, _ = oracle.fista_steps(s, i, v, c, lo, up, L, betas, x0, x0, threads=8)
    print("case: n=1e6 FP64 steps with automatic panels", flush=True)
    with hg.Box(s, i, v, c, lo, up, L) as box:
        assert box.info()["column_panels"] == 4
        for b in betas:
            box.step(b)
        assert np.array_equal(box.primal(), x_cpu)
    print("case: n=1e6 FP32 refine restart Jacobi solve", flush=True)
    with hg.Box(s, i, v, c, lo, up, L, precision=hg.PRECISION_FP32_REFINE, momentum=hg.MOMENTUM_RESTART,
                reduce=hg.REDUCE_TREE, step_metric=hg.STEP_JACOBI) as box:
        info = box.info()
        assert info["column_panels"] == 4 and info["column_panels32"] == 2
        rep = box.solve(tolerance=KKT_TOL, iteration_limit=20000, time_limit=120)
        x = box.primal()
    print(rep, flush=True)
    assert rep["status"] == "optimal" and rep["optimality_verified"]
    assert rep["stationarity_residual"] <= KKT_TOL
    assert np.abs(x - d["optimum"]).max() < 1e-4


def test_fp32_one_panel_beside_fp64_panels(hg):
    """Planted n = 4e5: the size rule splits the FP64 passes into 2 column
    panels and keeps the FP32 passes on one, a layout of its own whose row
    blocks come from the host row starts; the FP32 solve on it is verified and
    recovers x*."""
    n = 400000
    d = hg.gen_planted(n, 8, 7)
    s, i, v, c, lo, up = d["start"], d["index"], d["value"], d["cost"], d["lower"], d["upper"]
    with hg.Box(s, i, v, c, lo, up, hg.lipschitz(s, v), precision=hg.PRECISION_FP32_REFINE,
                momentum=hg.MOMENTUM_RESTART, step_metric=hg.STEP_JACOBI, slabs=1) as box:
        info = box.info()
        assert info["column_panels"] == 2 and info["column_panels32"] == 1 and info["slabs32"] == 0
        rep = box.solve(tolerance=KKT_TOL, iteration_limit=20000, time_limit=120)
        x = box.primal()
    assert rep["status"] == "optimal" and rep["optimality_verified"]
    assert np.abs(x - d["optimum"]).max() < 1e-4


def test_column_panels_fall_back_to_one(hg, oracle):
    """Panels need each row's columns ascending (the split keeps stored order
    only then) and are never used for a diagonal Hessian (its separable bound
    reads each row's entry): both keep one panel even when more are asked for,
    and the FP64 steps still equal the CPU path on the stored order."""
    q, d = planted_qp(hg, 3000, 8, 11)
    s, i, v, c, lo, up, L = engine_inputs(oracle, q)
    # reverse each row's stored order: a valid CSR whose columns descend
    i2, v2 = i.copy(), v.copy()
    for r in range(len(s) - 1):
        i2[s[r]:s[r + 1]] = i[s[r]:s[r + 1]][::-1]
        v2[s[r]:s[r + 1]] = v[s[r]:s[r + 1]][::-1]
    betas = oracle.betas(20)
    x0 = np.clip(np.zeros(q.n), lo, up)
    x_cpu, _ = oracle.fista_steps(s, i2, v2, c, lo, up, L, betas, x0, x0)
    with hg.Box(s, i2, v2, c, lo, up, L, column_panels=4) as box:
        assert box.info()["column_panels"] == 1
        for b in betas:
            box.step(b)
        assert np.array_equal(box.primal(), x_cpu)
    n = 500
    ds = np.arange(n + 1, dtype=np.int32)
    with hg.Box(ds, np.arange(n, dtype=np.int32), np.full(n, 2.0), np.ones(n), -np.ones(n), np.ones(n), 2.0,
                column_panels=4) as box:
        assert box.info()["column_panels"] == 1
