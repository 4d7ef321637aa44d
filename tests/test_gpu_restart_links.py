"""Restart decisions made inside the next step launch against decisions made by
a finalize launch.

Under the restart schedule a run of back-to-back step launches hands each
step's restart decision to the next launch: every workgroup of that launch sums
the step's partials in the finalize launch's order and reaches the same
decision, and workgroup 0 stores it. Only the run's last step is followed by a
finalize launch. An FP32 phase that ends on the steps' own residual runs the
same steps whatever its period lengths, so solves whose periods start at one
step (the first decision from a finalize launch, then alternating) and at 25
steps (nearly all decisions inside step launches) must agree bit for bit, with
and without graph replay, on every report field and the primal. Layouts whose
step cannot decide inside its own workgroups (one-group slabs, several FP32
column panels) take a finalize launch before the step instead and must agree
too. Every solve is also held to the oracle's check.
Every call goes through the C ABI of libhgqp.so.
"""
import numpy as np
import pytest

from helpers import planted_qp

pytestmark = pytest.mark.gpu

KKT_TOL = 1e-6
REPORT_KEYS = ("status", "restarts", "iterations", "fp32_iterations", "checks", "objective", "primal_residual",
               "stationarity_residual", "dual_bound", "relative_gap", "optimality_verified", "finite")


def solve(hg, data, check_every, use_graph, **opt):
    s, i, v, c, lo, up, L = data
    with hg.Box(s, i, v, c, lo, up, L, use_graph=use_graph, **opt) as box:
        info = box.info()
        # the FP64 phase after an FP32 phase is decided by checks every polish_every
        # steps; it is fixed so that only the FP32 phase's periods change
        rep = box.solve(tolerance=KKT_TOL, iteration_limit=20000, time_limit=120, check_every=check_every,
                        polish_every=5)
        x = box.primal()
    return info, rep, x


def assert_same(a, b):
    for k in REPORT_KEYS:
        va, vb = a[1][k], b[1][k]
        assert va == vb or (isinstance(va, float) and np.isnan(va) and np.isnan(vb)), (k, va, vb)
    assert np.array_equal(a[2], b[2]), np.abs(a[2] - b[2]).max()


def layouts(hg):
    """(n, draws, seed, box options, what the layout must be)"""
    tree = dict(reduce=hg.REDUCE_TREE)
    return {
        # CSR step launches decide for their predecessor (the single-instance C1 shape)
        "csr": (10000, 8, 11, dict(tree), lambda i: i["slabs32"] == 0 and i["column_panels32"] <= 1),
        # split slabs: the sum launch decides, four groups
        "split4": (200000, 8, 12, dict(tree, slabs=2, slab_groups=4), lambda i: i["slab_groups"] == 4),
        # one-group slabs under ORDERED: a finalize launch before each step
        "one_group": (200000, 8, 13, dict(reduce=hg.REDUCE_ORDERED, slabs=2), lambda i: i["slab_groups"] == 1),
        # two FP32 column panels: the carry launch picks the source point, so a finalize launch comes first
        "panels": (100000, 8, 14, dict(tree, slabs=1, column_panels=2), lambda i: i["column_panels32"] == 2),
    }


@pytest.mark.parametrize("precision", ["refine", "polish"])
@pytest.mark.parametrize("layout", ["csr", "split4", "one_group", "panels"])
def test_linked_decisions_equal_finalized(hg, oracle, layout, precision):
    n, draws, seed, opt, expect = layouts(hg)[layout]
    q, d = planted_qp(hg, n, draws, seed)
    start, index, value = oracle.build_csr(q)
    data = (start, index, value, q.cost, q.lower, q.upper, oracle.lipschitz(start, value))
    prec = hg.PRECISION_FP32_REFINE if precision == "refine" else hg.PRECISION_FP32_POLISH
    opt = dict(opt, precision=prec, momentum=hg.MOMENTUM_RESTART, step_metric=hg.STEP_JACOBI)
    print(f"case=linked_restart layout={layout} precision={precision} n={n}", flush=True)
    base = solve(hg, data, 1, 1, **opt)
    assert expect(base[0]), base[0]
    rep = base[1]
    print(rep, flush=True)
    assert rep["status"] == "optimal" and rep["optimality_verified"] and rep["restarts"] > 0
    chk = oracle.check(q, base[2], KKT_TOL)
    assert chk["optimality_verified"] and chk["stationarity_residual"] <= KKT_TOL
    assert np.abs(base[2] - d["optimum"]).max() < 1e-4
    for check_every, use_graph in ((25, 1), (25, 0), (1, 0), (7, 1)):
        assert_same(base, solve(hg, data, check_every, use_graph, **opt))
