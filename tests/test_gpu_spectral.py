"""The spectral step metric: the Jacobi denominators D scaled by
theta = min(1, 1.05 lambda), lambda the device's power-iteration estimate of
the largest eigenvalue of D^-1 H, made when the handle is created.

The estimate is pinned by what the mathematics fixes, not by a restatement:
  - a Rayleigh quotient never exceeds the largest eigenvalue, so
    theta <= 1.05 lambda_max, with lambda_max from scipy's eigsh on
    D^-1/2 H D^-1/2 (the same spectrum); on these instances 40 iterations come
    within 5%, so theta >= lambda_max and theta D - H is positive semidefinite;
  - instances whose lambda_max is 1 to within the 2^-20 margin of D
    (tridiagonal, diagonally dominant rows) keep theta = 1 exactly, and their
    spectral solves equal the Jacobi solves bit for bit;
  - under ORDERED the estimate does not depend on the column panel count (each
    row still sums its entries in stored order).
The steps are pinned as the Jacobi steps are, with theta D in place of D: the
FP64 step with momentum 0 is clip(x - (Hx + c) / (theta D)), and FP32 steps
over the CSR pass and the four-group split layout equal the FP32 oracle with
denominators theta D bit for bit. Solves meet the contract (the oracle's check,
the planted optimum) in fewer iterations than the Jacobi metric.
Every call goes through the C ABI of libhgqp.so.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from helpers import planted_qp, random_csr_qp, tridiagonal_qp
from oracle import fista_steps_np, row_abs_sums

pytestmark = pytest.mark.gpu

KKT_TOL = 1e-6
MARGIN = 1.05
REPORT_KEYS = ("status", "restarts", "iterations", "fp32_iterations", "checks", "objective", "primal_residual",
               "stationarity_residual", "dual_bound", "relative_gap", "optimality_verified", "finite")


def inputs(oracle, q):
    start, index, value = oracle.build_csr(q)
    return start, index, value, q.cost, q.lower, q.upper, oracle.lipschitz(start, value)


def jacobi_d(start, value, L):
    d = row_abs_sums(start, value) * (1 + 2.0 ** -20)
    d[np.diff(start) == 0] = L
    return d


def lambda_max(start, index, value, L):
    """Largest eigenvalue of D^-1 H (as that of D^-1/2 H D^-1/2)."""
    n = len(start) - 1
    h = sp.csr_matrix((value, index, start), shape=(n, n))
    s = sp.diags(1.0 / np.sqrt(jacobi_d(start, value, L)))
    a = s @ h @ s
    if n <= 4000:
        return float(np.linalg.eigvalsh(a.toarray()).max())
    return float(sla.eigsh(a, k=1, which="LA", tol=1e-10)[0][0])


def spectral_info(hg, data, **opt):
    s, i, v, c, lo, up, L = data
    with hg.Box(s, i, v, c, lo, up, L, step_metric=hg.STEP_SPECTRAL, **opt) as box:
        return box.info()


@pytest.mark.parametrize("case", [(3000, 8, 11), (20000, 8, 12), (20000, 16, 3)])
def test_scale_brackets_the_largest_eigenvalue(hg, oracle, case):
    n, draws, seed = case
    q = planted_qp(hg, n, draws, seed)[0]
    data = inputs(oracle, q)
    lam = lambda_max(*data[:3], data[6])
    info = spectral_info(hg, data, reduce=hg.REDUCE_TREE)
    theta = info["jacobi_scale"]
    print(f"case=planted n={n} draws={draws} lambda_max={lam:.6f} theta={theta:.6f}", flush=True)
    assert 0.5 < lam < 1.0
    assert lam <= theta <= MARGIN * lam * (1 + 1e-12)


@pytest.mark.parametrize("case", ["tridiagonal", "dominant_rows"])
def test_unit_spectrum_keeps_the_jacobi_metric(hg, oracle, case):
    """lambda_max = 1 - O(2^-20): theta = 1 and the solve equals the Jacobi solve."""
    if case == "tridiagonal":
        q = tridiagonal_qp(hg, 2000)[0]
    else:
        q = random_csr_qp(400, 5, long_rows=(3,), empty_rows=(7,))
    data = inputs(oracle, q)
    assert lambda_max(*data[:3], data[6]) > 1 - 1e-5
    s, i, v, c, lo, up, L = data
    out = []
    for metric in (hg.STEP_JACOBI, hg.STEP_SPECTRAL):
        with hg.Box(s, i, v, c, lo, up, L, step_metric=metric, momentum=hg.MOMENTUM_RESTART,
                    precision=hg.PRECISION_FP32_REFINE, reduce=hg.REDUCE_TREE) as box:
            info = box.info()
            rep = box.solve(tolerance=KKT_TOL, iteration_limit=20000, time_limit=120)
            out.append((info, rep, box.primal()))
    assert out[1][0]["jacobi_scale"] == 1.0 and out[0][0]["jacobi_scale"] == 1.0
    for k in REPORT_KEYS:
        assert out[0][1][k] == out[1][1][k], k
    assert np.array_equal(out[0][2], out[1][2])


def test_ordered_scale_ignores_column_panels(hg, oracle):
    q = planted_qp(hg, 100000, 8, 14)[0]
    data = inputs(oracle, q)
    got = [spectral_info(hg, data, reduce=hg.REDUCE_ORDERED, column_panels=p) for p in (1, 2, 3)]
    assert [g["column_panels"] for g in got] == [1, 2, 3]
    assert got[0]["jacobi_scale"] < 1.0
    assert got[0]["jacobi_scale"] == got[1]["jacobi_scale"] == got[2]["jacobi_scale"]


def test_fp64_step_is_the_scaled_projection(hg, oracle):
    """Two steps with momentum 0 from clip(0): x+ = clip(x - (Hx + c) / (theta D))."""
    q = planted_qp(hg, 3000, 8, 11)[0]
    s, i, v, c, lo, up, L = inputs(oracle, q)
    with hg.Box(s, i, v, c, lo, up, L, step_metric=hg.STEP_SPECTRAL) as box:
        theta = box.info()["jacobi_scale"]
        box.step(0.0)
        x1 = box.primal()
        box.step(0.0)
        x2 = box.primal()
    assert theta < 1.0
    n = len(c)
    rows = np.repeat(np.arange(n), np.diff(s))
    d = jacobi_d(s, v, L) * theta

    def step(x):
        return np.clip(x - (np.bincount(rows, weights=v * x[i], minlength=n) + c) / d, lo, up)

    x0 = np.clip(np.zeros(n), lo, up)
    for got, expect in ((x1, step(x0)), (x2, step(x1))):
        assert np.abs(got - expect).max() <= 1e-13 * (1 + np.abs(expect).max())
    # the unscaled step is a different point
    assert np.abs(x1 - np.clip(x0 - c / jacobi_d(s, v, L), lo, up)).max() > 1e-6


def widen(x32, lo, up):
    x = x32.astype(np.float64)
    x = np.where(x32 == lo.astype(np.float32), lo, x)
    return np.where(x32 == up.astype(np.float32), up, x)


@pytest.mark.parametrize("layout", ["csr", "split4"])
def test_fp32_steps_bit_exact(hg, oracle, layout):
    """12 FP32 steps (reference momentum, no restart) equal the FP32 oracle with
    denominators theta D bit for bit; the split layout sums each row over its
    four column ranges, then its diagonal product."""
    q = planted_qp(hg, 200000, 8, 21)[0]
    s, i, v, c, lo, up, L = data = inputs(oracle, q)
    if layout == "csr":
        opt, ranges = dict(reduce=hg.REDUCE_ORDERED, slabs=1), None
    else:
        opt = dict(reduce=hg.REDUCE_TREE, slabs=2, slab_groups=4)
        slabs = -(-q.n // 8192)
        ranges = [min(q.n, ((h * slabs + 3) // 4) * 8192) for h in range(4)] + [q.n]
    steps = 12
    with hg.Box(s, i, v, c, lo, up, L, precision=hg.PRECISION_FP32_POLISH, momentum=hg.MOMENTUM_REFERENCE,
                step_metric=hg.STEP_SPECTRAL, **opt) as box:
        info = box.info()
        rep = box.solve(tolerance=1e-6, iteration_limit=steps, check_every=steps, fp32_switch=1e-300,
                        time_limit=120)
        x = box.primal()
    print(f"layout={layout} theta={info['jacobi_scale']} ranges={ranges}", flush=True)
    assert rep["status"] == "iteration_limit" and rep["fp32_iterations"] == steps, rep
    assert info["slab_groups"] == (4 if ranges else 0)
    theta = info["jacobi_scale"]
    assert theta < 1.0
    x0 = np.clip(np.zeros(len(c)), lo, up)
    x32, _ = fista_steps_np(s, i, v, c, lo, up, jacobi_d(s, v, L) * theta, oracle.betas(steps), x0, x0,
                            dtype=np.float32, ranges=ranges, diagonal_last=ranges is not None)
    expect = widen(x32, lo, up)
    assert np.array_equal(x, expect), np.abs(x - expect).max()


@pytest.mark.parametrize("precision", ["fp64", "refine"])
@pytest.mark.parametrize("layout", ["csr", "split4", "panels"])
def test_spectral_solves_meet_the_contract(hg, oracle, layout, precision):
    """Restart solves to KKT <= 1e-6 on planted instances with 16 draws per row:
    verified by the device and the oracle, at the planted optimum, and in fewer
    iterations than the Jacobi metric on the same handle options."""
    n = {"csr": 20000, "split4": 200000, "panels": 100000}[layout]
    opt = {"csr": dict(reduce=hg.REDUCE_TREE, slabs=1),
           "split4": dict(reduce=hg.REDUCE_TREE, slabs=2, slab_groups=4),
           "panels": dict(reduce=hg.REDUCE_TREE, slabs=1, column_panels=2)}[layout]
    prec = hg.PRECISION_FP64 if precision == "fp64" else hg.PRECISION_FP32_REFINE
    q, d = planted_qp(hg, n, 16, 31)
    s, i, v, c, lo, up, L = inputs(oracle, q)
    reps = {}
    for name, metric in (("jacobi", hg.STEP_JACOBI), ("spectral", hg.STEP_SPECTRAL)):
        with hg.Box(s, i, v, c, lo, up, L, precision=prec, momentum=hg.MOMENTUM_RESTART, step_metric=metric,
                    **opt) as box:
            info = box.info()
            # a check after every step: the iteration counts are not rounded up to a check period
            rep = box.solve(tolerance=KKT_TOL, iteration_limit=20000, time_limit=120, check_every=1,
                            polish_every=1)
            x = box.primal()
        print(f"layout={layout} precision={precision} metric={name} theta={info['jacobi_scale']} "
              f"iterations={rep['iterations']} restarts={rep['restarts']}", flush=True)
        assert rep["status"] == "optimal" and rep["optimality_verified"], rep
        chk = oracle.check(q, x, KKT_TOL)
        assert chk["optimality_verified"] and chk["stationarity_residual"] <= KKT_TOL
        assert np.abs(x - d["optimum"]).max() < 1e-4
        reps[name] = rep
    assert reps["spectral"]["iterations"] < reps["jacobi"]["iterations"]

