"""The device verdict on every branch of check_qp_solution, the reference's own
box-QP regression cases through the HIP engine, and repeated solves of one
handle at different tolerances.

The device check restates check_qp_solution (qp_admm.cpp:63-117) for a box
QP: objective in the original sense (sign, offset), stationarity, scaled
violation, the tangent dual bound or, for a positive diagonal Hessian, the
separable bound (qp_admm.cpp:106-112), the relative gap and the verdict rule
(qp_admm.cpp:115-116). Each field is compared with the oracle's check at the
same iterate. Every call goes through the C ABI of libhgqp.so.
"""
import numpy as np
import pytest

from helpers import planted_qp
from oracle import BoxQP

pytestmark = pytest.mark.gpu


def engine_inputs(oracle, q):
    """The sign-normalized mirrored CSR, c and L exactly as the reference assembles them."""
    start, index, value = oracle.build_csr(q)
    sign = -1.0 if q.maximize else 1.0
    return start, index, value, sign * q.cost, q.lower, q.upper, oracle.lipschitz(start, value)


def two_var():
    """test_solver.cpp:85-89: H = diag(2, 2), c = (-2, -4), x in [-10, 10]^2; optimum (1, 2), objective -5."""
    return BoxQP(2, [0, 1], [0, 1], [2.0, 2.0], [-2.0, -4.0], [-10, -10], [10, 10])


def diag_maximize():
    """test_v2.cpp:28-33 without its row: maximize -(x^2 + 2y^2 + 3z^2) + 4x + 10y + 20z + 1.25
    over [0, 10]^3 (a positive diagonal after the sign change), optimum (2, 2.5, 10/3)."""
    return BoxQP(3, [0, 1, 2], [0, 1, 2], [-2.0, -4.0, -6.0], [4.0, 10.0, 20.0], [0, 0, 0], [10, 10, 10],
                 offset=1.25, maximize=True)


def planted_maximize(hg):
    """A planted instance stated as the maximization of its negation, with an offset."""
    q, _ = planted_qp(hg, 3000, 8, 17)
    return BoxQP(q.n, q.hrow, q.hcol, -q.hval, -q.cost, q.lower, q.upper, offset=-7.5, maximize=True)


def tiny_curvature():
    """The box analogue of test_v2.cpp:53-56: H = diag(1e-16, 1), c = (-1e-8, 0). One
    step from 0 lands at (1e-8, 0), where the projected gradient is about 1e-8 but the
    separable bound is -0.5 and the value about 0, so the gap blocks verification."""
    return BoxQP(2, [0, 1], [0, 1], [1e-16, 1.0], [-1e-8, 0.0], [-1e12, -1], [1e12, 1])


CASES = {
    "two_var": (two_var, 3, "diagonal"),
    "diag_maximize": (diag_maximize, 4, "diagonal"),
    "planted_maximize": (None, 30, "tangent"),
    "tiny_curvature": (tiny_curvature, 1, "diagonal"),
}

FIELDS = ("objective", "primal_residual", "stationarity_residual", "dual_bound", "relative_gap")


def assert_same_check(dev, ref):
    assert dev["optimality_verified"] == ref["optimality_verified"], (dev, ref)
    assert dev["complementarity_residual"] == ref["complementarity_residual"] == 0.0
    for k in FIELDS:
        a, b = dev[k], ref[k]
        if np.isinf(b):
            assert a == b, (k, a, b)
        else:
            assert abs(a - b) <= 1e-12 * (1 + abs(b)), (k, a, b)


@pytest.mark.parametrize("case", list(CASES))
def test_check_branches_match_oracle(hg, oracle, case):
    """After a few steps and after a full solve, every field of the device check
    equals the oracle's at the same iterate: the separable bound of a positive
    diagonal Hessian, maximization with an offset (objective in the original
    sense, bound in the minimization sense), and the tiny-curvature gap rule."""
    make, steps, kind = CASES[case]
    q = make() if make else planted_maximize(hg)
    s, i, v, c, lo, up, L = engine_inputs(oracle, q)
    betas = np.zeros(steps) if case == "tiny_curvature" else oracle.betas(steps)
    with hg.Box(s, i, v, c, lo, up, L, maximize=int(q.maximize), offset=q.offset) as box:
        for b in betas:
            box.step(b)
        x = box.primal()
        dev = box.check(1e-7)
        ref = oracle.check(q, x, 1e-7)
        print(case, "after steps", dev, ref, flush=True)
        assert ref["positive_diagonal"] == (kind == "diagonal")
        assert_same_check(dev, ref)
        if case == "tiny_curvature":
            assert np.array_equal(x, [1e-8, 0.0])
            assert dev["stationarity_residual"] < 1e-7 and dev["relative_gap"] > 0.1
            assert not dev["optimality_verified"]
            return
        box.reset()
        rep = box.solve(tolerance=1e-9, iteration_limit=20000, time_limit=60)
        x = box.primal()
        dev = box.check(1e-9)
    ref = oracle.check(q, x, 1e-9)
    print(case, "after solve", rep, ref, flush=True)
    assert rep["status"] == "optimal"
    assert_same_check(dev, ref)
    if case == "two_var":
        assert np.abs(x - [1, 2]).max() < 1e-8 and abs(dev["objective"] + 5) < 1e-12
        assert abs(dev["dual_bound"] + 5) < 1e-12
    if case == "diag_maximize":
        assert np.abs(x - [2, 2.5, 10 / 3]).max() < 1e-8
        assert abs(dev["objective"] - (4 + 12.5 + 100 / 3 + 1.25)) < 1e-10  # original sense with the offset


def block_psd(n=300):
    """test_v2.cpp:18-24: 2x2 blocks [[1,2],[2,5]] (PSD, not diagonally dominant), c = -1, x in [0,10]."""
    hr, hc, hv = [], [], []
    for j in range(n):
        hr.append(j); hc.append(j); hv.append(5.0 if j % 2 else 1.0)
        if j % 2:
            hr.append(j); hc.append(j - 1); hv.append(2.0)
    return BoxQP(n, hr, hc, hv, -np.ones(n), np.zeros(n), 10 * np.ones(n))


def test_block_psd_through_the_engine(hg, reference):
    """The reference's n = 300 case (test_v2.cpp:18-24) through the whole-model entry:
    the convexity check takes the LDL^T path, and the solve equals the compiled
    reference's CPU solve bit for bit (status, iterations, objective, primal)."""
    q = block_psd()
    ref = reference.solve_box(q, time_limit=10)
    got = hg.solve_qp(q.n, q.hrow, q.hcol, q.hval, q.cost, q.lower, q.upper, time_limit=10)
    print(got["status"], got["iterations"], got["objective"], ref["iterations"], ref["objective"], flush=True)
    assert ref["status"] == got["status"] == "optimal"
    assert got["optimality_verified"] and got["algorithm"].startswith("hip_fista_box")
    assert got["iterations"] == ref["iterations"]
    assert got["objective"] == ref["objective"]
    assert got["stationarity_residual"] == ref["stationarity_residual"]
    assert np.array_equal(got["primal"], ref["primal"])
    assert abs(got["objective"] + 75) < 1e-6


def test_two_var_through_the_engine(hg, reference):
    """test_solver.cpp:85-89 as a box QP through the whole-model entry: objective -5,
    the positive-diagonal verdict (gap 0), bit for bit with the reference CPU solve."""
    q = two_var()
    ref = reference.solve_box(q, time_limit=10)
    got = hg.solve_qp(q.n, q.hrow, q.hcol, q.hval, q.cost, q.lower, q.upper, time_limit=10)
    assert ref["status"] == got["status"] == "optimal"
    assert got["iterations"] == ref["iterations"] and got["objective"] == ref["objective"]
    assert np.array_equal(got["primal"], ref["primal"])
    assert abs(got["objective"] + 5) < 1e-7 and got["relative_gap"] == ref["relative_gap"]


@pytest.mark.parametrize("mode", ["fp64_reference", "fp32_polish"])
def test_resolve_at_another_tolerance(hg, oracle, mode):
    """One handle solved at 1e-6, reset and solved at 1e-9 (and the reverse) gives
    the same report and x as a fresh handle at each tolerance: the replayed check
    periods carry the tolerance of the solve that runs them."""
    q, _ = planted_qp(hg, 10000, 8, 31)
    s, i, v, c, lo, up, L = engine_inputs(oracle, q)
    opt = dict(use_graph=1)
    sopt = dict(iteration_limit=20000, time_limit=60)
    if mode == "fp32_polish":
        opt.update(precision=hg.PRECISION_FP32_POLISH, momentum=hg.MOMENTUM_RESTART, reduce=hg.REDUCE_TREE,
                   step_metric=hg.STEP_JACOBI)
        sopt.update(fp32_switch=1e-5, check_every=10, polish_every=5)

    def fresh(tol):
        with hg.Box(s, i, v, c, lo, up, L, **opt) as box:
            rep = box.solve(tolerance=tol, **sopt)
            return rep, box.primal()

    expect = {tol: fresh(tol) for tol in (1e-6, 1e-9)}
    for order in ((1e-6, 1e-9), (1e-9, 1e-6)):
        with hg.Box(s, i, v, c, lo, up, L, **opt) as box:
            for k, tol in enumerate(order):
                if k:
                    box.reset()
                rep = box.solve(tolerance=tol, **sopt)
                x = box.primal()
                print(mode, order, tol, rep["iterations"], rep["stationarity_residual"], flush=True)
                assert rep["status"] == "optimal" and rep["stationarity_residual"] <= tol
                assert rep["iterations"] == expect[tol][0]["iterations"]
                assert np.array_equal(x, expect[tol][1])
    assert expect[1e-9][0]["iterations"] > expect[1e-6][0]["iterations"]


def convexity_corpus(seed):
    """Small models for the convexity gate: dominant, dominant up to an exact tie
    whose radius rounds differently by order, one ulp short of that tie, negative
    and zero diagonals, PSD but not dominant (LDL^T), indefinite, maximization."""
    rng = np.random.default_rng(seed)
    out = []
    vals = [0.1, 0.2, 0.3, 0.7, 1e-17, 3.3]
    for case in range(40):
        n = int(rng.integers(3, 9))
        hr, hc, hv = [], [], []
        for i in range(1, n):
            for j in range(i):
                if rng.random() < 0.6:
                    hr.append(i); hc.append(j); hv.append(float(rng.choice(vals)) * float(rng.choice([-1, 1])))
        radius = np.zeros(n)
        for i, j, v in zip(hr, hc, hv):  # term order, as the reference adds them
            radius[i] += abs(v)
            radius[j] += abs(v)
        kind = case % 8
        diag = radius.copy()                                   # an exact tie in term order
        if kind == 1:
            diag = radius + rng.uniform(0.1, 1.0, n)           # dominant
        elif kind == 2:
            diag = np.nextafter(radius, -np.inf)               # one ulp short: LDL^T decides
        elif kind == 3:
            diag[0] = -0.5                                     # negative diagonal
        elif kind == 4:
            diag[int(rng.integers(0, n))] = 0.0                # zero diagonal
        elif kind == 5:
            diag = 0.6 * radius + 1e-3                         # not dominant, maybe PSD
        elif kind == 6:
            diag = 0.2 * radius                                # likely indefinite
        maximize = kind == 7
        sign = -1.0 if maximize else 1.0
        d = list(range(n))
        q = BoxQP(n, hr + d, hc + d, [sign * v for v in hv] + [sign * x for x in diag],
                  rng.normal(0, 1, n), -np.ones(n), np.ones(n), maximize=maximize)
        out.append((case, q))
    return out


def test_device_convexity_gate_equals_reference(hg, reference):
    """hgqp_solve_qp decides curvature with the device's dominance pass (exact
    diagonal sums, order-robust radius bounds) and the host LDL^T pass where the
    device cannot decide; the outcome equals the compiled reference's
    check_convex_hessian on every model of the corpus."""
    seen = set()
    for seed in (1, 2, 3):
        for case, q in convexity_corpus(seed):
            expect = reference.check_convex(q)
            try:
                got = hg.solve_qp(q.n, q.hrow, q.hcol, q.hval, q.cost, q.lower, q.upper, maximize=q.maximize,
                                  iteration_limit=0, time_limit=10)["status"]
            except hg.Unsupported:  # a zero Hessian: rejected after the gate passed it (accelerated_qp.cpp:28)
                assert not np.any(q.hval), (seed, case)
                got = "zero_hessian"
            assert (got != "unsupported_curvature") == expect, (seed, case, got, expect)
            seen.add((case % 8, expect))
    # exact ties accepted, dominant accepted, negative diagonals rejected, and the
    # one-ulp-short rows decided by the LDL^T pass one way or the other
    assert {(0, True), (1, True), (3, False)} <= seen
    assert (2, True) in seen or (2, False) in seen
