"""Dense multi-RHS engine (BF16 MFMA iterations, FP64 verification) against the oracle."""
import numpy as np
import pytest

from oracle import BoxQP

pytestmark = pytest.mark.gpu

KKT_TOL = 1e-6
OBJ_REL = 1e-8


def triplets(Q):
    r, c = np.tril_indices(Q.shape[0])
    return r.astype(np.int64), c.astype(np.int64), Q[r, c]


@pytest.mark.parametrize("n,k", [(200, 40), (384, 64)])
def test_dense_rhs_meet_contract(hg, oracle, n, k):
    Q, Cr, lo, up = hg.gen_dense(n, k, seed=n)
    with hg.Dense(Q, Cr, lo, up) as d:
        rep = d.solve(tolerance=KKT_TOL)
        x, checks, status, iters = d.results()
    print(rep, flush=True)
    assert rep["verified"] == k and all(s == "optimal" for s in status)
    lam = np.linalg.eigvalsh(Q)[-1]
    assert lam <= rep["lipschitz"] <= 1.2 * lam
    hr, hc, hv = triplets(Q)
    for j in range(k):
        q = BoxQP(n, hr, hc, hv, Cr[j], lo, up)
        chk = oracle.check(q, x[j], KKT_TOL)
        assert chk["optimality_verified"] and chk["stationarity_residual"] <= KKT_TOL, (j, chk)
        assert abs(checks[j]["objective"] - chk["objective"]) <= 1e-12 * (1 + abs(chk["objective"]))
    for j in range(0, k, max(1, k // 4)):
        q = BoxQP(n, hr, hc, hv, Cr[j], lo, up)
        ref = oracle.solve_box(q, tolerance=1e-10, iteration_limit=200000, time_limit=120)
        assert ref["status"] == "optimal"
        assert abs(checks[j]["objective"] - ref["objective"]) <= OBJ_REL * (1 + abs(ref["objective"]))
        assert np.abs(x[j] - ref["primal"]).max() < 1e-4


def test_dense_padding_and_bounds(hg, oracle):
    """n and K not multiples of the tiles, per-variable bounds, a RHS optimal at x0."""
    n, k = 150, 7
    Q, Cr, lo, up = hg.gen_dense(n, k, seed=3)
    rng = np.random.default_rng(4)
    lo = -rng.uniform(0.2, 2, n)
    up = rng.uniform(0.2, 2, n)
    Cr[0] = 0.0  # x0 = 0 is optimal: verified at iteration 0
    with hg.Dense(Q, Cr, lo, up) as d:
        d.solve(tolerance=KKT_TOL)
        x, checks, status, iters = d.results()
    assert status[0] == "optimal" and iters[0] == 0 and np.all(x[0] == 0)
    hr, hc, hv = triplets(Q)
    for j in range(k):
        chk = oracle.check(BoxQP(n, hr, hc, hv, Cr[j], lo, up), x[j], KKT_TOL)
        assert status[j] == "optimal" and chk["optimality_verified"], (j, chk)
