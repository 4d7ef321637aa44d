"""The pool's iterates do not depend on its period length.

A pool period is `check_every` FP32 steps of every active slot, each followed
by the slots' restart and stopping decisions, then the FP64 checks of the slots
that are new or waiting. A slot's FP32 cycle ends on its own step residual and
then waits, frozen, for the period's FP64 check, and a new cycle starts from
the checked point with fresh momentum, so a slot's iterates do not depend on
where the period boundaries fall. A batch solved with periods of 2, 10 and 24
steps must agree bit for bit on every result field and the primal, over row
blocks and over slabs, and each result must meet the contract against the
oracle.
Every call goes through the C ABI of libhgqp.so.
"""
import numpy as np
import pytest

from test_gpu_pool import check_against_oracle

pytestmark = pytest.mark.gpu

KKT_TOL = 1e-6
FIELDS = ("status", "iterations", "restarts", "checks", "cycles", "objective", "primal_residual",
          "stationarity_residual", "optimality_verified")


def solve_batch(hg, batch, check_every, slabs):
    rows = sum(d["n"] + 1 for d in batch)
    nnz = sum(len(d["value"]) for d in batch)
    with hg.Pool(max_rows=rows, max_nnz=nnz, max_instances=len(batch), check_every=check_every,
                 slabs=slabs) as p:
        for k, d in enumerate(batch):
            assert p.admit(d, tag=k, tolerance=KKT_TOL) is not None
        return {r["tag"]: r for r in p.solve_all()}


@pytest.mark.parametrize("slabs", [0, 1])
def test_period_length_does_not_change_results(hg, oracle, slabs):
    batch = hg.gen_batch(12, seed=23, n_min=500, n_max=30000)
    print(f"case=pool_decisions slabs={slabs} sizes={[d['n'] for d in batch]}", flush=True)
    base = solve_batch(hg, batch, 2, slabs)
    assert sorted(base) == list(range(len(batch)))
    assert sum(r["restarts"] for r in base.values()) > 0
    for k, d in enumerate(batch):
        check_against_oracle(hg, oracle, d, base[k])
    for every in (10, 24):
        other = solve_batch(hg, batch, every, slabs)
        for k in range(len(batch)):
            for f in FIELDS:
                assert base[k][f] == other[k][f], (every, k, f, base[k][f], other[k][f])
            assert np.array_equal(base[k]["primal"], other[k]["primal"]), (every, k)
