"""Device CSR assembly (hgqp_csr_assemble, the path of hgqp_solve_qp and
hgqp_server_submit_problem) against the oracle's restatement of the reference
assembly (accelerated_qp.cpp:20-24 with sparse_matrix.hpp:10-17), bit for bit:
duplicates summed in term order, exact zeros and cancellations dropped,
upper-triangle terms mirrored, the maximization sign, empty and long rows.
The oracle's assembly is itself pinned bitwise to the compiled reference
(tests/test_oracle.py)."""
import numpy as np
import pytest

from helpers import planted_qp, random_csr_qp, tridiagonal_qp
from oracle import BoxQP

pytestmark = pytest.mark.gpu


def awkward_qp(n, seed, maximize):
    """Terms in shuffled order over both triangles, with duplicates, exact
    cancellations, zeros and a wide dynamic range."""
    rng = np.random.default_rng(seed)
    k = 6 * n
    r = rng.integers(0, n, size=k)
    c = rng.integers(0, n, size=k)
    v = rng.normal(0, 1, size=k) * 10.0 ** rng.uniform(-3, 3, size=k)
    dup = rng.choice(k, size=k // 5)
    r, c, v = np.concatenate([r, r[dup]]), np.concatenate([c, c[dup]]), np.concatenate([v, rng.normal(size=len(dup))])
    can = rng.choice(len(r), size=k // 10, replace=False)
    r, c, v = np.concatenate([r, c[can]]), np.concatenate([c, r[can]]), np.concatenate([v, -v[can]])
    z = rng.integers(0, n, size=50)
    r, c, v = np.concatenate([r, z]), np.concatenate([c, z[::-1]]), np.concatenate([v, np.zeros(50)])
    long_row = int(rng.integers(0, n))
    r = np.concatenate([r, np.full(n, long_row)])
    c = np.concatenate([c, np.arange(n)])
    v = np.concatenate([v, rng.normal(size=n)])
    perm = rng.permutation(len(r))
    r, c, v = r[perm], c[perm], v[perm]
    keep = (r != 7) & (c != 7)  # an empty row and column
    r, c, v = r[keep], c[keep], v[keep]
    cost = rng.normal(size=n)
    return BoxQP(n, r, c, v, cost, -np.ones(n), np.ones(n), maximize=maximize)


def cases(hg):
    return {
        "awkward_min": awkward_qp(3000, 1, False),
        "awkward_max": awkward_qp(2500, 2, True),
        "planted": planted_qp(hg, 20000, 16, 5)[0],
        "long_and_empty_rows": random_csr_qp(1500, 5, long_rows=(3, 700, 1499), density=0.004, empty_rows=(10, 11)),
        "tridiagonal": tridiagonal_qp(hg, 5000)[0],
    }


@pytest.mark.parametrize("case", ["awkward_min", "awkward_max", "planted", "long_and_empty_rows", "tridiagonal"])
def test_device_csr_is_reference_assembly(hg, oracle, case):
    q = cases(hg)[case]
    print(f"case={case} n={q.n} terms={len(q.hval)}", flush=True)
    s, i, v, L = hg.csr_assemble(q.n, q.hrow, q.hcol, q.hval, q.cost, q.lower, q.upper, maximize=q.maximize)
    s0, i0, v0 = oracle.build_csr(q)
    assert np.array_equal(s, s0)
    assert np.array_equal(i, i0)
    assert np.array_equal(v, v0)
    assert L == oracle.lipschitz(s0, v0)


def test_full_size_assembly_equals_generator_csr(hg):
    """n = 1e6, 16 draws (33M stored entries): the device assembly of the lower
    triplets equals the generator's CSR, which the CPU tests pin to the
    reference assembly."""
    d = hg.gen_planted(1000000, 16, 2026)
    hr, hc, hv = hg.lower_triplets(d["start"], d["index"], d["value"])
    print(f"terms={len(hv)}", flush=True)
    s, i, v, L = hg.csr_assemble(1000000, hr, hc, hv, d["cost"], d["lower"], d["upper"])
    assert np.array_equal(s, d["start"]) and np.array_equal(i, d["index"]) and np.array_equal(v, d["value"])
    assert L == hg.lipschitz(d["start"], d["value"])


def test_whole_model_outcomes_decided_on_device(hg):
    # infeasible bounds come after the assembly (accelerated_qp.cpp:31)
    r = hg.solve_qp(2, [0, 1], [0, 1], [2.0, 2.0], [-2.0, -4.0], [-10, 3], [10, 2])
    assert r["status"] == "infeasible_bounds"
    # duplicates that cancel to a zero Hessian: L = 0 after the assembly (accelerated_qp.cpp:28)
    with pytest.raises(hg.HgqpError) as e:
        hg.solve_qp(2, [0, 0, 1, 1], [0, 0, 1, 1], [1.5, -1.5, 2.0, -2.0], [1.0, 1.0], [-1, -1], [1, 1])
    assert e.value.code == hg.UNSUPPORTED
    with pytest.raises(hg.InvalidArgument):
        hg.csr_assemble(3, [0, 5], [0, 1], [1.0, 1.0], [0.0] * 3, [-1.0] * 3, [1.0] * 3)
