"""The instance pool with FP32 steps over per-instance column slabs.

As with row blocks, every instance of a batch must meet the contract on its
own (oracle check at KKT <= 1e-6, objective within 1e-8 of the FP64 oracle
solve, planted optimum recovered), and its iterates must not depend on what
else is resident. The batches span one-slab instances, whose rows of up to 65
entries are runs longer than a chunk, and multi-slab instances cut into
several slab blocks."""
import numpy as np
import pytest

from oracle import BoxQP

pytestmark = pytest.mark.gpu

KKT_TOL = 1e-6
OBJ_REL = 1e-8


def as_qp(hg, d):
    hr, hc, hv = hg.lower_triplets(d["start"], d["index"], d["value"])
    return BoxQP(d["n"], hr, hc, hv, d["cost"], d["lower"], d["upper"])


def budget(batch):
    return sum(d["n"] + 1 for d in batch), sum(len(d["value"]) for d in batch)


def solve_batch(hg, batch, **opts):
    rows, nnz = budget(batch)
    with hg.Pool(max_rows=rows, max_nnz=nnz, max_instances=len(batch), **opts) as p:
        for k, d in enumerate(batch):
            assert p.admit(d, tag=k, tolerance=KKT_TOL) is not None
        return {r["tag"]: r for r in p.solve_all()}


def test_slab_batch_matches_oracle(hg, oracle):
    batch = hg.gen_batch(10, seed=5, n_min=800, n_max=30000)
    print("batch n:", [d["n"] for d in batch], flush=True)
    res = solve_batch(hg, batch, slabs=1)
    assert sorted(res) == list(range(len(batch)))
    for k, d in enumerate(batch):
        r = res[k]
        assert r["status"] == "optimal" and r["optimality_verified"], r
        q = as_qp(hg, d)
        chk = oracle.check(q, r["primal"], KKT_TOL)
        assert chk["optimality_verified"] and chk["stationarity_residual"] <= KKT_TOL
        ref = oracle.solve_box(q, tolerance=1e-9, iteration_limit=20000, time_limit=120)
        assert abs(chk["objective"] - ref["objective"]) <= OBJ_REL * (1 + abs(ref["objective"]))
        assert np.abs(r["primal"] - d["optimum"]).max() < 1e-4


def test_slab_alone_equals_mix(hg):
    batch = hg.gen_batch(8, seed=13, n_min=500, n_max=40000)
    mix = solve_batch(hg, batch, slabs=1)
    for k in (0, 3, 7):
        alone = solve_batch(hg, [batch[k]], slabs=1)[0]
        assert alone["iterations"] == mix[k]["iterations"]
        assert np.array_equal(alone["primal"], mix[k]["primal"])
        assert alone["objective"] == mix[k]["objective"]


def test_slab_pool_agrees_with_row_blocks(hg):
    """Slab steps add each row in stored order, the row-block steps in a tree,
    so the FP32 iterates differ in rounding; both verify and land on the same
    optimum."""
    batch = hg.gen_batch(8, seed=17, n_min=2000, n_max=50000)
    a = solve_batch(hg, batch, slabs=1)
    b = solve_batch(hg, batch, slabs=0)
    for k, d in enumerate(batch):
        assert a[k]["optimality_verified"] and b[k]["optimality_verified"]
        assert abs(a[k]["objective"] - b[k]["objective"]) <= OBJ_REL * (1 + abs(b[k]["objective"]))
        assert np.abs(a[k]["primal"] - d["optimum"]).max() < 1e-4
        assert np.abs(b[k]["primal"] - d["optimum"]).max() < 1e-4


def test_slab_pool_reuses_slots(hg):
    """Instances admitted into slots (and slab block ranges) that earlier
    instances freed solve exactly as in a fresh pool."""
    first = hg.gen_batch(4, seed=19, n_min=3000, n_max=20000)
    second = hg.gen_batch(4, seed=23, n_min=3000, n_max=20000)
    rows, nnz = budget(first + second)
    with hg.Pool(max_rows=rows, max_nnz=nnz, max_instances=4, slabs=1) as p:
        for k, d in enumerate(first):
            assert p.admit(d, tag=k, tolerance=KKT_TOL) is not None
        p.solve_all()
        for k, d in enumerate(second):
            assert p.admit(d, tag=10 + k, tolerance=KKT_TOL) is not None
        reused = {r["tag"] - 10: r for r in p.solve_all()}
    fresh = solve_batch(hg, second, slabs=1)
    for k in range(4):
        assert reused[k]["iterations"] == fresh[k]["iterations"]
        assert np.array_equal(reused[k]["primal"], fresh[k]["primal"])


def test_slab_pool_refusals(hg):
    with pytest.raises(hg.InvalidArgument):
        hg.Server(max_rows=1000, max_nnz=10000, max_instances=4, slabs=1)
    d = hg.gen_batch(1, seed=3, n_min=500, n_max=600)[0]
    i2, v2 = d["index"].copy(), d["value"].copy()
    s = d["start"]
    for r in range(d["n"]):  # descending columns: a valid CSR the slabs cannot take
        i2[s[r]:s[r + 1]] = d["index"][s[r]:s[r + 1]][::-1]
        v2[s[r]:s[r + 1]] = d["value"][s[r]:s[r + 1]][::-1]
    bad = dict(d, index=i2, value=v2)
    with hg.Pool(max_rows=d["n"] + 1, max_nnz=len(d["value"]), max_instances=1, slabs=1) as p:
        with pytest.raises(hg.HgqpError) as e:
            p.admit(bad, tag=0, tolerance=KKT_TOL)
        assert e.value.code == hg.UNSUPPORTED
        assert p.admit(d, tag=1, tolerance=KKT_TOL) is not None
        (r,) = p.solve_all()
        assert r["optimality_verified"]


def long_run_instance(hg, oracle, lengths, seed, n=4000, band=(200, 262, 20)):
    """One-slab instance (n <= 12,288, so each row is one run) whose rows
    `lengths` (row -> stored entries) have exactly that many, from disjoint
    neighbour sets of plain rows, plus a band of consecutive rows of `band[2]`
    entries each (three-chunk runs one after another, so the packing must move
    runs to the next group of 64 chunks). Diagonally dominant, bounds [-1, 1]."""
    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []
    long_rows = dict(lengths)
    for i in range(*band[:2]):
        long_rows[i] = band[2]
    plain = [j for j in range(n) if j not in long_rows]
    rng.shuffle(plain)
    at = 0
    for i, length in long_rows.items():
        nb = plain[at:at + length - 1]  # with the diagonal: `length` entries in row i
        at += length - 1
        assert len(nb) == length - 1, "not enough plain rows for disjoint neighbours"
        for j in nb:
            rows.append(max(i, j)); cols.append(min(i, j)); vals.append(float(rng.normal(0, 0.3)))
    rows = np.array(rows, dtype=np.int64); cols = np.array(cols, dtype=np.int64); vals = np.array(vals)
    radius = np.zeros(n)
    np.add.at(radius, rows, np.abs(vals)); np.add.at(radius, cols, np.abs(vals))
    diag = np.arange(n, dtype=np.int64)
    hr = np.concatenate([rows, diag]); hc = np.concatenate([cols, diag])
    hv = np.concatenate([vals, radius + rng.uniform(0.1, 1.0, size=n)])
    q = BoxQP(n, hr, hc, hv, rng.normal(0, 1, size=n), -np.ones(n), np.ones(n))
    start, index, value = oracle.build_csr(q)
    lens = np.diff(start)
    for i, length in long_rows.items():
        assert lens[i] == length
    d = dict(n=n, start=start, index=index, value=value, cost=q.cost, lower=q.lower, upper=q.upper)
    return q, d


@pytest.mark.parametrize("case", ["chained", "unchained"])
def test_slab_pool_long_runs(hg, oracle, case):
    """Runs of 16, 24, 64, 65, 200 and 512 entries (whole multiples of a chunk
    among them) and 62 consecutive three-chunk runs, summed by the chained
    pool step (a run's chunks added across lanes); with a run of 600 entries
    (more than 64 chunks) the layout stays unchained and the one-thread follow
    sums it. Either way: verified by the oracle's check, the FP64 oracle's
    objective, the row-block pool's objective, and alone equals mix."""
    lengths = {100: 16, 101: 24, 102: 64, 103: 65, 104: 200, 105: 512}
    if case == "unchained":
        lengths[106] = 600
    q, d = long_run_instance(hg, oracle, lengths, seed=31)
    print(f"case={case} n={d['n']} nnz={len(d['value'])} longest={int(np.diff(d['start']).max())}", flush=True)
    others = hg.gen_batch(3, seed=41, n_min=800, n_max=6000)
    mix = solve_batch(hg, [d] + others, slabs=1)
    alone = solve_batch(hg, [d], slabs=1)[0]
    blocks = solve_batch(hg, [d], slabs=0)[0]
    r = mix[0]
    assert r["status"] == "optimal" and r["optimality_verified"], r
    chk = oracle.check(q, r["primal"], KKT_TOL)
    assert chk["optimality_verified"] and chk["stationarity_residual"] <= KKT_TOL
    ref = oracle.solve_box(q, tolerance=1e-9, iteration_limit=20000, time_limit=120)
    assert abs(chk["objective"] - ref["objective"]) <= OBJ_REL * (1 + abs(ref["objective"]))
    assert abs(blocks["objective"] - r["objective"]) <= OBJ_REL * (1 + abs(r["objective"]))
    assert alone["iterations"] == r["iterations"] and np.array_equal(alone["primal"], r["primal"])
