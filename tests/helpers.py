"""Instance builders shared by the tests (inputs only; no solver logic)."""
import numpy as np

from oracle import BoxQP


def planted_qp(hg, n, draws, seed, lower=-1.0, upper=1.0):
    """Planted instance from the library's generator, as (BoxQP triplets, CSR dict)."""
    d = hg.gen_planted(n, draws, seed, lower, upper)
    hr, hc, hv = hg.lower_triplets(d["start"], d["index"], d["value"])
    q = BoxQP(n, hr, hc, hv, d["cost"], d["lower"], d["upper"])
    return q, d


def tridiagonal_qp(hg, n):
    d = hg.gen_tridiagonal(n)
    q = BoxQP(n, d["hrow"], d["hcol"], d["hval"], d["cost"], d["lower"], d["upper"])
    return q, d


def random_csr_qp(n, seed, long_rows=(), density=0.01, empty_rows=()):
    """Symmetric diagonally dominant instance with chosen long and empty rows."""
    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []
    for i in range(n):
        if i in empty_rows:
            continue
        cands = rng.choice(n, size=max(1, int(density * n)), replace=False)
        if i in long_rows:
            cands = np.arange(n)
        for j in cands:
            j = int(j)
            if j < i and j not in empty_rows:
                rows.append(i); cols.append(j); vals.append(float(rng.normal(0, 0.3)))
    rows = np.array(rows, dtype=np.int64); cols = np.array(cols, dtype=np.int64); vals = np.array(vals)
    radius = np.zeros(n)
    np.add.at(radius, rows, np.abs(vals)); np.add.at(radius, cols, np.abs(vals))
    diag_rows = np.array([i for i in range(n) if i not in empty_rows], dtype=np.int64)
    dvals = radius[diag_rows] + rng.uniform(0.1, 1.0, size=len(diag_rows))
    hr = np.concatenate([rows, diag_rows]); hc = np.concatenate([cols, diag_rows]); hv = np.concatenate([vals, dvals])
    cost = rng.normal(0, 1, size=n)
    lower = -rng.uniform(0.5, 2, size=n); upper = rng.uniform(0.5, 2, size=n)
    return BoxQP(n, hr, hc, hv, cost, lower, upper)
