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


def rows_qp(n, m_eq, m_ineq, seed, row_nnz=6, free=0.1):
    """Convex QP with linear rows, feasible by construction: sparse strictly
    diagonally dominant H (lower triplets), m_eq equality rows and m_ineq
    inequality rows (two- or one-sided) through a point x0 inside the box,
    bounds [-1, 1] with a fraction `free` of the variables unbounded."""
    from oracle import Rows
    rng = np.random.default_rng(seed)
    k = 3 * n
    r = rng.integers(0, n, size=k)
    c = rng.integers(0, n, size=k)
    keep = r > c
    r, c = r[keep], c[keep]
    v = rng.normal(0, 0.1, size=len(r))
    radius = np.zeros(n)
    np.add.at(radius, r, np.abs(v))
    np.add.at(radius, c, np.abs(v))
    diag = radius + rng.uniform(0.5, 2.0, size=n)
    hr = np.concatenate([r, np.arange(n)])
    hc = np.concatenate([c, np.arange(n)])
    hv = np.concatenate([v, diag])
    lower = -np.ones(n)
    upper = np.ones(n)
    free_idx = rng.random(n) < free
    lower[free_idx] = -np.inf
    upper[free_idx] = np.inf
    x0 = rng.uniform(-0.5, 0.5, size=n)
    m = m_eq + m_ineq
    start = np.zeros(m + 1, dtype=np.int64)
    cols, vals = [], []
    for i in range(m):
        cj = np.sort(rng.choice(n, size=row_nnz, replace=False))
        cols.append(cj)
        vals.append(rng.normal(0, 1, size=row_nnz))
        start[i + 1] = start[i] + row_nnz
    col = np.concatenate(cols) if m else np.zeros(0, dtype=np.int64)
    val = np.concatenate(vals) if m else np.zeros(0)
    act = np.array([vals[i] @ x0[cols[i]] for i in range(m)])
    rl = act.copy()
    ru = act.copy()
    # inequality rows are two-sided with finite bounds: the check takes the
    # bound on the side of the dual's sign, so a finite far side keeps the
    # complementarity term finite for an inactive row's small ADMM dual
    for i in range(m_eq, m):
        rl[i] = act[i] - rng.uniform(0.05, 1.0)
        ru[i] = act[i] + rng.uniform(0.05, 1.0)
    cost = rng.normal(0, 1, size=n)
    q = BoxQP(n, hr, hc, hv, cost, lower, upper)
    return q, Rows(start, col, val, rl, ru)


# min 1/2 (x1^2 + x2^2)  s.t.  x1 + x2 = 1,  -10 <= x <= 10  (x* = (1/2, 1/2), objective 1/4)
QCL = """twovar QCL minimize
2
1
2
1 1 1.0
2 2 1.0
0 0
0
2
1 1 1.0
1 2 1.0
1e30
1 0
1 0
-10 0
10 0
0 0
0 0
0 0
0
0
"""
