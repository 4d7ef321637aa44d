"""ctypes access to the oracle and to the compiled reference (oracle/_ref).

TEST INFRASTRUCTURE ONLY. Tests, the smoke check and the benchmark's
CPU-baseline leg use this module as the checker. The product never imports it.

``Oracle("oracle")`` loads ``oracle/liboracle_boxqp.so`` (the CPU restatement in
boxqp_oracle.cpp); ``Oracle("reference")`` loads ``oracle/_ref/libref_boxqp.so``
(the reference's own sources plus ref_shim.cpp). Both expose the same calls.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBS = {
    "oracle": (os.path.join(HERE, "liboracle_boxqp.so"), "orc_"),
    "reference": (os.path.join(HERE, "_ref", "libref_boxqp.so"), "ref_"),
}


class Problem(C.Structure):
    _fields_ = [
        ("n", C.c_int64), ("nterms", C.c_int64),
        ("hrow", C.POINTER(C.c_int64)), ("hcol", C.POINTER(C.c_int64)), ("hval", C.POINTER(C.c_double)),
        ("cost", C.POINTER(C.c_double)), ("lower", C.POINTER(C.c_double)), ("upper", C.POINTER(C.c_double)),
        ("offset", C.c_double), ("maximize", C.c_int32),
    ]


class RowsStruct(C.Structure):
    _fields_ = [("m", C.c_int64), ("start", C.POINTER(C.c_int64)), ("col", C.POINTER(C.c_int64)),
                ("val", C.POINTER(C.c_double)), ("lower", C.POINTER(C.c_double)), ("upper", C.POINTER(C.c_double))]


class Result(C.Structure):
    _fields_ = [
        ("status", C.c_char * 32), ("algorithm", C.c_char * 32),
        ("objective", C.c_double), ("primal_residual", C.c_double),
        ("stationarity_residual", C.c_double), ("complementarity_residual", C.c_double),
        ("dual_bound", C.c_double), ("relative_gap", C.c_double),
        ("iterations", C.c_uint64), ("optimality_verified", C.c_int32), ("positive_diagonal", C.c_int32),
    ]

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_}
        d["status"] = self.status.decode()
        d["algorithm"] = self.algorithm.decode()
        d["optimality_verified"] = bool(self.optimality_verified)
        return d


class Options(C.Structure):
    _fields_ = [("time_limit", C.c_double), ("tolerance", C.c_double),
                ("iteration_limit", C.c_uint64), ("threads", C.c_int32)]


class QplibView(C.Structure):
    _fields_ = [("n", C.c_int64), ("m", C.c_int64), ("nterms", C.c_int64),
                ("hrow", C.POINTER(C.c_int64)), ("hcol", C.POINTER(C.c_int64)), ("hval", C.POINTER(C.c_double)),
                ("cost", C.POINTER(C.c_double)), ("lower", C.POINTER(C.c_double)), ("upper", C.POINTER(C.c_double)),
                ("rstart", C.POINTER(C.c_int64)), ("rcol", C.POINTER(C.c_int64)), ("rval", C.POINTER(C.c_double)),
                ("row_lower", C.POINTER(C.c_double)), ("row_upper", C.POINTER(C.c_double)),
                ("offset", C.c_double), ("maximize", C.c_int32)]


class OracleError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class BoxQP:
    """A box QP in the reference's triplet form (qp.hpp:6-10): one entry per
    lower-triangle position with the full value; numpy arrays are kept alive here."""

    def __init__(self, n, hrow, hcol, hval, cost, lower, upper, offset=0.0, maximize=False):
        self.n = int(n)
        self.hrow = np.ascontiguousarray(hrow, dtype=np.int64)
        self.hcol = np.ascontiguousarray(hcol, dtype=np.int64)
        self.hval = np.ascontiguousarray(hval, dtype=np.float64)
        self.cost = np.ascontiguousarray(cost, dtype=np.float64)
        self.lower = np.ascontiguousarray(lower, dtype=np.float64)
        self.upper = np.ascontiguousarray(upper, dtype=np.float64)
        self.offset = float(offset)
        self.maximize = bool(maximize)

    def struct(self):
        return Problem(self.n, len(self.hval), _ptr(self.hrow, C.c_int64), _ptr(self.hcol, C.c_int64),
                       _ptr(self.hval, C.c_double), _ptr(self.cost, C.c_double), _ptr(self.lower, C.c_double),
                       _ptr(self.upper, C.c_double), self.offset, int(self.maximize))


class Rows:
    """Linear rows of a model (mps.hpp:13-14) in CSR form: row i holds
    (col, val)[start[i]:start[i+1]], with row_lower <= a_i x <= row_upper."""

    def __init__(self, start, col, val, lower, upper):
        self.start = np.ascontiguousarray(start, dtype=np.int64)
        self.col = np.ascontiguousarray(col, dtype=np.int64)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        self.lower = np.ascontiguousarray(lower, dtype=np.float64)
        self.upper = np.ascontiguousarray(upper, dtype=np.float64)
        self.m = len(self.lower)

    def struct(self):
        return RowsStruct(self.m, _ptr(self.start, C.c_int64), _ptr(self.col, C.c_int64), _ptr(self.val, C.c_double),
                          _ptr(self.lower, C.c_double), _ptr(self.upper, C.c_double))


class Oracle:
    def __init__(self, kind="oracle"):
        path, prefix = LIBS[kind]
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (make -C oracle {'ref' if kind == 'reference' else 'all'})")
        self.kind = kind
        self.lib = C.CDLL(path)
        self.p = prefix
        f = self._fn
        f("last_error").restype = C.c_char_p
        f("check_convex").argtypes = [C.POINTER(Problem)]
        f("check").argtypes = [C.POINTER(Problem), C.POINTER(C.c_double), C.c_double, C.POINTER(Result)]
        f("build_csr").argtypes = [C.POINTER(Problem), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                   C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        f("solve_box").argtypes = [C.POINTER(Problem), C.POINTER(Options), C.POINTER(Result), C.POINTER(C.c_double)]
        f("validate").argtypes = [C.POINTER(Problem)]
        f("check_rows").argtypes = [C.POINTER(Problem), C.POINTER(RowsStruct), C.POINTER(C.c_double),
                                    C.POINTER(C.c_double), C.c_double, C.POINTER(Result)]
        if kind == "reference":
            self.lib.ref_solve_admm.argtypes = [C.POINTER(Problem), C.POINTER(RowsStruct), C.POINTER(Options),
                                                C.POINTER(Result), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        if kind == "oracle":
            self.lib.orc_lipschitz.restype = C.c_double
            self.lib.orc_lipschitz.argtypes = [C.c_int64, C.POINTER(C.c_int32), C.POINTER(C.c_double)]
            self.lib.orc_spmv.argtypes = [C.c_int64, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                          C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int]
            self.lib.orc_betas.argtypes = [C.c_int64, C.POINTER(C.c_double)]
            self.lib.orc_fista_steps.argtypes = [C.c_int64, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                                 C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                                 C.POINTER(C.c_double), C.c_double, C.POINTER(C.c_double), C.c_int64,
                                                 C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int]
            self.lib.orc_gen_tridiagonal.restype = C.c_double
            self.lib.orc_gen_tridiagonal.argtypes = [C.c_int64] + [C.POINTER(C.c_int64)] * 2 + \
                [C.POINTER(C.c_double)] * 5

    def _fn(self, name):
        return getattr(self.lib, self.p + name)

    def _raise(self, rc):
        if rc != 0:
            raise OracleError(rc, self._fn("last_error")().decode())

    def validate(self, q):
        s = q.struct()
        self._raise(self._fn("validate")(C.byref(s)))

    def check_convex(self, q):
        s = q.struct()
        v = self._fn("check_convex")(C.byref(s))
        if v < 0:
            self._raise(-v)
        return bool(v)

    def check(self, q, x, tol=1e-7):
        s = q.struct()
        x = np.ascontiguousarray(x, dtype=np.float64)
        r = Result()
        self._raise(self._fn("check")(C.byref(s), _ptr(x, C.c_double), tol, C.byref(r)))
        return r.as_dict()

    def build_csr(self, q):
        s = q.struct()
        start = np.zeros(q.n + 1, dtype=np.int32)
        cap = max(1, 2 * len(q.hval))
        index = np.zeros(cap, dtype=np.int32)
        value = np.zeros(cap, dtype=np.float64)
        nnz = C.c_int64(0)
        self._raise(self._fn("build_csr")(C.byref(s), _ptr(start, C.c_int32), _ptr(index, C.c_int32),
                                          _ptr(value, C.c_double), C.byref(nnz)))
        return start, index[:nnz.value].copy(), value[:nnz.value].copy()

    def solve_box(self, q, time_limit=30.0, tolerance=1e-7, iteration_limit=200000, threads=1):
        s = q.struct()
        o = Options(time_limit, tolerance, iteration_limit, threads)
        r = Result()
        x = np.zeros(q.n, dtype=np.float64)
        self._raise(self._fn("solve_box")(C.byref(s), C.byref(o), C.byref(r), _ptr(x, C.c_double)))
        d = r.as_dict()
        d["primal"] = x
        return d

    # -- oracle-only helpers ------------------------------------------------
    def lipschitz(self, start, value):
        return self.lib.orc_lipschitz(len(start) - 1, _ptr(start, C.c_int32), _ptr(value, C.c_double))

    def spmv(self, start, index, value, x, threads=1):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(len(start) - 1, dtype=np.float64)
        self.lib.orc_spmv(len(start) - 1, _ptr(start, C.c_int32), _ptr(index, C.c_int32), _ptr(value, C.c_double),
                          _ptr(x, C.c_double), _ptr(y, C.c_double), threads)
        return y

    def betas(self, count):
        b = np.zeros(count, dtype=np.float64)
        self.lib.orc_betas(count, _ptr(b, C.c_double))
        return b

    def fista_steps(self, start, index, value, c, lower, upper, lipschitz, betas, x, w, threads=1):
        """Applies len(betas) CPU steps to copies of (x, w); returns the new pair."""
        x = np.array(x, dtype=np.float64)
        w = np.array(w, dtype=np.float64)
        betas = np.ascontiguousarray(betas, dtype=np.float64)
        c, lower, upper = (np.ascontiguousarray(a, dtype=np.float64) for a in (c, lower, upper))
        self.lib.orc_fista_steps(len(start) - 1, _ptr(start, C.c_int32), _ptr(index, C.c_int32),
                                 _ptr(value, C.c_double), _ptr(c, C.c_double), _ptr(lower, C.c_double),
                                 _ptr(upper, C.c_double), lipschitz, _ptr(betas, C.c_double), len(betas),
                                 _ptr(x, C.c_double), _ptr(w, C.c_double), threads)
        return x, w

    def gen_tridiagonal(self, n):
        m = 2 * n - 1
        hrow = np.zeros(m, dtype=np.int64)
        hcol = np.zeros(m, dtype=np.int64)
        hval = np.zeros(m)
        cost, lower, upper, opt = (np.zeros(n) for _ in range(4))
        f = self.lib.orc_gen_tridiagonal(n, _ptr(hrow, C.c_int64), _ptr(hcol, C.c_int64), _ptr(hval, C.c_double),
                                         _ptr(cost, C.c_double), _ptr(lower, C.c_double), _ptr(upper, C.c_double),
                                         _ptr(opt, C.c_double))
        return BoxQP(n, hrow, hcol, hval, cost, lower, upper), opt, f


    def check_rows(self, q, rows, x, row_dual, tol=1e-7):
        """check_qp_solution (qp_admm.cpp:63-117) of (x, row_dual) on a model with rows."""
        s, rs = q.struct(), rows.struct()
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(row_dual, dtype=np.float64)
        r = Result()
        self._raise(self._fn("check_rows")(C.byref(s), C.byref(rs), _ptr(x, C.c_double), _ptr(y, C.c_double),
                                           tol, C.byref(r)))
        return r.as_dict()

    def read_qplib(self, text):
        """The reference's own QPLIB reader (read_qplib, qp.cpp:67-117; reference
        library only): None when it rejects the text, else the parsed model as
        the dict that hgqp.read_qplib returns."""
        if self.kind != "reference":
            raise OracleError(3, "the QPLIB reader is pinned against the compiled reference only")
        if isinstance(text, str):
            text = text.encode()
        f = self.lib.ref_read_qplib
        f.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(QplibView)]
        v = QplibView()
        if f(text, len(text), C.byref(v)) != 0:
            return None

        def arr(p, k, t):
            return np.ctypeslib.as_array(p, shape=(k,)).astype(t, copy=True) if k else np.zeros(0, t)
        k = int(v.rstart[v.m]) if v.m else 0
        return dict(n=v.n, hrow=arr(v.hrow, v.nterms, np.int64), hcol=arr(v.hcol, v.nterms, np.int64),
                    hval=arr(v.hval, v.nterms, np.float64), cost=arr(v.cost, v.n, np.float64),
                    lower=arr(v.lower, v.n, np.float64), upper=arr(v.upper, v.n, np.float64),
                    offset=v.offset, maximize=bool(v.maximize), rows=v.m,
                    rstart=arr(v.rstart, v.m + 1, np.int64), rcol=arr(v.rcol, k, np.int64),
                    rval=arr(v.rval, k, np.float64), row_lower=arr(v.row_lower, v.m, np.float64),
                    row_upper=arr(v.row_upper, v.m, np.float64))

    def solve_admm(self, q, rows, tolerance=1e-7, iteration_limit=200000, time_limit=60.0, threads=1):
        """The reference's own solve_qp_admm (reference library only)."""
        if self.kind != "reference":
            raise OracleError(3, "the ADMM path is driven through the compiled reference only")
        s, rs = q.struct(), rows.struct()
        o = Options(time_limit, tolerance, iteration_limit, threads)
        r = Result()
        x = np.zeros(q.n)
        y = np.zeros(rows.m)
        self._raise(self.lib.ref_solve_admm(C.byref(s), C.byref(rs), C.byref(o), C.byref(r), _ptr(x, C.c_double),
                                            _ptr(y, C.c_double)))
        d = r.as_dict()
        d["primal"] = x
        d["row_dual"] = y
        return d


def available(kind):
    return os.path.exists(LIBS[kind][0])


# -- numpy restatement of the FISTA step in a chosen precision -------------------

def row_abs_sums(start, value):
    """Per-row sum of |H_ik| in stored order, accumulated in FP64 one term at a
    time (the Jacobi step denominators before their 1 + 2^-20 factor)."""
    return _ordered_row_sums(start, np.abs(np.asarray(value, dtype=np.float64)), None, np.float64)


def _ordered_row_sums(start, value, gathered, dtype):
    """sum_k value[k] * gathered[k] over each row, added one rounded term at a
    time in stored column order (sparse_matrix.hpp:26-33), in `dtype`
    arithmetic; with gathered None the terms are value[k] themselves. Rows are
    processed position by position, so the work is O(nnz)."""
    start = np.asarray(start, dtype=np.int64)
    n = len(start) - 1
    lens = np.diff(start)
    order = np.argsort(-lens, kind="stable")  # rows by decreasing length
    sorted_lens = lens[order]
    s = np.zeros(n, dtype=dtype)
    maxlen = int(lens.max()) if n else 0
    for j in range(maxlen):
        count = int(np.searchsorted(-sorted_lens, -j, side="left"))  # rows with more than j entries
        rows = order[:count]
        k = start[rows] + j
        term = value[k] if gathered is None else value[k] * gathered[k]
        s[rows] = s[rows] + term
    return s


def _column_range(start, index, lo, hi):
    """The CSR restricted to entries with lo <= column < hi, stored order kept:
    (start, positions of the kept entries)."""
    keep = (index >= lo) & (index < hi)
    n = len(start) - 1
    row_of = np.repeat(np.arange(n), np.diff(start))
    counts = np.bincount(row_of[keep], minlength=n)
    sub = np.zeros(len(start), dtype=np.int64)
    np.cumsum(counts, out=sub[1:])
    return sub, np.nonzero(keep)[0]


def fista_steps_np(start, index, value, c, lower, upper, denominator, betas, x, w, dtype=np.float64, ranges=None,
                   diagonal_last=False):
    """len(betas) projected gradient steps with extrapolation
    (accelerated_qp.cpp:47-57, CPU branch) in `dtype` arithmetic:

        g    = H w                         (each row in stored column order)
        next = max(l, min(u, w - (g + c) / d))
        w    = next + beta (next - x);  x = next

    d is the scalar L (a float) or an array of per-row denominators. Matrix,
    vectors, d and each beta are first rounded to `dtype`; every product,
    sum, quotient and clip is a single IEEE operation in `dtype`, with no
    fused multiply-add. In FP64 with the scalar L this is the oracle's
    orc_fista_steps bit for bit (tests/test_oracle.py); in FP32 it is the
    device's FP32 step.

    `ranges` (column boundaries 0 = b_0 < b_1 < ... < b_G = n) restates a
    row sum taken by column ranges: each range's entries are summed in stored
    order from 0, and the G range sums are added in range order,
    g = ((p_0 + p_1) + p_2) + ...; with `diagonal_last` the diagonal entries
    are left out of the ranges and added after them, g = (...) + h_ii w_i.
    Returns the new (x, w)."""
    f = np.dtype(dtype).type
    start = np.asarray(start, dtype=np.int64)
    index = np.asarray(index, dtype=np.int64)
    val = np.asarray(value, dtype=np.float64).astype(dtype)
    c, lo, up = (np.asarray(a, dtype=np.float64).astype(dtype) for a in (c, lower, upper))
    den = np.asarray(denominator, dtype=np.float64).astype(dtype)
    x = np.asarray(x, dtype=np.float64).astype(dtype)
    w = np.asarray(w, dtype=np.float64).astype(dtype)
    parts = None
    n = len(start) - 1
    dvals = np.zeros(n, dtype=dtype)
    if ranges is not None:
        if diagonal_last:
            row_of = np.repeat(np.arange(n), np.diff(start))
            on = index == row_of
            np.add.at(dvals, row_of[on], val[on])  # one entry per row (assembled CSR)
            index = np.where(on, -1, index)  # in no range
        parts = [_column_range(start, index, ranges[h], ranges[h + 1]) for h in range(len(ranges) - 1)]
    for b in betas:
        if parts is None:
            g = _ordered_row_sums(start, val, w[index], dtype)
        else:
            g = None
            for sub, pos in parts:
                p = _ordered_row_sums(sub, val[pos], w[index[pos]], dtype)
                g = p if g is None else g + p
            if diagonal_last:
                g = g + dvals * w
        v = w - (g + c) / den
        m = np.where(v < up, v, up)
        nxt = np.where(lo < m, m, lo)
        w = nxt + f(b) * (nxt - x)
        x = nxt
    return x, w
