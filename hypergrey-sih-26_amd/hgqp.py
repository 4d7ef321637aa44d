"""Python binding of the hgqp C ABI (include/hgqp.h) through ctypes.

The library is ``libhgqp.so`` next to this file, built for gfx950 by
``make -C hypergrey-sih-26_amd``. Loading fails loudly when it is missing:
there is no other implementation behind these calls.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "libhgqp.so")

OK, INVALID_ARGUMENT, NO_DEVICE, OUT_OF_MEMORY, DEVICE_ERROR, BUSY, UNSUPPORTED = range(7)
STATUS = ["optimal", "iteration_limit", "time_limit", "numerical_error", "infeasible_bounds",
          "unsupported_curvature", "running"]
PRECISION_FP64, PRECISION_FP32_POLISH, PRECISION_FP32_REFINE = 0, 1, 2
MOMENTUM_REFERENCE, MOMENTUM_RESTART = 0, 1
REDUCE_ORDERED, REDUCE_TREE = 0, 1
STEP_SCALAR, STEP_JACOBI, STEP_SPECTRAL = 0, 1, 2
NOALLOC_MAX_ROWS = 2097152  # HGQP_NOALLOC_MAX_ROWS


class HgqpError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"hgqp error {code}: {message}")
        self.code = code


class InvalidArgument(HgqpError, ValueError):
    pass


class Unsupported(HgqpError, ValueError):
    """A model the GPU engine does not take (a zero Hessian, accelerated_qp.cpp:28)."""


class BoxOptions(C.Structure):
    _fields_ = [("precision", C.c_int32), ("momentum", C.c_int32), ("reduce", C.c_int32),
                ("use_graph", C.c_int32), ("timing", C.c_int32), ("maximize", C.c_int32),
                ("offset", C.c_double), ("diagonal_hessian", C.c_int32), ("step_metric", C.c_int32),
                ("column_panels", C.c_int32), ("slabs", C.c_int32), ("slab_groups", C.c_int32)]


class CheckReport(C.Structure):
    _fields_ = [("objective", C.c_double), ("primal_residual", C.c_double),
                ("stationarity_residual", C.c_double), ("complementarity_residual", C.c_double),
                ("dual_bound", C.c_double), ("relative_gap", C.c_double),
                ("optimality_verified", C.c_int32), ("finite", C.c_int32)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["optimality_verified"] = bool(d["optimality_verified"])
        d["finite"] = bool(d["finite"])
        return d


class SolveOptions(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("time_limit", C.c_double), ("iteration_limit", C.c_uint64),
                ("check_every", C.c_int32), ("polish_every", C.c_int32), ("fp32_switch", C.c_double)]


class BoxReport(C.Structure):
    _fields_ = [("status", C.c_int32), ("restarts", C.c_int32), ("iterations", C.c_uint64),
                ("fp32_iterations", C.c_uint64), ("check", CheckReport), ("solve_seconds", C.c_double),
                ("step_kernel_ms", C.c_double), ("step_launches", C.c_uint64),
                ("step32_kernel_ms", C.c_double), ("step32_launches", C.c_uint64), ("checks", C.c_uint64)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "check"}
        d["status"] = STATUS[self.status]
        d.update(self.check.as_dict())
        return d


class BoxInfo(C.Structure):
    _fields_ = [("n", C.c_int64), ("nnz", C.c_int64), ("row_blocks", C.c_int64), ("device_bytes", C.c_int64),
                ("lipschitz", C.c_double), ("column_panels", C.c_int32), ("column_panels32", C.c_int32),
                ("slabs32", C.c_int32), ("slab_groups", C.c_int32), ("jacobi_scale", C.c_double),
                ("narrow_rows", C.c_int32), ("reserved", C.c_int32)]


class Problem(C.Structure):
    _fields_ = [("n", C.c_int64), ("nterms", C.c_int64),
                ("hrow", C.POINTER(C.c_int64)), ("hcol", C.POINTER(C.c_int64)), ("hval", C.POINTER(C.c_double)),
                ("cost", C.POINTER(C.c_double)), ("lower", C.POINTER(C.c_double)), ("upper", C.POINTER(C.c_double)),
                ("offset", C.c_double), ("maximize", C.c_int32), ("reserved", C.c_int32)]


class Rows(C.Structure):
    _fields_ = [("m", C.c_int64), ("start", C.POINTER(C.c_int64)), ("col", C.POINTER(C.c_int64)),
                ("val", C.POINTER(C.c_double)), ("row_lower", C.POINTER(C.c_double)),
                ("row_upper", C.POINTER(C.c_double))]


class QpOptions(C.Structure):
    _fields_ = [("time_limit", C.c_double), ("tolerance", C.c_double), ("iteration_limit", C.c_uint64),
                ("threads", C.c_int32), ("precision", C.c_int32), ("momentum", C.c_int32), ("reduce", C.c_int32),
                ("step_metric", C.c_int32), ("use_graph", C.c_int32)]


class QpResult(C.Structure):
    _fields_ = [("status", C.c_char * 32), ("algorithm", C.c_char * 64), ("objective", C.c_double),
                ("primal_residual", C.c_double), ("stationarity_residual", C.c_double),
                ("complementarity_residual", C.c_double), ("dual_bound", C.c_double), ("relative_gap", C.c_double),
                ("iterations", C.c_uint64), ("optimality_verified", C.c_int32), ("restarts", C.c_int32),
                ("fp32_iterations", C.c_uint64), ("solve_seconds", C.c_double), ("inner_iterations", C.c_uint64),
                ("host_check_seconds", C.c_double), ("assembly_seconds", C.c_double)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["status"] = self.status.decode()
        d["algorithm"] = self.algorithm.decode()
        d["optimality_verified"] = bool(self.optimality_verified)
        return d


class Planted(C.Structure):
    _fields_ = [("n", C.c_int64), ("draws", C.c_int32), ("reserved", C.c_int32), ("seed", C.c_uint64),
                ("lower", C.c_double), ("upper", C.c_double)]


class PoolOptions(C.Structure):
    _fields_ = [("max_rows", C.c_int64), ("max_nnz", C.c_int64), ("max_instances", C.c_int32),
                ("check_every", C.c_int32), ("block_loads", C.c_int32), ("block_rows", C.c_int32),
                ("use_graph", C.c_int32), ("momentum", C.c_int32), ("fp32_switch", C.c_double),
                ("step_metric", C.c_int32), ("slabs", C.c_int32), ("slab_block_entries", C.c_int32)]


class PoolResult(C.Structure):
    _fields_ = [("tag", C.c_int64), ("slot", C.c_int32), ("status", C.c_int32), ("iterations", C.c_uint64),
                ("restarts", C.c_int32), ("checks", C.c_int32), ("cycles", C.c_int32), ("reserved", C.c_int32),
                ("check", CheckReport)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("check", "reserved")}
        d["status"] = STATUS[self.status]
        d.update(self.check.as_dict())
        return d


class Served(C.Structure):
    _fields_ = [("result", PoolResult), ("queued_seconds", C.c_double), ("latency_seconds", C.c_double)]

    def as_dict(self):
        d = self.result.as_dict()
        d["queued_seconds"] = self.queued_seconds
        d["latency_seconds"] = self.latency_seconds
        return d


class DenseOptions(C.Structure):
    _fields_ = [("lipschitz", C.c_double), ("power_iterations", C.c_int32), ("check_every", C.c_int32),
                ("lipschitz_margin", C.c_double), ("fp32_switch", C.c_double)]


class DenseReport(C.Structure):
    _fields_ = [("verified", C.c_int64), ("max_iterations", C.c_uint64), ("bf16_gemms", C.c_uint64),
                ("fp64_gemms", C.c_uint64), ("worst_stationarity", C.c_double), ("lipschitz", C.c_double),
                ("solve_seconds", C.c_double)]


P_I32, P_I64, P_F64 = C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_double)
_ADMIT = [C.c_int64, C.c_int64, P_I32, P_I32, P_F64, P_F64, P_F64, P_F64, C.c_double, C.c_int32, C.c_double,
          C.c_double, C.c_uint64, C.c_int64, P_F64]

# Exported symbols and their argument types; tests check that every symbol
# declared in include/hgqp.h is exported.
SIGNATURES = {
    "hgqp_last_error": ([], C.c_char_p),
    "hgqp_device_name": ([C.c_char_p, C.c_size_t], C.c_int),
    "hgqp_box_options_default": ([C.POINTER(BoxOptions)], None),
    "hgqp_solve_options_default": ([C.POINTER(SolveOptions)], None),
    "hgqp_qp_options_default": ([C.POINTER(QpOptions)], None),
    "hgqp_box_create": ([C.c_int64, C.c_int64, P_I32, P_I32, P_F64, P_F64, P_F64, P_F64, C.c_double,
                         C.POINTER(BoxOptions), C.POINTER(C.c_void_p)], C.c_int),
    "hgqp_box_step": ([C.c_void_p, C.c_double], C.c_int),
    "hgqp_box_primal": ([C.c_void_p, P_F64, C.c_int64], C.c_int),
    "hgqp_box_check": ([C.c_void_p, C.c_double, C.POINTER(CheckReport)], C.c_int),
    "hgqp_box_solve": ([C.c_void_p, C.POINTER(SolveOptions), C.POINTER(BoxReport)], C.c_int),
    "hgqp_box_reset": ([C.c_void_p], C.c_int),
    "hgqp_box_destroy": ([C.c_void_p], C.c_int),
    "hgqp_box_info_get": ([C.c_void_p, C.POINTER(BoxInfo)], C.c_int),
    "hgqp_solve_qp": ([C.POINTER(Problem), C.POINTER(QpOptions), C.POINTER(QpResult), P_F64], C.c_int),
    "hgqp_check_qp_solution": ([C.POINTER(Problem), C.POINTER(Rows), P_F64, P_F64, C.c_double,
                                C.POINTER(QpResult)], C.c_int),
    "hgqp_solve_qp_admm": ([C.POINTER(Problem), C.POINTER(Rows), C.POINTER(QpOptions), C.POINTER(QpResult),
                            P_F64, P_F64], C.c_int),
    "hgqp_csr_lipschitz": ([C.c_int64, P_I32, P_F64], C.c_double),
    "hgqp_csr_assemble": ([C.POINTER(Problem), C.c_int64, P_I32, P_I32, P_F64, P_I64, P_F64], C.c_int),
    "hgqp_dense_options_default": ([C.POINTER(DenseOptions)], None),
    "hgqp_dense_create": ([C.c_int64, C.c_int64, P_F64, P_F64, P_F64, P_F64, C.POINTER(DenseOptions),
                           C.POINTER(C.c_void_p)], C.c_int),
    "hgqp_dense_solve": ([C.c_void_p, C.c_double, C.c_uint64, C.POINTER(DenseReport)], C.c_int),
    "hgqp_dense_results": ([C.c_void_p, P_F64, C.POINTER(CheckReport), P_I32, C.POINTER(C.c_uint64)], C.c_int),
    "hgqp_dense_time_gemm": ([C.c_void_p, C.c_int32, P_F64, P_F64], C.c_int),
    "hgqp_dense_destroy": ([C.c_void_p], C.c_int),
    "hgqp_qplib_parse": ([C.c_char_p, C.c_size_t, C.POINTER(C.c_void_p)], C.c_int),
    "hgqp_qplib_problem": ([C.c_void_p, C.POINTER(Problem), P_I64], C.c_int),
    "hgqp_qplib_rows": ([C.c_void_p, C.POINTER(Rows)], C.c_int),
    "hgqp_qplib_free": ([C.c_void_p], C.c_int),
    "hgqp_pool_options_default": ([C.POINTER(PoolOptions)], None),
    "hgqp_pool_bytes": ([C.POINTER(PoolOptions), P_I64], C.c_int),
    "hgqp_pool_create": ([C.POINTER(PoolOptions), C.POINTER(C.c_void_p)], C.c_int),
    "hgqp_pool_admit": ([C.c_void_p] + _ADMIT + [P_I32], C.c_int),
    "hgqp_pool_run": ([C.c_void_p, C.c_int32], C.c_int),
    "hgqp_pool_collect": ([C.c_void_p, C.POINTER(PoolResult), C.c_int32, P_I32, P_I32], C.c_int),
    "hgqp_pool_autotune": ([C.c_void_p, P_I32, P_I32], C.c_int),
    "hgqp_pool_time_step": ([C.c_void_p, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)], C.c_int),
    "hgqp_pool_destroy": ([C.c_void_p], C.c_int),
    "hgqp_server_start": ([C.POINTER(PoolOptions), C.POINTER(C.c_void_p)], C.c_int),
    "hgqp_server_submit": ([C.c_void_p] + _ADMIT[0:], C.c_int),
    "hgqp_server_poll": ([C.c_void_p, C.POINTER(Served), C.c_int32, P_I32], C.c_int),
    "hgqp_server_submit_problem": ([C.c_void_p, C.POINTER(Problem), C.c_double, C.c_uint64, C.c_int64, P_F64,
                                    P_I32], C.c_int),
    "hgqp_server_pending": ([C.c_void_p, P_I64], C.c_int),
    "hgqp_server_stop": ([C.c_void_p], C.c_int),
    "hgqp_gen_tridiagonal": ([C.c_int64, P_I64, P_I64, P_F64, P_F64, P_F64, P_F64, P_F64, P_F64], C.c_int),
    "hgqp_gen_planted": ([C.POINTER(Planted), P_I64, P_I32, P_I32, P_F64, P_F64, P_F64, P_F64, P_F64, P_F64],
                         C.c_int),
    # device pointers: start, index, value, cost, lower, upper, x; then tolerance, maximize, offset,
    # diagonal_hessian, stream, device report
    "hgqp_check_device_noalloc": ([C.c_int64, C.c_int64] + [C.c_void_p] * 7 +
                                  [C.c_double, C.c_int32, C.c_double, C.c_int32, C.c_void_p, C.c_void_p], C.c_int),
}

_lib = None


def lib():
    """Loads libhgqp.so once; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY):
            raise ImportError(f"{LIBRARY} is not built; run `make -C {HERE}` (hipcc, gfx950)")
        L = C.CDLL(LIBRARY)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _raise(rc):
    if rc != OK:
        msg = lib().hgqp_last_error().decode()
        if rc == INVALID_ARGUMENT:
            raise InvalidArgument(rc, msg)
        if rc == UNSUPPORTED:
            raise Unsupported(rc, msg)
        raise HgqpError(rc, msg)


def device_name():
    buf = C.create_string_buffer(256)
    _raise(lib().hgqp_device_name(buf, 256))
    return buf.value.decode()


def box_options(**kw):
    o = BoxOptions()
    lib().hgqp_box_options_default(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def solve_options(**kw):
    o = SolveOptions()
    lib().hgqp_solve_options_default(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Box:
    """A resident box engine (hgqp_box_*). H is the mirrored, sign-normalized CSR."""

    def __init__(self, start, index, value, cost, lower, upper, lipschitz, **options):
        self._keep = [np.ascontiguousarray(start, dtype=np.int32), np.ascontiguousarray(index, dtype=np.int32),
                      np.ascontiguousarray(value, dtype=np.float64), np.ascontiguousarray(cost, dtype=np.float64),
                      np.ascontiguousarray(lower, dtype=np.float64), np.ascontiguousarray(upper, dtype=np.float64)]
        s, i, v, c, l, u = self._keep
        self.n = len(c)
        self.handle = C.c_void_p()
        o = box_options(**options)
        _raise(lib().hgqp_box_create(self.n, len(v), _ptr(s, C.c_int32), _ptr(i, C.c_int32), _ptr(v, C.c_double),
                                     _ptr(c, C.c_double), _ptr(l, C.c_double), _ptr(u, C.c_double), float(lipschitz),
                                     C.byref(o), C.byref(self.handle)))
        self._keep = None  # the library copied everything

    def step(self, momentum):
        _raise(lib().hgqp_box_step(self.handle, float(momentum)))

    def primal(self):
        x = np.empty(self.n, dtype=np.float64)
        _raise(lib().hgqp_box_primal(self.handle, _ptr(x, C.c_double), self.n))
        return x

    def check(self, tolerance=1e-7):
        r = CheckReport()
        _raise(lib().hgqp_box_check(self.handle, tolerance, C.byref(r)))
        return r.as_dict()

    def solve(self, **kw):
        o = solve_options(**kw)
        r = BoxReport()
        _raise(lib().hgqp_box_solve(self.handle, C.byref(o), C.byref(r)))
        return r.as_dict()

    def reset(self):
        _raise(lib().hgqp_box_reset(self.handle))

    def info(self):
        i = BoxInfo()
        _raise(lib().hgqp_box_info_get(self.handle, C.byref(i)))
        return {k: getattr(i, k) for k, _ in i._fields_}

    def close(self):
        if self.handle:
            _raise(lib().hgqp_box_destroy(self.handle))
            self.handle = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def qp_options(**kw):
    o = QpOptions()
    lib().hgqp_qp_options_default(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _problem(n, hrow, hcol, hval, cost, lower, upper, offset, maximize):
    """hgqp_problem over contiguous copies; returns (view, arrays kept alive)."""
    arrs = [np.ascontiguousarray(hrow, dtype=np.int64), np.ascontiguousarray(hcol, dtype=np.int64),
            np.ascontiguousarray(hval, dtype=np.float64), np.ascontiguousarray(cost, dtype=np.float64),
            np.ascontiguousarray(lower, dtype=np.float64), np.ascontiguousarray(upper, dtype=np.float64)]
    hr, hc, hv, c, l, u = arrs
    if not (len(hr) == len(hc) == len(hv)) or not (len(c) == len(l) == len(u) == int(n)):
        raise InvalidArgument(INVALID_ARGUMENT, "problem arrays disagree in length")
    p = Problem(int(n), len(hv), _ptr(hr, C.c_int64), _ptr(hc, C.c_int64), _ptr(hv, C.c_double),
                _ptr(c, C.c_double), _ptr(l, C.c_double), _ptr(u, C.c_double), float(offset), int(bool(maximize)), 0)
    return p, arrs


def solve_qp(n, hrow, hcol, hval, cost, lower, upper, offset=0.0, maximize=False, **options):
    """hgqp_solve_qp: the whole-model entry (solve_accelerated_box with the HIP engine)."""
    p, keep = _problem(n, hrow, hcol, hval, cost, lower, upper, offset, maximize)
    o = qp_options(**options)
    r = QpResult()
    x = np.zeros(int(n), dtype=np.float64)
    _raise(lib().hgqp_solve_qp(C.byref(p), C.byref(o), C.byref(r), _ptr(x, C.c_double)))
    d = r.as_dict()
    d["primal"] = x
    return d


def solve_qp_admm(n, hrow, hcol, hval, cost, lower, upper, rstart, rcol, rval, row_lower, row_upper, offset=0.0,
                  maximize=False, **options):
    """hgqp_solve_qp_admm: a convex QP with linear rows (row i: entries
    rstart[i]:rstart[i+1] of (rcol, rval), row_lower <= a_i x <= row_upper),
    the reference's scaled ADMM with its iterations on the GPU."""
    p, keep = _problem(n, hrow, hcol, hval, cost, lower, upper, offset, maximize)
    rows, keep_rows = _rows(rstart, rcol, rval, row_lower, row_upper)
    m = rows.m
    o = qp_options(**options)
    r = QpResult()
    x = np.zeros(int(n), dtype=np.float64)
    y = np.zeros(max(m, 1), dtype=np.float64)
    _raise(lib().hgqp_solve_qp_admm(C.byref(p), C.byref(rows), C.byref(o), C.byref(r), _ptr(x, C.c_double),
                                    _ptr(y, C.c_double)))
    d = r.as_dict()
    d["primal"] = x
    d["row_dual"] = y[:m].copy()
    return d


def _rows(rstart, rcol, rval, row_lower, row_upper):
    rs = np.ascontiguousarray(rstart, dtype=np.int64)
    rc = np.ascontiguousarray(rcol, dtype=np.int64)
    rv = np.ascontiguousarray(rval, dtype=np.float64)
    rl = np.ascontiguousarray(row_lower, dtype=np.float64)
    ru = np.ascontiguousarray(row_upper, dtype=np.float64)
    m = len(rl)
    if len(ru) != m or len(rs) != m + 1:
        raise InvalidArgument(INVALID_ARGUMENT, "row arrays disagree in length")
    return Rows(m, _ptr(rs, C.c_int64), _ptr(rc, C.c_int64), _ptr(rv, C.c_double), _ptr(rl, C.c_double),
                _ptr(ru, C.c_double)), (rs, rc, rv, rl, ru)


def check_qp_solution(n, hrow, hcol, hval, cost, lower, upper, rstart, rcol, rval, row_lower, row_upper, x, row_dual,
                      tolerance=1e-7, offset=0.0, maximize=False):
    """hgqp_check_qp_solution: the reference's verdict on (x, row duals), on the host."""
    p, keep = _problem(n, hrow, hcol, hval, cost, lower, upper, offset, maximize)
    rows, keep_rows = _rows(rstart, rcol, rval, row_lower, row_upper)
    xv = np.ascontiguousarray(x, dtype=np.float64)
    yv = np.ascontiguousarray(row_dual, dtype=np.float64)
    if len(xv) != int(n) or len(yv) != rows.m:
        raise InvalidArgument(INVALID_ARGUMENT, "x and row duals must have lengths n and m")
    r = QpResult()
    _raise(lib().hgqp_check_qp_solution(C.byref(p), C.byref(rows), _ptr(xv, C.c_double),
                                        _ptr(yv, C.c_double) if rows.m else None, float(tolerance), C.byref(r)))
    return r.as_dict()


def csr_assemble(n, hrow, hcol, hval, cost, lower, upper, maximize=False):
    """hgqp_csr_assemble: the sign-normalized mirrored CSR and L, assembled on the device."""
    p, keep = _problem(n, hrow, hcol, hval, cost, lower, upper, 0.0, maximize)
    nnz, L = C.c_int64(), C.c_double()
    _raise(lib().hgqp_csr_assemble(C.byref(p), 0, None, None, None, C.byref(nnz), C.byref(L)))
    start = np.zeros(int(n) + 1, np.int32)
    index = np.zeros(max(nnz.value, 1), np.int32)
    value = np.zeros(max(nnz.value, 1))
    _raise(lib().hgqp_csr_assemble(C.byref(p), len(index), _ptr(start, C.c_int32), _ptr(index, C.c_int32),
                                   _ptr(value, C.c_double), C.byref(nnz), C.byref(L)))
    return start, index[:nnz.value].copy(), value[:nnz.value].copy(), L.value


def read_qplib(text):
    """Parses QPLIB text (hgqp_qplib_parse) into a dict of numpy arrays in the
    triplet form of solve_qp, plus 'rows' (the number of linear rows; the box
    engine needs 0) and the rows as CSR (rstart, rcol, rval, row_lower,
    row_upper), the form of solve_qp_admm."""
    if isinstance(text, str):
        text = text.encode()
    h = C.c_void_p()
    _raise(lib().hgqp_qplib_parse(text, len(text), C.byref(h)))
    try:
        view, rows = Problem(), C.c_int64()
        _raise(lib().hgqp_qplib_problem(h, C.byref(view), C.byref(rows)))
        n, m = view.n, view.nterms
        arr = lambda p, k, t: np.ctypeslib.as_array(p, shape=(k,)).astype(t, copy=True) if k else np.zeros(0, t)
        out = dict(n=n, hrow=arr(view.hrow, m, np.int64), hcol=arr(view.hcol, m, np.int64),
                   hval=arr(view.hval, m, np.float64), cost=arr(view.cost, n, np.float64),
                   lower=arr(view.lower, n, np.float64), upper=arr(view.upper, n, np.float64),
                   offset=view.offset, maximize=bool(view.maximize), rows=rows.value)
        rv = Rows()
        _raise(lib().hgqp_qplib_rows(h, C.byref(rv)))
        k = int(rv.start[rv.m]) if rv.m else 0
        out.update(rstart=arr(rv.start, rv.m + 1, np.int64), rcol=arr(rv.col, k, np.int64),
                   rval=arr(rv.val, k, np.float64), row_lower=arr(rv.row_lower, rv.m, np.float64),
                   row_upper=arr(rv.row_upper, rv.m, np.float64))
        return out
    finally:
        lib().hgqp_qplib_free(h)


def gen_tridiagonal(n):
    """The reference's seeded tridiagonal family; returns (triplets..., optimum, objective)."""
    m = 2 * n - 1
    hrow, hcol = np.zeros(m, np.int64), np.zeros(m, np.int64)
    hval = np.zeros(m)
    cost, lower, upper, opt = (np.zeros(n) for _ in range(4))
    f = C.c_double()
    _raise(lib().hgqp_gen_tridiagonal(n, _ptr(hrow, C.c_int64), _ptr(hcol, C.c_int64), _ptr(hval, C.c_double),
                                      _ptr(cost, C.c_double), _ptr(lower, C.c_double), _ptr(upper, C.c_double),
                                      _ptr(opt, C.c_double), C.byref(f)))
    return dict(n=n, hrow=hrow, hcol=hcol, hval=hval, cost=cost, lower=lower, upper=upper, optimum=opt,
                objective=f.value)


def gen_planted(n, draws, seed, lower=-1.0, upper=1.0):
    """Planted sparse box QP (full CSR plus the lower-triangle triplets)."""
    cap = n + 2 * n * draws
    start = np.zeros(n + 1, np.int32)
    index = np.zeros(cap, np.int32)
    value = np.zeros(cap)
    cost, lo, up, opt = (np.zeros(n) for _ in range(4))
    nnz, f = C.c_int64(), C.c_double()
    spec = Planted(n, draws, 0, seed, lower, upper)
    _raise(lib().hgqp_gen_planted(C.byref(spec), C.byref(nnz), _ptr(start, C.c_int32), _ptr(index, C.c_int32),
                                  _ptr(value, C.c_double), _ptr(cost, C.c_double), _ptr(lo, C.c_double),
                                  _ptr(up, C.c_double), _ptr(opt, C.c_double), C.byref(f)))
    index = index[:nnz.value].copy()
    value = value[:nnz.value].copy()
    return dict(n=n, start=start, index=index, value=value, cost=cost, lower=lo, upper=up, optimum=opt,
                objective=f.value)


def gen_batch(count, seed, n_min=1000, n_max=100000, draws_min=4, draws_max=32, lower=(-2.0, -0.5),
              upper=(0.5, 2.0)):
    """Uneven planted batch: n log-uniform in [n_min, n_max], draws per row uniform
    in {draws_min..draws_max}, per-instance scalar bounds lower ~ U(lower),
    upper ~ U(upper); instance k uses generator seed (seed, k)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(round(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
        draws = int(rng.integers(draws_min, draws_max + 1))
        lo = float(rng.uniform(*lower))
        up = float(rng.uniform(*upper))
        d = gen_planted(n, min(draws, n - 1), (int(seed) << 20) + k, lo, up)
        d["draws"] = draws
        out.append(d)
    return out


def check_device_noalloc(n, nnz, start, index, value, cost, lower, upper, x, tolerance, report, maximize=False,
                         offset=0.0, diagonal_hessian=0, stream=None):
    """hgqp_check_device_noalloc: every array argument and `report` are device
    addresses (ints); the report is written there in stream order."""
    _raise(lib().hgqp_check_device_noalloc(int(n), int(nnz), start, index, value, cost, lower, upper, x,
                                           float(tolerance), int(bool(maximize)), float(offset),
                                           int(diagonal_hessian), stream, report))


def lower_triplets(start, index, value):
    """Lower-triangle triplets (row, col, value), col <= row, of a symmetric CSR."""
    n = len(start) - 1
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(start))
    keep = index <= rows
    return rows[keep], index[keep].astype(np.int64), value[keep]


def lipschitz(start, value):
    """hgqp_csr_lipschitz: max_i sum_k |H_ik| summed in stored order (accelerated_qp.cpp:25-26)."""
    start = np.ascontiguousarray(start, dtype=np.int32)
    value = np.ascontiguousarray(value, dtype=np.float64)
    return lib().hgqp_csr_lipschitz(len(start) - 1, _ptr(start, C.c_int32), _ptr(value, C.c_double))


def pool_options(**kw):
    o = PoolOptions()
    lib().hgqp_pool_options_default(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def pool_bytes(**kw):
    b = C.c_int64()
    _raise(lib().hgqp_pool_bytes(C.byref(pool_options(**kw)), C.byref(b)))
    return b.value


def _instance_args(inst, lipschitz_value, maximize, offset, tolerance, iteration_limit, tag, x_out):
    """Arguments of hgqp_pool_admit / hgqp_server_submit for an instance dict with
    start, index, value (sign-normalized mirrored CSR), cost, lower, upper."""
    keep = [np.ascontiguousarray(inst["start"], dtype=np.int32), np.ascontiguousarray(inst["index"], dtype=np.int32),
            np.ascontiguousarray(inst["value"], dtype=np.float64), np.ascontiguousarray(inst["cost"], dtype=np.float64),
            np.ascontiguousarray(inst["lower"], dtype=np.float64), np.ascontiguousarray(inst["upper"], dtype=np.float64)]
    s, i, v, c, lo, up = keep
    n = len(c)
    args = [n, len(v), _ptr(s, C.c_int32), _ptr(i, C.c_int32), _ptr(v, C.c_double), _ptr(c, C.c_double),
            _ptr(lo, C.c_double), _ptr(up, C.c_double), float(lipschitz_value), int(bool(maximize)), float(offset),
            float(tolerance), int(iteration_limit), int(tag), _ptr(x_out, C.c_double)]
    return args, keep


class Pool:
    """An instance pool (hgqp_pool_*): many box QPs advanced together."""

    def __init__(self, **options):
        self.handle = C.c_void_p()
        _raise(lib().hgqp_pool_create(C.byref(pool_options(**options)), C.byref(self.handle)))
        self.out = {}  # tag -> x buffer (kept alive until collected)

    def admit(self, inst, tag, tolerance=1e-6, iteration_limit=20000, lipschitz_value=None, maximize=False,
              offset=0.0):
        """Returns the slot, or None when the pool is full (HGQP_BUSY)."""
        L = lipschitz_value if lipschitz_value is not None else lipschitz(inst["start"], inst["value"])
        x = np.zeros(len(inst["cost"]), dtype=np.float64)
        args, keep = _instance_args(inst, L, maximize, offset, tolerance, iteration_limit, tag, x)
        slot = C.c_int32(-1)
        rc = lib().hgqp_pool_admit(self.handle, *args, C.byref(slot))
        if rc == BUSY:
            return None
        _raise(rc)
        self.out[int(tag)] = x
        return slot.value

    def run(self, periods=1):
        _raise(lib().hgqp_pool_run(self.handle, int(periods)))

    def collect(self):
        """Finished instances as dicts with 'primal'; also returns the resident count."""
        out = []
        buf = (PoolResult * 256)()
        while True:
            cnt, active = C.c_int32(), C.c_int32()
            _raise(lib().hgqp_pool_collect(self.handle, buf, 256, C.byref(cnt), C.byref(active)))
            for k in range(cnt.value):
                d = buf[k].as_dict()
                d["primal"] = self.out.pop(d["tag"])
                out.append(d)
            if cnt.value < 256:
                return out, active.value

    def autotune(self):
        l, r = C.c_int32(), C.c_int32()
        _raise(lib().hgqp_pool_autotune(self.handle, C.byref(l), C.byref(r)))
        return l.value, r.value

    def time_step(self, reps=6):
        """Average FP32 step time (us) with every resident instance active, and
        the step's algorithmic bytes."""
        us, nbytes = C.c_double(), C.c_double()
        _raise(lib().hgqp_pool_time_step(self.handle, int(reps), C.byref(us), C.byref(nbytes)))
        return us.value, nbytes.value

    def solve_all(self, max_periods=100000):
        """Runs periods until nothing is resident; returns all results."""
        results = []
        for _ in range(max_periods):
            self.run(1)
            got, active = self.collect()
            results += got
            if active == 0:
                return results
        raise RuntimeError("pool did not finish")

    def close(self):
        if self.handle:
            _raise(lib().hgqp_pool_destroy(self.handle))
            self.handle = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """The serving loop (hgqp_server_*): submit from any thread, poll results."""

    def __init__(self, **options):
        self.handle = C.c_void_p()
        _raise(lib().hgqp_server_start(C.byref(pool_options(**options)), C.byref(self.handle)))
        self.out = {}

    def submit(self, inst, tag, tolerance=1e-6, iteration_limit=20000, lipschitz_value=None, maximize=False,
               offset=0.0):
        L = lipschitz_value if lipschitz_value is not None else lipschitz(inst["start"], inst["value"])
        x = np.zeros(len(inst["cost"]), dtype=np.float64)
        args, keep = _instance_args(inst, L, maximize, offset, tolerance, iteration_limit, tag, x)
        self.out[int(tag)] = x
        _raise(lib().hgqp_server_submit(self.handle, *args))

    def submit_problem(self, n, hrow, hcol, hval, cost, lower, upper, tag, offset=0.0, maximize=False,
                       tolerance=1e-6, iteration_limit=20000):
        """hgqp_server_submit_problem: a model in triplet form. Returns the
        status name: "running" when submitted (the result comes from poll),
        else the status the model ended with before any iteration."""
        p, keep = _problem(n, hrow, hcol, hval, cost, lower, upper, offset, maximize)
        x = np.zeros(int(n), dtype=np.float64)
        st = C.c_int32()
        self.out[int(tag)] = x
        try:
            _raise(lib().hgqp_server_submit_problem(self.handle, C.byref(p), float(tolerance), int(iteration_limit),
                                                    int(tag), _ptr(x, C.c_double), C.byref(st)))
        except Exception:
            self.out.pop(int(tag), None)
            raise
        if STATUS[st.value] != "running":
            self.out.pop(int(tag), None)
        return STATUS[st.value]

    def poll(self):
        out = []
        buf = (Served * 256)()
        while True:
            cnt = C.c_int32()
            _raise(lib().hgqp_server_poll(self.handle, buf, 256, C.byref(cnt)))
            for k in range(cnt.value):
                d = buf[k].as_dict()
                d["primal"] = self.out.pop(d["tag"])
                out.append(d)
            if cnt.value < 256:
                return out

    def pending(self):
        p = C.c_int64()
        _raise(lib().hgqp_server_pending(self.handle, C.byref(p)))
        return p.value

    def stop(self):
        if self.handle:
            _raise(lib().hgqp_server_stop(self.handle))
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass


class Dense:
    """Dense multi-RHS engine (hgqp_dense_*): one dense Q (n x n), K cost vectors (K x n)."""

    def __init__(self, Q, C_rhs, lower, upper, **options):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        C_rhs = np.ascontiguousarray(C_rhs, dtype=np.float64)
        lower = np.ascontiguousarray(lower, dtype=np.float64)
        upper = np.ascontiguousarray(upper, dtype=np.float64)
        self.n, self.k = Q.shape[0], C_rhs.shape[0]
        o = DenseOptions()
        lib().hgqp_dense_options_default(C.byref(o))
        for key, val in options.items():
            setattr(o, key, val)
        self.handle = C.c_void_p()
        _raise(lib().hgqp_dense_create(self.n, self.k, _ptr(Q, C.c_double), _ptr(C_rhs, C.c_double),
                                       _ptr(lower, C.c_double), _ptr(upper, C.c_double), C.byref(o),
                                       C.byref(self.handle)))

    def solve(self, tolerance=1e-6, iteration_limit=20000):
        r = DenseReport()
        _raise(lib().hgqp_dense_solve(self.handle, tolerance, iteration_limit, C.byref(r)))
        return {k: getattr(r, k) for k, _ in r._fields_}

    def results(self):
        x = np.zeros((self.k, self.n))
        checks = (CheckReport * self.k)()
        status = np.zeros(self.k, dtype=np.int32)
        iters = np.zeros(self.k, dtype=np.uint64)
        _raise(lib().hgqp_dense_results(self.handle, _ptr(x, C.c_double), checks, _ptr(status, C.c_int32),
                                        iters.ctypes.data_as(C.POINTER(C.c_uint64))))
        return x, [c.as_dict() for c in checks], [STATUS[s] for s in status], iters

    def time_gemm(self, reps=50):
        us, tf = C.c_double(), C.c_double()
        _raise(lib().hgqp_dense_time_gemm(self.handle, reps, C.byref(us), C.byref(tf)))
        return us.value, tf.value

    def close(self):
        if self.handle:
            _raise(lib().hgqp_dense_destroy(self.handle))
            self.handle = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def gen_dense(n, k, seed):
    """Q = M'M/n + 0.1 I with M an n x n N(0,1) matrix, K cost vectors N(0,1), bounds [-1, 1]."""
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((n, n))
    Q = M.T @ M / n + 0.1 * np.eye(n)
    Q = 0.5 * (Q + Q.T)
    C_rhs = rng.standard_normal((k, n))
    return Q, C_rhs, -np.ones(n), np.ones(n)
