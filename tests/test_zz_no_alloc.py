"""hgqp_check_device_noalloc (the check with no device allocation) against the
existing entry it varies, hgqp_box_check, and against the oracle's check.

The caller's arrays live in device memory: this test allocates them itself
through the HIP runtime libhgqp.so already loaded (ctypes on
libamdhip64.so.7), uploads the instance and an iterate, and reads the device
report back. Stationarity, violation and the verdict must equal
hgqp_box_check's exactly; objective, dual bound and gap (double-double sums
grouped by other workgroups) to 1e-13. Calls follow each other on one GPU, the
second and third on smaller inputs than the first, so each runs on the
scratch array the previous call left behind.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import planted_qp
from oracle import BoxQP

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the loaded HIP runtime (test plumbing)."""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.held = []

    def _ok(self, rc, what):
        assert rc == 0, f"{what} failed with HIP error {rc}"

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        self._ok(self.rt.hipMalloc(C.byref(p), max(a.nbytes, 8)), "hipMalloc")
        self.held.append(p)
        if a.nbytes:
            self._ok(self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")
        return p.value

    def empty(self, nbytes):
        p = C.c_void_p()
        self._ok(self.rt.hipMalloc(C.byref(p), nbytes), "hipMalloc")
        self.held.append(p)
        return p.value

    def read(self, ptr, nbytes):
        buf = C.create_string_buffer(nbytes)
        self._ok(self.rt.hipMemcpy(buf, ptr, nbytes, D2H), "hipMemcpy D2H")  # null-stream copy: waits
        return buf.raw

    def free(self):
        for p in self.held:
            self.rt.hipFree(p)
        self.held = []


def engine_inputs(oracle, q):
    start, index, value = oracle.build_csr(q)
    sign = -1.0 if q.maximize else 1.0
    return start, index, value, sign * q.cost, q.lower, q.upper, oracle.lipschitz(start, value)


def planted_max(hg, n, seed):
    q, _ = planted_qp(hg, n, 8, seed)
    return BoxQP(q.n, q.hrow, q.hcol, -q.hval, -q.cost, q.lower, q.upper, offset=3.25, maximize=True)


def diagonal(n, seed):
    rng = np.random.default_rng(seed)
    d = np.arange(n)
    return BoxQP(n, d, d, rng.uniform(0.5, 2.0, n), rng.normal(0, 1, n), -np.ones(n), np.ones(n), offset=-1.5)


def compare(hg, oracle, hip, q, steps, tol, diag):
    s, i, v, c, lo, up, L = engine_inputs(oracle, q)
    with hg.Box(s, i, v, c, lo, up, L, maximize=int(q.maximize), offset=q.offset) as box:
        for b in oracle.betas(steps):
            box.step(b)
        x = box.primal()
        ref = box.check(tol)
    args = [hip.upload(np.asarray(a, dtype=t)) for a, t in
            ((s, np.int32), (i, np.int32), (v, np.float64), (c, np.float64), (lo, np.float64), (up, np.float64),
             (x, np.float64))]
    rep_ptr = hip.empty(C.sizeof(hg.CheckReport))
    print(f"noalloc check: n={q.n} nnz={len(v)} maximize={q.maximize} diagonal={diag}", flush=True)
    hg.check_device_noalloc(q.n, len(v), *args, tol, rep_ptr, maximize=q.maximize, offset=q.offset,
                            diagonal_hessian=diag)
    got = hg.CheckReport.from_buffer_copy(hip.read(rep_ptr, C.sizeof(hg.CheckReport))).as_dict()
    orc = oracle.check(q, x, tol)
    print(got, ref, flush=True)
    for k in ("optimality_verified", "finite", "primal_residual", "stationarity_residual", "complementarity_residual"):
        assert got[k] == ref[k], (k, got[k], ref[k])
    for k in ("objective", "dual_bound"):
        assert abs(got[k] - ref[k]) <= 1e-13 * (1 + abs(ref[k])), (k, got[k], ref[k])
        assert abs(got[k] - orc[k]) <= 1e-12 * (1 + abs(orc[k])), (k, got[k], orc[k])
    assert abs(got["relative_gap"] - ref["relative_gap"]) <= 1e-13
    assert got["optimality_verified"] == orc["optimality_verified"]
    assert abs(got["stationarity_residual"] - orc["stationarity_residual"]) <= 1e-12 * (1 + orc["stationarity_residual"])
    return got


def test_noalloc_check_equals_box_check(hg, oracle):
    hip = Hip()
    try:
        # largest first: planted, maximize with an offset, long enough for
        # many 64-row windows; then smaller inputs over the same scratch
        a = compare(hg, oracle, hip, planted_max(hg, 300000, 5), 40, 1e-6, 0)
        b = compare(hg, oracle, hip, planted_max(hg, 3000, 6), 400, 1e-6, 0)
        assert b["optimality_verified"] and not a["optimality_verified"]
        compare(hg, oracle, hip, diagonal(777, 7), 30, 1e-7, 1)
    finally:
        hip.free()


def test_noalloc_rejects_rows_beyond_the_scratch(hg):
    with pytest.raises(hg.InvalidArgument):
        hg.check_device_noalloc(hg.NOALLOC_MAX_ROWS + 1, 10, 8, 8, 8, 8, 8, 8, 8, 1e-6, 8)
    with pytest.raises(hg.InvalidArgument):
        hg.check_device_noalloc(10, 10, 8, 8, 8, 8, 8, 8, 8, 1e-6, 8, diagonal_hessian=2)
