"""CPU checks of the C ABI library: it loads, exports what include/hgqp.h declares,
validates input before touching a device, and its host-side generators agree
with the oracle. No compute call here needs a GPU."""
import os
import re

import numpy as np
import pytest

from helpers import planted_qp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fname in os.listdir(os.path.join(ROOT, "include")):
        if fname.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fname)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(hgqp_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol(hg):
    L = hg.lib()
    names = declared_symbols()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert names <= set(hg.SIGNATURES), "binding lacks a declared symbol"


def test_struct_layouts_match_header(hg, tmp_path):
    """ctypes layouts equal the C compiler's for every struct of include/hgqp.h."""
    import ctypes as C
    import subprocess
    pairs = {"hgqp_box_options": hg.BoxOptions, "hgqp_check_report": hg.CheckReport,
             "hgqp_solve_options": hg.SolveOptions, "hgqp_box_report": hg.BoxReport, "hgqp_box_info": hg.BoxInfo,
             "hgqp_problem": hg.Problem, "hgqp_qp_options": hg.QpOptions, "hgqp_qp_result": hg.QpResult,
             "hgqp_planted": hg.Planted, "hgqp_pool_options": hg.PoolOptions, "hgqp_pool_result": hg.PoolResult,
             "hgqp_served": hg.Served}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "hgqp.h"', "int main(void){"]
    for cname, cls in pairs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                              check=True).stdout.split("\n") if line)
    for cname, cls in pairs.items():
        assert int(out[cname]) == C.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert int(out[f"{cname}.{f}"]) == getattr(cls, f).offset, (cname, f)


def test_no_device_is_reported_not_hidden(hg):
    """Without a GPU the engine refuses with HGQP_NO_DEVICE instead of computing elsewhere."""
    import ctypes as C
    if C.CDLL(hg.LIBRARY) and os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    with pytest.raises(hg.HgqpError) as e:
        hg.device_name()
    assert e.value.code == hg.NO_DEVICE
    d = hg.gen_planted(100, 4, 1)
    with pytest.raises(hg.HgqpError) as e:
        hg.Box(d["start"], d["index"], d["value"], d["cost"], d["lower"], d["upper"], 5.0)
    assert e.value.code == hg.NO_DEVICE


def test_invalid_input_rejected_before_device(hg):
    d = hg.gen_planted(100, 4, 1)
    args = [d["start"], d["index"], d["value"], d["cost"], d["lower"], d["upper"]]
    with pytest.raises(hg.HgqpError) as e:
        hg.Box(*args, 0.0)  # zero Lipschitz bound: the zero-Hessian LP path is not the GPU's
    assert e.value.code == hg.UNSUPPORTED
    bad = [a.copy() for a in args]
    bad[1][3] = 100  # column out of range
    with pytest.raises(hg.InvalidArgument):
        hg.Box(*bad, 5.0)
    bad = [a.copy() for a in args]
    bad[4][0], bad[5][0] = 1.0, 0.0  # lower > upper
    with pytest.raises(hg.InvalidArgument):
        hg.Box(*bad, 5.0)
    for opt in ({"column_panels": 17}, {"column_panels": -1}, {"step_metric": 7}, {"precision": 9}, {"slabs": 3},
                {"slabs": -1}, {"slab_groups": 3}, {"slab_groups": -1},
                {"slab_groups": 2, "reduce": hg.REDUCE_ORDERED}):
        with pytest.raises(hg.InvalidArgument):
            hg.Box(*args, 5.0, **opt)
    # the pool keeps one Jacobi metric for all its instances: no spectral scale
    for opt in ({"slabs": 2}, {"slabs": -1}, {"check_every": 3}, {"step_metric": hg.STEP_SPECTRAL}):
        with pytest.raises(hg.InvalidArgument):
            hg.Pool(**opt)
    for kw in ({"threads": 0}, {"tolerance": 0.0}, {"time_limit": -1.0}):
        with pytest.raises(hg.InvalidArgument):
            hg.solve_qp(2, [0, 1], [0, 1], [2.0, 2.0], [-2.0, -4.0], [-10, -10], [10, 10], **kw)
    with pytest.raises(hg.InvalidArgument):
        hg.solve_qp(2, [0, 1], [0, 1], [2.0, 2.0], [np.nan, -4.0], [-10, -10], [10, 10])
    with pytest.raises(hg.HgqpError) as e:  # zero Hessian (accelerated_qp.cpp:28)
        hg.solve_qp(2, [0], [0], [0.0], [-2.0, -4.0], [-10, -10], [10, 10])
    assert e.value.code == hg.UNSUPPORTED


def test_host_outcomes_without_device(hg):
    """Outcomes decided on the host before any device work: curvature (the
    convexity test runs on the host) and an all-zero Hessian."""
    r = hg.solve_qp(2, [0, 1, 1], [0, 1, 0], [1.0, 1.0, 1.5], [0, 0], [-1, -1], [1, 1])
    assert r["status"] == "unsupported_curvature" and not r["optimality_verified"]
    with pytest.raises(hg.HgqpError) as e:
        hg.solve_qp(3, [0, 2], [0, 1], [0.0, 0.0], [1.0, 0.0, 0.0], [-1, -1, -1], [1, 1, 1])
    assert e.value.code == hg.UNSUPPORTED


def test_tridiagonal_generator_equals_oracle(hg, oracle):
    for n in (2, 7, 10000):
        a = hg.gen_tridiagonal(n)
        q, opt, f = oracle.gen_tridiagonal(n)
        assert np.array_equal(a["hrow"], q.hrow) and np.array_equal(a["hcol"], q.hcol)
        for k, v in (("hval", q.hval), ("cost", q.cost), ("lower", q.lower), ("upper", q.upper), ("optimum", opt)):
            assert np.array_equal(a[k], v), k
        assert a["objective"] == f


def test_planted_generator_csr_is_reference_assembly(hg, oracle):
    """The generator's CSR is bit-identical to what the reference assembles from its triplets."""
    for n, k, seed in ((500, 8, 1), (2000, 16, 2), (300, 0, 3), (64, 63, 4)):
        q, d = planted_qp(hg, n, k, seed)
        s, i, v = oracle.build_csr(q)
        assert np.array_equal(s, d["start"]) and np.array_equal(i, d["index"]) and np.array_equal(v, d["value"])
        assert hg.lipschitz(d["start"], d["value"]) == oracle.lipschitz(s, v)
        assert oracle.check_convex(q)


def test_planted_generator_properties(hg):
    d = hg.gen_planted(20000, 8, 9)
    n, s, idx, val = d["n"], d["start"], d["index"], d["value"]
    rows = np.repeat(np.arange(n), np.diff(s))
    off = idx != rows
    radius = np.zeros(n)
    np.add.at(radius, rows[off], np.abs(val[off]))
    diag = val[~off]
    assert len(diag) == n
    margin = diag - radius
    assert margin.min() >= 0.1 - 1e-12 and margin.max() <= 1.0 + 1e-12
    # symmetric
    import scipy.sparse as sp
    H = sp.csr_matrix((val, idx, s), shape=(n, n))
    assert abs(H - H.T).max() == 0
    # planted split: about 30% / 30% / 40%
    x = d["optimum"]
    at_lo, at_up = np.mean(x == -1), np.mean(x == 1)
    assert abs(at_lo - 0.3) < 0.02 and abs(at_up - 0.3) < 0.02
    # KKT of x*: g = Hx* + c is >= 0.1 at lower, <= -0.1 at upper, ~0 inside
    g = H @ x + d["cost"]
    assert g[x == -1].min() >= 0.1 - 1e-12 and g[x == 1].max() <= -0.1 + 1e-12
    assert np.abs(g[(x > -1) & (x < 1)]).max() < 1e-12
    # same seed, same instance; other seed, other instance
    e = hg.gen_planted(20000, 8, 9)
    assert np.array_equal(e["value"], val) and e["objective"] == d["objective"]
    assert not np.array_equal(hg.gen_planted(20000, 8, 10)["value"][:100], val[:100])
