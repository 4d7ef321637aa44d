"""QPLIB input and the CLI contract.

CPU: the reader parses the real box QPs QPLIB 8792 and 8991 (public QPLIB data
kept in tests/golden), and the oracle solving the parsed models reproduces the
reference's own recorded runs of those files bit for bit (objective,
iterations, stationarity and the full primal vector of
reports/cpu-validation-2026-09-17/results/{8792,8991}.default.ours.json).
GPU: hgqp_cli with the HIP engine reproduces the same records."""
import gzip
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import BoxQP

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CLI = os.path.join(os.path.dirname(GOLDEN), "..", "hypergrey-sih-26_amd", "hgqp_cli")


def model_text(name):
    with gzip.open(os.path.join(GOLDEN, f"{name}.qplib.gz"), "rb") as f:
        return f.read()


def record(name):
    with open(os.path.join(GOLDEN, "reference_records.json")) as f:
        rec = json.load(f)["qplib"][name]
    rec["primal"] = np.load(os.path.join(GOLDEN, "qplib_primal.npz"))[f"x{name}"]
    return rec


@pytest.mark.parametrize("name", ["8792", "8991"])
def test_model_hash_and_shape(name):
    import hashlib
    rec = record(name)
    assert hashlib.sha256(model_text(name)).hexdigest() == rec["model_sha256"]
    assert hashlib.sha256(rec["primal"].tobytes()).hexdigest() == rec["primal_sha256"]


@pytest.mark.parametrize("name", ["8792", "8991"])
def test_oracle_on_parsed_model_reproduces_reference_run(hg, oracle, name):
    d = hg.read_qplib(model_text(name))
    rec = record(name)
    assert d["rows"] == 0 and d["n"] == rec["columns"]
    q = BoxQP(d["n"], d["hrow"], d["hcol"], d["hval"], d["cost"], d["lower"], d["upper"], d["offset"], d["maximize"])
    r = oracle.solve_box(q, tolerance=1e-7, iteration_limit=200000, time_limit=120)
    assert r["status"] == rec["status"] == "optimal"
    assert r["iterations"] == rec["iterations"]
    assert r["objective"] == rec["objective"]
    assert r["stationarity_residual"] == rec["stationarity_residual"]
    assert r["dual_bound"] == rec["dual_bound_minimization_sense"]
    assert np.array_equal(r["primal"], rec["primal"])


def test_reader_rejects_malformed_input(hg):
    good = model_text("8991").decode()
    with pytest.raises(hg.InvalidArgument):
        hg.read_qplib(good[: len(good) // 3])  # truncated
    with pytest.raises(hg.InvalidArgument):
        hg.read_qplib("name QCB minimize 0\n")  # empty model
    with pytest.raises(hg.InvalidArgument):
        hg.read_qplib("name QIB minimize 2\n")  # integer variables
    with pytest.raises(hg.InvalidArgument):
        hg.read_qplib("name QCB minimize 2 1 1 2 3.0\n")  # upper-triangle Hessian entry
    with pytest.raises(hg.InvalidArgument):
        hg.read_qplib("name QCB minimize 2 1 1 1 2x\n")  # malformed number


def test_reader_halves_off_diagonal_and_reads_bounds(hg):
    text = """tiny QCB maximize
2 # variables
3 # Hessian entries, lower triangle of Q in 1/2 x'Qx
1 1 -4
2 1 3
2 2 -6
0 # default cost
1 # one nondefault cost
2 5.5
0.25 # offset
1e20 # infinity
-1 # default lower
0
1e20 # default upper (infinite)
1
1 2
0 0
0 0
0
0
"""
    d = hg.read_qplib(text)
    assert d["maximize"] and d["offset"] == 0.25 and d["rows"] == 0
    assert list(d["hval"]) == [-4.0, 1.5, -6.0]  # off-diagonal halved (qp.cpp:84)
    assert list(d["cost"]) == [0.0, 5.5]
    assert list(d["lower"]) == [-1.0, -1.0] and d["upper"][0] == 2.0 and np.isinf(d["upper"][1])


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["8792", "8991"])
def test_cli_reproduces_reference_run(tmp_path, name):
    path = tmp_path / f"{name}.qplib"
    path.write_bytes(model_text(name))
    out = subprocess.run([CLI, str(path), "--json", "--backend", "hip"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    r = json.loads(out.stdout)
    rec = record(name)
    assert r["status"] == "optimal" and r["algorithm"] == "hip_fista_box" and r["optimality_verified"]
    assert r["iterations"] == rec["iterations"] and r["columns"] == rec["columns"] and r["rows"] == 0
    assert np.array_equal(np.array(r["primal"]), rec["primal"])
    assert abs(r["objective"] - rec["objective"]) <= 1e-14 * abs(rec["objective"])
    assert abs(r["stationarity_residual"] - rec["stationarity_residual"]) <= 1e-12 * (1 + rec["stationarity_residual"])
    assert r["row_dual"] == []


@pytest.mark.gpu
def test_cli_exit_codes(tmp_path):
    path = tmp_path / "8991.qplib"
    path.write_bytes(model_text("8991"))
    out = subprocess.run([CLI, str(path), "--json", "--iteration-limit", "1"], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 1 and json.loads(out.stdout)["status"] == "iteration_limit"
    out = subprocess.run([CLI, str(path), "--backend", "cuda"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 2 and out.stderr.startswith("error: ")
    out = subprocess.run([CLI, str(path), "--threads", "0"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 2


@pytest.mark.gpu
def test_cli_sends_linear_rows_to_admm(tmp_path):
    """main.cpp:145: a model with rows goes to solve_qp_admm; the JSON carries
    the row duals. min 1/2 (x1^2 + x2^2) s.t. x1 + x2 = 1: x* = (1/2, 1/2), y = -1/2."""
    from helpers import QCL
    path = tmp_path / "twovar.qplib"
    path.write_text(QCL)
    out = subprocess.run([CLI, str(path), "--json", "--tolerance", "1e-6"], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    r = json.loads(out.stdout)
    assert r["status"] == "optimal" and r["algorithm"] == "hip_admm_pcg" and r["rows"] == 1
    assert abs(r["objective"] - 0.25) <= 1e-6
    assert np.abs(np.array(r["primal"]) - 0.5).max() <= 1e-6
    assert abs(r["row_dual"][0] + 0.5) <= 1e-6


# ---- the reader against the reference's own reader (oracle/_ref) -------------

TINY = """tiny QCB maximize
2 # variables
3 # Hessian entries, lower triangle of Q in 1/2 x'Qx
1 1 -4
2 1 3
2 2 -6
0 # default cost
1 # one nondefault cost
2 5.5
0.25 # offset
1e20 # infinity
-1 # default lower
0
1e20 # default upper (infinite)
1
1 2
0 0
0 0
0
0
"""

POOL = ["0", "-1", "1", "2", "3", "1.5", "-0", "1e20", "-1e20", "1e400", "-1e400", "1e-5000", "4e-324", "nan",
        "inf", "-inf", "0x1p1", "+2", "2x", "abc", "100000", "100001", "5000001", "1e30", "QCB", "QCL", "QIB",
        "minimize", "maximize", "#", "1#2"]


def parse_or_none(hg, text):
    try:
        return hg.read_qplib(text)
    except hg.InvalidArgument:
        return None


def bits(a):
    a = np.asarray(a)
    return a.view(np.uint64) if a.dtype == np.float64 else a


def assert_same_model(a, b, label):
    assert (a is None) == (b is None), (label, "accepted by one reader only", a is None, b is None)
    if a is None:
        return
    for k in ("n", "rows", "maximize"):
        assert a[k] == b[k], (label, k)
    assert bits(np.float64(a["offset"])) == bits(np.float64(b["offset"])), (label, "offset")
    for k in ("hrow", "hcol", "hval", "cost", "lower", "upper", "rstart", "rcol", "rval", "row_lower", "row_upper"):
        assert len(a[k]) == len(b[k]) and np.array_equal(bits(a[k]), bits(b[k])), (label, k)


def corpus():
    from helpers import QCL
    bases = {"tiny": TINY, "qcl": QCL}
    out = [("8792", model_text("8792").decode()), ("8991", model_text("8991").decode())]
    for name, text in bases.items():
        out.append((name, text))
        words = text.split()
        out.append((name + ":one_line", " ".join(words)))
        out.append((name + ":crlf_tabs", text.replace("\n", "\r\n").replace(" ", "\t")))
        for k in range(len(words)):
            out.append((f"{name}:cut{k}", " ".join(words[:k])))
            out.append((f"{name}:drop{k}", " ".join(words[:k] + words[k + 1:])))
            for w in POOL:
                out.append((f"{name}:w{k}={w}", " ".join(words[:k] + [w] + words[k + 1:])))
        for tail in (" 7", "\n7", "\n# trailing comment\n\n", "\n\v\n", "\n \t\r\n", "\f", "\n\f\n", "\n#x 7\n",
                     "\n\x00\n"):
            out.append((f"{name}:tail{tail!r}", text.rstrip("\n") + tail))
    return out


def test_reader_equals_reference_reader(hg, reference):
    """Accept/reject and every parsed array (bit for bit, signed zeros included)
    equal the reference's own read_qplib on the two public models, two small
    models, and every truncation, single-word deletion and single-word
    replacement (from a pool of numbers, out-of-range values, non-finite and
    malformed words, type codes and comment marks) of the small ones, plus
    trailing content of every kind."""
    accepted = rejected = 0
    for label, text in corpus():
        mine, ref = parse_or_none(hg, text), reference.read_qplib(text)
        assert_same_model(mine, ref, label)
        accepted += mine is not None
        rejected += mine is None
    print("accepted", accepted, "rejected", rejected, flush=True)
    assert accepted > 100 and rejected > 500


def test_reader_random_mutations_equal_reference(hg, reference):
    """Seeded random edits of the small models (two to four word swaps,
    duplications and replacements at once) are accepted or rejected alike, and
    accepted ones parse to the same model."""
    from helpers import QCL
    rng = np.random.default_rng(2026)
    for base in (TINY, QCL):
        words = base.split()
        for trial in range(1500):
            w = list(words)
            for _ in range(int(rng.integers(2, 5))):
                op = int(rng.integers(0, 3))
                k = int(rng.integers(0, len(w)))
                if op == 0:
                    j = int(rng.integers(0, len(w)))
                    w[k], w[j] = w[j], w[k]
                elif op == 1:
                    w.insert(k, w[k])
                else:
                    w[k] = POOL[int(rng.integers(0, len(POOL)))]
            text = "\n".join(w)
            assert_same_model(parse_or_none(hg, text), reference.read_qplib(text), f"trial {trial}: {text!r}")
