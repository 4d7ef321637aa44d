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
