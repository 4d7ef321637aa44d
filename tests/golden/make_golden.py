"""Writes the golden fixtures of tests/golden from the reference's own recorded
results (run here only; the reference checkout is not needed afterwards).

Sources (paths relative to 02-software/solver/ of the reference):
  reports/kaggle-gpu-03/returned/results.json        seeded tridiagonal family,
      n = 1e4 .. 1e6, CPU (1 and 4 threads) and CUDA, 3 repeats each
  reports/cpu-validation-2026-09-17/results/{8792,8991}.default.ours.json
      QPLIB box QPs solved by the reference CPU path (cpu_fista_box)
  datasets/qplib/{8792,8991}.qplib                   the QPLIB models (public data)
Outputs: reference_records.json, qplib_primal.npz, 8792.qplib.gz, 8991.qplib.gz
"""
import gzip
import hashlib
import json
import os
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/02-software/solver"
OUT = os.path.dirname(os.path.abspath(__file__))


def main():
    rows = json.load(open(os.path.join(REF, "reports/kaggle-gpu-03/returned/results.json")))
    tri = {}
    for r in rows:
        key = (r["columns"], r["backend"], r["threads"])
        tri.setdefault(key, {k: r[k] for k in ("columns", "backend", "threads", "status", "passed", "iterations",
                                               "objective", "known_objective", "relative_objective_error",
                                               "max_solution_error", "stationarity_residual")})
    records = {"tridiagonal": sorted(tri.values(), key=lambda d: (d["columns"], d["backend"], d["threads"])),
               "tridiagonal_source": "reports/kaggle-gpu-03/returned/results.json",
               "qplib": {}}
    primals = {}
    for name in ("8792", "8991"):
        src = f"reports/cpu-validation-2026-09-17/results/{name}.default.ours.json"
        d = json.load(open(os.path.join(REF, src)))
        x = np.array(d["primal"], dtype=np.float64)
        primals[f"x{name}"] = x
        records["qplib"][name] = {k: d[k] for k in ("status", "algorithm", "optimality_verified", "iterations",
                                                     "columns", "objective", "primal_residual",
                                                     "stationarity_residual", "complementarity_residual",
                                                     "dual_bound_minimization_sense", "relative_gap")}
        records["qplib"][name]["primal_sha256"] = hashlib.sha256(x.tobytes()).hexdigest()
        records["qplib"][name]["source"] = src
        with open(os.path.join(REF, f"datasets/qplib/{name}.qplib"), "rb") as f:
            raw = f.read()
        records["qplib"][name]["model_sha256"] = hashlib.sha256(raw).hexdigest()
        with gzip.open(os.path.join(OUT, f"{name}.qplib.gz"), "wb", compresslevel=9) as g:
            g.write(raw)
    np.savez_compressed(os.path.join(OUT, "qplib_primal.npz"), **primals)
    with open(os.path.join(OUT, "reference_records.json"), "w") as f:
        json.dump(records, f, indent=1)


if __name__ == "__main__":
    main()
