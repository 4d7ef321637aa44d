#!/usr/bin/env python3
"""Box-QP benchmark on one MI355X.

One step = one complete solve, from x0 = clip(0, l, u) to a device-verified
KKT residual <= 1e-6, of the planted single instance n = 1,000,000 with 16
off-diagonal draws per row (about 33M stored entries), with the matrix and
vectors already resident in HBM. The headline value is solves/s of the fastest
mode whose results meet the contract; every mode is reported beside it with its
error against the FP64 reference-schedule solve (bit-identical to the CPU
reference path) and against the planted optimum.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--modes a,b] [--no-cpu-baseline]

Prints one JSON line. The roofline figure is the dominant kernel's algorithmic
bytes per launch divided by its average launch time from HIP events recorded
on the engine's stream around every step launch of the timed solves.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hypergrey-sih-26_amd"))

import hgqp  # noqa: E402

HBM_PEAK_GBS = 8000.0
KKT_TOL = 1e-6
METRIC = "box-QP solves/s + p50/p99 latency at KKT≤1e-6; SpMV HBM GB/s vs peak"

STEP = {"scalar": hgqp.STEP_SCALAR, "jacobi": hgqp.STEP_JACOBI}
MODES = {
    # name: (precision, momentum, reduce, step metric)
    "fp64_reference": (hgqp.PRECISION_FP64, hgqp.MOMENTUM_REFERENCE, hgqp.REDUCE_ORDERED, hgqp.STEP_SCALAR),
    "fp64_restart_jacobi": (hgqp.PRECISION_FP64, hgqp.MOMENTUM_RESTART, hgqp.REDUCE_TREE, hgqp.STEP_JACOBI),
    "fp32_polish_restart_jacobi": (hgqp.PRECISION_FP32_POLISH, hgqp.MOMENTUM_RESTART, hgqp.REDUCE_TREE,
                                   hgqp.STEP_JACOBI),
    "fp32_refine_restart": (hgqp.PRECISION_FP32_REFINE, hgqp.MOMENTUM_RESTART, hgqp.REDUCE_TREE, hgqp.STEP_SCALAR),
    "fp32_refine_restart_jacobi": (hgqp.PRECISION_FP32_REFINE, hgqp.MOMENTUM_RESTART, hgqp.REDUCE_TREE,
                                   hgqp.STEP_JACOBI),
    "fp32_refine_restart_spectral": (hgqp.PRECISION_FP32_REFINE, hgqp.MOMENTUM_RESTART, hgqp.REDUCE_TREE,
                                     hgqp.STEP_SPECTRAL),
}


def step_bytes(n, nnz, value_bytes, vector_bytes, jacobi=False):
    """Algorithmic HBM bytes of one fused step: values + int32 columns once, row
    starts once, the gathered vector once, x, c, lower, upper read, x and the
    extrapolated point written (7 vectors), plus the per-row step denominators
    read by the Jacobi metric (an 8th vector)."""
    return nnz * (value_bytes + 4) + 4 * (n + 1) + (8 if jacobi else 7) * n * vector_bytes


def measured_traffic(use32, mode, n, draws, nnz, panels, slabs, groups):
    """HBM bytes per FP32 step (all of its launches) from the committed
    rocprofv3 PMC summary (FETCH_SIZE x 2 + WRITE_SIZE, calibrated on plain
    streaming kernels), when this run's mode, instance, panel and slab counts
    are the ones profiled; otherwise None."""
    path = os.path.join(ROOT, "profiles", "traffic_planted_n1e6.json")
    if not use32 or not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    modes = t.get("modes", [t.get("mode")])  # modes whose FP32 step is the profiled one
    if (t.get("workload") != f"planted n={n} draws={draws} nnz={nnz}" or mode not in modes
            or t.get("column_panels32") != panels or t.get("slabs32", 0) != slabs
            or t.get("slab_groups", 1 if slabs else 0) != groups):
        return None
    return t["traffic_bytes_per_step"]


def run_mode(name, inst, L, steps, warmup, check_every, polish_every):
    precision, momentum, reduce, metric = MODES[name]
    # the reference schedule keeps the reference's check cadence of 25 iterations
    cadence = dict(check_every=25, polish_every=0) if name == "fp64_reference" else \
        dict(check_every=check_every, polish_every=polish_every)
    # solves/s and latency: graph-replayed check periods, no per-period events
    box = hgqp.Box(inst["start"], inst["index"], inst["value"], inst["cost"], inst["lower"], inst["upper"], L,
                   precision=precision, momentum=momentum, reduce=reduce, step_metric=metric, timing=0, use_graph=1)
    try:
        info = box.info()
        for _ in range(warmup):
            box.reset()
            box.solve(tolerance=KKT_TOL, iteration_limit=20000, time_limit=60, **cadence)
        lat, reps = [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            a = time.perf_counter()
            box.reset()
            rep = box.solve(tolerance=KKT_TOL, iteration_limit=20000, time_limit=60, **cadence)
            lat.append(time.perf_counter() - a)
            reps.append(rep)
        wall = time.perf_counter() - t0
        x = box.primal()
    finally:
        box.close()
    # kernel times: the same solves on a handle with HIP events around each period's steps
    timed = hgqp.Box(inst["start"], inst["index"], inst["value"], inst["cost"], inst["lower"], inst["upper"], L,
                     precision=precision, momentum=momentum, reduce=reduce, step_metric=metric, timing=1,
                     use_graph=0)
    try:
        treps = []
        for _ in range(max(1, min(steps, 3))):
            timed.reset()
            treps.append(timed.solve(tolerance=KKT_TOL, iteration_limit=20000, time_limit=60, **cadence))
    finally:
        timed.close()
    return dict(name=name, wall=wall, lat=lat, reps=reps, timed=treps, x=x,
                panels=(info["column_panels"], info["column_panels32"]), slabs=info["slabs32"],
                slab_groups=info["slab_groups"], jacobi_scale=info["jacobi_scale"])


def host_inclusive(name, inst, L, check_every, polish_every, reps=3):
    """The same solve with host buffers on both sides: create (upload of the CSR
    and vectors, FP32 copies), solve, read x back. Reported beside `value`,
    which times the solve with inputs already in HBM."""
    precision, momentum, reduce, metric = MODES[name]
    cadence = dict(check_every=25, polish_every=0) if name == "fp64_reference" else \
        dict(check_every=check_every, polish_every=polish_every)
    up, sv, rd = [], [], []
    for _ in range(reps):
        a = time.perf_counter()
        box = hgqp.Box(inst["start"], inst["index"], inst["value"], inst["cost"], inst["lower"], inst["upper"], L,
                       precision=precision, momentum=momentum, reduce=reduce, step_metric=metric, use_graph=0)
        b = time.perf_counter()
        try:
            box.solve(tolerance=KKT_TOL, iteration_limit=20000, time_limit=60, **cadence)
            c = time.perf_counter()
            box.primal()
            d = time.perf_counter()
        finally:
            box.close()
        up.append(b - a)
        sv.append(c - b)
        rd.append(d - c)
    total = np.median(np.array(up) + np.array(sv) + np.array(rd))
    return {"mode": name, "upload_ms": round(float(np.median(up)) * 1e3, 3),
            "solve_ms": round(float(np.median(sv)) * 1e3, 3), "readback_ms": round(float(np.median(rd)) * 1e3, 3),
            "solves_per_s": round(1.0 / total, 3), "reps": reps}


def whole_model(name, inst, check_every, polish_every, reps=2):
    """The whole-model entry (hgqp_solve_qp) from the reference's triplet form:
    host checks (validation, convexity), device assembly of the mirrored CSR and
    L, upload of the vectors, the solve and the read-back of x."""
    precision, momentum, reduce, metric = MODES[name]
    hr, hc, hv = hgqp.lower_triplets(inst["start"], inst["index"], inst["value"])
    n = len(inst["cost"])
    times, r = [], None
    for _ in range(reps):
        a = time.perf_counter()
        r = hgqp.solve_qp(n, hr, hc, hv, inst["cost"], inst["lower"], inst["upper"], tolerance=KKT_TOL,
                          iteration_limit=20000, time_limit=120, precision=precision, momentum=momentum,
                          reduce=reduce, step_metric=metric)
        times.append(time.perf_counter() - a)
    return {"mode": name, "terms": int(len(hv)), "seconds": round(float(np.median(times)), 4),
            "host_check_seconds": round(r["host_check_seconds"], 4),
            "assembly_seconds": round(r["assembly_seconds"], 4),
            "status": r["status"], "iterations": int(r["iterations"]), "algorithm": r["algorithm"],
            "max_abs_err_vs_planted_x": float(np.abs(r["primal"] - inst["optimum"]).max())}


def host_cpu():
    """The host's CPU model, its logical CPUs, the CPUs this process may run on
    and the physical cores among them (from /proc/cpuinfo)."""
    usable = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else list(range(os.cpu_count() or 1))
    model, cores, cur = None, set(), {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" not in line:
                    continue
                k, v = (s.strip() for s in line.split(":", 1))
                if k == "processor":
                    cur = {"cpu": int(v)}
                elif k == "model name":
                    model = model or v
                elif k == "physical id":
                    cur["package"] = v
                elif k == "core id" and cur.get("cpu") in usable:
                    cores.add((cur.get("package"), v))
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count(), "usable_cpus": len(usable),
            "usable_physical_cores": len(cores) or None}


def reference_cpu():
    """The reference's own CPU sources (oracle/_ref), else the oracle port."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle, available
    kind = "reference" if available("reference") else "oracle"
    source = "oracle/_ref (reference CPU sources)" if kind == "reference" else "oracle port"
    return Oracle(kind), ("reference" if kind == "reference" else "port"), source


def cpu_full_solve(o, q, tolerance, threads):
    """One full CPU solve at the reference's timing boundary (scale_benchmark.cpp:20-21:
    solve_accelerated_box from the triplet model, CSR assembly and the check every
    25 iterations included, generation excluded)."""
    a = time.perf_counter()
    r = o.solve_box(q, tolerance=tolerance, iteration_limit=20000, threads=threads, time_limit=600)
    return time.perf_counter() - a, r


def cpu_baseline(inst, threads):
    """The reference CPU path on the same instance, timed as full solves at the
    reference's boundary: once on every CPU this process may use, once on one
    thread; a setup-only run (iteration limit 0: assembly and the iteration-0
    check) gives the loop time beside it."""
    o, kind, source = reference_cpu()
    from oracle import BoxQP
    hr, hc, hv = hgqp.lower_triplets(inst["start"], inst["index"], inst["value"])
    q = BoxQP(inst["n"], hr, hc, hv, inst["cost"], inst["lower"], inst["upper"])

    def run(nthreads):
        full, r = cpu_full_solve(o, q, KKT_TOL, nthreads)
        a = time.perf_counter()
        o.solve_box(q, tolerance=KKT_TOL, iteration_limit=0, threads=nthreads, time_limit=600)
        setup = time.perf_counter() - a
        text = (f"{source}, {nthreads} OpenMP thread{'s' if nthreads > 1 else ''}: one full solve of this "
                f"instance from its triplet model, {r['iterations']} iterations ({r['status']}): {full:.2f}s, "
                f"of which setup (CSR assembly + iteration-0 check) {setup:.2f}s")
        return {"value": 1.0 / full, "unit": "solves/s", "cores": nthreads, "sample": text, "solve_seconds": full,
                "setup_seconds": setup, "loop_seconds": full - setup, "iterations": int(r["iterations"]),
                "status": r["status"], "objective": r["objective"]}

    out = run(threads)
    out["kind"] = kind
    out["host"] = host_cpu()
    if threads > 1:
        out["single_thread"] = run(1)
    return out


SLAB_BLOCK_CANDIDATES = (16384, 32768, 65536, 131072)  # stored entries per slab block


def batch_workload(args):
    """512 uneven planted instances solved together in one pool (FP32 correction
    cycles, FP64 verification), against the plain FP64 single-instance loop."""
    batch = hgqp.gen_batch(args.instances, args.seed, 1000, 100000, 4, 32)
    rows = sum(d["n"] + 1 for d in batch)
    nnz = sum(len(d["value"]) for d in batch)
    Ls = [hgqp.lipschitz(d["start"], d["value"]) for d in batch]
    print(f"# batch: {len(batch)} instances, rows {rows}, nnz {nnz}", file=sys.stderr, flush=True)

    def make_pool(entries=0):
        return hgqp.Pool(max_rows=rows, max_nnz=nnz, max_instances=len(batch), block_loads=0, block_rows=0,
                         slabs=args.pool_slabs, slab_block_entries=entries,
                         check_every=args.check_every, step_metric=STEP[args.pool_step],
                         fp32_switch=args.pool_switch)

    # slab pools: the slab-block size, timed (all instances active, six FP32 steps) for each candidate
    slab_tuning, entries = None, 0
    if args.pool_slabs:
        slab_tuning = {}
        for cand in SLAB_BLOCK_CANDIDATES:
            with make_pool(cand) as probe:
                for k, d in enumerate(batch):
                    assert probe.admit(d, tag=k, tolerance=KKT_TOL, lipschitz_value=Ls[k]) is not None
                slab_tuning[cand] = round(probe.time_step(6)[0], 1)
            print(f"# slab blocks of {cand} entries: {slab_tuning[cand]} us per step", file=sys.stderr, flush=True)
        entries = min(slab_tuning, key=slab_tuning.get)
    pool = make_pool(entries)
    times, results, admits = [], None, []
    tuned = None
    try:
        for rep in range(args.warmup + args.steps):
            a = time.perf_counter()
            for k, d in enumerate(batch):
                assert pool.admit(d, tag=k, tolerance=KKT_TOL, lipschitz_value=Ls[k]) is not None
            admits.append(time.perf_counter() - a)
            if tuned is None:
                tuned = pool.autotune()
                step_us, step_bytes_all = pool.time_step(6)  # all 512 instances active, before the solve
            t0 = time.perf_counter()
            res = pool.solve_all()
            dt = time.perf_counter() - t0
            if rep >= args.warmup:
                times.append(dt)
                results = res
    finally:
        pool.close()
    res = {r["tag"]: r for r in results}
    ok = all(r["status"] == "optimal" and r["optimality_verified"] for r in res.values())
    worst_x = max(float(np.abs(res[k]["primal"] - d["optimum"]).max()) for k, d in enumerate(batch))
    worst_obj = max(abs(res[k]["objective"] - d["objective"]) / (1 + abs(d["objective"])) for k, d in enumerate(batch))
    iters = np.array([res[k]["iterations"] for k in range(len(batch))])
    wall = float(np.median(times))
    # plain FP64 single-instance form, one instance at a time, same instances
    plain_n = min(len(batch), args.plain_sample)
    boxes = [hgqp.Box(d["start"], d["index"], d["value"], d["cost"], d["lower"], d["upper"], Ls[k])
             for k, d in enumerate(batch[:plain_n])]
    try:
        for b in boxes[:4]:
            b.solve(tolerance=KKT_TOL, iteration_limit=20000, time_limit=60)
            b.reset()
        t0 = time.perf_counter()
        plain = [b.solve(tolerance=KKT_TOL, iteration_limit=20000, time_limit=60) for b in boxes]
        plain_wall = time.perf_counter() - t0
    finally:
        for b in boxes:
            b.close()
    plain_rate = plain_n / plain_wall if plain_n else None  # --plain-sample 0: no plain loop
    return {
        "metric": METRIC, "value": round(len(batch) / wall, 2), "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic (seeded planted batch, admitted to HBM before timing)",
        "config": {"workload": f"batched: {len(batch)} planted box QPs, n log-uniform in [1e3, 1e5], "
                               f"draws/row uniform in 4..32, bounds l~U(-2,-0.5), u~U(0.5,2), KKT<=1e-6, "
                               f"one pool, graph-replayed check periods; one step = the whole batch",
                   "rows": rows, "nnz": nnz, "check_every": args.check_every, "step_metric": args.pool_step,
                   "first_cycle_target": f"{args.pool_switch} x tol",
                   "slabs": args.pool_slabs,
                   # what was timed: the slab-block size of a slab pool, the row-block shape of the others
                   # (a slab pool's row blocks serve its FP64 checks only and keep a fixed shape)
                   **({"autotuned_slab_block_entries": entries, "slab_block_step_us": slab_tuning,
                       "check_block_loads": tuned[0], "check_block_rows": tuned[1]} if args.pool_slabs else
                      {"autotuned_block_loads": tuned[0], "autotuned_block_rows": tuned[1]})},
        "roofline": {"bound": "hbm", "achieved": round(step_bytes_all / step_us / 1e3, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(step_bytes_all / step_us / 1e3 / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": (f"pool_slab_step (FP32 step over per-instance column slabs, all {len(batch)} "
                                f"instances active)" if args.pool_slabs else
                                f"pool_step<{tuned[0]}> (FP32 step, all {len(batch)} instances active)"),
                     "algorithmic_bytes": step_bytes_all, "avg_launch_us": round(step_us, 1)},
        "all_verified": ok, "worst_max_abs_err_vs_planted_x": worst_x, "worst_objective_rel_err_vs_planted": worst_obj,
        "iterations": {"min": int(iters.min()), "median": float(np.median(iters)), "max": int(iters.max())},
        "admission_ms": round(float(np.median(admits)) * 1e3, 1),
        "plain_fp64_single_instance": {"solves_per_s": round(plain_rate, 2) if plain_n else None,
                                       "instances_timed": plain_n,
                                       "all_verified": all(p["status"] == "optimal" for p in plain)},
        "speedup_vs_plain_fp64": round(len(batch) / wall / plain_rate, 2) if plain_n else None,
        "device": hgqp.device_name(),
    }


def serve_workload(args):
    """Poisson arrival trace served by the continuous-batching loop, against the
    plain one-at-a-time FP64 loop on the same trace (both in real time). Every
    request is its own planted instance (seeded by its index), generated ahead
    of its arrival by a pool of host threads and dropped once submitted; with
    --distinct K > 0 the requests draw from K instances instead."""
    import threading
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(args.seed)
    total = max(args.requests, args.saturate)
    sizes = np.round(np.exp(rng.uniform(np.log(1000), np.log(200000), size=total))).astype(np.int64)
    arrivals = np.cumsum(rng.exponential(1.0 / args.rate, size=args.requests))
    pick = rng.integers(0, args.distinct, size=total) if args.distinct > 0 else np.arange(total)

    def make(j):
        """Request j's instance and Lipschitz bound (the same for every replay of the trace)."""
        k = int(pick[j])
        d = hgqp.gen_planted(int(sizes[k]), 8, (args.seed << 20) + k, -1.0, 1.0)
        return d, hgqp.lipschitz(d["start"], d["value"])

    makers = ThreadPoolExecutor(max_workers=8)
    lookahead = 256

    class Ahead:
        """Instances of requests first .. count-1 in order, generated `lookahead` ahead."""

        def __init__(self, count):
            self.count, self.next, self.q = count, 0, {}
            for j in range(min(lookahead, count)):
                self.q[j] = makers.submit(make, j)
            self.next = min(lookahead, count)

        def get(self, j):
            inst = self.q.pop(j).result()
            if self.next < self.count:
                self.q[self.next] = makers.submit(make, self.next)
                self.next += 1
            return inst

    sample = set(range(0, args.requests, max(1, args.requests // 200)))
    print(f"# serve: {args.requests} requests at {args.rate}/s over {arrivals[-1]:.1f}s; "
          f"{'a distinct instance per request' if args.distinct <= 0 else f'{args.distinct} distinct instances'}",
          file=sys.stderr, flush=True)

    def pool_server(arrivals, count, producers=1, ready=None):
        """ready: instances generated beforehand (the capacity run); else generated ahead of arrival."""
        srv = hgqp.Server(max_rows=args.budget_rows, max_nnz=args.budget_nnz, max_instances=args.slots,
                          check_every=args.check_every, use_graph=0 if args.no_graph else 1,
                          fp32_switch=args.pool_switch,
                          step_metric=STEP[args.pool_step])
        done, optimum = {}, {}
        ahead = None if ready is not None else Ahead(count)
        t0 = time.perf_counter()
        try:
            failed = []

            def producer(first):
                try:
                    for j in range(first, count, producers):
                        d, L = ready[j] if ready is not None else ahead.get(j)
                        delay = t0 + arrivals[j] - time.perf_counter()
                        if delay > 0:
                            time.sleep(delay)
                        srv.submit(d, tag=j, tolerance=KKT_TOL, lipschitz_value=L)
                        if j in sample:
                            optimum[j] = d["optimum"]
                except BaseException as e:  # surfaced by the poll loop below
                    failed.append(e)
            threads = [threading.Thread(target=producer, args=(i,)) for i in range(producers)]
            for th in threads:
                th.start()
            while len(done) < count:
                if failed:
                    raise RuntimeError(f"request producer failed: {failed[0]!r}")
                for r in srv.poll():
                    done[r["tag"]] = r
                time.sleep(0.0005)
            for th in threads:
                th.join()
            wall = time.perf_counter() - t0
        finally:
            srv.stop()
        lat = np.array([done[j]["latency_seconds"] for j in range(count)]) * 1e3
        ok = all(done[j]["status"] == "optimal" for j in range(count))
        err = max(float(np.abs(done[j]["primal"] - optimum[j]).max()) for j in optimum) if optimum else None
        return {"requests": count, "solves_per_s": count / wall, "p50_ms": float(np.percentile(lat, 50)),
                "p99_ms": float(np.percentile(lat, 99)), "all_verified": ok, "wall_s": wall,
                "sampled_max_abs_err_vs_planted_x": err, "sampled": len(optimum)}

    def plain_loop(arrivals, count, ready=None):
        """One request at a time in arrival order: upload, FP64 reference-schedule solve, read x."""
        lat = []
        ok = True
        ahead = None if ready is not None else Ahead(count)
        t0 = time.perf_counter()
        for j in range(count):
            d, L = ready[j] if ready is not None else ahead.get(j)
            now = time.perf_counter() - t0
            if now < arrivals[j]:
                time.sleep(arrivals[j] - now)
            with hgqp.Box(d["start"], d["index"], d["value"], d["cost"], d["lower"], d["upper"], L) as b:
                r = b.solve(tolerance=KKT_TOL, iteration_limit=20000, time_limit=60)
                b.primal()
            ok &= r["status"] == "optimal"
            lat.append(time.perf_counter() - t0 - arrivals[j])
        wall = time.perf_counter() - t0
        lat = np.array(lat) * 1e3
        return {"requests": count, "solves_per_s": count / wall, "p50_ms": float(np.percentile(lat, 50)),
                "p99_ms": float(np.percentile(lat, 99)), "all_verified": bool(ok), "wall_s": wall}

    served = pool_server(arrivals, args.requests)
    print(f"# served (real time): {served}", file=sys.stderr, flush=True)
    plain = plain_loop(arrivals, args.requests) if not args.no_plain else None
    print(f"# plain (real time): {plain}", file=sys.stderr, flush=True)
    # capacity: the first requests of the trace, all queued at t=0
    saturated = None
    if args.saturate:
        at_once = np.zeros(args.saturate)
        ready = list(makers.map(make, range(args.saturate)))  # generated before the clock starts
        sat_served = pool_server(at_once, args.saturate, producers=4, ready=ready)  # as the HTTP route submits
        print(f"# served (saturated): {sat_served}", file=sys.stderr, flush=True)
        sat_plain = plain_loop(at_once, args.saturate, ready=ready) if not args.no_plain else None
        del ready
        print(f"# plain (saturated): {sat_plain}", file=sys.stderr, flush=True)
        saturated = {"requests": args.saturate, "served": sat_served, "plain_fp64_one_at_a_time": sat_plain,
                     "speedup_solves_per_s": (sat_served["solves_per_s"] / sat_plain["solves_per_s"]
                                              if sat_plain else None)}
    makers.shutdown(wait=False)
    return {
        "metric": METRIC, "value": round(served["solves_per_s"], 2), "unit": "solves/s", "n_gpus": 1,
        "steps": args.requests, "warmup": 0, "ms_per_step": round(served["wall_s"] / args.requests * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic (seeded planted instances; Poisson arrival trace replayed in real time)",
        "config": {"workload": f"serving: {args.requests} requests, Poisson lambda={args.rate}/s, n log-uniform in "
                               f"[1e3, 2e5], 8 draws/row, bounds [-1,1], KKT<=1e-6, continuous batching in one pool",
                   "slots": args.slots, "budget_rows": args.budget_rows, "budget_nnz": args.budget_nnz,
                   "distinct_instances": args.distinct if args.distinct > 0 else "one per request",
                   "check_every": args.check_every, "first_cycle_target": f"{args.pool_switch} x tol",
                   "step_metric": args.pool_step,
                   "graph_replay": not args.no_graph},
        "p50_ms": round(served["p50_ms"], 3), "p99_ms": round(served["p99_ms"], 3),
        "served": served, "plain_fp64_one_at_a_time": plain, "saturated": saturated,
        "device": hgqp.device_name(),
    }


def dense_workload(args):
    """One dense Q = M'M/n + 0.1 I (n = 4096) with 256 cost vectors: BF16 MFMA
    correction cycles, FP64 MFMA verification; every RHS re-checked by the
    oracle's FP64 check on the CPU."""
    Q, Cr, lo, up = hgqp.gen_dense(args.dense_n, args.dense_k, args.seed)
    print(f"# dense: n={args.dense_n} K={args.dense_k}", file=sys.stderr, flush=True)
    d = hgqp.Dense(Q, Cr, lo, up)
    times, rep = [], None
    try:
        for i in range(args.warmup + args.steps):
            t0 = time.perf_counter()
            rep = d.solve(tolerance=KKT_TOL, iteration_limit=20000)
            if i >= args.warmup:
                times.append(time.perf_counter() - t0)
        x, checks, status, iters = d.results()
        gemm_us, tflops = d.time_gemm(50)
    finally:
        d.close()
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import BoxQP, Oracle
    o = Oracle("oracle")
    r, c = np.tril_indices(args.dense_n)
    hr, hc, hv = r.astype(np.int64), c.astype(np.int64), Q[r, c]
    worst_kkt, worst_obj = 0.0, 0.0
    sample = range(0, args.dense_k, max(1, args.dense_k // args.dense_check))
    for j in sample:
        chk = o.check(BoxQP(args.dense_n, hr, hc, hv, Cr[j], lo, up), x[j], KKT_TOL)
        worst_kkt = max(worst_kkt, chk["stationarity_residual"])
        worst_obj = max(worst_obj, abs(chk["objective"] - checks[j]["objective"]) / (1 + abs(chk["objective"])))
    wall = float(np.median(times))
    return {
        "metric": METRIC, "value": round(args.dense_k / wall, 1), "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16+f32+f64",
        "data": "synthetic (seeded Q = M'M/n + 0.1 I and N(0,1) costs, resident in HBM before timing)",
        "config": {"workload": f"dense multi-RHS: n={args.dense_n}, {args.dense_k} cost vectors, bounds [-1,1], "
                               f"KKT<=1e-6; one step = all RHS; BF16 MFMA iterations, FP64 MFMA verification"},
        "roofline": {"bound": "mfma", "achieved": round(tflops, 1), "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": round(tflops / 2500.0, 4), "traffic": None,
                     "kernel": "BF16 GEMM Q W' (gemm_bf16_wide or gemm_bf16, whichever the handle timed faster)",
                     "avg_launch_us": round(gemm_us, 2),
                     "flops_per_launch": 2 * args.dense_n ** 2 * args.dense_k},
        "all_verified": bool(rep["verified"] == args.dense_k), "max_iterations": int(rep["max_iterations"]),
        "bf16_gemms": int(rep["bf16_gemms"]), "fp64_gemms": int(rep["fp64_gemms"]),
        "lipschitz": rep["lipschitz"], "oracle_checked_rhs": len(list(sample)),
        "oracle_worst_kkt": worst_kkt, "oracle_worst_objective_rel_diff": worst_obj,
        "device": hgqp.device_name(),
    }


# The reference's published runs of its seeded tridiagonal family (median of
# three, seconds): CPU 1 thread, CPU 4 threads, Tesla T4
# (reports/kaggle-gpu-03/REPORT.md:19-23).
PUBLISHED_TRIDIAGONAL = {10000: (0.1159, 0.0838, 0.3821), 50000: (0.5729, 0.3955, 0.5139),
                         200000: (2.5509, 1.6158, 0.9799), 500000: (7.3144, 4.0860, 1.8996),
                         1000000: (15.923, 8.342, 3.501)}
TRIDIAGONAL_MODES = ("fp64_reference", "fp64_restart_jacobi", "fp32_refine_restart_jacobi")


def slab_groups_workload(args):
    """FP32 step time of the slab layouts on the planted instance: a fixed run
    of FP32 steps (no phase switch, one check at the end) on handles with one
    (ORDERED), two and four (TREE) workgroups per row block, timed with the
    handle's HIP events around each period's step launches."""
    inst = hgqp.gen_planted(args.n, args.draws, args.seed)
    L = hgqp.lipschitz(inst["start"], inst["value"])
    steps = args.group_steps
    out = {}
    for groups in (1, 2, 4):
        reduce = hgqp.REDUCE_ORDERED if groups == 1 else hgqp.REDUCE_TREE
        box = hgqp.Box(inst["start"], inst["index"], inst["value"], inst["cost"], inst["lower"], inst["upper"], L,
                       precision=hgqp.PRECISION_FP32_POLISH, momentum=hgqp.MOMENTUM_REFERENCE, reduce=reduce,
                       step_metric=hgqp.STEP_JACOBI, slabs=2, slab_groups=groups, timing=1, use_graph=0)
        try:
            info = box.info()
            per = []
            for rep in range(3):
                box.reset()
                r = box.solve(tolerance=KKT_TOL, iteration_limit=steps, check_every=steps, fp32_switch=1e-300,
                              time_limit=60)
                per.append(r["step32_kernel_ms"] / r["step32_launches"] * 1e3)
        finally:
            box.close()
        out[str(groups)] = {"us_per_step": [round(u, 2) for u in per], "slabs32": info["slabs32"],
                            "slab_groups": info["slab_groups"], "row_blocks": info["row_blocks"]}
        print(f"# {groups} group(s): {out[str(groups)]}", file=sys.stderr, flush=True)
    nbytes = step_bytes(args.n, len(inst["value"]), 4, 4, jacobi=True)
    best = min(min(v["us_per_step"]) for v in out.values())
    return {"metric": "fp32_step_us", "value": round(best, 2), "unit": "us", "n_gpus": 1, "steps": steps,
            "warmup": 0, "higher_is_better": False, "dtype": "f32", "data": "synthetic (seeded planted instance)",
            "config": {"workload": f"planted n={args.n}, {args.draws} draws/row, nnz={len(inst['value'])}: "
                                   f"{steps} FP32 Jacobi steps per run, three runs per layout"},
            "algorithmic_bytes_per_step": nbytes, "groups": out, "device": hgqp.device_name()}


def tridiagonal_workload(args):
    """The reference's own scale benchmark (scale_benchmark.cpp:13-25): the
    seeded tridiagonal family H = tridiag(-1, 2.02, -1), x in [0, 3], tolerance
    1e-7, timed from the triplet model to the verified result (assembly, upload,
    iterations, checks and the read-back of x), generation excluded, median of
    three. Each n runs the FP64 reference schedule (bit-identical to the CPU path:
    1700 iterations) and the fast modes through hgqp_solve_qp, the reference's CPU
    sources on every usable CPU and on one thread, and holds every result to the
    reference's pass rule (optimal and verified, objective within 1e-8 of the
    planted one, max |x - x*| < 1e-4)."""
    o, kind, source = reference_cpu()
    from oracle import BoxQP
    threads = args.cpu_threads or host_cpu()["usable_cpus"]
    sizes = [int(s) for s in args.tri_sizes.split(",")]
    out = []
    for n in sizes:
        d = hgqp.gen_tridiagonal(n)
        f_star = d["objective"]

        def assess(r, seconds):
            err = float(np.abs(r["primal"] - d["optimum"]).max())
            obj_err = abs(r["objective"] - f_star) / (1 + abs(f_star))
            return {"seconds": seconds, "iterations": int(r["iterations"]), "status": r["status"],
                    "objective": r["objective"], "objective_rel_err": obj_err, "max_abs_err": err,
                    "passed": bool(r["status"] == "optimal" and r["optimality_verified"] and obj_err <= 1e-8
                                   and err < 1e-4)}
        row = {"n": n, "gpu": {}}
        for name in TRIDIAGONAL_MODES:
            precision, momentum, reduce, metric = MODES[name]
            kw = dict(tolerance=1e-7, iteration_limit=20000, time_limit=120, precision=precision, momentum=momentum,
                      reduce=reduce, step_metric=metric)
            hgqp.solve_qp(n, d["hrow"], d["hcol"], d["hval"], d["cost"], d["lower"], d["upper"], **kw)  # warm
            times, r = [], None
            for _ in range(3):
                a = time.perf_counter()
                r = hgqp.solve_qp(n, d["hrow"], d["hcol"], d["hval"], d["cost"], d["lower"], d["upper"], **kw)
                times.append(time.perf_counter() - a)
            row["gpu"][name] = assess(r, float(np.median(times)))
            row["gpu"][name]["algorithm"] = r["algorithm"]
            if name == "fp64_reference":
                gpu_x = r["primal"]
        q = BoxQP(n, d["hrow"], d["hcol"], d["hval"], d["cost"], d["lower"], d["upper"])
        t_all, r_all = cpu_full_solve(o, q, 1e-7, threads)
        t_one, r_one = cpu_full_solve(o, q, 1e-7, 1)
        row["cpu"] = {"threads": threads, "all_threads": assess(r_all, t_all), "one_thread": assess(r_one, t_one)}
        row["fp64_reference_bit_identical_to_cpu"] = bool(
            np.array_equal(gpu_x, r_all["primal"]) and row["gpu"]["fp64_reference"]["iterations"] == r_all["iterations"])
        pub = PUBLISHED_TRIDIAGONAL.get(n)
        if pub:
            row["published"] = {"cpu_1_thread_s": pub[0], "cpu_4_threads_s": pub[1], "t4_gpu_s": pub[2],
                                "source": "reports/kaggle-gpu-03/REPORT.md:19-23 (median of three)"}
            row["speedup_vs_published_t4"] = {k: round(pub[2] / v["seconds"], 2) for k, v in row["gpu"].items()}
        print(f"# tridiagonal n={n}: " + ", ".join(f"{k} {v['seconds'] * 1e3:.1f} ms ({v['iterations']} it)"
                                                 for k, v in row["gpu"].items())
              + f"; cpu {threads} threads {t_all:.3f}s, 1 thread {t_one:.3f}s", file=sys.stderr, flush=True)
        out.append(row)
    # roofline of the FP64 reference step at the largest n (HIP events around the step launches)
    n = sizes[-1]
    d = hgqp.gen_tridiagonal(n)
    start, index, value, L = hgqp.csr_assemble(n, d["hrow"], d["hcol"], d["hval"], d["cost"], d["lower"], d["upper"])
    with hgqp.Box(start, index, value, d["cost"], d["lower"], d["upper"], L, timing=1, use_graph=0) as box:
        rep = box.solve(tolerance=1e-7, iteration_limit=20000, time_limit=120)
    us = rep["step_kernel_ms"] / rep["step_launches"] * 1e3
    nbytes = step_bytes(n, len(value), 8, 8)
    big = out[-1]
    head = max((k for k, v in big["gpu"].items() if v["passed"]), key=lambda k: 1 / big["gpu"][k]["seconds"],
               default="fp64_reference")
    sec = big["gpu"][head]["seconds"]
    return {
        "metric": METRIC, "value": round(1.0 / sec, 3), "unit": "solves/s", "n_gpus": 1, "steps": 3, "warmup": 1,
        "ms_per_step": round(sec * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if MODES[head][0] == hgqp.PRECISION_FP64 else "f32+f64",
        "data": "synthetic (the reference's seeded tridiagonal family, scale_benchmark.cpp:13-18)",
        "config": {"workload": f"the reference's scale benchmark: tridiagonal planted box QP, n in {sizes}, "
                               f"tolerance 1e-7, one solve from the triplet model per step (host arrays in and out); "
                               f"value = fastest passing mode at n={n}", "mode": head},
        "roofline": {"bound": "hbm", "achieved": round(nbytes / us / 1e3, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(nbytes / us / 1e3 / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "step_kernel<double> (FP64 reference step, one panel, one row per thread)", "algorithmic_bytes": nbytes,
                     "avg_launch_us": round(us, 2)},
        "cpu_baseline": {"value": round(1.0 / big["cpu"]["all_threads"]["seconds"], 4), "unit": "solves/s",
                         "cores": threads, "kind": kind,
                         "sample": f"{source}: one full solve per n at the reference's boundary, n={n} shown",
                         "host": host_cpu()},
        "sizes": out, "device": hgqp.device_name(),
    }


def gen_rows_qp(n, m_eq, m_ineq, seed, row_nnz=6, free=0.1):
    """Convex QP with linear rows, feasible by construction: a sparse strictly
    diagonally dominant H (lower triplets, ~1.5 off-diagonals per row), m_eq
    equality rows and m_ineq two-sided inequality rows with row_nnz entries
    each through a point x0 inside the box, bounds [-1, 1] with a fraction
    `free` of the variables unbounded. Rows as CSR (start, col, val)."""
    rng = np.random.default_rng(seed)
    r = rng.integers(0, n, size=3 * n)
    c = rng.integers(0, n, size=3 * n)
    keep = r > c
    r, c = r[keep], c[keep]
    v = rng.normal(0, 0.1, size=len(r))
    radius = np.zeros(n)
    np.add.at(radius, r, np.abs(v))
    np.add.at(radius, c, np.abs(v))
    diag = radius + rng.uniform(0.5, 2.0, size=n)
    lower, upper = -np.ones(n), np.ones(n)
    free_idx = rng.random(n) < free
    lower[free_idx], upper[free_idx] = -np.inf, np.inf
    x0 = rng.uniform(-0.5, 0.5, size=n)
    m = m_eq + m_ineq
    col = np.sort(np.array([rng.choice(n, size=row_nnz, replace=False) for _ in range(m)]), axis=1)
    val = rng.normal(0, 1, size=(m, row_nnz))
    act = np.einsum("ij,ij->i", val, x0[col])
    rl, ru = act.copy(), act.copy()
    rl[m_eq:] -= rng.uniform(0.05, 1.0, size=m_ineq)
    ru[m_eq:] += rng.uniform(0.05, 1.0, size=m_ineq)
    return dict(n=n, hrow=np.concatenate([r, np.arange(n)]), hcol=np.concatenate([c, np.arange(n)]),
                hval=np.concatenate([v, diag]), cost=rng.normal(0, 1, size=n), lower=lower, upper=upper,
                rstart=np.arange(m + 1, dtype=np.int64) * row_nnz, rcol=col.reshape(-1).astype(np.int64),
                rval=val.reshape(-1), row_lower=rl, row_upper=ru, m=m)


def admm_workload(args):
    """Convex QP with linear rows through the GPU ADMM (hgqp_solve_qp_admm),
    against the reference's own CPU ADMM (oracle/_ref) on the same model."""
    n = args.admm_n
    p = gen_rows_qp(n, n // 20, 3 * n // 20, args.seed)
    nnz_h = 2 * (len(p["hval"]) - n) + n
    nnz_a = len(p["rval"]) + n  # the identity rows of the variable bounds join A
    M = p["m"] + n
    keys = ("n", "hrow", "hcol", "hval", "cost", "lower", "upper", "rstart", "rcol", "rval", "row_lower", "row_upper")

    def solve():
        return hgqp.solve_qp_admm(*(p[k] for k in keys), tolerance=KKT_TOL, iteration_limit=20000, time_limit=300,
                                  use_graph=0 if args.no_graph else 1)

    for _ in range(args.warmup):
        solve()
    lat, res = [], []
    for _ in range(args.steps):
        a = time.perf_counter()
        res.append(solve())
        lat.append(time.perf_counter() - a)
    r = res[-1]
    ok = all(x["status"] == "optimal" and x["optimality_verified"] for x in res)
    same = all(np.array_equal(x["primal"], r["primal"]) for x in res)
    wall = float(np.median(lat))
    # one PCG iteration: t = rho (A d), H d, kd = H d + sigma d + A' t, the
    # vector updates (next, r, pre, d) - reads and writes of each array once
    pcg_bytes = (12 * nnz_a + 4 * (M + 1) + 8 * n + 16 * M) + (12 * nnz_h + 4 * (n + 1) + 16 * n) + \
                (12 * nnz_a + 4 * (n + 1) + 8 * M + 24 * n) + 88 * n
    inner = int(r["inner_iterations"])
    achieved = inner * pcg_bytes / r["solve_seconds"] / 1e9
    out = {
        "metric": METRIC, "value": round(1.0 / wall, 3), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded linearly constrained QP, host arrays in)",
        "config": {"workload": f"linearly constrained QP: n={n}, {n // 20} equality + {3 * n // 20} two-sided "
                               f"inequality rows (6 entries each), sparse diagonally dominant H, bounds [-1,1] "
                               f"(10% free), KKT<=1e-6; one step = one complete ADMM solve from host arrays",
                   "nnz_h": nnz_h, "nnz_a_with_bound_rows": nnz_a, "rows": p["m"]},
        "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3), "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "PCG iterations (operator_half, kd_product, pcg_update, pcg_direction): "
                               "algorithmic bytes of all of them over the whole solve time",
                     "algorithmic_bytes_per_pcg_iteration": pcg_bytes, "pcg_iterations": inner},
        "iterations": int(r["iterations"]), "objective": r["objective"], "all_verified": ok,
        "repeat_runs_identical": same, "algorithm": r["algorithm"], "device": hgqp.device_name(),
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle import BoxQP, Oracle, Rows, available
        if available("reference"):
            threads = args.cpu_threads or min(16, os.cpu_count() or 1)
            q = BoxQP(n, p["hrow"], p["hcol"], p["hval"], p["cost"], p["lower"], p["upper"])
            rows = Rows(p["rstart"], p["rcol"], p["rval"], p["row_lower"], p["row_upper"])
            a = time.perf_counter()
            rr = Oracle("reference").solve_admm(q, rows, tolerance=KKT_TOL, time_limit=300, threads=threads)
            t = time.perf_counter() - a
            out["cpu_baseline"] = {
                "value": 1.0 / t, "unit": "solves/s", "cores": threads, "kind": "reference",
                "sample": f"oracle/_ref (the reference's solve_qp_admm), {threads} OpenMP threads, one complete "
                          f"solve of the same model: {rr['status']} ({rr['algorithm']}), {rr['iterations']} "
                          f"iterations, {t:.2f}s",
                "objective_rel_diff_vs_gpu": abs(rr["objective"] - r["objective"]) / (1 + abs(rr["objective"]))}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--draws", type=int, default=16)
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--modes", default=",".join(MODES))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--check-every", type=int, default=10, help="check cadence of the non-reference modes")
    ap.add_argument("--polish-every", type=int, default=5, help="FP64 check cadence after an FP32 phase")
    ap.add_argument("--workload", choices=["single", "batch", "serve", "dense", "admm", "tridiagonal", "slab-groups"],
                    default="single")
    ap.add_argument("--group-steps", type=int, default=200, help="slab-groups: FP32 steps per timed run")
    ap.add_argument("--tri-sizes", default="10000,50000,200000,500000,1000000",
                    help="tridiagonal: the n of the reference's scale benchmark")
    ap.add_argument("--admm-n", type=int, default=100_000, help="admm: variables (rows: n/20 equality, 3n/20 inequality)")
    ap.add_argument("--dense-n", type=int, default=4096)
    ap.add_argument("--dense-k", type=int, default=256)
    ap.add_argument("--dense-check", type=int, default=256, help="dense: RHS re-checked by the oracle")
    ap.add_argument("--instances", type=int, default=512, help="batch: instances")
    ap.add_argument("--plain-sample", type=int, default=512, help="batch: instances timed one at a time in FP64")
    ap.add_argument("--requests", type=int, default=20000, help="serve: requests in the trace")
    ap.add_argument("--rate", type=float, default=200.0, help="serve: Poisson arrival rate per second")
    ap.add_argument("--distinct", type=int, default=0,
                    help="serve: distinct instances the trace draws from; 0 = a distinct instance per request")
    ap.add_argument("--slots", type=int, default=256, help="serve: pool slots")
    ap.add_argument("--budget-rows", type=int, default=8_000_000, help="serve: pool row budget")
    ap.add_argument("--budget-nnz", type=int, default=160_000_000, help="serve: pool entry budget")
    ap.add_argument("--no-plain", action="store_true", help="serve: skip the one-at-a-time FP64 loop")
    ap.add_argument("--pool-slabs", type=int, choices=(0, 1), default=1,
                    help="batch: 1 = FP32 steps over per-instance column slabs, 0 = row blocks")
    ap.add_argument("--pool-switch", type=float, default=None,
                    help="batch/serve: a pool's first FP32 cycle ends at this multiple of the tolerance "
                         "(default: 10 for batch, the library's default, best for throughput; 2 for serve, "
                         "where one cycle more often verifies and latency drops)")
    ap.add_argument("--pool-step", choices=sorted(STEP), default="jacobi",
                    help="batch/serve: step metric of the pool (per-row Jacobi or the scalar 1/L)")
    ap.add_argument("--no-graph", action="store_true",
                    help="serve, admm: launch each kernel directly instead of replaying captured graphs")
    ap.add_argument("--saturate", type=int, default=2000,
                    help="serve: also replay this many requests queued at once (capacity); 0 = skip")
    args = ap.parse_args()
    if args.pool_switch is None:
        args.pool_switch = 2.0 if args.workload == "serve" else 10.0
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (instances would shard "
              f"across GPUs with no exchange; see the design notes).", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch":
        print(json.dumps(batch_workload(args)))
        return
    if args.workload == "serve":
        print(json.dumps(serve_workload(args)))
        return
    if args.workload == "dense":
        print(json.dumps(dense_workload(args)))
        return
    if args.workload == "admm":
        print(json.dumps(admm_workload(args)))
        return
    if args.workload == "tridiagonal":
        print(json.dumps(tridiagonal_workload(args)))
        return
    if args.workload == "slab-groups":
        print(json.dumps(slab_groups_workload(args)))
        return

    device = hgqp.device_name()
    t = time.perf_counter()
    inst = hgqp.gen_planted(args.n, args.draws, args.seed, -1.0, 1.0)
    gen_s = time.perf_counter() - t
    n, nnz = inst["n"], len(inst["value"])
    L = hgqp.lipschitz(inst["start"], inst["value"])
    print(f"# device {device}; instance n={n} nnz={nnz} L={L:.6f} generated in {gen_s:.1f}s", file=sys.stderr,
          flush=True)

    results = {}
    for name in args.modes.split(","):
        r = run_mode(name, inst, L, args.steps, args.warmup, args.check_every, args.polish_every)
        results[name] = r
        rep = r["reps"][-1]
        print(f"# {name}: {args.steps / r['wall']:.2f} solves/s, iterations {rep['iterations']} "
              f"(fp32 {rep['fp32_iterations']}), status {rep['status']}, kkt {rep['stationarity_residual']:.3e}",
              file=sys.stderr, flush=True)

    ref_x = results["fp64_reference"]["x"] if "fp64_reference" in results else None
    modes = {}
    for name, r in results.items():
        reps = r["reps"]
        ok = all(p["status"] == "optimal" and p["optimality_verified"] and p["stationarity_residual"] <= KKT_TOL
                 for p in reps + r["timed"])
        same = all(p["iterations"] == reps[-1]["iterations"] and p["objective"] == reps[-1]["objective"]
                   for p in reps + r["timed"])
        obj = reps[-1]["objective"]
        timed = r["timed"]  # the timed handle's solves (same arithmetic, events around each period)
        f64_ms = sum(p["step_kernel_ms"] for p in timed)
        f64_n = sum(p["step_launches"] for p in timed)
        f32_ms = sum(p["step32_kernel_ms"] for p in timed)
        f32_n = sum(p["step32_launches"] for p in timed)
        lat = np.array(r["lat"]) * 1e3
        m = {
            "solves_per_s": len(reps) / r["wall"],
            "ms_per_solve": r["wall"] / len(reps) * 1e3,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "iterations": int(reps[-1]["iterations"]), "fp32_iterations": int(reps[-1]["fp32_iterations"]),
            "restarts": int(reps[-1]["restarts"]), "checks": int(reps[-1]["checks"]),
            "all_verified": bool(ok), "graph_and_timed_runs_agree": bool(same),
            "kkt": reps[-1]["stationarity_residual"],
            "objective": obj,
            "objective_rel_err_vs_planted": abs(obj - inst["objective"]) / (1 + abs(inst["objective"])),
            "max_abs_err_vs_planted_x": float(np.abs(r["x"] - inst["optimum"]).max()),
            "fp64_step_us": f64_ms / f64_n * 1e3 if f64_n else None,
            "fp32_step_us": f32_ms / f32_n * 1e3 if f32_n else None,
            "column_panels": r["panels"][0], "column_panels32": r["panels"][1], "slabs32": r["slabs"],
            "slab_groups": r["slab_groups"], "jacobi_scale": r["jacobi_scale"],
        }
        if ref_x is not None:
            m["max_abs_err_vs_fp64_reference_x"] = float(np.abs(r["x"] - ref_x).max())
            m["objective_rel_err_vs_fp64_reference"] = (
                abs(obj - results["fp64_reference"]["reps"][-1]["objective"]) / (1 + abs(obj)))
        modes[name] = m

    valid = {k: v for k, v in modes.items() if v["all_verified"] and v["objective_rel_err_vs_planted"] <= 1e-8}
    head = max(valid, key=lambda k: valid[k]["solves_per_s"]) if valid else max(modes, key=lambda k: modes[k]["solves_per_s"])
    hm = modes[head]
    # roofline of the head mode's dominant kernel: its step kernel in the precision that ran most
    use32 = hm["fp32_step_us"] is not None and hm["fp32_iterations"] >= hm["iterations"] - hm["fp32_iterations"]
    us = hm["fp32_step_us"] if use32 else hm["fp64_step_us"]
    nbytes = step_bytes(n, nnz, 4 if use32 else 8, 4 if use32 else 8, jacobi=MODES[head][3] != hgqp.STEP_SCALAR)
    achieved = nbytes / (us * 1e-6) / 1e9
    panels = hm["column_panels32"] if use32 else hm["column_panels"]
    slabs = hm["slabs32"] if use32 else 0
    groups = hm["slab_groups"] if use32 else 0
    tname = "float" if use32 else "double"
    if slabs and groups > 1:
        kernel = (f"slab_sum_kernel + slab_finish_kernel (one FP32 step over {slabs} column slabs staged in LDS, "
                  f"{groups} workgroups per row block)")
    elif slabs:
        kernel = f"slab_step_kernel (one FP32 step over {slabs} column slabs staged in LDS)"
    elif panels == 1:
        kernel = f"step_kernel<{tname}>"
    else:
        kernel = f"step_kernel<{tname}> + {panels - 1} carry_kernel<{tname}> (one step over {panels} column panels)"
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(use32, head, n, args.draws, nnz, panels, slabs, groups),
                "kernel": kernel, "algorithmic_bytes": nbytes, "avg_launch_us": round(us, 2),
                "column_panels": panels, "slabs": slabs, "slab_groups": groups}
    # the FP64 reference-schedule step beside it
    if "fp64_reference" in modes and modes["fp64_reference"]["fp64_step_us"]:
        b64 = step_bytes(n, nnz, 8, 8)
        u64 = modes["fp64_reference"]["fp64_step_us"]
        roofline["fp64_reference_step"] = {"algorithmic_bytes": b64, "avg_launch_us": round(u64, 2),
                                           "achieved": round(b64 / (u64 * 1e-6) / 1e9, 1)}

    pcie = host_inclusive(head, inst, L, args.check_every, args.polish_every)
    whole = whole_model(head, inst, args.check_every, args.polish_every)
    print(f"# whole-model entry: {whole}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or host_cpu()["usable_cpus"]
        cpu = cpu_baseline(inst, threads)
        print(f"# cpu baseline: {cpu['sample']}; {cpu.get('single_thread', {}).get('sample')}", file=sys.stderr,
              flush=True)

    line = {
        "metric": METRIC, "value": round(hm["solves_per_s"], 3), "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(hm["ms_per_solve"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if MODES[head][0] != hgqp.PRECISION_FP64 else "f64",
        "data": "synthetic (seeded planted instance, generated on the host, resident in HBM before timing)",
        "config": {"workload": f"planted box QP, n={n}, {args.draws} off-diagonal draws/row, nnz={nnz}, "
                               f"bounds [-1,1], KKT<=1e-6, one complete solve per step",
                   "mode": head, "seed": args.seed, "check_every": args.check_every,
                   "polish_every": args.polish_every},
        "p50_ms": round(hm["p50_ms"], 3), "p99_ms": round(hm["p99_ms"], 3),
        "roofline": roofline, "cpu_baseline": cpu, "host_inclusive": pcie, "whole_model": whole, "modes": modes,
        "device": device,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
