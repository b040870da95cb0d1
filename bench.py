#!/usr/bin/env python3
"""SpGEMM benchmark on one MI355X.

Default workload: C = A*A for the 27-point stencil on a 96^3 grid
(n = 884,736, nnz(A) = 23,393,656, FP64 values U(-1,1), seed 1), the
banded PDE-like configuration of the headline metric.  A "step" is one full
engine call (pre-pass, symbolic, scans, numeric) with A resident in HBM; C is
released after each call, as the reference does (src/spgemm_engine.cuh:199-209).

Prints one JSON line:
  value      GFLOP/s = 2 * products / (ms per call)
  roofline   dominant kernel (numeric phase): algorithmic bytes
             2*bytes(A) + bytes(C) per call / its HIP-event duration, vs 8 TB/s
  cpu_baseline  the CPU Gustavson oracle on a bounded row sample, all host cores
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def make_workload(name: str):
    import appt_spgemm_amd as ap
    if name == "stencil27_96":
        return ap.stencil27(96, 0.0, 1), "27-point stencil 96^3, U(-1,1), seed 1"
    if name == "stencil27_96_drop":
        return ap.stencil27(96, 0.1, 1), "27-point stencil 96^3, off-diagonals dropped p=0.1, seed 1"
    if name == "laplace2d_64":
        return ap.laplace2d(64), "2D 5-point Laplacian 64x64 (4/-1)"
    if name.startswith("rmat"):
        scale = int(name[4:] or 16)
        return ap.rmat(scale, 16, seed=1), f"R-MAT scale {scale}, edge factor 16, (.57,.19,.19,.05), seed 1"
    if name.startswith("er"):
        n, deg = er_params(name)
        return ap.erdos_renyi(n, float(deg), 3), f"Erdos-Renyi n={n}, mean row nnz {deg}, seed 3"
    raise SystemExit(f"unknown workload {name}")


def er_params(name: str):
    """er<n>_<mean row nnz>, e.g. er200000_16."""
    n, deg = name[2:].split("_")
    return int(n), int(deg)


GEN_SPEC = {"stencil27_96": "gen:stencil27:96:0:1", "stencil27_96_drop": "gen:stencil27:96:0.1:1",
            "laplace2d_64": "gen:laplace2d:64"}


def rocsparse_context(workload: str, op: str, reps: int, timeout_s: float = 240.0):
    """rocSPARSE generic SpGEMM on the same matrix, timed as the reference
    times its vendor path (per call: transpose if any, all stages and
    allocations); context only, never the measured value."""
    import subprocess
    spec = GEN_SPEC.get(workload)
    if spec is None and workload.startswith("rmat"):
        spec = f"gen:rmat:{int(workload[4:] or 16)}:16:1"
    if spec is None and workload.startswith("er"):
        n, deg = er_params(workload)
        spec = f"gen:er:{n}:{deg}:3"
    exe = os.path.join(ROOT, "appt_spgemm_amd", "bin", "appt_rocsparse")
    if spec is None or not os.path.exists(exe):
        return None
    try:
        r = subprocess.run([exe, "ctx", spec, op, str(reps)], capture_output=True, text=True, timeout=timeout_s)
    except subprocess.TimeoutExpired:
        return {"error": f"timed out after {timeout_s:.0f} s"}
    line = [l for l in r.stdout.splitlines() if l.startswith("RESULT,")]
    if r.returncode != 0 or not line:
        return {"error": (r.stderr.strip().splitlines() or ["failed"])[-1][:200]}
    f = line[-1].split(",")
    return {"ms": float(f[6]), "nnz_c": int(f[5]), "reps": reps,
            "scope": "csr2csc (A*A^T) + buffer size + nnz + compute + allocations, per call"}


def pmc_traffic(workload: str, op: str):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary for this workload (each
    counter scaled by its reading for a device copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}_{op}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return int(round(1024 * (t["fetch_kib_per_dispatch"] * t["fetch_scale"] +
                             t["write_kib_per_dispatch"] * t["write_scale"])))


def bytes_csr(n: int, nnz: int, rowptr_bytes: int = 4) -> int:
    return rowptr_bytes * (n + 1) + 12 * nnz


def cpu_baseline(m, op: str, target_s: float) -> dict:
    """Oracle (CPU Gustavson port, OpenMP over rows) on all host cores: the
    full product, repeated until about target_s seconds of CPU work."""
    import oracle
    host = host_cpu()
    threads = host["usable_cores"]
    b = None
    if op != "AA":
        b = oracle.transpose(m.n, m.rowptr, m.col, m.val)
    f_all = oracle.row_products(m.n, m.rowptr, m.col, m.rowptr if op == "AA" else b[0])
    prods = int(f_all.sum())
    t0 = time.perf_counter()
    oracle.rows_range(m.n, m.rowptr, m.col, m.val, 0, m.n, op, threads, b)
    dt = max(time.perf_counter() - t0, 1e-6)
    reps = max(1, int(target_s / dt))
    t0 = time.perf_counter()
    for _ in range(reps):
        oracle.rows_range(m.n, m.rowptr, m.col, m.val, 0, m.n, op, threads, b)
    dt = time.perf_counter() - t0
    # one thread on a leading row sample of about a tenth of the work
    r1 = int(np.searchsorted(np.cumsum(f_all), max(prods // 10, 1))) + 1
    r1 = min(max(r1, 1), m.n)
    p1 = int(f_all[:r1].sum())
    t1 = time.perf_counter()
    oracle.rows_range(m.n, m.rowptr, m.col, m.val, 0, r1, op, 1, b)
    d1 = max(time.perf_counter() - t1, 1e-6)
    return {"value": round(2.0 * prods * reps / dt / 1e9, 4), "unit": "GFLOP/s", "cores": threads,
            "kind": "port", "ms_per_call": round(dt * 1e3 / reps, 3),
            "value_1thread": round(2.0 * p1 / d1 / 1e9, 4), "cpu_model": host.get("model", "unknown"),
            "lscpu": host,
            "sample": f"the whole {op} product ({m.n} rows, {prods} products) x {reps} calls, {dt:.2f} s, "
                      f"oracle/oracle.c Gustavson (symbolic + numeric passes), OpenMP dynamic rows; "
                      f"1 thread: rows [0, {r1}) ({p1} products), {d1:.2f} s"}


def cpu_model() -> str:
    """CPU model name as lscpu reports it (/proc/cpuinfo if lscpu is absent)."""
    info = host_cpu()
    return info.get("model", "unknown")


def host_cpu() -> dict:
    """lscpu's model name, CPU and socket counts, and the cores this process
    may run on (its affinity mask: a container's share of the host)."""
    import subprocess
    out = {}
    try:
        r = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10)
        for line in r.stdout.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "Model name":
                out["model"] = v
            elif k == "CPU(s)":
                out["lscpu_cpus"] = int(v)
            elif k == "Socket(s)":
                out["sockets"] = int(v)
            elif k == "Core(s) per socket":
                out["cores_per_socket"] = int(v)
    except (OSError, ValueError, subprocess.TimeoutExpired):
        pass
    if "model" not in out:
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    out["model"] = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
    try:
        out["usable_cores"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        out["usable_cores"] = os.cpu_count() or 1
    return out


def copy_bandwidth(mib: int = 1024, reps: int = 5) -> float:
    """Achievable HBM bandwidth: a device-to-device hipMemcpy of `mib` MiB
    (read + write counted), timed with HIP events on the null stream; GB/s.
    Calls the HIP runtime the engine library already loaded."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    nbytes = mib * (1 << 20)
    x, y, e0, e1 = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
    ok = hip.hipMalloc(C.byref(x), C.c_size_t(nbytes)) == 0 and hip.hipMalloc(C.byref(y), C.c_size_t(nbytes)) == 0
    ms = C.c_float(0.0)
    if ok:
        hip.hipMemset(x, 0, C.c_size_t(nbytes))
        hip.hipEventCreate(C.byref(e0))
        hip.hipEventCreate(C.byref(e1))
        hip.hipMemcpy(y, x, C.c_size_t(nbytes), 3)  # device to device
        hip.hipEventRecord(e0, None)
        for _ in range(reps):
            hip.hipMemcpyAsync(y, x, C.c_size_t(nbytes), 3, None)
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        hip.hipEventElapsedTime(C.byref(ms), e0, e1)
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
    if x:
        hip.hipFree(x)
    if y:
        hip.hipFree(y)
    return 2.0 * nbytes / (ms.value / reps * 1e-3) / 1e9 if ms.value > 0 else float("nan")


def products_of(m, op: str) -> int:
    deg = np.diff(np.asarray(m.rowptr, np.int64))
    if op == "AA":
        return int(deg[np.asarray(m.col, np.int64)].sum()) if m.nnz else 0
    cc = np.bincount(np.asarray(m.col, np.int64), minlength=m.n)
    return int((cc.astype(np.int64) ** 2).sum())


def bench_batch(args) -> dict:
    """Batched uneven workload: 4096 matrices, n log-uniform [64, 16384], half
    random-banded, half Erdos-Renyi, each A*A or A*A^T.  Plain per-matrix calls
    against one planned, graph-replayed batch (autotuned bins)."""
    import appt_spgemm_amd as ap
    from appt_spgemm_amd.workloads import batch_specs, generate
    count = args.batch_count
    specs, ops = batch_specs(count, args.seed)
    t0 = time.perf_counter()
    mats = generate(specs)
    print(f"# generated {count} matrices in {time.perf_counter() - t0:.1f} s, "
          f"sum n {sum(m.n for m in mats)}, sum nnz {sum(m.nnz for m in mats)}", file=sys.stderr, flush=True)
    products = sum(products_of(m, op) for m, op in zip(mats, ops))
    eng = ap.Engine(0)
    dev = [eng.upload(m) for m in mats]
    # plain: one engine call per matrix (inputs resident), C released per call
    nnz_c = 0
    for d, op in zip(dev, ops):
        R = eng.compute(d, op)
        nnz_c += R.nnz
        R.release()
    plain_steps = max(1, args.plain_steps)
    t0 = time.perf_counter()
    for _ in range(plain_steps):
        for d, op in zip(dev, ops):
            R = eng.compute(d, op)
            R.release()
    plain_ms = (time.perf_counter() - t0) * 1e3 / plain_steps
    print(f"# plain: {plain_ms:.2f} ms per pass over {count} matrices", file=sys.stderr, flush=True)
    b = ap.Batch(eng, [m.n for m in mats], [m.nnz for m in mats], ops)
    b.gather(dev)
    b.plan(True)
    tuning = b.autotune(2) if args.autotune else b.info()["tuning"]
    info = b.info()
    for _ in range(max(args.warmup, 1)):
        b.execute()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.execute()
    batch_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    assert info["nnz_c"] == nnz_c, (info["nnz_c"], nnz_c)
    assert info["products"] == products, (info["products"], products)
    b_alg = sum(2 * bytes_csr(m.n, m.nnz) for m in mats) + 4 * (sum(m.n for m in mats) + count) + 12 * nnz_c
    b.close()
    for d in dev:
        d.free()
    eng.close()
    gfl_b = 2.0 * products / (batch_ms * 1e-3) / 1e9
    gfl_p = 2.0 * products / (plain_ms * 1e-3) / 1e9
    achieved = b_alg / (batch_ms * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        import oracle
        threads = host_cpu()["usable_cores"]
        sample = list(range(0, count, max(1, count // 256)))
        sp = sum(products_of(mats[k], ops[k]) for k in sample)
        t0 = time.perf_counter()
        for k in sample:
            m = mats[k]
            oracle.spgemm(m.n, m.rowptr, m.col, m.val, ops[k], threads, False)
        dt = time.perf_counter() - t0
        cpu = {"value": round(2.0 * sp / dt / 1e9, 4), "unit": "GFLOP/s", "cores": threads, "kind": "port",
               "sample": f"{len(sample)} of the {count} matrices (every {max(1, count // 256)}th), full products, "
                         f"{dt:.2f} s, oracle/oracle.c Gustavson, OpenMP dynamic rows"}
    return {
        "metric": "batched SpGEMM aggregate GFLOP/s (FP64, 4096 uneven matrices, one graph-replayed launch set)",
        "value": round(gfl_b, 3), "unit": "GFLOP/s", "ms_per_step": round(batch_ms, 4),
        "dtype": "f64",
        "config": {"workload": f"{count} matrices, n log-uniform [64,16384], even banded (bw U{{1..32}}, fill 0.7), "
                               f"odd Erdos-Renyi (mean deg U(2,20)), op A*A or A*A^T p=1/2, seed {args.seed}",
                   "sum_n": sum(m.n for m in mats), "sum_nnz_a": sum(m.nnz for m in mats), "nnz_c": nnz_c,
                   "products": products, "bytes_alg": b_alg, "plain_ms_per_pass": round(plain_ms, 3),
                   "plain_GFLOPs": round(gfl_p, 3), "batched_over_plain": round(plain_ms / batch_ms, 3),
                   "tuning": tuning, "bins_used": info["bins_used"], "graph": info["graph"]},
        "roofline": {"bound": "hbm", "kernel": "batched launch set (whole graph)", "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None},
        "cpu_baseline": cpu,
    }


def extras(args):
    """The batched and served forms beside the headline: the 4096-matrix batch
    (plain per-matrix calls against the graph-replayed batch) and the served
    stream at two fixed offered rates, 1000 jobs each."""
    import copy
    a = copy.copy(args)
    a.steps, a.warmup, a.no_cpu = 10, 2, True
    t0 = time.perf_counter()
    b = bench_batch(a)
    bc = b["config"]
    batched = {"matrices": args.batch_count, "batched_GFLOPs": b["value"], "batched_ms_per_pass": b["ms_per_step"],
               "plain_GFLOPs": bc["plain_GFLOPs"], "plain_ms_per_pass": bc["plain_ms_per_pass"],
               "batched_over_plain": bc["batched_over_plain"], "products": bc["products"], "nnz_c": bc["nnz_c"],
               "roofline_frac": b["roofline"]["frac"], "seconds": round(time.perf_counter() - t0, 1)}
    print(f"# batched: {json.dumps(batched)}", file=sys.stderr, flush=True)
    s = copy.copy(args)
    s.serve_jobs, s.serve_qps = 1000, "10,20"
    t0 = time.perf_counter()
    r = bench_serve(s)
    served = {"slots": r["config"]["slots"], "hbm_budget_GB": r["config"]["hbm_budget_GB"],
              "pool": r["config"]["workload"], "rates": r["config"]["rates"],
              "seconds": round(time.perf_counter() - t0, 1)}
    print(f"# served: {json.dumps(served)}", file=sys.stderr, flush=True)
    return batched, served


def bench_panels(args) -> dict:
    """R-MAT in row panels: exact row pointers from the symbolic phase, then
    C in panels of at most --panel-gb each, every panel timed, spot-checked
    against the oracle on sampled row ranges, and released."""
    import appt_spgemm_amd as ap
    import oracle
    scale = int(args.workload[4:].split("_")[0])
    t0 = time.perf_counter()
    m = ap.rmat(scale, 16, seed=1)
    print(f"# R-MAT scale {scale}: n={m.n} nnz={m.nnz} generated in {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    op = args.op
    eng = ap.Engine(0)
    A = eng.upload(m)
    t0 = time.perf_counter()
    rp = eng.symbolic_rowptr(A, op)
    sym_ms = (time.perf_counter() - t0) * 1e3
    nnz_c = int(rp[-1])
    parts = ap.panels(rp, int(args.panel_gb * 1e9))
    print(f"# nnz(C)={nnz_c} ({12 * nnz_c / 1e9:.1f} GB), {len(parts)} panels, symbolic {sym_ms:.1f} ms",
          file=sys.stderr, flush=True)
    b_host = None if op == "AA" else oracle.transpose(m.n, m.rowptr, m.col, m.val)
    rng = np.random.default_rng(args.seed)
    eng.set_timing(True)
    panel_ms, products, checked, worst = [], 0, 0, 0.0
    for k, (r0, r1) in enumerate(parts):
        t1 = time.perf_counter()
        R = eng.compute_rows(A, op, r0, r1)
        ms = (time.perf_counter() - t1) * 1e3
        pt = eng.phase_times()
        products += int(pt["products"])
        panel_ms.append(ms)
        assert R.nnz == int(rp[r1] - rp[r0]), (k, R.nnz)
        # spot check: 4 ranges of up to 8 rows inside the panel
        for _ in range(4):
            a = int(rng.integers(r0, r1))
            b = min(r1, a + 8)
            got = R.download_rows(a - r0, b - r0)
            ro, rc, rv, rs = oracle.rows_range(m.n, m.rowptr, m.col, m.val, a, b, op, 0, b_host, want_scale=True)
            assert np.array_equal(got.rowptr, ro) and np.array_equal(got.col, rc), f"structure differs, rows {a}:{b}"
            err = np.abs(got.val - rv)
            assert np.all(err <= 1e-12 * rs), f"values differ, rows {a}:{b}"
            worst = max(worst, float(np.max(err / np.maximum(rs, 1e-300))) if len(err) else 0.0)
            checked += b - a
        R.release()
        print(f"# panel {k}: rows [{r0},{r1}) nnz {int(rp[r1] - rp[r0])} {ms:.1f} ms "
              f"(numeric {pt['numeric_ms']:.1f}, symbolic {pt['symbolic_ms']:.1f})", file=sys.stderr, flush=True)
    A.free()
    eng.close()
    total_ms = sum(panel_ms)
    b_alg = 2 * bytes_csr(m.n, m.nnz) * len(parts) + 8 * (m.n + len(parts)) + 12 * nnz_c
    return {
        "metric": "SpGEMM GFLOP/s (FP64, structural) over row panels, with HBM GB/s vs roofline",
        "value": round(2.0 * products / (total_ms * 1e-3) / 1e9, 3), "unit": "GFLOP/s",
        "ms_per_step": round(total_ms, 3), "dtype": "f64",
        "config": {"workload": f"R-MAT scale {scale}, edge factor 16, (.57,.19,.19,.05), seed 1, in row panels",
                   "op": op, "n": m.n, "nnz_a": m.nnz, "nnz_c": nnz_c, "products": products,
                   "panels": len(parts), "panel_GB_max": args.panel_gb, "panel_ms": [round(x, 2) for x in panel_ms],
                   "symbolic_plan_ms": round(sym_ms, 2), "rows_checked": checked,
                   "worst_err_over_scale": worst, "bytes_alg": b_alg},
        "roofline": {"bound": "hbm", "kernel": "all panels (whole calls)",
                     "achieved": round(b_alg / (total_ms * 1e-3) / 1e9, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(b_alg / (total_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "traffic": None},
        "cpu_baseline": None,
    }


def bench_serve(args) -> dict:
    """Served stream: jobs from the Laplacian / stencil / R-MAT generators, n
    log-uniform in [1e3, 2e6], A*A or A*A^T, resident on the device; a closed
    burst measures saturation, then Poisson arrivals at 50 % and 80 % of it."""
    import appt_spgemm_amd as ap
    from appt_spgemm_amd.workloads import generate, serve_job
    rng = np.random.default_rng(args.seed)
    pool_n = args.serve_pool
    jobs = generate([lambda r=np.random.default_rng(args.seed * 7919 + k), k=k:
                         serve_job(r, 1000, args.serve_nmax, seed=k) for k in range(pool_n)], workers=8)
    eng = ap.Engine(0)
    dev = [eng.upload(m) for m, _, _ in jobs]
    print(f"# serve pool: {pool_n} matrices, n " + ",".join(str(m.n) for m, _, _ in jobs), file=sys.stderr,
          flush=True)

    # one job sequence and one set of unit inter-arrival gaps for every run,
    # so the burst and the Poisson runs serve the same mix
    picks = rng.integers(0, pool_n, args.serve_jobs)
    gaps = rng.exponential(1.0, args.serve_jobs)
    ops = [jobs[k][1] for k in picks]

    def run(qps: float | None, count: int) -> dict:
        import queue
        import threading
        with ap.Server(0, slots=args.slots, hbm_budget_bytes=200 * 10 ** 9) as srv:
            # a client releases each result as soon as its job is done, so
            # results do not pile up while later jobs arrive
            done_q: queue.Queue = queue.Queue()

            def releaser():
                n_done = 0
                while True:
                    j = done_q.get()
                    if j is None:
                        return
                    srv.wait(j)
                    srv.release(j)
                    n_done += 1
                    if n_done % 250 == 0:
                        print(f"# serve {qps or 'burst'}: {n_done}/{count} jobs done", file=sys.stderr, flush=True)

            th = threading.Thread(target=releaser)
            th.start()
            t0 = time.perf_counter()
            t_next = t0
            try:
                for n_k, k in enumerate(picks[:count]):
                    if qps:
                        t_next += gaps[n_k] / qps
                        dt = t_next - time.perf_counter()
                        if dt > 0:
                            time.sleep(dt)
                    done_q.put(srv.submit(dev[k], ops[n_k]))
            finally:
                done_q.put(None)
                th.join()
            st = srv.stats()
        return {"offered_qps": qps, "jobs": count, "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3),
                "mean_ms": round(st["mean_ms"], 3), "jobs_per_s": round(st["jobs_per_s"], 3),
                "GFLOPs": round(st["gflops"], 3), "refused": st["refused"], "failed": st["failed"],
                "groups_run": st["groups_run"], "peak_reserved_GB": round(st["peak_bytes_reserved"] / 1e9, 3),
                "peak_device_GB": round(st["peak_device_bytes"] / 1e9, 3)}

    if args.serve_qps:
        # fixed offered rates (jobs/s)
        sat = None
        rates = [run(float(q), args.serve_jobs) for q in args.serve_qps.split(",")]
    else:
        sat = run(None, args.serve_jobs)
        rates = [run(f * sat["jobs_per_s"], args.serve_jobs) for f in (0.5, 0.8)]
    for d in dev:
        d.free()
    eng.close()
    return {
        "metric": "served SpGEMM p50/p99 latency at fixed offered QPS, with throughput",
        "value": rates[-1]["p99_ms"], "unit": f"ms (p99 at {rates[-1]['offered_qps']:.1f} jobs/s offered)",
        "higher_is_better": False,
        "dtype": "f64",
        "config": {"workload": f"pool of {pool_n} resident jobs, Laplacian / 27-pt stencil / R-MAT, n log-uniform "
                               f"[1e3, {args.serve_nmax:g}], A*A or A*A^T, seed {args.seed}",
                   "slots": args.slots, "hbm_budget_GB": 200, "saturation": sat, "rates": rates},
        "roofline": None,
        "cpu_baseline": None,
    }


def main() -> int:
    ap_ = argparse.ArgumentParser()
    ap_.add_argument("--gpus", type=int, default=1)
    ap_.add_argument("--steps", type=int, default=20)
    ap_.add_argument("--warmup", type=int, default=3)
    ap_.add_argument("--workload", default="stencil27_96")
    ap_.add_argument("--op", default="AA", choices=["AA", "AAt", "AAtS"])
    ap_.add_argument("--cpu-seconds", type=float, default=10.0)
    ap_.add_argument("--no-cpu", action="store_true")
    ap_.add_argument("--seed", type=int, default=1)
    ap_.add_argument("--batch-count", type=int, default=4096)
    ap_.add_argument("--plain-steps", type=int, default=2)
    ap_.add_argument("--autotune", type=int, default=1)
    ap_.add_argument("--slots", type=int, default=4)
    ap_.add_argument("--serve-pool", type=int, default=24)
    ap_.add_argument("--serve-jobs", type=int, default=200)
    ap_.add_argument("--serve-nmax", type=float, default=2e6)
    ap_.add_argument("--serve-qps", default="", help="fixed offered rates, e.g. 50,100 (default: 50 %% and 80 %% "
                                                    "of the measured saturation)")
    ap_.add_argument("--extras", type=int, default=1,
                     help="default workload: also measure the batched (4096 matrices) and served (2 fixed rates, "
                          "1000 jobs each) forms and report them under config")
    ap_.add_argument("--copy-probe", type=int, default=0, help="MiB; run only a device copy (counter calibration)")
    ap_.add_argument("--rocsparse", type=int, default=1, help="also time rocSPARSE on the same matrix (context)")
    ap_.add_argument("--panel-gb", type=float, default=60.0, help="C bytes per row panel (rmatN_panels)")
    args = ap_.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "this engine has no multi-GPU path", file=sys.stderr)
        return 2
    if args.copy_probe:
        # calibration for rocprofv3 FETCH_SIZE / WRITE_SIZE: one device copy of
        # a known size (reads and writes copy_probe MiB each)
        import torch
        n = args.copy_probe * (1 << 20) // 8
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_MiB": args.copy_probe, "dispatches": 3}), flush=True)
        return 0
    if args.workload in ("batch4096", "serve") or args.workload.endswith("_panels"):
        res = (bench_batch(args) if args.workload == "batch4096" else
               bench_serve(args) if args.workload == "serve" else bench_panels(args))
        out = {"metric": res["metric"], "value": res["value"], "unit": res["unit"], "n_gpus": 1,
               "steps": args.steps, "warmup": args.warmup, "ms_per_step": res.get("ms_per_step"),
               "higher_is_better": res.get("higher_is_better", True), "scaling": "weak", "vs_baseline": None,
               "dtype": res["dtype"], "data": "synthetic", "config": res["config"], "roofline": res["roofline"],
               "cpu_baseline": res["cpu_baseline"]}
        print(json.dumps(out), flush=True)
        return 0

    import appt_spgemm_amd as ap
    m, desc = make_workload(args.workload)
    print(f"# workload {args.workload} op {args.op}: n={m.n} nnz={m.nnz}", file=sys.stderr, flush=True)
    eng = ap.Engine(0)
    A = eng.upload(m)
    # warmup (also discovers nnz(C) and products)
    for _ in range(max(args.warmup, 1)):
        R = eng.compute(A, args.op)
        nnz_c = R.nnz
        R.release()
    eng.set_timing(True)
    phase_sum: dict[str, float] = {}
    step_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        R = eng.compute(A, args.op)
        R.release()
        step_ms.append((time.perf_counter() - ts) * 1e3)
        pt = eng.phase_times()
        for k in ("transpose_ms", "prepass_ms", "symbolic_ms", "scan_ms", "numeric_ms", "mirror_ms",
                  "total_ms"):
            phase_sum[k] = phase_sum.get(k, 0.0) + pt[k]
    wall = time.perf_counter() - t0
    products = int(pt["products"])
    eng.set_timing(False)
    ms = wall * 1e3 / args.steps
    phases = {k: round(v / args.steps, 4) for k, v in phase_sum.items()}
    b_alg = 2 * bytes_csr(m.n, m.nnz) + bytes_csr(m.n, nnz_c, 4 if nnz_c < 2 ** 31 else 8)
    dom = max(("symbolic_ms", "numeric_ms", "prepass_ms", "transpose_ms", "mirror_ms"),
              key=lambda k: phases.get(k, 0.0))
    dom_ms = phases[dom]
    achieved = b_alg / (dom_ms * 1e-3) / 1e9
    gflops = 2.0 * products / (ms * 1e-3) / 1e9
    call_gbs = b_alg / (ms * 1e-3) / 1e9
    A.free()
    eng.close()
    roc = rocsparse_context(args.workload, args.op, 5) if args.rocsparse else None
    if roc and "ms" in roc:
        roc["GFLOPs"] = round(2.0 * products / (roc["ms"] * 1e-3) / 1e9, 3)
        roc["nnz_c_matches"] = roc["nnz_c"] == nnz_c
    cpu = None if args.no_cpu else cpu_baseline(m, args.op, args.cpu_seconds)
    copy_gbs = copy_bandwidth()
    batched = served = None
    n_rows, nnz_a = m.n, m.nnz
    if args.extras and args.workload == "stencil27_96":
        del m
        batched, served = extras(args)
    out = {
        "metric": "SpGEMM GFLOP/s (FP64, structural C=A*A) with HBM GB/s vs roofline",
        "value": round(gflops, 3),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": desc, "op": args.op, "n": n_rows, "nnz_a": nnz_a, "nnz_c": nnz_c,
                   "median_ms_per_step": round(float(np.median(step_ms)), 4),
                   "products": products, "flops": 2 * products, "bytes_alg": b_alg,
                   "call_GBps": round(call_gbs, 2), "call_roofline_frac": round(call_gbs / HBM_PEAK_GBS, 4),
                   "phase_ms": phases, "rocsparse": roc, "batched": batched, "served": served},
        "roofline": {"bound": "hbm", "kernel": dom.replace("_ms", ""), "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": pmc_traffic(args.workload, args.op) if dom == "numeric_ms" else None,
                     "copy_peak": round(copy_gbs, 1), "frac_of_copy": round(achieved / copy_gbs, 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
