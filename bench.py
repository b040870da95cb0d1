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
    raise SystemExit(f"unknown workload {name}")


def bytes_csr(n: int, nnz: int, rowptr_bytes: int = 4) -> int:
    return rowptr_bytes * (n + 1) + 12 * nnz


def cpu_baseline(m, op: str, target_s: float) -> dict:
    """Oracle (CPU Gustavson port, OpenMP over rows) on all host cores: the
    full product, repeated until about target_s seconds of CPU work."""
    import oracle
    threads = os.cpu_count() or 1
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
    return {"value": round(2.0 * prods * reps / dt / 1e9, 4), "unit": "GFLOP/s", "cores": threads,
            "kind": "port", "ms_per_call": round(dt * 1e3 / reps, 3),
            "sample": f"the whole {op} product ({m.n} rows, {prods} products) x {reps} calls, {dt:.2f} s, "
                      f"oracle/oracle.c Gustavson (symbolic + numeric passes), OpenMP dynamic rows"}


def main() -> int:
    ap_ = argparse.ArgumentParser()
    ap_.add_argument("--gpus", type=int, default=1)
    ap_.add_argument("--steps", type=int, default=20)
    ap_.add_argument("--warmup", type=int, default=3)
    ap_.add_argument("--workload", default="stencil27_96")
    ap_.add_argument("--op", default="AA", choices=["AA", "AAt", "AAtS"])
    ap_.add_argument("--cpu-seconds", type=float, default=10.0)
    ap_.add_argument("--no-cpu", action="store_true")
    args = ap_.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "this engine has no multi-GPU path", file=sys.stderr)
        return 2

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
    t0 = time.perf_counter()
    for _ in range(args.steps):
        R = eng.compute(A, args.op)
        R.release()
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
    cpu = None if args.no_cpu else cpu_baseline(m, args.op, args.cpu_seconds)
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
        "config": {"workload": desc, "op": args.op, "n": m.n, "nnz_a": m.nnz, "nnz_c": nnz_c,
                   "products": products, "flops": 2 * products, "bytes_alg": b_alg,
                   "call_GBps": round(call_gbs, 2), "call_roofline_frac": round(call_gbs / HBM_PEAK_GBS, 4),
                   "phase_ms": phases},
        "roofline": {"bound": "hbm", "kernel": dom.replace("_ms", ""), "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
