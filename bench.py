#!/usr/bin/env python3
"""Benchmark of the TaiChi binary SpMV hot path on one MI355X.

Headline workload: plain SpMV on R-MAT n=2^22, edge factor 16 (67.1M draws),
(a,b,c,d)=(0.499,0.001,0.001,0.499), seed 1, self loops kept, duplicates
removed; x FP32 uniform [0,1).  One step = one y = A x over the resident
format.  Prints one JSON line.

  python bench.py [--steps 50] [--warmup 10] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "GNNZ/s + % HBM roofline per SpMV; served req/s + p99 latency vs plain spmv"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(rp, col, x, seconds=10.0):
    """Reference CPU CSR SpMV (main.cu:61-69), single core, from the oracle."""
    import oracle
    nnz = len(col)
    times = []
    t_end = time.perf_counter() + seconds
    while time.perf_counter() < t_end or len(times) < 5:
        t0 = time.perf_counter()
        oracle.csr_spmv_f32(rp, col, x)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    return {"value": nnz / med / 1e9, "unit": "GNNZ/s", "cores": 1, "kind": "port",
            "sample": f"full R-MAT 2^22 matrix ({nnz} nnz), {len(times)} sequential FP32 "
                      f"CSR SpMVs, median {med * 1e3:.2f} ms"}


def read_traffic(name):
    """HBM bytes per launch measured with rocprofv3 --pmc (profiles/<name>)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--autotune", type=int, default=1)
    ap.add_argument("--config", default="",
                    help="rows_per_tile,lds_values,grid[,flags] (skips the autotuner)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after timing, run 5 plain 256 MiB device copies (profiler calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (this build has no multi-GPU path); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    import torch
    import taichi_amd as T

    dev = torch.device("cuda:0")
    n = 1 << args.scale
    ef = 16
    t0 = time.time()
    rp, col = T.gen_rmat(n, args.scale, ef * n, 0.499, 0.001, 0.001, 1)
    nnz = len(col)
    log(f"generated R-MAT 2^{args.scale}: nnz={nnz} in {time.time() - t0:.1f}s")
    t0 = time.time()
    fmt, _ = T.partition_shapes(n, n, rp, col)
    log(f"partition (GPU thumbnail + host): {time.time() - t0:.2f}s S={fmt.dia_shapes} "
        f"D={fmt.num_negs} dense={fmt.dense_nnz} Z={fmt.zero_elements} P={fmt.sparse_nnz}")
    A = T.Matrix(fmt, nnz=nnz)
    st = A.stats()
    x = T.gen_vector(n, 0, 12345)
    xd = torch.from_numpy(x).to(dev)
    yd = torch.empty(n, dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()
    if args.config:
        r, c, g, fl = (list(int(v) for v in args.config.split(",")) + [0])[:4]
        A.set_config(1, T.F32, r, c, g, fl)
        log(f"config: rows_per_tile={r} lds_values={c} grid={g} flags={fl}")
    elif args.autotune:
        cfg, us = A.autotune(xd, yd)
        log(f"autotune: rows_per_tile={cfg[0]} lds_values={cfg[1]} grid={cfg[2]} "
            f"flags={cfg[3]} -> {us:.2f} us")

    # correctness spot check against the FP64 CSR product
    A.spmv(xd, yd)
    torch.cuda.synchronize()
    import oracle
    y64 = oracle.csr_spmv_f64(rp, col, x.astype(np.float64))
    absrow = oracle.csr_absrow_f64(rp, col, x.astype(np.float64))
    err = oracle.rel_err(yd.cpu().numpy(), y64, absrow)
    log(f"max componentwise relative error vs FP64: {err:.3e}")

    for _ in range(args.warmup):
        A.spmv(xd, yd)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        A.spmv(xd, yd)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms = e0.elapsed_time(e1) / args.steps
    t_s = ms * 1e-3

    # the same SpMVs replayed from a hipGraph (launch overhead removed)
    g = T.Graph(A, xd, yd, iters=args.steps, stream=stream.cuda_stream)
    g.launch()
    torch.cuda.synchronize()
    e0.record(stream)
    g.launch()
    e1.record(stream)
    torch.cuda.synchronize()
    ms_graph = e0.elapsed_time(e1) / args.steps
    g.close()

    # cold: a 1 GiB scrub between SpMVs evicts the format from the MALL
    scrub = torch.empty(1 << 28, dtype=torch.float32, device=dev)
    cold = []
    for _ in range(10):
        scrub.add_(1.0)
        e0.record(stream)
        A.spmv(xd, yd)
        e1.record(stream)
        torch.cuda.synchronize()
        cold.append(e0.elapsed_time(e1))
    ms_cold = float(np.median(cold))
    del scrub
    if args.calibrate_copy:
        src = torch.ones(1 << 26, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    B = A.bytes(1, T.F32)
    achieved = B / t_s / 1e9
    gnnz = nnz / t_s / 1e9
    out = {
        "metric": METRIC,
        "value": round(gnnz, 3),
        "unit": "GNNZ/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded R-MAT generator, no dataset download)",
        "config": {"workload": "rmat22_plain_spmv", "n": n, "edge_factor": ef,
                   "abcd": [0.499, 0.001, 0.001, 0.499], "seed": 1, "nnz": nnz,
                   "dia_shapes": fmt.dia_shapes, "num_negs": fmt.num_negs,
                   "remainder": fmt.remainder, "format_bytes": st["format_bytes"],
                   "offset_bytes": st["offset_bytes"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": read_traffic("spmv_rmat22_traffic.txt"),
                     "algorithmic_bytes": B},
        "graph_ms_per_step": round(ms_graph, 5),
        "graph_gnnz_per_s": round(nnz / (ms_graph * 1e-3) / 1e9, 3),
        "cold_ms_per_step": round(ms_cold, 5),
        "max_rel_err_vs_fp64": err,
        "wall_s_timed": round(wall, 4),
    }
    if not args.no_cpu:
        log("timing the CPU baseline")
        out["cpu_baseline"] = cpu_baseline(rp, col, x, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
