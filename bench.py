#!/usr/bin/env python3
"""Headline benchmark: plain GPU FTE top-k queries/s on one MI355X.

Workload (BASELINE "Plain GPU FTE + radix select"): N = 2^26 objects,
d = 8 FP32 attributes, independent U[0,1) data (seed 0) generated in HBM,
query on all m = 8 attributes with weights 1.0, k = 64.  One step = one
query (prefix pass + stream pass + final select), outputs left in HBM.

Prints ONE JSON line (rank 0).  Roofline: the dominant kernel is the stream
pass; its algorithmic bytes per launch are 4*m bytes per streamed object
times the objects it streams, divided by its average duration measured with
HIP events on the engine's stream.  cpu_baseline: the oracle's SIMD+threads
TPAc restatement on the same data on this host's cores.
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
N_DEFAULT = 1 << 26
D = 8
K = 64


def load_profile_traffic():
    """HBM bytes per stream-pass launch from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel
    calibration), if one is present."""
    path = os.path.join(ROOT, "profiles", "fte_stream_pmc.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch ="):
                    return float(line.split("=", 1)[1].strip())
    except (OSError, ValueError):
        pass
    return None


def calibrate_copy(nbytes):
    """A plain device copy of nbytes (torch's copy kernel), the reference the
    PMC byte counters are scaled by."""
    import torch
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"calibration_copy_bytes": nbytes, "launches": 3}), flush=True)


def cpu_baseline(x, n, m, k, threads, budget_s, gpu_ids, gpu_scores):
    import oracle
    w = np.ones(D, np.float32)
    attr = list(range(D - m, D))
    oracle.tpac_threads(x, n, k, w, attr, threads)  # warm-up (reference: one warm-up call)
    times = []
    ids = sc = None
    t_start = time.time()
    while not times or (time.time() - t_start < budget_s and len(times) < 50):
        t0 = time.perf_counter()
        ids, sc, _ = oracle.tpac_threads(x, n, k, w, attr, threads)
        times.append(time.perf_counter() - t0)
    mism = int(np.sum(ids != gpu_ids) + np.sum(sc.view(np.uint32) != gpu_scores.view(np.uint32)))
    return {"value": 1.0 / float(np.mean(times)), "unit": "queries/s", "cores": threads, "kind": "port",
            "sample": f"full workload (N={n}, d={D}, m={m}, k={k}, same data); "
                      f"{len(times)} timed queries after 1 warm-up, mean; oracle or_tpac_threads "
                      f"(AVX2 16-wide + {threads} threads, cpu/TPAc.h:256-358 restated)",
            "median_s": float(np.median(times))}, mism


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=N_DEFAULT)
    ap.add_argument("--m", type=int, default=D)
    ap.add_argument("--k", type=int, default=K)
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--lanes", type=int, default=2,
                    help="pipeline lanes for back-to-back queries (1 = no overlap between queries)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run a plain copy of BYTES bytes (PMC calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); "
              "the sharded design is described in DESIGN.md", file=sys.stderr)
        sys.exit(2)

    import torch
    import topk_amd

    n, m, k = args.n, args.m, args.k
    attr = list(range(D - m, D))
    w = np.ones(D, np.float32)
    eng = topk_amd.Engine.generated(n, D, topk_amd.INDEPENDENT, 0)
    eng.set_tuning("lanes", args.lanes)
    ids_d = torch.empty(k, dtype=torch.int64, device="cuda")
    sc_d = torch.empty(k, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()

    for _ in range(args.warmup):
        eng.fte_async(k, w, attr, ids_d.data_ptr(), sc_d.data_ptr())
    eng.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.fte_async(k, w, attr, ids_d.data_ptr(), sc_d.data_ptr())
    eng.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    qps = args.steps / elapsed

    # dominant kernel duration with HIP events on the engine's stream (timing
    # mode runs every query on lane 0, so launches do not overlap)
    eng.timing(True)
    eng.timing_reset()
    for _ in range(args.steps):
        eng.fte_async(k, w, attr, ids_d.data_ptr(), sc_d.data_ptr())
    eng.sync()
    stream_ms, launches = eng.timing_get("fte_stream")
    eng.timing(False)
    avg_ms = stream_ms / max(1, launches)
    # the sample-seeded pipeline streams every object (the 128 x 1024 sampled
    # objects are scored again by the stream pass)
    streamed = n
    alg_bytes = 4.0 * m * streamed
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    traffic = load_profile_traffic()

    # single-query latency through the synchronous C ABI (host buffers)
    lat = []
    for _ in range(10):
        t1 = time.perf_counter()
        gpu_ids, gpu_sc, _ = eng.fte(k, w, attr)
        lat.append(time.perf_counter() - t1)
    st = eng.stats()

    out = {
        "metric": "queries/s + HBM GB/s, N=2^26 d=8 k=64: learned/batched vs plain GPU FTE",
        "value": qps,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (independent U[0,1), seed 0, generated in HBM)",
        "config": {"workload": "plain GPU FTE top-k", "n": n, "d": D, "m": m, "k": k,
                   "weights": "1.0", "distribution": "independent"},
        "query_hbm_gbs": 4.0 * m * n / (ms_per_step * 1e-3) / 1e9,
        "query_roofline_frac": 4.0 * m * n / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "fte_stream_kernel", "avg_launch_ms": avg_ms,
                     "alg_bytes_per_launch": alg_bytes},
        "fallback": st["fallback"],
        "lanes": args.lanes,
        "sync_query_latency_ms": float(np.median(lat)) * 1e3,
        "candidates_per_query": st["candidates"],
    }

    if not args.no_cpu:
        x = topk_amd.generate(n, D, topk_amd.INDEPENDENT, 0, threads=min(os.cpu_count() or 1, 16))
        threads = min(os.cpu_count() or 1, 16)
        cpu, mism = cpu_baseline(x, n, m, k, threads, args.cpu_budget, gpu_ids, gpu_sc)
        out["cpu_baseline"] = cpu
        out["mismatches_vs_oracle"] = mism
        del x
    print(json.dumps(out), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
