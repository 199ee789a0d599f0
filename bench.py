#!/usr/bin/env python3
"""Headline benchmark on one MI355X: learned-threshold batched top-k
queries/s against plain GPU FTE queries/s, N = 2^26, d = 8, k = 64.

Store: N = 2^26 objects, d = 8 FP32 attributes, independent U[0,1) data
(seed 0) generated in HBM.

* plain FTE (the comparison point): back-to-back single queries on all 8
  attributes (k = 64) through two pipeline lanes; its stream kernel is the
  HBM-bound kernel (4*m bytes per object), timed with HIP events on the
  engine's stream -> ``plain_fte.roofline``.
* learned + batched (``value``): partition layout (``topk_init`` PTA), a
  k-th score predictor trained on random-weight queries labelled by the
  exact batch path, then timed batches of 1024 held-out queries (weights
  U(0.1, 1) on all 8 attributes, k = 64).  One step = one batch; two
  batches are kept in flight (topk_batch_submit / topk_batch_wait), and the
  rate of one synchronous topk_batch call per batch is reported beside it.  Every
  held-out answer is judged against the plain GPU FTE path (``mismatches``);
  the fallback rate is reported.  The dominant kernel is the batch scoring
  kernel, bound by FP32 VALU issue (2*m flops per scored (object, query)
  pair) -> ``roofline``.
* cpu_baseline: the oracle's SIMD+threads TPAc restatement on the same store
  and held-out queries, on the host's cores, within a time budget.

``--mode fte`` runs the plain FTE part alone (one step = one query); it is
the command the FTE rocprofv3 summaries under profiles/ were taken with.
``--mode ingest`` times data ingest alone (one step = one file load),
``--mode stream`` a store in host memory streamed through HBM, ``--mode
kernels`` BTA, GPA, radix select, normalisation and reorder_attr, and
``--mode configs`` the other workloads (PTA at N = 2^28, batched and learned
d = 16, ingest, host-memory store); ``--mode train`` the predictor's training
alone on the headline store (the command of the training kernel summary),
then the same training at mini-batches of 1024, 2048 and 4096 samples.
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
FP32_PEAK_TFLOPS = 157.3      # published MI355X FP32 vector peak (packed FMA)
FP32_ISSUE_TFLOPS = 39.3      # 256 CUs x 64 lanes x 2.4 GHz: one non-packed mul or add per lane-cycle
N_DEFAULT = 1 << 26
D = 8
K = 64
METRIC = "queries/s + HBM GB/s, N=2^26 d=8 k=64: learned/batched vs plain GPU FTE"


def load_profile_traffic(name):
    """HBM bytes per launch from a committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration)."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch ="):
                    return float(line.split("=", 1)[1].strip())
    except (OSError, ValueError):
        pass
    return None


def calibrate_copy(nbytes):
    """A plain device copy of nbytes, the reference the PMC byte counters are
    scaled by."""
    import torch
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"calibration_copy_bytes": nbytes, "launches": 3}), flush=True)


def plain_fte(eng, n, m, k, steps, warmup, lanes):
    import torch
    attr = list(range(D - m, D))
    w = np.ones(D, np.float32)
    eng.set_tuning("lanes", lanes)
    ids_d = torch.empty(k, dtype=torch.int64, device="cuda")
    sc_d = torch.empty(k, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    for _ in range(warmup):
        eng.fte_async(k, w, attr, ids_d.data_ptr(), sc_d.data_ptr())
    eng.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.fte_async(k, w, attr, ids_d.data_ptr(), sc_d.data_ptr())
    eng.sync()
    elapsed = time.perf_counter() - t0
    # stream-kernel duration with HIP events on the engine's stream (timing
    # mode runs every query on lane 0, so launches do not overlap)
    eng.timing(True)
    eng.timing_reset()
    for _ in range(max(steps // 2, 10)):
        eng.fte_async(k, w, attr, ids_d.data_ptr(), sc_d.data_ptr())
    eng.sync()
    stream_ms, launches = eng.timing_get("fte_stream")
    eng.timing(False)
    avg_ms = stream_ms / max(1, launches)
    alg_bytes = 4.0 * m * n  # the sample-seeded pipeline streams every object
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    lat = []
    for _ in range(10):
        t1 = time.perf_counter()
        eng.fte(k, w, attr)
        lat.append(time.perf_counter() - t1)
    qps = steps / elapsed
    return {
        "qps": qps, "ms_per_query": elapsed * 1e3 / steps, "steps": steps, "lanes": lanes, "m": m, "k": k,
        "query_hbm_gbs": 4.0 * m * n / (elapsed / steps) / 1e9,
        "sync_query_latency_ms": float(np.median(lat)) * 1e3,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": load_profile_traffic("fte_stream_pmc.md"),
                     "kernel": "fte_stream_kernel", "avg_launch_ms": avg_ms, "alg_bytes_per_launch": alg_bytes},
    }


def train_step_stats(rep):
    """Mini-batch steps of the predictor's training: count, device time per
    step (HIP events around all launches, gaps included) and the rate over
    6 x 10433 flops per sample."""
    ms = rep["ms_steps"]
    return {"steps": rep["steps"], "step_us": ms * 1e3 / max(1, rep["steps"]),
            "step_tflops": rep["flops"] / (ms * 1e-3) / 1e12 if ms > 0 else 0.0}


def random_queries(rng, count, k):
    return [(k, rng.uniform(0.1, 1.0, D).astype(np.float32), np.arange(D, dtype=np.uint8)) for _ in range(count)]


def learned_batched(eng, n, k, args):
    import torch
    import topk_amd
    rng = np.random.default_rng(2026)
    nb = args.batch
    t0 = time.perf_counter()
    eng.init(topk_amd.TOPK_INIT_PTA)
    tt_layout = time.perf_counter() - t0
    train = random_queries(rng, args.train_queries, k)
    rep = eng.learned_train(train)
    held = [eng.prepare_batch(random_queries(rng, nb, k)) for _ in range(args.held_batches)]
    eng.set_tuning("batch_mode", 2)
    tuning = None
    if args.tune:
        for name, val in zip(("batch_vpt", "batch_bs", "batch_qpi", "batch_digit", "batch_grid", "batch_item"),
                             (int(t) for t in args.tune.split(","))):
            eng.set_tuning(name, val)
        tuning = {"fixed": args.tune}
    elif not args.no_autotune:
        tuning = eng.batch_autotune(held[0])

    def timed(steps, warmup, mode):
        """Two batches in flight (topk_batch_submit / topk_batch_wait): the
        host stages batch i+1 and completes batch i-1 while batch i runs.
        The statistics (fallbacks, scored pairs) of the same sequence of
        batches are read afterwards, one synchronous call per batch, so
        their reads stay out of the timed loop (batches are deterministic)."""
        eng.set_tuning("batch_mode", mode)
        for i in range(warmup):
            eng.batch_run(held[i % len(held)])
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        pending = []
        for i in range(steps):
            pending.append(eng.batch_submit(held[i % len(held)]))
            if len(pending) == 2:
                eng.batch_wait(pending.pop(0))
        for t in pending:
            eng.batch_wait(t)
        torch.cuda.synchronize()
        el = time.perf_counter() - t1
        fb = ff = pairs = objs = cand = items = 0
        for i in range(steps):
            eng.batch_run(held[i % len(held)])
            st = eng.stats()
            fb += st["fallback"]
            ff += st["fte_fallback"]
            pairs += st["partitions_scored"]
            objs += st["tuple_count"]
            cand += st["candidates"]
            items += st["items"]
        return el, fb, ff, pairs, objs, cand, items

    def timed_sync(steps):
        """One synchronous topk_batch call per batch (no overlap)."""
        t1 = time.perf_counter()
        for i in range(steps):
            eng.batch_run(held[i % len(held)])
        return (time.perf_counter() - t1) / steps

    # seed thresholds (the method without the network), then learned
    s_el, s_fb, _, s_pairs, s_objs, _, _ = timed(max(2, args.steps // 4), 1, 1)
    el, fb, ff, pairs, objs, cand, items = timed(args.steps, args.warmup, 2)
    steps = args.steps
    qps = steps * nb / el
    sync_s = timed_sync(steps)

    # scoring-kernel duration with HIP events (eager launches, one stream)
    eng.timing(True)
    eng.timing_reset()
    tsteps = min(steps, 10)
    tterms = 0
    for i in range(tsteps):
        eng.batch_run(held[i % len(held)])
        tterms += eng.stats()["terms"]
    score_ms, launches = eng.timing_get("batch_score")
    score2_ms, launches2 = eng.timing_get("batch_score_second")
    score_ms, launches = score_ms + score2_ms, launches + launches2  # both passes of every batch
    seed_ms, _ = eng.timing_get("batch_tau_seed")
    lrn_ms, _ = eng.timing_get("batch_tau_learned")
    eng.timing(False)
    avg_ms = score_ms / max(1, launches)
    # mul + add per (object, query, attribute) term the scoring kernel
    # scored (counted on the device; chunk-pruned objects not included),
    # over all its launches (first and second passes)
    flops = 2.0 * tterms / max(1, launches)
    tflops = 2.0 * tterms / (score_ms * 1e-3) / 1e12 if score_ms > 0 else 0.0

    # every held-out answer judged by the plain GPU FTE path
    mism = 0
    eng.set_tuning("batch_mode", 2)
    for pb in held:
        eng.batch_run(pb)
        for i, (kq, w, attr) in enumerate(pb.queries):
            gi, gs, gt = pb.result(i, n)
            ri, rs, rt = eng.fte(kq, w, attr)
            mism += int(not (np.array_equal(gi, ri) and np.array_equal(gs.view(np.uint32), rs.view(np.uint32))
                             and gt == rt))
    eng.set_tuning("batch_mode", 0)
    total_pairs = n * nb
    # which scoring kernel the installed launch shape runs (the autotuner's choice)
    query_major = bool(tuning and tuning.get("fused"))
    return {
        "qps": qps, "ms_per_batch": el * 1e3 / steps, "steps": steps, "queries_per_batch": nb,
        "pipelining": "two batches in flight (topk_batch_submit / topk_batch_wait)",
        "sync_call_qps": nb / sync_s, "sync_call_ms_per_batch": sync_s * 1e3,
        "seed_batched_qps": max(2, args.steps // 4) * nb / s_el,
        # learned threshold above the k-th score: re-run exactly in the same launch
        "fallback_rate": fb / (steps * nb),
        # answered by the plain FTE path after both device passes
        "fte_fallback_rate": ff / (steps * nb),
        "mismatches_vs_plain_fte": mism, "judged_queries": len(held) * nb,
        "scored_pair_frac": objs / (steps * total_pairs),
        "scored_pair_frac_seed": s_objs / (max(2, args.steps // 4) * total_pairs),
        "candidates_per_query": cand / (steps * nb),
        "work_items_per_batch": items / steps, "scored_partition_query_pairs_per_batch": pairs / steps,
        "autotune": tuning,
        "threshold_stage_ms": {"learned_mlp": lrn_ms / max(1, tsteps), "seed_pass_in_second_passes": seed_ms / max(1, tsteps)},
        "layout_s": tt_layout,
        "training": {"queries": len(train), "n_train": rep["n_train"], "n_calib": rep["n_calib"],
                     "train_mse": rep["train_mse"], "calib_mse": rep["calib_mse"], "margin": rep["margin"],
                     "calib_fallback": rep["calib_fallback"], "label_s": rep["ms_label"] / 1e3,
                     "train_s": rep["ms_train"] / 1e3, **train_step_stats(rep)},
        "roofline": {"bound": "valu", "achieved": tflops, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": tflops / FP32_PEAK_TFLOPS,
                     "traffic": load_profile_traffic("batch_chunk_pmc.md" if query_major else "batch_score_pmc.md"),
                     "kernel": "batch_chunk_kernel" if query_major else "batch_score_kernel",
                     "avg_launch_ms": avg_ms, "flops_per_launch": flops,
                     "issue_peak": FP32_ISSUE_TFLOPS, "issue_frac": tflops / FP32_ISSUE_TFLOPS},
        "held": held,
    }


def host_info():
    """Logical CPUs and CPU model of the host the CPU baselines run on."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def c1_bench(threads):
    """The reference's CPU-runnable case: N = 2^20, d = 8, independent (seed
    0), attributes {4, 5, 6, 7}, weight 1, k = 16 (validation/bench_cpu.h
    row qq = 4).  The oracle's SIMD+threads TPAc restatement on the host's
    cores (1 warm-up + 10 timed calls, mean and median, as cpu/AA.h:202) and
    the GPU paths on the same store."""
    import oracle
    import topk_amd
    n, d, k = 1 << 20, 8, 16
    w = np.ones(d, np.float32)
    attr = [4, 5, 6, 7]
    x = topk_amd.generate(n, d, topk_amd.INDEPENDENT, 0, threads=threads)
    ref = oracle.tpac_threads(x, n, k, w, attr, threads)
    ts = []
    for _ in range(10):
        t0 = time.perf_counter()
        oracle.tpac_threads(x, n, k, w, attr, threads)
        ts.append(time.perf_counter() - t0)
    with topk_amd.Engine(x) as e:
        e.fte(k, w, attr)
        t0 = time.perf_counter()
        for _ in range(100):
            got = e.fte(k, w, attr)
        fte_q = (time.perf_counter() - t0) / 100
        e.init(topk_amd.TOPK_INIT_PTA)
        e.pta(k, w, attr)
        t0 = time.perf_counter()
        for _ in range(100):
            gp = e.pta(k, w, attr)
        pta_q = (time.perf_counter() - t0) / 100
    same = bool(np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]) and
                np.array_equal(gp[0], ref[0]) and np.array_equal(gp[1], ref[1]))
    return {"cpu_tpac_threads_ms_mean": float(np.mean(ts)) * 1e3, "cpu_tpac_threads_ms_median": float(np.median(ts)) * 1e3,
            "cpu_threads": threads, "host": host_info(), "gpu_fte_ms": fte_q * 1e3, "gpu_pta_ms": pta_q * 1e3,
            "gpu_matches_cpu": same}


def cpu_baseline(x, n, k, threads, budget_s, held, gpu_check):
    """Oracle TPAc (SIMD + threads) on held-out queries within budget_s."""
    import oracle
    qs = held[0].queries
    oracle.tpac_threads(x, n, k, qs[0][1], qs[0][2], threads)  # warm-up
    times, mism = [], 0
    t_start = time.time()
    i = 0
    while not times or (time.time() - t_start < budget_s and i < len(qs)):
        kq, w, attr = qs[i]
        t0 = time.perf_counter()
        ids, sc, thr = oracle.tpac_threads(x, n, kq, w, attr, threads)
        times.append(time.perf_counter() - t0)
        gi, gs, gt = gpu_check(i)
        mism += int(not (np.array_equal(ids, gi) and np.array_equal(sc.view(np.uint32), gs.view(np.uint32))))
        i += 1
    return {"value": 1.0 / float(np.mean(times)), "unit": "queries/s", "cores": threads, "kind": "port",
            "sample": f"{len(times)} held-out queries of the benchmark (N={n}, d={D}, all 8 attributes, "
                      f"k={k}, same store), 1 warm-up, mean; oracle or_tpac_threads (AVX2 16-wide + "
                      f"{threads} threads, cpu/TPAc.h:256-358 restated)",
            "median_s": float(np.median(times)), "host": host_info()}, mism


def subset_queries(rng, count, d, ks):
    """Random attribute subsets of 2..d attributes (ascending), weights
    U(0.1, 1) by attribute id, k drawn from ks."""
    out = []
    for _ in range(count):
        m = int(rng.integers(2, d + 1))
        attr = np.sort(rng.choice(d, size=m, replace=False)).astype(np.uint8)
        out.append((int(rng.choice(ks)), rng.uniform(0.1, 1.0, d).astype(np.float32), attr))
    return out


def pipelined(eng, pbs, steps):
    """Batches per second with two in flight; returns (q/s, fallback, fte
    fallback).  The fallback counts of the same batches are read afterwards,
    one synchronous call per batch (batches are deterministic)."""
    import torch
    for pb in pbs[:2]:
        eng.batch_run(pb)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pend, nq = [], 0
    for i in range(steps):
        pend.append(eng.batch_submit(pbs[i % len(pbs)]))
        nq += pbs[i % len(pbs)].nq
        if len(pend) == 2:
            eng.batch_wait(pend.pop(0))
    for t in pend:
        eng.batch_wait(t)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    fb = ff = 0
    for pb in pbs:
        eng.batch_run(pb)
        st = eng.stats()
        fb, ff = fb + st["fallback"], ff + st["fte_fallback"]
    nb = len(pbs) * pbs[0].nq
    return nq / el, fb / nb, ff / nb


def pta_bench(threads, cpu=True, queries=20):
    """PTA (C3): N = 2^28, d = 8 anticorrelated (seed 3), 4096-object
    partitions, weights U(0.1, 1) on all 8 attributes (query seed 100), k =
    128, one synchronous topk_pta call per query.

    Touched bytes per query = the bound pass over every partition ((8d + 12)
    bytes: per-attribute max and min, attribute-sum bounds, size) + 4*m bytes
    per object the stream scored (counted on the device) -> roofline against
    HBM; the FTE-equivalent rate is 4*N*m bytes / query time.  CPU baseline:
    the oracle's threaded PTA (cpu/PTA.h:569-671 restated: partitions in
    descending bound order, strict skip, per-thread heaps) over the engine's
    own layout, 1 warm-up + 10 timed calls on the first 10 queries, each
    compared with the GPU answer."""
    import oracle
    import topk_amd
    n3, d, k = 1 << 28, 8, 128
    rng = np.random.default_rng(100)
    ws = [rng.uniform(0.1, 1.0, d).astype(np.float32) for _ in range(queries)]
    attr = np.arange(d, dtype=np.uint8)
    with topk_amd.Engine.generated(n3, d, topk_amd.ANTICORRELATED, 3) as eng:
        eng.set_partition_size(4096)
        t0 = time.perf_counter()
        eng.init(topk_amd.TOPK_INIT_PTA)
        init_s = time.perf_counter() - t0
        for w in ws[:3]:
            eng.pta(k, w, attr)
        lat, parts, objs, gpu = [], [], [], []
        for w in ws:
            t0 = time.perf_counter()
            got = eng.pta(k, w, attr)
            lat.append(time.perf_counter() - t0)
            st = eng.stats()
            parts.append(st["partitions_scored"])
            objs.append(st["tuple_count"])
            gpu.append(got)
        P = st["partitions_total"]
        t0 = time.perf_counter()
        for w in ws[:5]:
            eng.fte(k, w, attr)
        fte_s = (time.perf_counter() - t0) / 5
        layout = eng.pta_layout() if cpu else None
    t_q = float(np.mean(lat))
    bound_bytes = (8 * d + 12) * P
    touched = bound_bytes + 4.0 * d * float(np.mean(objs))
    achieved = touched / t_q / 1e9
    out = {
        "pta_qps": 1.0 / t_q, "ms_per_query_mean": t_q * 1e3, "ms_per_query_median": float(np.median(lat)) * 1e3,
        "fte_qps": 1.0 / fte_s, "partitions": P, "partitions_scored_mean": float(np.mean(parts)),
        "partitions_scored_frac": float(np.mean(parts)) / P, "objects_scored_mean": float(np.mean(objs)),
        "objects_scored_frac": float(np.mean(objs)) / n3, "layout_build_s": init_s,
        "fte_equivalent_GBps": 4.0 * n3 * d / t_q / 1e9,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "touched_bytes_per_query": touched,
                     "bound_pass_bytes": bound_bytes, "scored_bytes": 4.0 * d * float(np.mean(objs)),
                     "time": "whole synchronous topk_pta call"},
    }
    if cpu:
        pid, poff, psz, pmax, pmin = layout
        x = topk_amd.generate(n3, d, topk_amd.ANTICORRELATED, 3, threads=threads)
        valid = pid != 0xFFFFFFFF
        pidv = np.where(valid, pid, 0).astype(np.int64)
        pdata = np.empty((d, pid.size), np.float32)
        for m in range(d):
            np.take(x[m], pidv, out=pdata[m])
        del x, pidv
        oracle.pta_topk(pdata, pid, poff, psz, pmax, pmin, k, ws[0], attr, threads)  # warm-up
        ts, mism, cscored = [], 0, []
        for i in range(10):
            t0 = time.perf_counter()
            ri, rs, rt, sc = oracle.pta_topk(pdata, pid, poff, psz, pmax, pmin, k, ws[i], attr, threads)
            ts.append(time.perf_counter() - t0)
            cscored.append(sc)
            gi, gs, gt = gpu[i]
            mism += int(not (np.array_equal(gi, ri) and np.array_equal(gs.view(np.uint32), rs.view(np.uint32))
                             and gt == rt))
        del pdata
        out["cpu_baseline"] = {
            "kind": "port", "cores": threads, "value": 1.0 / float(np.mean(ts)), "unit": "queries/s",
            "ms_mean": float(np.mean(ts)) * 1e3, "ms_median": float(np.median(ts)) * 1e3,
            "partitions_scored_mean": float(np.mean(cscored)), "host": host_info(),
            "sample": "1 warm-up + 10 timed calls on the first 10 queries; oracle or_pta_threads (partition "
                      "box bounds in descending order, strict skip, per-thread heaps) over the engine's own "
                      "partition layout"}
        out["mismatches_vs_cpu_pta"] = mism
    return out


def batch_kernel_view(eng, pbs, steps=10):
    """Eager batches with HIP-event timing of the scoring kernel (both
    passes): device-counted (object, query) pairs and terms per batch, the
    scoring kernel's time per batch and its FP32 rate (2 flops per term)."""
    eng.timing(True)
    eng.timing_reset()
    pairs = terms = 0
    for i in range(steps):
        eng.batch_run(pbs[i % len(pbs)])
        st = eng.stats()
        pairs += st["tuple_count"]
        terms += st["terms"]
    s1, l1 = eng.timing_get("batch_score")
    s2, l2 = eng.timing_get("batch_score_second")
    seed_ms, _ = eng.timing_get("batch_tau_seed")
    lrn_ms, _ = eng.timing_get("batch_tau_learned")
    eng.timing(False)
    score_ms = s1 + s2
    tflops = 2.0 * terms / (score_ms * 1e-3) / 1e12 if score_ms > 0 else 0.0
    return {"pairs_per_batch": pairs / steps, "terms_per_batch": terms / steps,
            "score_ms_per_batch": score_ms / steps, "score_launches_per_batch": (l1 + l2) / steps,
            "avg_launch_ms": score_ms / max(1, l1 + l2), "flops_per_launch": 2.0 * terms / max(1, l1 + l2),
            "tflops": tflops, "tau_ms_per_batch": (seed_ms + lrn_ms) / steps}


def valu_roofline(view, kernel):
    t = view["tflops"]
    return {"bound": "valu", "achieved": t, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": t / FP32_PEAK_TFLOPS,
            "issue_peak": FP32_ISSUE_TFLOPS, "issue_frac": t / FP32_ISSUE_TFLOPS, "kernel": kernel,
            "avg_launch_ms": view["avg_launch_ms"], "flops_per_launch": view["flops_per_launch"],
            "flops": "2 per (object, query, attribute) term the scoring kernel scored, counted on the device"}


def batched_bench(threads, steps=20, cpu=True):
    """Batched (C4): N = 2^25, d = 16 mixed (seed 4), 1024 queries per batch
    on random subsets of 2-16 attributes, weights U(0.1, 1), k from
    {16..256} (query seed 200), seed thresholds, two batches in flight.
    Full-scan flops = 2*N*sum(m) per batch (what one scan per query
    computes); executed flops = 2 per device-counted term of the scoring
    kernel.  CPU baseline: the oracle's findTopKsimdMQ restatement (one scan
    per query, queries over the host threads) on the first 32 queries."""
    import oracle
    import topk_amd
    n4, d = 1 << 25, 16
    ks = [16, 32, 64, 128, 256]
    with topk_amd.Engine.generated(n4, d, topk_amd.MIXED, 4) as eng:
        eng.init(topk_amd.TOPK_INIT_PTA)
        qlists = [subset_queries(np.random.default_rng(200 + i), 1024, d, ks) for i in range(4)]
        pbs = [eng.prepare_batch(ql) for ql in qlists]
        eng.set_tuning("batch_mode", 1)
        tune = eng.batch_autotune(pbs[0])
        qps, fb, ff = pipelined(eng, pbs, steps)
        view = batch_kernel_view(eng, pbs)
        sample = qlists[0][:32]
        gpu_sample = eng.batch(sample)
    sum_m = float(np.mean([sum(len(q[2]) for q in ql) for ql in qlists]))
    t_batch = 1024.0 / qps
    full_flops = 2.0 * n4 * sum_m
    out = {"qps": qps, "ms_per_batch": t_batch * 1e3, "fallback_rate": fb, "fte_fallback_rate": ff,
           "autotune": tune, "threshold": "seed (no predictor)",
           "full_scan_flops_per_batch": full_flops, "full_scan_equivalent_tflops": full_flops / t_batch / 1e12,
           "executed_flops_per_batch": 2.0 * view["terms_per_batch"],
           "executed_tflops_whole_batch": 2.0 * view["terms_per_batch"] / t_batch / 1e12,
           "scored_pair_frac": view["pairs_per_batch"] / (1024.0 * n4), "kernel_view": view,
           "roofline": valu_roofline(view, "batch scoring kernel (batch_score timing)")}
    if cpu:
        x = topk_amd.generate(n4, d, topk_amd.MIXED, 4, threads=threads)
        amax = max(len(q[2]) for q in sample)
        attrs = np.zeros((len(sample), amax), np.uint8)
        for i, q in enumerate(sample):
            attrs[i, :len(q[2])] = q[2]
        t0 = time.perf_counter()
        ref = oracle.mq_threads(x, n4, [q[0] for q in sample], [len(q[2]) for q in sample],
                                np.stack([q[1] for q in sample]), attrs, threads)
        cpu_s = time.perf_counter() - t0
        del x
        out["mismatches_vs_oracle_sample"] = sum(
            int(not (np.array_equal(g[0], r[0]) and np.array_equal(g[1], r[1]))) for g, r in zip(gpu_sample, ref))
        out["cpu_baseline"] = {"kind": "port", "cores": threads, "value": len(sample) / cpu_s, "unit": "queries/s",
                               "host": host_info(), "sample": "first 32 queries of the first batch, one call"}
    return out


def learned16_bench(steps=20, train_queries=100_000):
    """Learned (C5): N = 2^26, d = 16 mixed (seed 5), 100k training queries
    (seed 300), 4096 held-out queries (seed 301) in batches of 1024; learned
    and seed thresholds on the same batches (threshold-stage time and
    device-counted scored pairs of each), every held-out answer judged
    against the plain GPU FTE path."""
    import topk_amd
    ks = [16, 32, 64, 128, 256]
    n5, d = 1 << 26, 16
    with topk_amd.Engine.generated(n5, d, topk_amd.MIXED, 5) as eng:
        train = subset_queries(np.random.default_rng(300), train_queries, d, ks)
        t0 = time.perf_counter()
        rep = eng.learned_train(train)
        train_s = time.perf_counter() - t0
        held = subset_queries(np.random.default_rng(301), 4096, d, ks)
        pbs = [eng.prepare_batch(held[i:i + 1024]) for i in range(0, 4096, 1024)]
        eng.set_tuning("batch_mode", 2)
        tune = eng.batch_autotune(pbs[0])
        qps, fb, ff = pipelined(eng, pbs, steps)
        lview = batch_kernel_view(eng, pbs)
        eng.set_tuning("batch_mode", 1)
        sqps, _, _ = pipelined(eng, pbs, steps)
        sview = batch_kernel_view(eng, pbs)
        eng.set_tuning("batch_mode", 2)
        mism = 0
        for pb in pbs:
            eng.batch_run(pb)
            for i, q in enumerate(pb.queries):
                gi, gs, gt = pb.result(i, n5)
                ri, rs, rt = eng.fte(*q)
                mism += int(not (np.array_equal(gi, ri) and np.array_equal(gs.view(np.uint32), rs.view(np.uint32))
                                 and gt == rt))
        t0 = time.perf_counter()
        for q in held[:50]:
            eng.fte(*q)
        fte_qps = 50 / (time.perf_counter() - t0)
    return {
        "learned_qps": qps, "seed_qps": sqps, "learned_over_seed": qps / sqps, "plain_fte_qps": fte_qps,
        "speedup_vs_plain_fte": qps / fte_qps, "fallback_rate": fb, "fte_fallback_rate": ff,
        "mismatches_vs_plain_fte": mism, "judged_queries": 4096,
        "threshold_stage_ms_per_batch": {"learned": lview["tau_ms_per_batch"], "seed": sview["tau_ms_per_batch"]},
        "scored_pairs_per_batch": {"learned": lview["pairs_per_batch"], "seed": sview["pairs_per_batch"]},
        "kernel_view": {"learned": lview, "seed": sview},
        "roofline": valu_roofline(lview, "batch scoring kernel (batch_score timing), learned thresholds"),
        "training": {"queries": len(train), "margin": rep["margin"], "calib_fallback": rep["calib_fallback"],
                     "saturated": rep.get("saturated"), "train_mse": rep["train_mse"], "wall_s": train_s,
                     **train_step_stats(rep)},
        "autotune": tune}


def other_configs(threads):
    """The other workloads of the scope table, measured once each (one JSON
    line; the headline line is the default mode): C1 (c1_bench), PTA at
    N = 2^28 (pta_bench), batched d = 16 (batched_bench), learned d = 16
    (learned16_bench), BTA / GPA / radix select / normalisation
    (kernels_bench), ingest at N = 2^24 (ingest_bench) and a host-memory
    store at N = 2^28 (stream_bench)."""
    res = {}
    res["c1_n2e20_d8_m4_k16"] = c1_bench(threads)
    res["pta_n2e28_anticorrelated_k128"] = pta_bench(threads)
    res["batch_n2e25_d16_mixed"] = batched_bench(threads)
    res["learned_n2e26_d16_mixed"] = learned16_bench()
    res["kernels_n2e26_d8"] = kernels_bench(threads)
    res["ingest_n2e24_d8"] = ingest_bench(1 << 24, 8, 3)
    res["stream_n2e28_d8_host"] = stream_bench(1 << 28, 8, 1 << 26, 64)
    print(json.dumps({"configs": res}), flush=True)


def ingest_bench(n, d, steps):
    """Data ingest: an N x d independent store (seed 0) written as text rows
    of "%f" values (the reference's writer, input/File.h:415-430) and as a
    binary column file in a temporary directory (page cache warm), each
    loaded into a fresh column store ``steps`` times.  The CPU baseline is the
    oracle's fscanf row loop (input/File.h:264-295, one thread) on the first
    2^20 rows; parity: the loaded columns equal the oracle's on that sample
    and the column file loads unchanged."""
    import shutil
    import tempfile

    import oracle
    import topk_amd
    x = topk_amd.generate(n, d, topk_amd.INDEPENDENT, 0)
    tmp = tempfile.mkdtemp(prefix="topk_ingest_")
    try:
        csv = os.path.join(tmp, "d_%d_%d_i.csv" % (n, d))
        col = os.path.join(tmp, "d_%d_%d_i.col" % (n, d))
        t0 = time.perf_counter()
        topk_amd.write_csv(csv, x)
        write_s = time.perf_counter() - t0
        topk_amd.write_colfile(col, x)
        text_runs, col_runs = [], []
        for _ in range(steps):
            with topk_amd.Engine.from_file(csv) as e:
                text_runs.append(e.ingest)
        sample = min(n, 1 << 20)
        with topk_amd.Engine.from_file(csv) as e:
            cols = e.read_columns()
        t0 = time.perf_counter()
        want, r = oracle.csv_read(csv, sample, d)
        cpu_s = time.perf_counter() - t0
        text_ok = r == sample and np.array_equal(cols[:, :sample].view(np.uint32), want.view(np.uint32))
        text_close = float(np.max(np.abs(cols - x)))
        for _ in range(steps):
            with topk_amd.Engine.from_file(col) as e:
                col_runs.append(e.ingest)
        with topk_amd.Engine.from_file(col) as e:
            col_ok = np.array_equal(e.read_columns(), x)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    tb = sorted(text_runs, key=lambda r: r["ms_total"])[len(text_runs) // 2]
    cb = sorted(col_runs, key=lambda r: r["ms_total"])[len(col_runs) // 2]
    sample_bytes = tb["bytes"] * sample / n
    return {
        "text": {"bytes": tb["bytes"], "rows": tb["rows"], "ms_total": tb["ms_total"], "ms_read": tb["ms_read"],
                 "ms_parse": tb["ms_parse"], "GB_per_s": tb["bytes"] / tb["ms_total"] / 1e6,
                 "rows_per_s": tb["rows"] / tb["ms_total"] * 1e3, "slow_lines": tb["slow_lines"],
                 "matches_oracle_sample": bool(text_ok), "max_abs_vs_source": text_close,
                 "write_s": write_s},
        "column_file": {"bytes": cb["bytes"], "ms_total": cb["ms_total"],
                        "GB_per_s": cb["bytes"] / cb["ms_total"] / 1e6, "loads_unchanged": bool(col_ok)},
        "cpu_baseline": {"kind": "port", "cores": 1, "sample": "first %d rows, fscanf row loop" % sample,
                         "rows_per_s": sample / cpu_s, "GB_per_s": sample_bytes / cpu_s / 1e9},
        "speedup_vs_cpu_rows": (tb["rows"] / tb["ms_total"] * 1e3) / (sample / cpu_s),
        "steps": steps,
    }


PCIE_PEAK_GBS = 64.0   # PCIe 5.0 x16, one direction


def kernels_bench(threads):
    """The reference's GPU kernel sketches and preprocessing on N = 2^26,
    d = 8 (independent, seed 0), each timed with HIP events on the engine's
    stream against its algorithmic bytes, checked against the oracle (or the
    plain FTE answer), with the oracle's CPU restatement timed beside it:
    BTA and GPA top-k (k = 64, all 8 attributes, weight 1), radix select
    (2^26 random 32-bit keys, rank N - 64), normalisation and reorder_attr."""
    import torch

    import oracle
    import topk_amd
    n, d, k = 1 << 26, 8, 64
    w = np.ones(d, np.float32)
    attr = list(range(d))
    res = {}

    def per_launch(e, name):
        ms, cnt = e.timing_get(name)
        return ms / max(1, cnt)

    def same(a, b):
        return bool(np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2])

    x = topk_amd.generate(n, d, topk_amd.INDEPENDENT, 0, threads=threads)
    t0 = time.perf_counter()
    cpu_ref = oracle.tpac_threads(x, n, k, w, attr, threads)
    cpu_q = time.perf_counter() - t0
    with topk_amd.Engine.generated(n, d, topk_amd.INDEPENDENT, 0) as e:
        e.timing(True)
        ref = e.fte(k, w, attr)
        for name, fn, kern in (("bta", e.bta, "bta_block"), ("gpa", e.gpa, "gpa")):
            fn(k, w, attr)
            e.timing_reset()
            t0 = time.perf_counter()
            for _ in range(5):
                got = fn(k, w, attr)
            q_s = (time.perf_counter() - t0) / 5
            kms = per_launch(e, kern)
            res[name] = {"qps": 1.0 / q_s, "kernel": kern, "kernel_ms": kms,
                         "scoring_GB_per_s": 4.0 * n * d / kms / 1e6,
                         "hbm_frac": 4.0 * n * d / kms / 1e6 / HBM_PEAK_GBS,
                         "matches_fte": same(got, ref),
                         "cpu_tpac_threads_qps": 1.0 / cpu_q, "cpu_threads": threads,
                         "matches_cpu": bool(np.array_equal(got[0], cpu_ref[0]) and np.array_equal(got[1], cpu_ref[1]))}
        keys = torch.randint(-2 ** 31, 2 ** 31 - 1, (n,), dtype=torch.int32, device="cuda",
                             generator=torch.Generator(device="cuda").manual_seed(7))
        hk = keys.cpu().numpy().view(np.uint32)
        r = n - k
        e.radix_select_u32(keys.data_ptr(), n, r)
        e.timing_reset()
        for _ in range(5):
            got = e.radix_select_u32(keys.data_ptr(), n, r)
        kms = per_launch(e, "radix_select")
        t0 = time.perf_counter()
        want = oracle.radix_select_u32(hk, r)
        cpu_s = time.perf_counter() - t0
        res["radix_select_u32"] = {"keys": n, "rank": r, "ms": kms, "GB_per_s_3_passes": 12.0 * n / kms / 1e6,
                                   "hbm_frac": 12.0 * n / kms / 1e6 / HBM_PEAK_GBS,
                                   "matches_oracle": bool(got == want and want == np.partition(hk, r)[r]),
                                   "cpu_ms": cpu_s * 1e3, "cpu_cores": 1}
        del keys
        e.init(topk_amd.TOPK_INIT_NORMALIZE)  # the first call also loads the kernels
        # timed warm: normalising normalised columns (minimum 0, range 1
        # exactly) changes no value
        e.timing_reset()
        e.init(topk_amd.TOPK_INIT_NORMALIZE)
        nms = per_launch(e, "normalize")
        t0 = time.perf_counter()
        xn = oracle.normalize(x, n)
        cpu_n = time.perf_counter() - t0
        norm_ok = bool(np.array_equal(e.read_columns().view(np.uint32), xn.view(np.uint32)))
        res["normalize"] = {"ms": nms, "GB_per_s": 12.0 * n * d / nms / 1e6,
                            "hbm_frac": 12.0 * n * d / nms / 1e6 / HBM_PEAK_GBS, "matches_oracle": norm_ok,
                            "cpu_ms": cpu_n * 1e3, "cpu_cores": 1}
        e.timing_reset()
        e.reorder_attr()
        rms = per_launch(e, "reorder_attr")
        t0 = time.perf_counter()
        xr = oracle.reorder_attr(xn, n)
        cpu_r = time.perf_counter() - t0
        res["reorder_attr"] = {"ms": rms, "GB_per_s": 8.0 * n * d / rms / 1e6,
                               "hbm_frac": 8.0 * n * d / rms / 1e6 / HBM_PEAK_GBS,
                               "matches_oracle": bool(np.array_equal(e.read_columns().view(np.uint32),
                                                                     xr.view(np.uint32))),
                               "cpu_ms": cpu_r * 1e3, "cpu_cores": 1}
    return res


def stream_bench(n, d, chunk, nbatch):
    """Stores larger than HBM: an N x d independent store (seed 0) in pinned
    host memory streamed through HBM in chunks.  Single queries (k = 64, all
    d attributes, weight 1) and one batch of ``nbatch`` queries (weights
    U(0.1, 1), k = 64) that streams each chunk once; the in-HBM FTE path on the
    same store is the comparison and the parity check."""
    import topk_amd
    rng = np.random.default_rng(500)
    attr = list(range(d))
    w1 = np.ones(d, np.float32)
    batch = [(K, rng.uniform(0.1, 1.0, d).astype(np.float32), attr) for _ in range(nbatch)]
    with topk_amd.HostStore(n, d, chunk) as hs:
        t0 = time.perf_counter()
        hs.generate(topk_amd.INDEPENDENT, 0, threads=16)
        fill_s = time.perf_counter() - t0
        hs.fte(K, w1, attr)  # warm-up
        t0 = time.perf_counter()
        single = [hs.fte(K, w1, attr) for _ in range(3)]
        single_s = (time.perf_counter() - t0) / 3
        rep1 = hs.report()
        t0 = time.perf_counter()
        got = hs.batch(batch)
        batch_s = time.perf_counter() - t0
        repb = hs.report()
    with topk_amd.Engine.generated(n, d, topk_amd.INDEPENDENT, 0) as e:
        ref = e.fte(K, w1, attr)
        e.fte(K, w1, attr)
        t0 = time.perf_counter()
        for _ in range(10):
            e.fte(K, w1, attr)
        hbm_qps = 10 / (time.perf_counter() - t0)
        mism = int(not all(np.array_equal(a, b) for a, b in zip(single[0], ref)))
        for q, g in zip(batch[:8], got[:8]):
            r = e.fte(*q)
            mism += int(not (np.array_equal(g[0], r[0]) and np.array_equal(g[1], r[1]) and g[2] == r[2]))
    pcie1 = rep1["bytes"] / single_s / 1e9
    return {
        "store": {"n": n, "d": d, "bytes": n * d * 4, "chunk_objects": chunk, "chunks": rep1["chunks"],
                  "fill_s": fill_s},
        "single_query": {"qps": 1.0 / single_s, "ms": single_s * 1e3, "bytes_streamed": rep1["bytes"],
                         "pcie_GB_per_s": pcie1, "pcie_frac": pcie1 / PCIE_PEAK_GBS},
        "batch": {"queries": nbatch, "qps": nbatch / batch_s, "ms": batch_s * 1e3,
                  "bytes_streamed": repb["bytes"], "pcie_GB_per_s": repb["bytes"] / batch_s / 1e9},
        "in_hbm_fte_qps": hbm_qps,
        "mismatches_vs_in_hbm_fte": mism, "judged_queries": 9,
        "pcie_peak_GB_per_s": PCIE_PEAK_GBS,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps: batches (default 200; 20 with --mode batched / learned16), queries with "
                         "--mode fte (200), file loads with --mode ingest (3)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--mode", choices=["learned", "fte", "configs", "ingest", "stream", "pta", "kernels", "c1", "train",
                                       "batched", "learned16"],
                    default="learned",
                    help="learned: the headline line; fte: plain FTE alone; configs: the other workloads; "
                         "ingest: data ingest alone; stream: a host-memory store streamed through HBM")
    ap.add_argument("--n", type=int, default=N_DEFAULT)
    ap.add_argument("--m", type=int, default=D, help="--mode fte: query width")
    ap.add_argument("--k", type=int, default=K)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--train-queries", type=int, default=None,
                    help="training queries (default 16384; 100000 with --mode learned16)")
    ap.add_argument("--held-batches", type=int, default=4)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tune", default="", metavar="VPT,BS,QPI,DIGIT,GRID[,ITEM]",
                    help="fixed batch launch shape instead of the autotuner's choice")
    ap.add_argument("--fte-steps", type=int, default=200)
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--lanes", type=int, default=2)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = {"batched": 20, "learned16": 20, "ingest": 3}.get(args.mode, 200)
    if args.train_queries is None and args.mode != "learned16":
        args.train_queries = 16384
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): a multi-GPU run would split "
              "the store by object range and merge each query's per-GPU top-k", file=sys.stderr)
        sys.exit(2)

    import torch

    torch.cuda.init()  # torch's HIP runtime first; the engine shares it
    import topk_amd

    if args.mode == "configs":
        other_configs(min(os.cpu_count() or 1, 16))
        return
    if args.mode == "c1":
        print(json.dumps({"c1_n2e20_d8_m4_k16": c1_bench(16)}), flush=True)
        return
    if args.mode == "kernels":
        print(json.dumps({"kernels_n2e26_d8": kernels_bench(16)}), flush=True)
        return
    if args.mode == "pta":
        print(json.dumps({"pta_n2e28_anticorrelated_k128": pta_bench(16, cpu=not args.no_cpu)}), flush=True)
        return
    if args.mode == "batched":
        print(json.dumps({"batch_n2e25_d16_mixed": batched_bench(16, args.steps, cpu=not args.no_cpu)}), flush=True)
        return
    if args.mode == "learned16":
        print(json.dumps({"learned_n2e26_d16_mixed": learned16_bench(args.steps, args.train_queries or 100_000)}),
              flush=True)
        return
    if args.mode == "stream":
        n = args.n if args.n != N_DEFAULT else 1 << 28
        print(json.dumps({"stream_n%d_d8" % n: stream_bench(n, D, 1 << 26, 64)}), flush=True)
        return
    if args.mode == "ingest":
        n = args.n if args.n != N_DEFAULT else 1 << 24
        print(json.dumps({"ingest_n%d_d8" % n: ingest_bench(n, D, max(1, args.steps))}), flush=True)
        return

    n, k = args.n, args.k
    eng = topk_amd.Engine.generated(n, D, topk_amd.INDEPENDENT, 0)
    if args.mode == "train":
        eng.init(topk_amd.TOPK_INIT_PTA)
        train = random_queries(np.random.default_rng(2026), args.train_queries, k)
        rep = eng.learned_train(train)
        # the same training at other mini-batch sizes: step rate against
        # the fit (calibration error and margin) it buys
        sweep = {}
        for b in (1024, 2048, 4096):
            r = eng.learned_train(train, batch=b)
            sweep[str(b)] = {"margin": r["margin"], "calib_mse": r["calib_mse"], "train_s": r["ms_train"] / 1e3,
                             **train_step_stats(r)}
        print(json.dumps({"train_n%d_d8_k%d" % (n, k): dict(rep, queries=args.train_queries,
                                                             **train_step_stats(rep), batch_sweep=sweep)}),
              flush=True)
        eng.close()
        return
    base = {"metric": METRIC, "unit": "queries/s", "n_gpus": 1, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic (independent U[0,1), seed 0, generated in HBM)"}

    if args.mode == "fte":
        fte = plain_fte(eng, n, args.m, k, args.steps, args.warmup, args.lanes)
        out = dict(base, value=fte["qps"], steps=args.steps, warmup=args.warmup, ms_per_step=fte["ms_per_query"],
                   config={"workload": "plain GPU FTE top-k", "n": n, "d": D, "m": args.m, "k": k,
                           "weights": "1.0", "distribution": "independent"},
                   roofline=fte.pop("roofline"), plain_fte=fte)
        print(json.dumps(out), flush=True)
        eng.close()
        return

    fte = plain_fte(eng, n, D, k, args.fte_steps, 20, args.lanes)
    lb = learned_batched(eng, n, k, args)
    held = lb.pop("held")
    out = dict(base, value=lb["qps"], steps=args.steps, warmup=args.warmup, ms_per_step=lb["ms_per_batch"],
               config={"workload": "learned-threshold batched top-k (1024 queries per step)", "n": n, "d": D,
                       "k": k, "queries_per_batch": args.batch, "weights": "U(0.1,1) on all 8 attributes",
                       "distribution": "independent", "train_queries": args.train_queries},
               roofline=lb.pop("roofline"), speedup_vs_plain_fte=lb["qps"] / fte["qps"], plain_fte=fte,
               learned_batched=lb)
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        x = topk_amd.generate(n, D, topk_amd.INDEPENDENT, 0, threads=threads)
        eng.set_tuning("batch_mode", 2)
        eng.batch_run(held[0])
        cpu, mism = cpu_baseline(x, n, k, threads, args.cpu_budget, held, lambda i: held[0].result(i, n))
        out["cpu_baseline"] = cpu
        out["mismatches_vs_oracle"] = mism
        del x
    print(json.dumps(out), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
