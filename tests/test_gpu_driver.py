"""The reference-style C++ driver (topk_amd/bin/bench_topk) against the
oracle: every (k, qq) report of its sweeps (validation/bench_cpu.h:155-237,
:323-402) carries the stdout keys gather.py:13-26 collects, and every printed
threshold equals the oracle's k-th score on the same generated, normalised
store -- for IMP 0..2 (one query per call) over TPAc and PTA, QM 0 and 1,
and for the multiple-query modes IMP 3 (random qq) and IMP 4 (same qq)."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
import topk_amd

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "topk_amd", "bin", "bench_topk")
HEAD = re.compile(r"^Benchmark <<<-------------(\d+),(\d+),(\d+)------------->>> $")
KEYS = ("tt_init", "tt_procesing", "tuple_count", "tuples_per_second")


def _run(args):
    r = subprocess.run([EXE] + [str(a) for a in args], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


def _reports(out):
    """Splits the output at each Benchmark header: [(n, x, k, {key: [values]})]."""
    reps = []
    for line in out.splitlines():
        m = HEAD.match(line)
        if m:
            reps.append((int(m.group(1)), int(m.group(2)), int(m.group(3)), {}))
            continue
        if reps and ":" in line:
            key, val = line.split(":", 1)
            reps[-1][3].setdefault(key, []).append(val.strip())
    return reps


def _attrs(d, qq, qm):
    return list(range(d - qq, d)) if qm == 0 else list(range(qq))


def _bits(v):
    return np.float32(v).view(np.uint32)


@pytest.fixture(scope="module")
def store():
    n, d, distr, seed = 200_003, 6, topk_amd.INDEPENDENT, 3
    x = topk_amd.generate(n, d, distr, seed)
    return n, d, distr, seed, oracle.normalize(x, n)


@pytest.mark.parametrize("algo,imp,qm", [("fte", 2, 0), ("fte", 0, 1), ("pta", 1, 0), ("pta", 2, 1)])
def test_driver_sweep_thresholds_match_oracle(gpu, store, algo, imp, qm):
    n, d, distr, seed, xn = store
    out = _run(["-n", n, "-d", d, "-s", 16, "-e", 64, "-i", 2, "-t", distr, "-r", seed, "-a", algo, "-p", imp,
                "-m", qm])
    reps = _reports(out)
    assert [(r[1], r[2]) for r in reps] == [(qq, k) for k in (16, 32, 64) for qq in range(2, d + 1)]
    w = np.ones(d, np.float32)
    for rn, qq, k, kv in reps:
        assert rn == n
        for key in ("tt_init", "tt_procesing", "tuple_count"):
            assert len(kv[key]) == 1 and float(kv[key][0]) >= 0, (key, kv)
        _, _, rthr = oracle.fte_topk(xn, n, k, w, _attrs(d, qq, qm))
        assert _bits(float(kv["threshold"][0])) == _bits(rthr), (algo, imp, qm, qq, k)


@pytest.mark.parametrize("qm", [0, 1])
def test_driver_random_multi_query_matches_oracle(gpu, store, qm):
    """IMP 3: one report per k; the last batch's thresholds per query."""
    n, d, distr, seed, xn = store
    wl = 48
    out = _run(["-n", n, "-d", d, "-s", 16, "-e", 32, "-i", 2, "-t", distr, "-r", seed, "-a", "mq", "-w", wl,
                "-m", qm, "--mq-thresholds"])
    reps = _reports(out)
    assert [(r[1], r[2]) for r in reps] == [(d, 16), (d, 32)]
    w = np.ones(d, np.float32)
    cache = {}
    for _, _, k, kv in reps:
        for key in KEYS:
            assert len(kv[key]) == 1 and float(kv[key][0]) >= 0, (key, kv)
        lines = kv["mq_threshold"]
        assert len(lines) == wl
        qqs = set()
        for j, line in enumerate(lines):
            jj, qq, kk, thr = line.split()
            assert int(jj) == j and int(kk) == k and 2 <= int(qq) <= d
            qqs.add(int(qq))
            key = (int(qq), k)
            if key not in cache:
                cache[key] = oracle.fte_topk(xn, n, k, w, _attrs(d, int(qq), qm))[2]
            assert _bits(float(thr)) == _bits(cache[key]), (qm, j, qq, k)
        assert len(qqs) > 1  # the workload draws several query widths


def test_driver_same_attribute_multi_query_matches_oracle(gpu, store):
    """IMP 4: one report per (k, qq), every query of the workload on attr[qq-2]."""
    n, d, distr, seed, xn = store
    wl = 8
    out = _run(["-n", n, "-d", d, "-s", 16, "-e", 32, "-q", 2, "-i", 1, "-t", distr, "-r", seed, "-a", "mq4", "-w",
                wl, "--mq-thresholds"])
    reps = _reports(out)
    assert [(r[1], r[2]) for r in reps] == [(qq, k) for k in (16, 32) for qq in range(2, d + 1, 2)]
    w = np.ones(d, np.float32)
    for _, qq, k, kv in reps:
        for key in KEYS:
            assert len(kv[key]) == 1 and float(kv[key][0]) >= 0, (key, kv)
        rthr = oracle.fte_topk(xn, n, k, w, _attrs(d, qq, 0))[2]
        lines = kv["mq_threshold"]
        assert len(lines) == wl
        for line in lines:
            _, q2, k2, thr = line.split()
            assert int(q2) == qq and int(k2) == k
            assert _bits(float(thr)) == _bits(rthr), (qq, k)
