"""Writes tests/golden/golden.npz: small fixtures computed with pure numpy
(float32 arithmetic element by element, no FMA), independent of the C oracle
and of the engine library.

* gen_<distr>: the first 300 objects (d = 8, seed 7) of each synthetic
  distribution, as column-major float32 bit patterns.  The generator is the
  engine's counter-based one (SplitMix64 finaliser, 24-bit uniforms,
  Irwin-Hall normals); restated here from its definition.
* c1_*: a reduced C1 configuration (N = 2^16, d = 8, independent, seed 0;
  attr = {4, 5, 6, 7} as bench_cpu.h:41-44 row qq = 4, weights 1.0, k = 16)
  and a weighted, unordered query on the correlated distribution (ties at
  the clipped value 1.0): ids and score bits of the top k by (score desc,
  id asc), scores summed in the query's attribute order (cpu/TPAc.h:129-142).

Run: python3 tests/golden/make_golden.py  (a few seconds)."""
import os

import numpy as np

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def mix64(z):
    z = np.asarray(z, np.uint64)
    with np.errstate(over="ignore"):
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def rbits24(sk, i, slot, tag):
    with np.errstate(over="ignore"):
        c = (i * np.uint64(64) + np.uint64(slot)) * np.uint64(16) + np.uint64(tag)
        return (mix64(sk + c) >> np.uint64(40)).astype(np.int64)


TWO_M24 = np.float32(2.0 ** -24)


def runif(sk, i, slot, tag):
    return rbits24(sk, i, slot, tag).astype(np.float32) * TWO_M24


def rnorm(sk, i, slot):
    s = np.zeros(i.shape, np.int64)
    for t in range(1, 13):
        s += rbits24(sk, i, slot, t)
    s -= 6 * 16777216
    return s.astype(np.float32) * TWO_M24


def clip01(x):
    return np.where(x < 0, np.float32(0), np.where(x > 1, np.float32(1), x)).astype(np.float32)


def generate(n, d, distr, seed, count=None):
    """(d, count) float32 column-major; distr 0 correlated, 1 independent,
    2 anticorrelated, 3 mixed (object-range thirds)."""
    count = n if count is None else count
    sk = mix64(np.uint64(seed) ^ np.uint64(0x5DEECE66D))
    i = np.arange(count, dtype=np.uint64)
    out = np.zeros((d, count), np.float32)
    kinds = np.full(count, distr)
    if distr == 3:
        idx = np.arange(count)
        kinds = np.where(idx < n // 3, 0, np.where(idx < 2 * (n // 3), 1, 2))
    for kind in (0, 1, 2):
        sel = kinds == kind
        if not sel.any():
            continue
        ii = i[sel]
        if kind == 1:
            for m in range(d):
                out[m, sel] = runif(sk, ii, m, 0)
        elif kind == 0:
            u = runif(sk, ii, 63, 0)
            for m in range(d):
                z = rnorm(sk, ii, m) * np.float32(0.05)
                out[m, sel] = clip01(u + z)
        else:
            v = [runif(sk, ii, m, 0) for m in range(d)]
            s = np.zeros(ii.shape, np.float32)
            for m in range(d):
                s = s + v[m]
            z = rnorm(sk, ii, 63) * np.float32(0.05)
            target = np.float32(d) * np.float32(0.5) + z
            c = (target - s) / np.float32(d)
            for m in range(d):
                out[m, sel] = clip01(v[m] + c)
    return out


def topk(x, k, weights, attr):
    s = np.zeros(x.shape[1], np.float32)
    for a in attr:
        s = s + x[a] * np.float32(weights[a])
    order = np.lexsort((np.arange(s.size), -s.astype(np.float64)))[:k]
    return order.astype(np.uint64), s[order]


def main():
    out = {}
    for distr in range(4):
        out[f"gen_{distr}"] = generate(300, 8, distr, 7).view(np.uint32)
    n, d = 1 << 16, 8
    x = generate(n, d, 1, 0)
    ids, sc = topk(x, 16, np.ones(d, np.float32), [4, 5, 6, 7])
    out["c1_ids"], out["c1_scores"] = ids, sc.view(np.uint32)
    xc = generate(n, d, 0, 2)
    w = np.array([0.5, 0.25, 0.75, 1.0, 0.125, 0.375, 0.625, 0.875], np.float32)
    attr = np.array([6, 1, 3, 0], np.uint8)
    ids, sc = topk(xc, 64, w, attr)
    out["cw_weights"], out["cw_attr"] = w, attr
    out["cw_ids"], out["cw_scores"] = ids, sc.view(np.uint32)
    np.savez(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.npz"), **out)


if __name__ == "__main__":
    main()
