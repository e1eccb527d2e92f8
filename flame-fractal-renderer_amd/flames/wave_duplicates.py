"""In-wave duplicate rate of plotted cells (measurement for the accumulation
design; numpy only, runs on the CPU).

A wavefront of the walker kernel runs 64 walkers in lock step, so one
iteration issues up to 64 cell updates at once.  Merging the updates of
lanes that hit the same cell before the global atomic (ballot / readlane)
saves at most the share of those updates that repeat a cell another lane of
the same wave-iteration already holds.  This script measures that share on
the benchmark flames by simulating 64-walker groups with numpy random
numbers (the same distributions as the walker streams: the xform chosen by
cumulative weight, the point from the flame's maps) and counting, per
iteration, the updates whose cell repeats within the group: over all
plotted samples, and over the samples that miss the per-workgroup LDS cell
cache (its slots, hash and 4-slot probe modelled after the kernel's), which
are the ones that reach the global atomics.

    python wave_duplicates.py [--groups 64] [--iters 2000]
"""

from __future__ import annotations

import argparse
import json
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
EPS = 1e-20


def load(path):
    with open(path) as f:
        return json.loads(re.sub(r"^\s*//.*$", "", f.read(), flags=re.M))


def affine(d, dims, key):
    if key in d:
        return np.array(d[key]["A"], float), np.array(d[key]["b"], float)
    return np.eye(dims), np.zeros(dims)


def var2d(name, x, y):
    r2 = x * x + y * y
    if name == "linear":
        return x, y
    if name == "sinusoidal":
        return np.sin(x), np.sin(y)
    if name == "spherical":
        k = 1.0 / (r2 + EPS)
        return x * k, y * k
    if name == "swirl":
        s, c = np.sin(r2), np.cos(r2)
        return x * s - y * c, x * c + y * s
    if name == "horseshoe":
        k = 1.0 / (np.sqrt(r2) + EPS)
        return (x - y) * (x + y) * k, 2.0 * x * y * k
    if name == "polar":
        return np.arctan2(y, x) / np.pi, np.sqrt(r2) - 1.0
    raise ValueError(f"variation {name} is not simulated here")


def cache_slots(flame, lds_kib=144):
    """slots of the walker kernel's LDS cell cache for this flame: entries of
    a key, a count and the colour words (8 bytes each in the default build),
    a power of two within the LDS budget"""
    r = int(flame.get("color_dimensions", 0))
    slot_bytes = 8 + 8 * r
    s = 1
    while 2 * s * slot_bytes <= lds_kib * 1024:
        s *= 2
    return s


class Cache:
    """the LDS cell cache of one workgroup: a cell claims an empty slot of
    its 4-slot probe window on first touch and keeps it; later touches of a
    held cell are hits, cells whose window is full miss"""

    def __init__(self, slots, cells):
        self.direct = slots >= cells
        self.slots = slots
        self.keys = np.full(slots, -1, np.int64)
        self.log2 = slots.bit_length() - 1

    def lookup(self, cell):
        if self.direct:
            return True
        h = ((cell * 0x9E3779B1) & 0xFFFFFFFF) >> (32 - self.log2)
        for w in range(4):
            s = (h + w) & (self.slots - 1)
            k = self.keys[s]
            if k == cell:
                return True
            if k < 0:
                self.keys[s] = cell
                return True
        return False


def simulate(flame, groups, iters, settle=53, seed=1, cached=False):
    dims = flame["dimensions"]
    xfs = [x for x in flame["xforms"] if x.get("weight", 1.0) > 0]
    w = np.array([x.get("weight", 1.0) for x in xfs])
    cw = np.cumsum(w / w.sum())
    cw[-1] = 1.0
    pre = [affine(x, dims, "pre_affine") for x in xfs]
    post = [affine(x, dims, "post_affine") for x in xfs]
    size = np.array(flame["size"][:dims])
    lo = np.array([b[0] for b in flame["bounds"]], float)
    hi = np.array([b[1] for b in flame["bounds"]], float)
    mult = size / (hi - lo)
    rng = np.random.default_rng(seed)
    n = groups * 64
    p = rng.uniform(-1, 1, (n, dims))
    repeats = plotted = 0
    # one cache per 16 waves (a 1024-lane workgroup)
    caches = ([Cache(cache_slots(flame), int(np.prod(size))) for _ in range((groups + 15) // 16)]
              if cached else None)
    for it in range(settle + iters):
        k = np.searchsorted(cw, rng.uniform(0, 1, n))
        out = np.empty_like(p)
        for j, x in enumerate(xfs):
            m = k == j
            if not m.any():
                continue
            t = p[m] @ pre[j][0].T + pre[j][1]
            v = np.zeros_like(t)
            for var in x["variations"]:
                if dims == 2:
                    ox, oy = var2d(var["name"], t[:, 0], t[:, 1])
                    v[:, 0] += var["weight"] * ox
                    v[:, 1] += var["weight"] * oy
                else:
                    v += var["weight"] * t          # linear only in 3-D flames here
            out[m] = v @ post[j][0].T + post[j][1]
        p = out
        bad = ~np.isfinite(p).all(axis=1) | (np.abs(p) > 1e20).any(axis=1)
        p[bad] = rng.uniform(-1, 1, (int(bad.sum()), dims))
        if it < settle:
            continue
        inside = ((p >= lo) & (p <= hi)).all(axis=1)
        idx = np.zeros(n, np.int64)
        stride = 1
        for i in range(dims):
            c = np.clip(((p[:, i] - lo[i]) * mult[i]).astype(np.int64), 0, size[i] - 1)
            idx += c * stride
            stride *= int(size[i])
        idx[~inside] = -1 - np.arange(n)[~inside]      # never equal to another lane
        if cached:
            # only the misses of the LDS cache reach the global atomics
            for lane in np.nonzero(inside)[0]:
                if caches[lane // 1024].lookup(int(idx[lane])):
                    idx[lane] = -1 - lane
            inside = idx >= 0
        g = idx.reshape(groups, 64)
        gs = np.sort(g, axis=1)
        dup = (gs[:, 1:] == gs[:, :-1]) & (gs[:, 1:] >= 0)
        repeats += int(dup.sum())
        plotted += int(inside.sum())
    return repeats / max(plotted, 1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--groups", type=int, default=64)
    ap.add_argument("--iters", type=int, default=2000)
    args = ap.parse_args()
    names = ["random_color_3840x2160.json", "sierpinski_1024.json", "tetrahedron_512.json"]
    batch = sorted(os.listdir(os.path.join(HERE, "batch")))
    big = [n for n in batch if load(os.path.join(HERE, "batch", n))["size"][0] == 2048][:4]
    small = [n for n in batch if load(os.path.join(HERE, "batch", n))["size"][0] == 256][:2]
    paths = [os.path.join(HERE, n) for n in names]
    paths += [os.path.join(HERE, "batch", n) for n in big + small]
    out = {}
    for path in paths:
        f = load(path)
        rate = simulate(f, args.groups, args.iters)
        miss = simulate(f, args.groups, args.iters, cached=True)
        key = os.path.relpath(path, HERE)
        out[key] = {"cells": int(np.prod(f["size"])), "repeat_share": rate,
                    "repeat_share_of_cache_misses": miss}
        print(f"{key:38s} {out[key]['cells']:>10d} cells  in-wave repeats {100 * rate:6.2f} %"
              f", among LDS-cache misses {100 * miss:6.2f} %", flush=True)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
