"""Benchmark: buffer-render throughput on one MI355X.

Headline workload (one "step"): one render pass of 2^32 samples of the
random 2-D colour flame (seed 2, 6 xforms x 3 variations from linear /
sinusoidal / spherical / swirl / horseshoe / polar) into its 3840x2160
buffer of uint64 counts + 3 double colour sums per cell (the reference's
default buffer layout, what ffr-buf writes), walkers chosen by the
renderer.  Measured with FP32 trajectories (the reported `value`) and with
FP64 trajectories; `l1_fp32_vs_fp64` is the L1 distance between the two
normalised histograms.

Also measured (sub-objects of `workloads`):
  - the Sierpinski triangle, 1024x1024 uint64 counts, 2^26 samples from
    1024 walkers of 65536 with 20 settle iterations, FP64 and FP32;
  - the 3-D Sierpinski tetrahedron, 512^3 uint32 counts, 2^33 samples,
    FP32 trajectories;
  - a batch of 256 seeded random colour flames (flames/batch: 2-8 xforms,
    256^2 to 2048^2 cells, 1000 samples per cell, 3.6e11 samples in all)
    rendered in one batched pass (persistent launch, HIP graph, launch shape
    from the tuner), FP32, FP64 and under the precision policy, with the
    per-flame L1 against FP64;
  - serving: 200 jobs arriving as a Poisson process at 200 per second (90 %
    from the batch set, 10 % 3-D 256^3 u32 jobs), 4 chunks each, submitted
    to the render server at their arrival times; throughput and p50/p99 job
    latency, FP32 and FP64.  `p99_latency_s` is the FP32 serving p99.

`image` is the ffr-img stage on the FP64 headline buffer (the reference's
default build): the device tonemap kernel's time and bytes for colour at 8
bits and gray at 16 bits, the wall time of the call that returns host
pixels, and the CPU oracle's single-thread time for the same image.

Prints one JSON line on stdout.  `roofline` is for the walker kernel:
algorithmic bytes = plotted samples x the bytes each plot adds into its
cell (32 B: an 8-byte count and three 8-byte colour sums; 8 B for the
Sierpinski counts, 4 B for the tetrahedron counts) over the kernel's
average duration from HIP events recorded on the renderer's stream.
`traffic` is the walker kernel's HBM bytes per launch from the profiler's
FETCH_SIZE and WRITE_SIZE counters on this workload (the profile summary
under profiles/ names the command), scaled by the same counters on
kernels of known traffic.  `cpu_baseline` times the CPU oracle (the
reference loop restated in C, threads = host cores) on a bounded sample
of the same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
                       [--only c2|c1|c3] [--fp32-only] [--calibrate]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "flame-fractal-renderer_amd")
sys.path.insert(0, PKG)
sys.path.insert(0, ROOT)

FLAMES = os.path.join(PKG, "flames")
C2_FLAME = os.path.join(FLAMES, "random_color_3840x2160.json")
C1_FLAME = os.path.join(FLAMES, "sierpinski_1024.json")
C3_FLAME = os.path.join(FLAMES, "tetrahedron_512.json")
C2_SAMPLES = 1 << 32
C1_SAMPLES = 1 << 26
C3_SAMPLES = 1 << 33
C2_SEED, C1_SEED, C3_SEED = 0x5EED0002, 0x5EED0001, 0x5EED0003
C4_DIR = os.path.join(FLAMES, "batch")
C4_SEED = 0x5EED0004
C4_SAMPLES_PER_CELL = 1000
C4_TUNE_FRACTION = 1.0 / 256
# serving trace: Poisson arrivals, 90 % jobs drawn from the 256-flame batch
# set (1000 samples per cell), 10 % 3-D jobs at 256^3 u32 (2^30 samples,
# 64 per cell), 4 chunks each, 200 GB of job buffers at most
C5_DIR = os.path.join(FLAMES, "serve3d")
C5_JOBS = 200
C5_RATE = 200.0
C5_SEED = 5
C5_3D_SHARE = 0.1
C5_3D_SAMPLES = 1 << 30
C5_CHUNKS = 4
C5_BUDGET = 200_000_000_000
HBM_PEAK_GBS = 8000.0
# walker kernel HBM bytes per launch on the headline workload, from two
# counter runs (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch, the timed
# step's dispatch): FP32 (native trajectories, table-driven kernel) FETCH
# 901.125 x 2.0 + WRITE 18719063.3125 x 1.0; FP64 (the flame's compiled
# kernel, which the probe keeps) FETCH 1001.1875 x 2.0 + WRITE
# 18871357.5625 x 1.0.  The factors come from kernels of known traffic on
# the same buffer (the add reads 2 x 265420800 B and reports 259236 KiB of
# FETCH_SIZE, the histogram pass reads the buffer once and reports half of
# it; the clear and the add each write 265420800 B and report exactly
# 259200 KiB of WRITE_SIZE).  Global atomics count 32 B each in WRITE_SIZE.
# The raw values are in profiles/walk_kernel_random_color_3840x2160_hbm.md.
C2_FP32_TRAFFIC_BYTES = 901.125 * 1024 * 2.0 + 18719063.3125 * 1024 * 1.0
C2_FP64_TRAFFIC_BYTES = 1001.1875 * 1024 * 2.0 + 18871357.5625 * 1024 * 1.0
# HBM bytes of one binned tetrahedron render (32 passes of 2^28 records):
# per pass walker 1.10 GB, tile histogram 1.09, column scan 0.03, scatter
# 2.53, tile accumulation 1.11 (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0)
C3_BINNED_TRAFFIC_BYTES = 32 * 5.86e9
# VALU instructions (SQ_INSTS_VALU, wave64) per headline launch of 2^32
# samples, by precision and kernel (table-driven or the flame's compiled
# one), from counter runs of the same launches
# (profiles/walk_kernel_random_color_3840x2160_valu.md); the compiled FP32
# kernel was counted only on 2^29-sample runs (x 8 here).  The VALU peak is
# one wave64 instruction per 4 cycles per SIMD (16 lanes a cycle, FP32 and
# FP64 alike on this device): 256 CUs x 4 SIMDs x 2.4 GHz / 4.
C2_VALU_PER_LAUNCH = {("fp32", False): 35077692829, ("fp32", True): 4069185530 * 8,
                      ("fp64", False): 51868535255, ("fp64", True): 46629032406}
VALU_PEAK_G = 256 * 4 * 2.4 / 4


def valu_roofline(precision, res):
    """the headline launch's VALU issue rate against the device's peak"""
    n = C2_VALU_PER_LAUNCH[(precision, bool(res["compiled"]))]
    achieved = n / (res["kernel_ms"] * 1e-3) / 1e9
    return {"achieved": achieved, "peak": VALU_PEAK_G, "unit": "G VALU instructions/s",
            "frac": achieved / VALU_PEAK_G, "instructions_per_launch": n}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read(path):
    with open(path) as fh:
        return fh.read()


# FP32 renders compared with FP64 ones take the FP64 settle count, so both
# draw the same words on every iteration and only the precision differs
# (the reference's defaults are 53 for double and 24 for float)
SETTLE_FP64 = 53


# the image stage (ffr-img) on the rendered headline buffer: gamma 2.2,
# colour at 8 bits (reads 32 B per cell, writes 3 B) and gray at 16 bits
# (reads the 8-byte count of each 32-byte cell, writes 2 B)
IMAGE_LEGS = (("color_8bit", 2, 8, 32), ("gray_16bit", 1, 16, 8))
IMAGE_GAMMA = 2.2


def image_leg(ffr, r, words, cpu):
    """Device tonemap kernel time (HIP events on the renderer's stream, pixels
    left in device memory), the host-pixel call's wall time (kernel, bounds
    pass and the copy to host), and, unless cpu is False, the oracle's
    single-thread image of the same buffer (the reference's render_image is
    one thread) with the count of differing pixels."""
    import numpy as np

    out = {}
    for name, mode, bits, read_b in IMAGE_LEGS:
        n = r.image_bytes(mode, bits)
        dev = ffr.DeviceMemory(n)
        ms = []
        for i in range(6):
            info = r.render_image_device(dev.ptr, n, mode, IMAGE_GAMMA, bits)
            if i:
                ms.append(info.kernel_ms)
        k = sum(ms) / len(ms)
        alg = r.cells * read_b + n
        t0 = time.perf_counter()
        pix, info = r.render_image(mode, IMAGE_GAMMA, bits)
        host_ms = 1e3 * (time.perf_counter() - t0)
        res = {
            "kernel_ms": k,
            "algorithmic_bytes": alg,
            "achieved_GBs": alg / (k * 1e-3) / 1e9,
            "frac": alg / (k * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "host_pixels_ms": host_ms,
        }
        if cpu:
            from oracle import oracle
            w, h = int(info.width), int(info.height)
            t0 = time.perf_counter()
            st, ref, _, _ = oracle.tonemap(words, w, h, r.cell_words, mode, IMAGE_GAMMA, bits)
            res["cpu_oracle_ms"] = 1e3 * (time.perf_counter() - t0)
            res["pixels_differing_from_oracle"] = int(np.count_nonzero(pix != ref))
        out[name] = res
        dev.close()
    return out


_atomic_peak = {}


def atomic_roofline(ffr, atomics, kernel_ms, cells, cell_words, word_bytes, flush_atomics=None):
    """The walker kernel's global atomic adds per second (counted by the
    kernel: misses of the LDS cell cache and its flushes) against the rate
    the device sustains for cell updates scattered uniformly over a buffer
    of the same shape (ffr_atomic_rate, 2^28 cell updates)."""
    key = (cells, cell_words, word_bytes)
    if key not in _atomic_peak:
        _atomic_peak[key] = ffr.atomic_rate(cells, cell_words, word_bytes, 1 << 28)
    peak = _atomic_peak[key] / 1e9
    achieved = atomics / (kernel_ms * 1e-3) / 1e9
    out = {"achieved": achieved, "peak": peak, "unit": "G atomics/s",
           "frac": achieved / peak if peak > 0 else None, "atomics_per_launch": atomics,
           "peak_buffer": f"{cells} cells x {cell_words} words of {word_bytes} B"}
    if flush_atomics is not None:
        # the LDS cache's flushes, reported apart: one 32-B atomic per word
        # of every occupied slot at the end of each task
        out["flush_atomics_per_launch"] = flush_atomics
        out["flush_bytes_per_launch"] = flush_atomics * 32
    return out


def run_gpu(ffr, text, *, precision, samples, steps, warmup, bytes_per_plot, seed,
            word=64, batch=0, settle=-1, keep=False, image=False, cpu=True):
    r = ffr.BufferRenderer(text, precision=precision, word=word, seed=seed,
                           batch_size=batch, settle=settle)
    for _ in range(warmup):
        if not r.render(samples):
            raise RuntimeError("render stopped at the bad-value limit")
    plot0 = r.samples_plotted
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        if not r.render(samples):
            raise RuntimeError("render stopped at the bad-value limit")
        kms.append(r.kernel_ms)
    t1 = time.perf_counter()
    plotted = (r.samples_plotted - plot0) / steps
    atomics = int(r.last_stats.atomics)
    flush_atomics = int(r.last_stats.flush_atomics)
    wall = t1 - t0
    kavg = sum(kms) / len(kms)
    achieved = plotted * bytes_per_plot / (kavg * 1e-3) / 1e9
    out = {
        "value": steps * samples / wall,
        "ms_per_step": 1e3 * wall / steps,
        "kernel_ms": kavg,
        "walkers": int(r.last_stats.walkers),
        "batch": int(r.last_stats.batch_size),
        "binned": bool(r.last_stats.binned),
        "compiled": bool(r.last_stats.compiled),
        "plotted_per_step": plotted,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
        },
        "atomic_roofline": atomic_roofline(ffr, atomics, kavg, r.cells, r.cell_words,
                                           r.word_bytes, flush_atomics),
    }
    hist = None
    if keep:
        # the buffer's trip to host memory (ffr_read_buffer: device to host
        # copy into pageable memory); not part of `value`
        t0 = time.perf_counter()
        words = r.read_buffer()
        out["read_buffer_ms"] = 1e3 * (time.perf_counter() - t0)

        class _Sink:
            def write(self, b):
                return len(b)
        # the streamed read: pinned chunks, the copy of one overlapping the
        # caller's write of the last (here a sink that keeps nothing)
        r.write_buffer_stream(_Sink())   # allocates the pinned chunks
        t0 = time.perf_counter()
        r.write_buffer_stream(_Sink())
        out["read_buffer_stream_ms"] = 1e3 * (time.perf_counter() - t0)
        if image:
            out["image"] = image_leg(ffr, r, words, cpu)
        hist = r.counts(words).astype("float64")
        del words
        hist /= max(hist.sum(), 1.0)
    r.close()
    return out, hist


def read_batch_flames():
    import re
    texts = []
    for name in sorted(os.listdir(C4_DIR)):
        if name.endswith(".json"):
            # the files open with a // comment line, which the renderer accepts
            texts.append(read(os.path.join(C4_DIR, name)))
    cells = []
    for t in texts:
        d = json.loads(re.sub(r"^\s*//.*$", "", t, flags=re.M))
        cells.append(d["size"][0] * d["size"][1])
    return texts, cells


def run_batch(ffr, texts, samples, precision, keep, seed0=C4_SEED):
    """One batched pass over all flames (after a tuning pass); returns the
    rates and, when keep, each flame's normalised histogram.  `precision` is
    one name for all flames or a list of one per flame."""
    precs = precision if isinstance(precision, list) else [precision] * len(texts)
    rs = [ffr.BufferRenderer(t, precision=p, seed=seed0 + i, settle=SETTLE_FP64)
          for i, (t, p) in enumerate(zip(texts, precs))]
    b = ffr.Batch(rs)
    t0 = time.perf_counter()
    b.tune(samples, C4_TUNE_FRACTION)
    tune_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    ok = b.render(samples)
    wall = time.perf_counter() - t0
    if not all(ok):
        raise RuntimeError("a batched flame stopped at the bad-value limit")
    st = b.last_stats
    plotted = sum(r.samples_plotted for r in rs)
    atomics = sum(int(r.last_stats.atomics) for r in rs)
    flush_atomics = sum(int(r.last_stats.flush_atomics) for r in rs)
    big = max(rs, key=lambda r: r.cells)
    hists = None
    if keep:
        hists = []
        for r in rs:
            h = r.counts().astype("float64")
            hists.append(h / max(h.sum(), 1.0))
    out = {
        "value": sum(samples) / wall,
        "ms_per_step": 1e3 * wall,
        "kernel_ms": st.kernel_ms,
        "flames": len(rs),
        "samples": int(sum(samples)),
        "tasks": int(st.tasks),
        "launches": int(st.launches),
        "graph": bool(st.graph),
        "shapes": [list(x) for x in b.shapes()],
        "tune_s": tune_s,
        "roofline": {
            "bound": "hbm",
            "achieved": plotted * 32 / (st.kernel_ms * 1e-3) / 1e9,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": plotted * 32 / (st.kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "traffic": None,
        },
        # priced at the largest flame's buffer shape (most of the samples)
        "atomic_roofline": atomic_roofline(ffr, atomics, st.kernel_ms, big.cells,
                                           big.cell_words, big.word_bytes, flush_atomics),
    }
    b.close()
    for r in rs:
        r.close()
    return out, hists


# per-flame check of the batch's FP32 renders against the FP64 oracle: a
# reduced sample count per flame, the same walker streams and settle count
C4_ORACLE_SAMPLES = 1 << 20


def batch_vs_oracle(ffr, texts):
    """Every batch flame rendered with FP32 trajectories on the GPU (one
    batched pass) and with FP64 on the CPU oracle, C4_ORACLE_SAMPLES each, on
    the same streams: per-flame L1 of the normalised histograms, next to the
    L1 between two FP64 oracle renders on different streams (the sampling
    floor at this count)."""
    import numpy as np
    from oracle import oracle

    n = C4_ORACLE_SAMPLES
    batch = int(ffr.lib().ffr_auto_batch(None, n))
    cores = (len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity")
             else os.cpu_count())
    rs = [ffr.BufferRenderer(t, precision="fp32", seed=C4_SEED + i, batch_size=batch,
                             settle=SETTLE_FP64) for i, t in enumerate(texts)]
    with ffr.Batch(rs) as b:
        if not all(b.render([n] * len(rs))):
            raise RuntimeError("a batched flame stopped at the bad-value limit")
    d, floor = [], []
    t0 = time.perf_counter()
    for i, (t, r) in enumerate(zip(texts, rs)):
        g = r.counts().astype(np.float64)
        r.close()
        f = oracle.OracleFlame(t, oracle.F64_U64, settle=SETTLE_FP64)
        a = f.render(n, batch, seed=C4_SEED + i, threads=cores).counts.astype(np.float64)
        o = f.render(n, batch, seed=C4_SEED + 0x10000 + i, threads=cores).counts.astype(
            np.float64)
        d.append(float(np.abs(g / g.sum() - a / a.sum()).sum()))
        floor.append(float(np.abs(o / o.sum() - a / a.sum()).sum()))
    cpu_s = time.perf_counter() - t0
    return {
        "samples_per_flame": n,
        "l1_fp32_gpu_vs_fp64_oracle": {"mean": sum(d) / len(d), "max": max(d),
                                       "over_1e-3": int(sum(1 for x in d if x > 1e-3))},
        "l1_fp64_oracle_other_streams": {"mean": sum(floor) / len(floor),
                                         "min": min(floor)},
        "per_flame": [round(x, 7) for x in d],
        "oracle_s": cpu_s, "oracle_threads": cores,
    }


def serving_trace(texts2d, cells2d):
    import numpy as np
    texts3d = [read(os.path.join(C5_DIR, n)) for n in sorted(os.listdir(C5_DIR))
               if n.endswith(".json")]
    rng = np.random.default_rng(C5_SEED)
    at = np.cumsum(rng.exponential(1.0 / C5_RATE, C5_JOBS))
    jobs = []
    for i in range(C5_JOBS):
        if rng.uniform() < C5_3D_SHARE:
            jobs.append((float(at[i]), texts3d[int(rng.integers(len(texts3d)))], C5_3D_SAMPLES,
                         True))
        else:
            k = int(rng.integers(len(texts2d)))
            jobs.append((float(at[i]), texts2d[k], C4_SAMPLES_PER_CELL * cells2d[k], False))
    return jobs


def run_serving(ffr, trace, precision):
    """Submit the trace at its arrival times to a server (the client thread
    sleeps until each arrival), wait for every job, report throughput and
    job latency percentiles."""
    import numpy as np
    with ffr.Server(hbm_budget=C5_BUDGET) as srv:
        t0 = time.perf_counter()
        ids = []
        for i, (at, text, n, is3d) in enumerate(trace):
            delay = at - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            word = 32 if (is3d and precision == "fp32") else 64
            ids.append(srv.submit(text, n, C5_CHUNKS, precision=precision, word=word,
                                  seed=C5_SEED * 1000 + i,
                                  settle=SETTLE_FP64 if precision == "fp32" else -1))
        infos = []
        for k, j in enumerate(ids):
            infos.append(srv.wait(j, timeout=900))
            if (k + 1) % 50 == 0:
                log(f"serving {precision}: {k + 1}/{len(ids)} jobs done")
        wall = time.perf_counter() - t0
        st = srv.stats()
        for j in ids:
            srv.release(j)
    if any(x.state != ffr.JOB_DONE for x in infos):
        raise RuntimeError("a served job failed")
    lat = np.array([x.done_s - x.submit_s for x in infos])
    first = np.array([x.first_chunk_s - x.submit_s for x in infos])
    total = sum(n for _, _, n, _ in trace)
    return {
        "value": total / wall,
        "unit": "samples/s",
        "jobs": len(trace),
        "samples": int(total),
        "wall_s": wall,
        "arrivals_s": trace[-1][0],
        "latency_p50_s": float(np.percentile(lat, 50)),
        "latency_p99_s": float(np.percentile(lat, 99)),
        "first_chunk_p50_s": float(np.percentile(first, 50)),
        "first_chunk_p99_s": float(np.percentile(first, 99)),
        "passes": int(st.passes),
        "device_busy_frac": st.busy_ms / 1e3 / wall,
        "peak_buffer_bytes": int(st.peak_bytes),
    }


def l1(a, b):
    import numpy as np
    return float(np.abs(a - b).sum())


def calibration_pass(ffr, text):
    """Kernels of known traffic on the headline buffer (8294400 cells x 32 B):
    the clear writes it once, the histogram statistics read it once, the add
    reads it twice and writes it once.  Their FETCH_SIZE / WRITE_SIZE
    counters scale the walker kernel's into bytes."""
    import numpy as np

    r = ffr.BufferRenderer(text, precision="fp32", seed=C2_SEED)
    zeros = np.zeros(r.cells * r.cell_words, dtype=np.uint64)
    r.add_buffer(zeros)
    r.histogram_stats()
    log(f"calibration: clear writes {r.buffer_bytes} bytes; histogram pass reads "
        f"{r.buffer_bytes} bytes; add reads {2 * r.buffer_bytes}, writes {r.buffer_bytes}")
    r.close()


def cpu_rate(text, mode, seconds, seed, batch, settle=-1, fixed=None, cores=None):
    """Samples/s of the CPU oracle on all host threads (or `cores`): either
    `fixed` samples, or about `seconds` of work sized from a short probe."""
    from oracle import oracle

    if cores is None:
        cores = (len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity")
                 else os.cpu_count())
    f = oracle.OracleFlame(text, mode, settle=settle)
    if fixed is None:
        n = 1 << 24
        t0 = time.perf_counter()
        f.render(n, batch, seed=seed, threads=cores)
        rate = n / (time.perf_counter() - t0)
        n = int(min(max(rate * seconds, 1 << 20), 1 << 31)) // batch * batch
    else:
        n = fixed
    t0 = time.perf_counter()
    f.render(n, batch, seed=seed, threads=cores)
    dt = time.perf_counter() - t0
    return n / dt, n, dt, cores


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--only", choices=["c1", "c2", "c3", "c4", "c5"], default=None,
                    help="run one workload only (profiling)")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run kernels of known traffic (counter scaling)")
    ap.add_argument("--fp32-only", action="store_true")
    ap.add_argument("--float-build", action="store_true",
                    help="config 2 in FP32 with 4-byte words only: the reference's float "
                         "build (float trajectories, FP64 literals promote, u32 counts and "
                         "f32 colour sums), for counter comparisons with the native FP32 "
                         "trajectories")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the renderer has no multi-GPU path "
              "(walker ranges would shard across GPUs with one integer sum of the "
              "buffers at the end)", file=sys.stderr)
        return 2
    import ffr
    from oracle import oracle

    S, W = args.steps, args.warmup
    want = (lambda k: args.only in (None, k))
    c2_text = read(C2_FLAME)
    if args.calibrate:
        calibration_pass(ffr, c2_text)

    fp32 = fp64 = None
    l1_c2 = floor_c2 = None
    if args.float_build:
        res, _ = run_gpu(ffr, c2_text, precision="fp32", word=32, samples=C2_SAMPLES, steps=S,
                         warmup=W, bytes_per_plot=16, seed=C2_SEED, settle=SETTLE_FP64)
        log("c2 float build", json.dumps(res))
        return 0
    if want("c2"):
        log(f"bench c2: {C2_SAMPLES} samples/step, warmup {W}, steps {S}")
        fp32, h32 = run_gpu(ffr, c2_text, precision="fp32", samples=C2_SAMPLES, steps=S,
                            warmup=W, bytes_per_plot=32, seed=C2_SEED, settle=SETTLE_FP64,
                            keep=not args.fp32_only)
        fp32["roofline"]["traffic"] = C2_FP32_TRAFFIC_BYTES
        fp32["valu_roofline"] = valu_roofline("fp32", fp32)
        log("c2 fp32", json.dumps(fp32))
        if not args.fp32_only:
            fp64, h64 = run_gpu(ffr, c2_text, precision="fp64", samples=C2_SAMPLES, steps=S,
                                warmup=W, bytes_per_plot=32, seed=C2_SEED, keep=True,
                                image=True, cpu=not args.no_cpu)
            fp64["roofline"]["traffic"] = C2_FP64_TRAFFIC_BYTES
            fp64["valu_roofline"] = valu_roofline("fp64", fp64)
            l1_c2 = l1(h32, h64)
            del h32
            # the sampling floor: FP64 against FP64 on other streams
            _, h64b = run_gpu(ffr, c2_text, precision="fp64", samples=C2_SAMPLES, steps=1,
                              warmup=0, bytes_per_plot=32, seed=C2_SEED + 0x100, keep=True)
            floor_c2 = l1(h64, h64b)
            del h64, h64b
            log("c2 fp64", json.dumps(fp64), "L1", l1_c2, "floor", floor_c2)

    workloads = {}
    if want("c1") and not args.fp32_only:
        text = read(C1_FLAME)
        for prec in ("fp64", "fp32"):
            log(f"bench c1 {prec}: {C1_SAMPLES} samples/step")
            res, h = run_gpu(ffr, text, precision=prec, samples=C1_SAMPLES, steps=S, warmup=W,
                             bytes_per_plot=8, seed=C1_SEED, batch=65536, settle=20, keep=True)
            res["hist"] = h
            workloads[f"sierpinski_1024_{prec}"] = res
        h64 = workloads["sierpinski_1024_fp64"].pop("hist")
        h32 = workloads["sierpinski_1024_fp32"].pop("hist")
        workloads["sierpinski_1024_fp32"]["l1_vs_fp64"] = l1(h32, h64)
        _, h64b = run_gpu(ffr, text, precision="fp64", samples=C1_SAMPLES, steps=1, warmup=0,
                          bytes_per_plot=8, seed=C1_SEED + 0x100, batch=65536, settle=20,
                          keep=True)
        workloads["sierpinski_1024_fp32"]["l1_fp64_vs_fp64_other_seed"] = l1(h64b, h64)
        # what the precision policy picks for this render (its probe renders
        # 1/64 of the samples in both precisions on the same streams)
        choice, probe_l1 = ffr.choose_precision(text, C1_SAMPLES, seed=C1_SEED, batch_size=65536,
                                                settle=20)
        workloads["sierpinski_1024_fp32"]["policy"] = {"choice": choice, "probe_l1": probe_l1}
        if not args.no_cpu:
            rate, n, dt, cores = cpu_rate(text, oracle.F64_U64, 0, C1_SEED, 65536, settle=20,
                                          fixed=C1_SAMPLES)
            workloads["sierpinski_1024_fp64"]["cpu"] = {
                "value": rate, "cores": cores, "kind": "port",
                "sample": f"the full {n} samples in FP64, {dt:.2f} s"}
        for k in ("sierpinski_1024_fp64", "sierpinski_1024_fp32"):
            log(k, json.dumps(workloads[k]))
    if want("c3") and not args.fp32_only:
        text = read(C3_FLAME)
        log(f"bench c3: {C3_SAMPLES} samples/step")
        res, _ = run_gpu(ffr, text, precision="fp32", word=32, samples=C3_SAMPLES, steps=S,
                         warmup=W, bytes_per_plot=4, seed=C3_SEED)
        if res["binned"]:
            # binned accumulation moves records, not atomics: per sample 4 B
            # written by the walker, the tile histogram's read, the scatter's
            # read and write, the accumulation's read.  The roofline prices
            # those bytes (the render's passes together); the measured HBM
            # traffic of the passes is 5.86 GB per pass of 2^28 records
            # (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 of each kernel,
            # profiles/bin_kernels_tetrahedron_512.md), 32 passes
            res["record_bytes_per_sample"] = 20
            res["record_traffic_GBs"] = (20 * C3_SAMPLES / (res["kernel_ms"] * 1e-3) / 1e9)
            res["roofline"] = {"bound": "hbm", "achieved": res["record_traffic_GBs"],
                               "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": res["record_traffic_GBs"] / HBM_PEAK_GBS,
                               "traffic": C3_BINNED_TRAFFIC_BYTES}
        workloads["tetrahedron_512_fp32"] = res
        log("tetrahedron_512_fp32", json.dumps(res))

    if want("c4") and not args.fp32_only:
        texts, cells = read_batch_flames()
        samples = [C4_SAMPLES_PER_CELL * c for c in cells]
        log(f"bench c4: {len(texts)} flames, {sum(samples)} samples in one batched pass")

        def l1_summary(hs, ref):
            d = [l1(a, b) for a, b in zip(hs, ref)]
            return d, {"mean": sum(d) / len(d), "max": max(d),
                       "over_1e-3": int(sum(1 for x in d if x > 1e-3))}

        res64, h64 = run_batch(ffr, texts, samples, "fp64", keep=True)
        log("batch_256_fp64", json.dumps(res64))
        res32, h32 = run_batch(ffr, texts, samples, "fp32", keep=True)
        d32, res32["l1_vs_fp64"] = l1_summary(h32, h64)
        del h32
        # the sampling floor per flame: FP64 against FP64 on other streams
        _, h64b = run_batch(ffr, texts, samples, "fp64", keep=True, seed0=C4_SEED + 0x10000)
        _, res32["l1_fp64_vs_fp64_other_seed"] = l1_summary(h64b, h64)
        del h64b
        log("batch_256_fp32", json.dumps(res32))
        # the precision policy: a short probe per flame on the same streams;
        # FP64 where FP32 and FP64 drift apart
        t0 = time.perf_counter()
        policy, probe = ffr.choose_precisions(texts, samples,
                                              [C4_SEED + i for i in range(len(texts))])
        probe_s = time.perf_counter() - t0
        resp, hp = run_batch(ffr, texts, samples, policy, keep=True)
        _, resp["l1_vs_fp64"] = l1_summary(hp, h64)
        del hp, h64
        flagged = [i for i, p in enumerate(policy) if p == "fp64"]
        resp["policy"] = {
            "fp64_flames": len(flagged), "probe_s": probe_s,
            "target": ffr.POLICY_L1_TARGET, "probe_fraction": ffr.POLICY_PROBE_FRACTION,
            "fp32_l1_of_fp32_flames_max": max([d32[i] for i in range(len(texts))
                                               if policy[i] == "fp32"] or [0.0]),
            "fp32_l1_over_1e-3_caught": int(sum(1 for i in flagged if d32[i] > 1e-3)),
            "fp32_l1_over_1e-3": int(sum(1 for x in d32 if x > 1e-3)),
            "probe_l1_max_fp32_flames": max([probe[i] for i in range(len(texts))
                                             if policy[i] == "fp32"] or [0.0]),
        }
        if not args.no_cpu:
            # the FP32 renders against the FP64 oracle (not the GPU's FP64)
            res32["vs_fp64_oracle"] = batch_vs_oracle(ffr, texts)
            log("batch_256_fp32 vs the FP64 oracle",
                json.dumps({k: v for k, v in res32["vs_fp64_oracle"].items()
                            if k != "per_flame"}))
        workloads["batch_256_fp32"] = res32
        workloads["batch_256_fp64"] = res64
        workloads["batch_256_policy"] = resp
        log("batch_256_policy", json.dumps(resp))
        c1cpu = workloads.get("sierpinski_1024_fp64", {}).get("cpu")
        if c1cpu:
            for k in ("batch_256_fp32", "batch_256_policy"):
                workloads[k]["vs_cpu_sierpinski_1024"] = workloads[k]["value"] / c1cpu["value"]

    if want("c5") and not args.fp32_only:
        texts, cells = read_batch_flames()
        trace = serving_trace(texts, cells)
        log(f"bench c5: {len(trace)} jobs over {trace[-1][0]:.2f} s of arrivals")
        for prec in ("fp32", "fp64"):
            res = run_serving(ffr, trace, prec)
            workloads[f"serving_{prec}"] = res
            log(f"serving_{prec}", json.dumps(res))

    if fp32 is None:
        if workloads:
            print(json.dumps({"workloads": workloads}), flush=True)
        return 0
    cpu = None
    if not args.no_cpu and not args.fp32_only:
        rate, n, dt, cores = cpu_rate(c2_text, oracle.F64_U64, args.cpu_seconds, C2_SEED, 8192)
        cpu = {
            "value": rate,
            "unit": "samples/s",
            "cores": cores,
            "kind": "port",
            "sample": f"{n} samples of the same flame in FP64 (walkers of 8192), "
                      f"{dt:.1f} s on {cores} host threads",
        }
        rate1, n1, dt1, _ = cpu_rate(c2_text, oracle.F64_U64, args.cpu_seconds / 3, C2_SEED,
                                     8192, cores=1)
        cpu["one_thread"] = {"value": rate1, "cores": 1,
                             "sample": f"{n1} samples, {dt1:.1f} s on one thread"}
        log("cpu", json.dumps(cpu))
    line = {
        "metric": "samples/sec and fraction of atomic/HBM roofline at FP32 vs FP64, 1 GPU; p99 latency",
        "value": fp32["value"],
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": S,
        "warmup": W,
        "ms_per_step": fp32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded random flame (flames/gen_flames.py), walker streams Philox4x32-10",
        "config": {
            "workload": "random 2-D colour flame (seed 2, 6 xforms x 3 variations), "
                        "3840x2160 cells of u64 count + 3 f64 colour sums, 2^32 samples per step",
            "precision": "fp32 trajectories, exact u64 counts",
            "walkers": fp32["walkers"],
            "samples_per_walker": fp32["batch"],
        },
        "roofline": fp32["roofline"],
        "atomic_roofline": fp32["atomic_roofline"],
        "valu_roofline": fp32["valu_roofline"],
        "kernel_ms": fp32["kernel_ms"],
        "compiled": fp32["compiled"],
        "fp64": None if fp64 is None else {
            "value": fp64["value"],
            "ms_per_step": fp64["ms_per_step"],
            "kernel_ms": fp64["kernel_ms"],
            "compiled": fp64["compiled"],
            "roofline": fp64["roofline"],
            "atomic_roofline": fp64["atomic_roofline"],
            "valu_roofline": fp64["valu_roofline"],
        },
        "image": None if fp64 is None else fp64.get("image"),
        "l1_fp32_vs_fp64": l1_c2,
        "l1_fp64_vs_fp64_other_seed": floor_c2,
        "p99_latency_s": workloads.get("serving_fp32", {}).get("latency_p99_s"),
        "workloads": workloads,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
