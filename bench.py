"""Benchmark: buffer-render throughput on one MI355X.

Workload (one "step"): one render pass of 2^32 samples of the random 2-D
colour flame (seed 2, 6 xforms x 3 variations from linear / sinusoidal /
spherical / swirl / horseshoe / polar) into its 3840x2160 buffer of
uint64 counts + 3 double colour sums per cell (the reference's default
buffer layout), walkers chosen by the renderer.  Measured with FP32
trajectories (the reported `value`) and with FP64 trajectories.

Prints one JSON line on stdout.  `roofline` is for the walker kernel:
algorithmic bytes = plotted samples x 32 B (the 8-byte count and three
8-byte colour words each sample adds into its cell) over the kernel's
average duration from HIP events on its stream.  `cpu_baseline` times the
CPU oracle (the reference loop restated in C, threads = host cores) on a
bounded sample of the same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
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

FLAME = os.path.join(PKG, "flames", "random_color_3840x2160.json")
SAMPLES = 1 << 32
SEED = 0x5EED0002
BYTES_PER_PLOT = 32   # u64 count + 3 x f64 colour, added into one cell
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_gpu(ffr, text, precision, steps, warmup):
    r = ffr.BufferRenderer(text, precision=precision, seed=SEED)
    for _ in range(warmup):
        if not r.render(SAMPLES):
            raise RuntimeError("render stopped at the bad-value limit")
    plot0 = r.samples_plotted
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        if not r.render(SAMPLES):
            raise RuntimeError("render stopped at the bad-value limit")
        kms.append(r.kernel_ms)
    t1 = time.perf_counter()
    plotted = (r.samples_plotted - plot0) / steps
    wall = t1 - t0
    kavg = sum(kms) / len(kms)
    achieved = plotted * BYTES_PER_PLOT / (kavg * 1e-3) / 1e9
    out = {
        "value": steps * SAMPLES / wall,
        "ms_per_step": 1e3 * wall / steps,
        "kernel_ms": kavg,
        "walkers": int(r.last_stats.walkers),
        "batch": int(r.last_stats.batch_size),
        "plotted_per_step": plotted,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
        },
    }
    r.close()
    return out


def calibration_pass(ffr, text):
    """One buffer add of known size: the kernel reads 2 x and writes 1 x the
    buffer (8294400 cells x 32 B); its FETCH_SIZE / WRITE_SIZE counters scale
    the walker kernel's into bytes."""
    import numpy as np

    r = ffr.BufferRenderer(text, precision="fp32", seed=SEED)
    zeros = np.zeros(r.cells * r.cell_words, dtype=np.uint64)
    r.add_buffer(zeros)
    r.histogram_stats()  # hist_stats_kernel reads every line of the buffer once
    log(f"calibration: clear writes {r.buffer_bytes} bytes; histogram pass reads "
        f"{r.buffer_bytes} bytes; add reads {2 * r.buffer_bytes}, writes {r.buffer_bytes}")
    r.close()


def cpu_baseline(text, seconds):
    from oracle import oracle

    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    f = oracle.OracleFlame(text, oracle.F64_U64)
    n = 1 << 20
    t0 = time.perf_counter()
    f.render(n, 8192, seed=SEED, threads=cores)
    rate = n / (time.perf_counter() - t0)
    n = int(min(max(rate * seconds, 1 << 20), 1 << 31)) // 8192 * 8192
    t0 = time.perf_counter()
    f.render(n, 8192, seed=SEED, threads=cores)
    dt = time.perf_counter() - t0
    return {
        "value": n / dt,
        "unit": "samples/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{n} samples of the same flame in FP64 (walkers of 8192), "
                  f"{dt:.1f} s on {cores} host threads",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--calibrate", action="store_true",
                    help="also run one buffer add of known size (for counter scaling)")
    ap.add_argument("--fp32-only", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(walker ranges would shard across GPUs with one integer sum of the "
              "buffers at the end; see DESIGN.md)", file=sys.stderr)
        return 2
    import ffr

    with open(FLAME) as fh:
        text = fh.read()
    log(f"bench: {SAMPLES} samples/step, warmup {args.warmup}, steps {args.steps}")
    if args.calibrate:
        calibration_pass(ffr, text)
    fp32 = run_gpu(ffr, text, "fp32", args.steps, args.warmup)
    log("fp32", json.dumps(fp32))
    if args.fp32_only:
        return 0
    fp64 = run_gpu(ffr, text, "fp64", args.steps, args.warmup)
    log("fp64", json.dumps(fp64))
    cpu = None if args.no_cpu else cpu_baseline(text, args.cpu_seconds)
    if cpu:
        log("cpu", json.dumps(cpu))
    line = {
        "metric": "samples/sec and fraction of atomic/HBM roofline at FP32 vs FP64, 1 GPU; p99 latency",
        "value": fp32["value"],
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
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
        "kernel_ms": fp32["kernel_ms"],
        "fp64": {
            "value": fp64["value"],
            "ms_per_step": fp64["ms_per_step"],
            "kernel_ms": fp64["kernel_ms"],
            "roofline": fp64["roofline"],
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
