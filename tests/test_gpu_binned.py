"""Binned accumulation on the GPU: count-only renders whose plotted samples
go to records sorted by tile and added from LDS, instead of one global atomic
add per sample (src/renderers/buffer_renderer.hpp:211-215).

Counts are integers, so the binned buffer must equal the atomic one bit for
bit, and the oracle's, with the same statistics.  Renders are split into
passes of iterations of every walker (walkers carry their point and stream
position from pass to pass); FFR_BIN_PASS_RECORDS shrinks a pass here so
that small renders take many passes.
"""

import json
import os

import numpy as np
import pytest

import flames as TF
from conftest import read_flame

pytestmark = pytest.mark.gpu

SEED = 0x5EED0009
MODES = {"fp64": 0, "fp32": 1, "fp32u64": 2}
THREADS = min(16, os.cpu_count() or 8)


def render(ffr, text, samples, batch, mode, binned, settle=20, **kw):
    precision = "fp64" if mode == "fp64" else "fp32"
    word = 32 if mode == "fp32" else 64
    r = ffr.BufferRenderer(text, precision=precision, word=word, seed=SEED, batch_size=batch,
                           settle=settle, binned=binned, **kw)
    ok = r.render(samples)
    return r, ok


def same(a, b, dims):
    assert np.array_equal(a.read_buffer(), b.read_buffer())
    assert a.samples_iterated == b.samples_iterated
    assert a.samples_plotted == b.samples_plotted
    assert np.array_equal(a.xform_distribution(), b.xform_distribution())
    ea, eb = a.point_extremes, b.point_extremes
    for i in range(dims):
        assert ea[i] == eb[i]


def vs_oracle(oracle, r, text, samples, batch, mode, dims, settle=20, **kw):
    f = oracle.OracleFlame(text, MODES[mode], settle=settle)
    ref = f.render(samples, batch, seed=SEED, threads=THREADS, **kw)
    assert np.array_equal(r.read_buffer(), ref.buffer)
    assert r.samples_iterated == ref.s_iter and r.samples_plotted == ref.s_plot
    assert np.array_equal(r.xform_distribution(), ref.xf_dist)
    ext = r.point_extremes
    for i in range(dims):
        assert ext[i] == (ref.pt_min[i], ref.pt_max[i])
    return ref


@pytest.fixture
def small_passes():
    """passes of 2^15 records: renders below take tens of passes"""
    old = os.environ.get("FFR_BIN_PASS_RECORDS")
    os.environ["FFR_BIN_PASS_RECORDS"] = str(1 << 15)
    yield
    if old is None:
        del os.environ["FFR_BIN_PASS_RECORDS"]
    else:
        os.environ["FFR_BIN_PASS_RECORDS"] = old


def sierpinski(size):
    return read_flame("sierpinski_1024.json").replace("[1024, 1024]", f"[{size}, {size}]")


CASES = {
    # name: (maker, dims, samples, batch)
    "tetrahedron_64": (lambda: TF.tetrahedron(size=64), 3, 600_000, 1500),
    "sierpinski_600": (lambda: sierpinski(600), 2, 700_001, 2000),
    "fern": (lambda: TF.fern(size=300), 2, 500_000, 999),
    "rotations_final": (TF.rotations, 2, 400_000, 1000),
    "one_d": (lambda: TF.one_d(size=5000), 1, 300_000, 700),
}


@pytest.mark.parametrize("mode", ["fp64", "fp32", "fp32u64"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_binned_equals_atomic_and_oracle(ffr_mod, oracle_mod, small_passes, name, mode):
    maker, dims, samples, batch = CASES[name]
    text = maker()
    print(f"binned {name} {mode}", flush=True)
    a, oka = render(ffr_mod, text, samples, batch, mode, 1)
    b, okb = render(ffr_mod, text, samples, batch, mode, 0)
    assert oka and okb
    assert a.last_stats.binned == 1 and b.last_stats.binned == 0
    same(a, b, dims)
    vs_oracle(oracle_mod, a, text, samples, batch, mode, dims)
    # a second render onto the same buffer continues the walkers' streams
    assert a.render(samples // 3) and b.render(samples // 3)
    same(a, b, dims)
    a.close()
    b.close()


def test_binned_bad_values(ffr_mod, oracle_mod, small_passes):
    """walkers that meet bad values start again mid-pass; the records of
    every bad value and the counts equal the oracle's"""
    text = TF.occasional_bad(size=256)
    samples, batch = 400_000, 1600
    r, ok = render(ffr_mod, text, samples, batch, "fp64", 1, settle=-1, bad_value_limit=10 ** 9)
    assert ok and r.last_stats.binned == 1
    ref = vs_oracle(oracle_mod, r, text, samples, batch, "fp64", 2, settle=-1,
                    bv_limit=10 ** 9, bv_cap=4096)
    assert ref.n_bad > 200 and r.last_stats.n_bad == ref.n_bad
    r.close()


def test_binned_bad_value_limit_stops(ffr_mod, small_passes):
    r, ok = render(ffr_mod, TF.expanding(size=64), 200_000, 1000, "fp64", 1, bad_value_limit=5)
    assert not ok and r.last_stats.binned == 1 and r.last_stats.n_bad >= 6
    # the buffer holds only samples of walkers that ran: no stale records
    s, _, _ = r.histogram_stats()
    assert s == r.samples_plotted
    r.close()


def test_binned_is_chosen_for_large_count_buffers(ffr_mod):
    """chosen by default for count-only buffers of 2^22 cells and more when
    the lane-per-walker kernel runs; not for small buffers, nor for a colour
    flame's first render; a colour flame takes it when asked for"""
    big = TF.tetrahedron(size=256)              # 2^24 cells
    # 32768 walkers: beyond what the wave-per-walker kernel takes
    r, ok = render(ffr_mod, big, 1 << 24, 512, "fp32", -1)
    assert ok and r.last_stats.binned == 1 and r.last_stats.wave_per_walker == 0
    r.close()
    small = TF.tetrahedron(size=64)             # 2^18 cells
    r, ok = render(ffr_mod, small, 1 << 24, 512, "fp32", -1)
    assert ok and r.last_stats.binned == 0
    r.close()
    colour = TF.color_flame(3, size=2048)
    r, ok = render(ffr_mod, colour, 1 << 22, 2048, "fp64", -1)
    assert ok and r.last_stats.binned == 0
    r.close()
    r, ok = render(ffr_mod, colour, 1 << 22, 2048, "fp64", 1)
    assert ok and r.last_stats.binned == 1
    r.close()


def test_tetrahedron_512_full_size(ffr_mod):
    """the benchmark's 3-D workload at its size (512^3 u32 counts, 2^33
    samples, FP32): the counts sum to the samples plotted (all but a few
    thousand land inside the unit cube), and the binned buffer equals the
    atomic one on the first 2^30 samples"""
    from conftest import FLAMES
    with open(os.path.join(FLAMES, "tetrahedron_512.json")) as f:
        text = f.read()
    r, ok = render(ffr_mod, text, 1 << 33, 0, "fp32", -1, settle=-1)
    assert ok and r.last_stats.binned == 1
    s, _, _ = r.histogram_stats()
    print(f"tetrahedron 512^3: {r.last_stats.kernel_ms:.1f} ms for 2^33 samples", flush=True)
    assert s == r.samples_plotted and r.samples_iterated == 1 << 33
    assert r.samples_plotted > (1 << 33) - 100_000
    r.close()
    a, oka = render(ffr_mod, text, 1 << 30, 0, "fp32", 1, settle=-1)
    b, okb = render(ffr_mod, text, 1 << 30, 0, "fp32", 0, settle=-1)
    assert oka and okb and a.last_stats.binned == 1 and b.last_stats.binned == 0
    print(f"2^30 samples: binned {a.last_stats.kernel_ms:.1f} ms, atomic "
          f"{b.last_stats.kernel_ms:.1f} ms", flush=True)
    same(a, b, 3)
    a.close()
    b.close()


def test_binned_over_4096_tiles_and_size_limit(ffr_mod):
    """count-only buffers of more than 4096 tiles of 2^15 cells (a 12000^2
    fern, 4395 tiles: the scatter's chunks of 16 records a lane) equal the
    atomic path; buffers of 2^28 cells and more keep the atomics even when
    binned accumulation is asked for"""
    text = TF.fern(size=12000)
    a, oka = render(ffr_mod, text, 1 << 22, 2048, "fp32", 1)
    b, okb = render(ffr_mod, text, 1 << 22, 2048, "fp32", 0)
    assert oka and okb and a.last_stats.binned == 1 and b.last_stats.binned == 0
    same(a, b, 2)
    a.close()
    b.close()
    r, ok = render(ffr_mod, TF.fern(size=16385), 1 << 20, 1024, "fp32", 1)  # > 2^28 cells
    assert ok and r.last_stats.binned == 0
    r.close()
