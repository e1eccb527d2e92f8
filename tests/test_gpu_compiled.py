"""Per-flame compiled kernels on the GPU: the walker code built at run time
with hiprtc for one flame, its xform count, table layout, variation types
and bounds as constants (the reference's per-flame objects with virtual
calc calls, src/variations/variations.hpp:38-59, src/types/xform.hpp:
211-227, become straight-line code for the flame's own variations).

The compiled kernel runs the same arithmetic in the same order as the
table-driven kernel, so counts and statistics must be equal bit for bit
(colour sums to the rounding of their atomic order), for linear,
transcendental and word-drawing flames, with and without the LDS cell
cache, in all three precision modes; linear flames also against the oracle.
"""

import json
import os

import numpy as np
import pytest

import flames as TF
from conftest import read_flame

pytestmark = pytest.mark.gpu

SEED = 0x5EED000C
MODES = {"fp64": 0, "fp32": 1, "fp32u64": 2}


def render(ffr, text, samples, batch, mode, compiled, cache=-1, settle=-1):
    precision = "fp64" if mode == "fp64" else "fp32"
    word = 32 if mode == "fp32" else 64
    r = ffr.BufferRenderer(text, precision=precision, word=word, seed=SEED, batch_size=batch,
                           settle=settle, cell_cache=cache, compiled=compiled)
    ok = r.render(samples)
    return r, ok


def same(a, b, dims):
    wa, wb = a.read_buffer(), b.read_buffer()
    assert np.array_equal(a.counts(wa), b.counts(wb))
    if a.color_dims:
        ca, cb = a.colors(wa), b.colors(wb)
        tol = 1e-12 if a.word_bytes == 8 else 2e-5
        assert np.allclose(ca, cb, rtol=tol, atol=tol * max(1.0, float(np.abs(cb).max())))
    assert a.samples_iterated == b.samples_iterated
    assert a.samples_plotted == b.samples_plotted
    assert np.array_equal(a.xform_distribution(), b.xform_distribution())
    ea, eb = a.point_extremes, b.point_extremes
    for i in range(dims):
        assert ea[i] == eb[i]
    assert a.last_stats.n_bad == b.last_stats.n_bad


def headline(size=(384, 216)):
    text = read_flame("random_color_3840x2160.json").replace(
        "3840,\n  2160", f"{size[0]},\n  {size[1]}")
    assert f'"size": [\n  {size[0]}' in text
    return text


def catalogue(seed):
    """three xforms of two variations drawn from the whole catalogue
    (drawing variations included), with colour and a final xform"""
    import oracle.model as M
    rng = np.random.default_rng(seed)
    names = [n for n, _ in M.VARIATIONS]
    xfs = []
    for k in range(3):
        vs = [TF.variation(str(n), float(rng.uniform(0.2, 0.8)), 2)
              for n in rng.choice(names, 2, replace=False)]
        A = (rng.uniform(-1, 1, (2, 2)) * 0.5).tolist()
        xfs.append({"weight": float(rng.uniform(0.5, 2)), "variations": vs,
                    "pre_affine": {"A": A, "b": rng.uniform(-0.5, 0.5, 2).tolist()},
                    "color": rng.uniform(0, 1, 2).tolist()})
    return json.dumps({"dimensions": 2, "size": [96, 96], "bounds": [[-3, 3], [-3, 3]],
                       "color_dimensions": 2, "xforms": xfs,
                       "final_xform": {"variations": [{"name": "linear", "weight": 0.9}]}})


FLAMES = {
    # name: (maker, dims, samples, batch)
    "sierpinski": (lambda: read_flame("sierpinski_1024.json").replace(
        "[1024, 1024]", "[256, 256]"), 2, 1_000_000, 1000),
    "fern": (TF.fern, 2, 600_000, 600),
    "rotations_final": (TF.rotations, 2, 500_000, 500),
    "tetrahedron": (TF.tetrahedron, 3, 500_000, 500),
    "one_d": (TF.one_d, 1, 300_000, 300),
    "colour_linear": (lambda: TF.color_flame(3), 2, 600_000, 600),
    "headline": (headline, 2, 1_000_000, 1000),
    "catalogue_a": (lambda: catalogue(3), 2, 300_000, 300),
    "catalogue_b": (lambda: catalogue(4), 2, 300_000, 300),
}


# each case compiles its own kernel (1-4 s with hiprtc): every flame once,
# the three precision modes and both accumulation paths spread over them
CASES = [
    ("sierpinski", "fp64", 0), ("sierpinski", "fp32", 1),
    ("fern", "fp32u64", 0), ("rotations_final", "fp64", 1),
    ("tetrahedron", "fp32", 0), ("tetrahedron", "fp64", 1), ("one_d", "fp32u64", 1),
    ("colour_linear", "fp64", 1), ("colour_linear", "fp32", 0),
    ("headline", "fp32u64", 1), ("headline", "fp64", 1), ("headline", "fp32", 0),
    ("catalogue_a", "fp64", 0), ("catalogue_b", "fp32u64", 1),
]


@pytest.mark.parametrize("name,mode,cache", CASES)
def test_compiled_equals_table_kernel(ffr_mod, name, mode, cache):
    maker, dims, samples, batch = FLAMES[name]
    text = maker()
    print(f"compiled {name} {mode} cache {cache}", flush=True)
    a, oka = render(ffr_mod, text, samples, batch, mode, 1, cache)
    b, okb = render(ffr_mod, text, samples, batch, mode, 0, cache)
    assert oka == okb
    assert a.last_stats.compiled == 1 and b.last_stats.compiled == 0
    same(a, b, dims)
    # the compiled kernel is kept: a second render reuses it
    assert a.render(samples // 2) == b.render(samples // 2)
    assert a.last_stats.compiled == 1
    same(a, b, dims)
    a.close()
    b.close()


@pytest.mark.parametrize("name,mode", [("sierpinski", "fp64"), ("rotations_final", "fp32"),
                                       ("tetrahedron", "fp32u64"), ("one_d", "fp64")])
def test_compiled_linear_flames_match_oracle(ffr_mod, oracle_mod, name, mode):
    maker, dims, samples, batch = FLAMES[name]
    text = maker()
    r, ok = render(ffr_mod, text, samples, batch, mode, 1, 0, settle=20)
    assert ok and r.last_stats.compiled == 1
    ref = oracle_mod.OracleFlame(text, MODES[mode], settle=20).render(
        samples, batch, seed=SEED, threads=8)
    assert np.array_equal(r.read_buffer(), ref.buffer)
    assert r.samples_iterated == ref.s_iter and r.samples_plotted == ref.s_plot
    assert np.array_equal(r.xform_distribution(), ref.xf_dist)
    r.close()


def test_compiled_not_taken_by_the_wave_per_walker_kernel(ffr_mod):
    """the wave-per-walker path keeps its own kernel"""
    sier = read_flame("sierpinski_1024.json").replace("[1024, 1024]", "[64, 64]")
    r, ok = render(ffr_mod, sier, 64_000, 1000, "fp64", 1)
    assert ok and r.last_stats.wave_per_walker == 1 and r.last_stats.compiled == 0
    r.close()


@pytest.fixture
def small_passes():
    """passes of 2^15 records: the renders below take tens of passes"""
    old = os.environ.get("FFR_BIN_PASS_RECORDS")
    os.environ["FFR_BIN_PASS_RECORDS"] = str(1 << 15)
    yield
    if old is None:
        del os.environ["FFR_BIN_PASS_RECORDS"]
    else:
        os.environ["FFR_BIN_PASS_RECORDS"] = old


def without_colour(text):
    """the flame with its colours removed (a count-only buffer)"""
    d = json.loads(text)
    d.pop("color_dimensions", None)
    for x in d["xforms"]:
        x.pop("color", None)
    return json.dumps(d)


def binned_render(ffr, text, samples, batch, mode, compiled):
    precision = "fp64" if mode == "fp64" else "fp32"
    word = 32 if mode == "fp32" else 64
    r = ffr.BufferRenderer(text, precision=precision, word=word, seed=SEED, batch_size=batch,
                           settle=20, binned=1, compiled=compiled)
    return r, r.render(samples)


BINNED = {
    # name: (maker, dims, samples, batch)
    "tetrahedron": (lambda: TF.tetrahedron(size=64), 3, 600_000, 1500),
    "fern": (lambda: TF.fern(size=300), 2, 500_000, 999),
    "catalogue_c": (lambda: without_colour(catalogue(5)), 2, 300_000, 300),
}


@pytest.mark.parametrize("name,mode", [("tetrahedron", "fp32"), ("tetrahedron", "fp64"),
                                       ("fern", "fp32u64"), ("catalogue_c", "fp64")])
def test_compiled_binned_walker(ffr_mod, oracle_mod, small_passes, name, mode):
    """binned accumulation with the flame's compiled walker: the same buffer
    and statistics as the table-driven walker, pass after pass; linear
    flames equal the oracle"""
    maker, dims, samples, batch = BINNED[name]
    text = maker()
    a, oka = binned_render(ffr_mod, text, samples, batch, mode, 1)
    b, okb = binned_render(ffr_mod, text, samples, batch, mode, 0)
    assert oka and okb
    assert a.last_stats.binned == 1 and a.last_stats.compiled == 1
    assert b.last_stats.binned == 1 and b.last_stats.compiled == 0
    assert np.array_equal(a.read_buffer(), b.read_buffer())
    assert a.samples_iterated == b.samples_iterated
    assert a.samples_plotted == b.samples_plotted
    assert np.array_equal(a.xform_distribution(), b.xform_distribution())
    for i in range(dims):
        assert a.point_extremes[i] == b.point_extremes[i]
    if name != "catalogue_c":
        ref = oracle_mod.OracleFlame(text, MODES[mode], settle=20).render(
            samples, batch, seed=SEED, threads=8)
        assert np.array_equal(a.read_buffer(), ref.buffer)
    a.close()
    b.close()


def test_compiled_binned_auto_probe(ffr_mod, small_passes, low_probe_threshold):
    """compiled -1 on the binned path: a render in steady use times its first
    pass with the table-driven walker and its second with the compiled one,
    keeps the faster, and the buffer equals a renderer's that never compiles"""
    maker, dims, samples, batch = BINNED["tetrahedron"]
    text = maker()
    a, oka = binned_render(ffr_mod, text, samples, batch, "fp32", -1)
    b, okb = binned_render(ffr_mod, text, samples, batch, "fp32", 0)
    assert oka and okb and a.last_stats.compiled == 0
    kept = []
    for _ in range(3):
        assert a.render(samples) and b.render(samples)
        kept.append(a.last_stats.compiled)
        assert np.array_equal(a.read_buffer(), b.read_buffer())
        assert a.samples_plotted == b.samples_plotted
    assert kept[0] == kept[1] == kept[2]
    a.close()
    b.close()


@pytest.fixture
def low_probe_threshold():
    old = os.environ.get("FFR_COMPILED_PROBE_SAMPLES")
    os.environ["FFR_COMPILED_PROBE_SAMPLES"] = "1000000"
    yield
    if old is None:
        del os.environ["FFR_COMPILED_PROBE_SAMPLES"]
    else:
        os.environ["FFR_COMPILED_PROBE_SAMPLES"] = old


@pytest.mark.parametrize("mode", ["fp64", "fp32"])
def test_compiled_auto_probe(ffr_mod, low_probe_threshold, mode):
    """compiled -1: the first render takes the table-driven kernel; a later
    render, once the renders have asked for enough samples, probes both
    kernels on a scratch buffer and keeps the faster.  The probe leaves the
    buffer, statistics and stream positions as they were: the counts equal a
    render with the compiled kernel off."""
    maker, dims, samples, batch = FLAMES["headline"]
    text = maker()
    a, oka = render(ffr_mod, text, samples, batch, mode, -1, 1)
    b, okb = render(ffr_mod, text, samples, batch, mode, 0, 1)
    assert oka and okb and a.last_stats.compiled == 0
    same(a, b, dims)
    choices = []
    for _ in range(3):
        assert a.render(samples) == b.render(samples)
        choices.append(a.last_stats.compiled)
        same(a, b, dims)
    # the choice is made once and kept
    assert choices[0] == choices[1] == choices[2]
    print(f"auto probe {mode}: compiled kernel {'kept' if choices[0] else 'not kept'}",
          flush=True)
    a.close()
    b.close()
