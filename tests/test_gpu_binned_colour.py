"""Binned accumulation of colour buffers on the GPU: every plotted sample
writes a record of its cell and its colour (after the final xform's blend)
instead of one atomic add per word (src/renderers/buffer_renderer.hpp:
211-229); the records are sorted by tile of 4096 cells and each tile's
counts and colour sums are added from LDS.

Counts are integers: the binned buffer's counts must equal the atomic
path's and the oracle's bit for bit, with the same statistics.  Colour sums
are added in another order (per sample into the buffer's words on the
atomic path; in f64 per tile here, once into the words): 8-byte words agree
to 1e-12 relative, the float build's f32 words to 1e-4 relative of the
largest sum (its atomic path rounds to f32 on every sample's add).
FFR_BIN_PASS_RECORDS shrinks a pass so that the renders take many passes,
with walkers' points, colours and stream positions carried between them.
"""

import os

import numpy as np
import pytest

import flames as TF

pytestmark = pytest.mark.gpu

SEED = 0x5EED000E
MODES = {"fp64": 0, "fp32": 1, "fp32u64": 2}
THREADS = min(16, os.cpu_count() or 8)


@pytest.fixture
def small_passes():
    old = os.environ.get("FFR_BIN_PASS_RECORDS")
    os.environ["FFR_BIN_PASS_RECORDS"] = str(1 << 15)
    yield
    if old is None:
        del os.environ["FFR_BIN_PASS_RECORDS"]
    else:
        os.environ["FFR_BIN_PASS_RECORDS"] = old


def render(ffr, text, samples, batch, mode, binned, compiled=0, settle=20, **kw):
    precision = "fp64" if mode == "fp64" else "fp32"
    word = 32 if mode == "fp32" else 64
    r = ffr.BufferRenderer(text, precision=precision, word=word, seed=SEED, batch_size=batch,
                           settle=settle, binned=binned, compiled=compiled, **kw)
    return r, r.render(samples)


def colours_close(a, b, mode):
    ca, cb = a.colors(a.read_buffer()), b.colors(b.read_buffer())
    tol = 1e-12 if mode != "fp32" else 1e-4
    scale = max(1.0, float(np.abs(cb).max()))
    assert np.allclose(ca, cb, rtol=tol, atol=tol * scale), float(np.abs(ca - cb).max())


def same(a, b, dims, mode):
    assert np.array_equal(a.counts(a.read_buffer()), b.counts(b.read_buffer()))
    colours_close(a, b, mode)
    assert a.samples_iterated == b.samples_iterated
    assert a.samples_plotted == b.samples_plotted
    assert np.array_equal(a.xform_distribution(), b.xform_distribution())
    for i in range(dims):
        assert a.point_extremes[i] == b.point_extremes[i]
    assert a.last_stats.n_bad == b.last_stats.n_bad


CASES = {
    # name: (maker, dims, samples, batch)
    "colour3_final": (lambda: TF.color_flame(3, size=200), 2, 600_000, 1500),
    "colour1": (lambda: TF.color_flame(1, final_color=False, size=150), 2, 400_000, 999),
    "colour4": (lambda: TF.color_flame(4, size=100), 2, 300_000, 700),
}


@pytest.mark.parametrize("mode", ["fp64", "fp32", "fp32u64"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_binned_colour_equals_atomic_and_oracle(ffr_mod, oracle_mod, small_passes, name, mode):
    maker, dims, samples, batch = CASES[name]
    text = maker()
    print(f"binned colour {name} {mode}", flush=True)
    a, oka = render(ffr_mod, text, samples, batch, mode, 1)
    b, okb = render(ffr_mod, text, samples, batch, mode, 0)
    assert oka and okb
    assert a.last_stats.binned == 1 and b.last_stats.binned == 0
    same(a, b, dims, mode)
    ref = oracle_mod.OracleFlame(text, MODES[mode], settle=20).render(
        samples, batch, seed=SEED, threads=THREADS)
    assert np.array_equal(a.counts(a.read_buffer()), ref.counts)
    assert a.samples_iterated == ref.s_iter and a.samples_plotted == ref.s_plot
    # a second render continues the walkers' streams
    assert a.render(samples // 3) and b.render(samples // 3)
    same(a, b, dims, mode)
    a.close()
    b.close()


def test_binned_colour_bad_values(ffr_mod, small_passes):
    """restarts draw new colour words; counts, colours and records equal the
    atomic path's"""
    import json
    d = json.loads(TF.occasional_bad(size=256))
    d["color_dimensions"] = 3
    for k, x in enumerate(d["xforms"]):
        x["color"] = [0.2 * k, 0.5, 1.0 - 0.1 * k]
    text = json.dumps(d)
    a, oka = render(ffr_mod, text, 400_000, 1600, "fp64", 1, settle=-1,
                    bad_value_limit=10 ** 9)
    b, okb = render(ffr_mod, text, 400_000, 1600, "fp64", 0, settle=-1,
                    bad_value_limit=10 ** 9)
    assert oka and okb and a.last_stats.binned == 1
    assert a.last_stats.n_bad > 200
    same(a, b, 2, "fp64")
    a.close()
    b.close()


@pytest.mark.parametrize("mode", ["fp64", "fp32u64"])
def test_binned_colour_compiled_walker(ffr_mod, small_passes, mode):
    """the flame's compiled binned walker writes the same records"""
    maker, dims, samples, batch = CASES["colour3_final"]
    text = maker()
    a, oka = render(ffr_mod, text, samples, batch, mode, 1, compiled=1)
    b, okb = render(ffr_mod, text, samples, batch, mode, 1, compiled=0)
    assert oka and okb and a.last_stats.compiled == 1 and b.last_stats.compiled == 0
    assert np.array_equal(a.counts(a.read_buffer()), b.counts(b.read_buffer()))
    colours_close(a, b, mode)
    a.close()
    b.close()


def _with_colours(text, r, seed):
    import json
    d = json.loads(text)
    rng = np.random.default_rng(seed)
    d["color_dimensions"] = r
    for k, x in enumerate(d["xforms"]):
        if k != 1:          # xform 1 leaves the colour alone
            x["color"] = rng.uniform(0, 1, r).tolist()
    return json.dumps(d)


@pytest.mark.parametrize("mode", ["fp64", "fp32u64"])
@pytest.mark.parametrize("name", ["tetrahedron_2colours", "one_d_1colour"])
def test_binned_colour_1d_3d(ffr_mod, oracle_mod, small_passes, name, mode):
    """colour records of 1-D and 3-D flames: the buffer equals the atomic
    path's, and the counts the oracle's"""
    if name == "tetrahedron_2colours":
        text, dims = _with_colours(TF.tetrahedron(size=48), 2, 11), 3
    else:
        text, dims = _with_colours(TF.one_d(size=300, nonlinear=True), 1, 12), 1
    print(f"binned colour {name} {mode}", flush=True)
    a, oka = render(ffr_mod, text, 500_000, 1000, mode, 1)
    b, okb = render(ffr_mod, text, 500_000, 1000, mode, 0)
    assert oka and okb and a.last_stats.binned == 1 and b.last_stats.binned == 0
    same(a, b, dims, mode)
    ref = oracle_mod.OracleFlame(text, MODES[mode], settle=20).render(
        500_000, 1000, seed=SEED, threads=THREADS)
    assert np.array_equal(a.counts(a.read_buffer()), ref.counts)
    a.close()
    b.close()


@pytest.mark.parametrize("mode", ["fp32u64", "fp32"])
def test_binned_colour_many_tiles(ffr_mod, small_passes, mode):
    """a colour buffer of more than 4096 tiles of 4096 cells (4200^2 cells,
    4307 tiles): the scatter sorts its f32 colours in chunks of 8 records a
    lane (16 up to 4096 tiles); the buffer equals the atomic path's"""
    text = TF.color_flame(3, size=4200)
    print(f"binned colour 4200^2 {mode}", flush=True)
    a, oka = render(ffr_mod, text, 1 << 21, 1024, mode, 1)
    b, okb = render(ffr_mod, text, 1 << 21, 1024, mode, 0)
    assert oka and okb and a.last_stats.binned == 1 and b.last_stats.binned == 0
    same(a, b, 2, mode)
    a.close()
    b.close()


def test_binned_colour_spread_flame_full_size(ffr_mod):
    """a batch flame whose samples spread beyond the LDS cache, at its size
    (2048^2, 3 colours, FP32, 8-byte words): the binned buffer equals the
    atomic one (counts exactly, colours to 1e-12), and the time of both is
    printed"""
    from conftest import FLAMES
    with open(os.path.join(FLAMES, "batch", "flame_1069.json")) as f:
        text = f.read()
    samples = 1 << 30
    a, oka = render(ffr_mod, text, samples, 0, "fp32u64", 1, settle=53)
    b, okb = render(ffr_mod, text, samples, 0, "fp32u64", 0, settle=53)
    assert oka and okb and a.last_stats.binned == 1 and b.last_stats.binned == 0
    print(f"flame_1069 2^30: binned {a.last_stats.kernel_ms:.1f} ms, atomic "
          f"{b.last_stats.kernel_ms:.1f} ms", flush=True)
    same(a, b, 2, "fp32u64")
    a.close()
    b.close()


def test_colour_flame_switches_to_binned_after_many_misses(ffr_mod):
    """binned -1: a colour flame renders on the atomic path first; after a
    render of 2^26 samples or more that made at least 0.6 global atomic
    adds per sample (LDS-cache misses, four words each here), its later
    renders are binned.  flame_1069 (2048^2) misses on about 0.6 of its
    samples (2.45 atomics per sample) and switches; flame_1013, whose
    samples stay in the cache, does not.  Either way the buffer equals one
    rendered on the atomic path throughout."""
    from conftest import FLAMES
    for name, switch in (("flame_1069", True), ("flame_1013", False)):
        with open(os.path.join(FLAMES, "batch", name + ".json")) as f:
            text = f.read()
        print(f"auto binned {name}", flush=True)
        a, oka = render(ffr_mod, text, 1 << 26, 0, "fp32u64", -1, settle=53)
        b, okb = render(ffr_mod, text, 1 << 26, 0, "fp32u64", 0, settle=53)
        assert oka and okb and a.last_stats.binned == 0
        per = (a.last_stats.atomics - a.last_stats.flush_atomics) / a.last_stats.s_iter
        assert (per >= 0.6) == switch, per
        assert a.render(1 << 26) and b.render(1 << 26)
        assert a.last_stats.binned == (1 if switch else 0)
        same(a, b, 2, "fp32u64")
        a.close()
        b.close()


@pytest.mark.parametrize("samples,batch", [(1, 0), (37, 5), (1000, 7)])
def test_binned_colour_ragged_renders(ffr_mod, oracle_mod, samples, batch):
    """renders of one sample, of fewer walkers than a task's 256 lanes and
    of a last walker shorter than the others: counts and statistics equal
    the oracle's, colours the atomic path's; a render of 0 samples leaves
    the buffer alone"""
    text = TF.color_flame(3, size=64)
    a, oka = render(ffr_mod, text, samples, batch, "fp64", 1)
    b, okb = render(ffr_mod, text, samples, batch, "fp64", 0)
    assert oka and okb and a.last_stats.binned == 1
    same(a, b, 2, "fp64")
    ref = oracle_mod.OracleFlame(text, MODES["fp64"], settle=20).render(
        samples, batch or a.last_stats.batch_size, seed=SEED, threads=1)
    assert np.array_equal(a.counts(a.read_buffer()), ref.counts)
    assert a.samples_iterated == ref.s_iter and a.samples_plotted == ref.s_plot
    before = a.read_buffer().copy()
    assert a.render(0)
    assert np.array_equal(a.read_buffer(), before)
    a.close()
    b.close()


def test_binned_colour_size_limit(ffr_mod):
    """colour records address tiles of 4096 cells, at most 8192 of them: a
    colour buffer of 2^25 cells or more keeps the atomics even when binned
    accumulation is asked for; one just under (8191 tiles, the scatter's
    largest LDS) takes it and equals the atomic path"""
    over = TF.color_flame(3, size=5793)     # 33,558,849 cells > 2^25
    r, ok = render(ffr_mod, over, 100_000, 1000, "fp32u64", 1)
    assert ok and r.last_stats.binned == 0
    r.close()
    under = TF.color_flame(3, size=5792)    # 33,547,264 cells < 2^25, 8191 tiles
    a, oka = render(ffr_mod, under, 100_000, 1000, "fp32u64", 1)
    b, okb = render(ffr_mod, under, 100_000, 1000, "fp32u64", 0)
    assert oka and okb and a.last_stats.binned == 1 and b.last_stats.binned == 0
    same(a, b, 2, "fp32u64")
    a.close()
    b.close()
