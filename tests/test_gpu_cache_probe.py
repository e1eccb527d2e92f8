"""The single render's cell-cache probe on the GPU: with cell_cache -1 a
renderer in steady use times its cached launch against the same launch
without the LDS cell cache, on a scratch buffer, and keeps the faster per
walker count (the batch tuner's cache/no-cache choice, for one flame).

Counts are integers and the walkers' streams do not depend on the launch,
so whichever launch the probe keeps, the buffer and statistics must equal a
renderer's whose cache is forced on, render after render; the probe itself
must leave no trace in the buffer or the statistics.
"""

import os

import numpy as np
import pytest

import flames as TF
from conftest import read_flame

pytestmark = pytest.mark.gpu

SEED = 0x5EED000D


@pytest.fixture
def low_probe_threshold():
    old = os.environ.get("FFR_CACHE_PROBE_SAMPLES")
    os.environ["FFR_CACHE_PROBE_SAMPLES"] = "1000000"
    yield
    if old is None:
        del os.environ["FFR_CACHE_PROBE_SAMPLES"]
    else:
        os.environ["FFR_CACHE_PROBE_SAMPLES"] = old


def same(a, b, dims):
    assert np.array_equal(a.counts(a.read_buffer()), b.counts(b.read_buffer()))
    assert a.samples_iterated == b.samples_iterated
    assert a.samples_plotted == b.samples_plotted
    assert np.array_equal(a.xform_distribution(), b.xform_distribution())
    for i in range(dims):
        assert a.point_extremes[i] == b.point_extremes[i]


CASES = {
    # name: (flame, dims, precision, walkers, samples per walker)
    # 128 walkers per CU of a linear flame on a 1024^2 buffer (beyond the
    # wave-per-walker kernel's limit): the cache lengthens each walker's
    # chain
    "sierpinski_1024": (lambda: read_flame("sierpinski_1024.json"), 2, "fp64", 32768, 256),
    # a concentrated colour flame: the cache takes most cell updates
    "headline_384": (lambda: read_flame("random_color_3840x2160.json").replace(
        "3840,\n  2160", "384,\n  216"), 2, "fp32", 32768, 256),
    "tetrahedron_64": (lambda: TF.tetrahedron(size=64), 3, "fp32", 32768, 128),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_cache_probe_keeps_results(ffr_mod, low_probe_threshold, name):
    maker, dims, precision, walkers, batch = CASES[name]
    text = maker()
    samples = walkers * batch
    a = ffr_mod.BufferRenderer(text, precision=precision, seed=SEED, batch_size=batch,
                               cell_cache=-1, compiled=0)
    b = ffr_mod.BufferRenderer(text, precision=precision, seed=SEED, batch_size=batch,
                               cell_cache=1, compiled=0)
    assert a.render(samples) and b.render(samples)
    assert a.last_stats.cache_slots > 0 and a.last_stats.wave_per_walker == 0
    same(a, b, dims)
    kept = []
    for _ in range(3):
        assert a.render(samples) and b.render(samples)
        kept.append(a.last_stats.cache_slots > 0)
        same(a, b, dims)
    # decided once per walker count and kept
    assert kept[0] == kept[1] == kept[2]
    print(f"cache probe {name}: cache {'kept' if kept[0] else 'dropped'}", flush=True)
    a.close()
    b.close()


def test_cache_probe_not_when_forced(ffr_mod, low_probe_threshold):
    """a cache forced on or off is never probed"""
    text = read_flame("sierpinski_1024.json")
    for cache, want in ((1, True), (0, False)):
        r = ffr_mod.BufferRenderer(text, seed=SEED, batch_size=256, cell_cache=cache, compiled=0)
        for _ in range(3):
            assert r.render(32768 * 256)
            assert (r.last_stats.cache_slots > 0) == want
        r.close()
