"""Batched rendering on the GPU: several renderers rendered together in one
persistent launch per kernel kind (replayed as a HIP graph) give each
renderer exactly what its own render gives: the same walkers on the same
streams.  Counts and statistics are compared exactly with twin renderers
rendered one at a time, colours within 1e-12 relative (atomic order).  The
graph is replayed several times with eager renders and synchronisations in
between, and every replay is checked.
"""

import json

import numpy as np
import pytest

import flames as TF
from conftest import read_flame

pytestmark = pytest.mark.gpu

SEED = 0x5EED0004


def mixed_flames():
    """(flame text, renderer options, samples): several kinds in one batch."""
    sier = read_flame("sierpinski_1024.json").replace("[1024, 1024]", "[128, 128]")
    c2 = read_flame("random_color_3840x2160.json").replace("3840,\n  2160", "192,\n  108")
    assert '"size": [\n  192' in c2
    return [
        (sier, dict(precision="fp64", settle=20), 300_000),
        (sier, dict(precision="fp32", word=32, settle=20), 200_000),
        (TF.color_flame(3), dict(precision="fp64"), 250_000),
        (TF.color_flame(3, size=96, seed=9), dict(precision="fp32"), 150_000),
        (c2, dict(precision="fp32"), 400_000),
        (c2, dict(precision="fp64"), 120_000),
        (TF.tetrahedron(), dict(precision="fp32", word=32), 90_000),
        (TF.one_d(), dict(precision="fp64"), 50_000),
        (TF.color_flame(5), dict(precision="fp64"), 60_000),
    ]


def make(ffr, flames, seed=SEED, **extra):
    out = []
    for i, (text, opt, _) in enumerate(flames):
        o = dict(opt)
        o.update(extra)
        out.append(ffr.BufferRenderer(text, seed=seed + i, batch_size=1000, **o))
    return out


def assert_same(a, b):
    wa, wb = a.read_buffer(), b.read_buffer()
    assert np.array_equal(a.counts(wa), b.counts(wb))
    if a.color_dims:
        ca, cb = a.colors(wa), b.colors(wb)
        tol = 1e-12 if a.word_bytes == 8 else 2e-6
        assert np.allclose(ca, cb, rtol=tol, atol=tol * max(1.0, float(np.abs(cb).max())))
    assert a.samples_iterated == b.samples_iterated
    assert a.samples_plotted == b.samples_plotted
    assert np.array_equal(a.xform_distribution(), b.xform_distribution())
    assert a.point_extremes == b.point_extremes


@pytest.mark.parametrize("graph", [True, False])
@pytest.mark.parametrize("cache", [-1, 1])
def test_batch_equals_one_at_a_time(ffr_mod, graph, cache):
    flames = mixed_flames()
    batched = make(ffr_mod, flames, cell_cache=cache)
    alone = make(ffr_mod, flames, cell_cache=cache)
    b = ffr_mod.Batch(batched)
    b.set_graph(graph)
    full = [s for _, _, s in flames]
    # the second pass leaves every third renderer out
    partial = [0 if i % 3 == 1 else s for i, s in enumerate(full)]
    for rep, samples in enumerate([full, partial, full]):
        print(f"batch pass {rep} graph={graph} cache={cache}", flush=True)
        ok = b.render(samples)
        assert all(ok)
        assert b.last_stats.graph == (1 if graph else 0)
        assert b.last_stats.samples == sum(samples)
        for r, s in zip(alone, samples):
            assert r.render(s)
        for x, y in zip(batched, alone):
            assert_same(x, y)
    b.close()
    for r in batched + alone:
        r.close()


def test_tuning_leaves_results_alone(ffr_mod):
    flames = mixed_flames()[:6]
    batched = make(ffr_mod, flames)
    alone = make(ffr_mod, flames)
    b = ffr_mod.Batch(batched)
    samples = [s for _, _, s in flames]
    b.tune(samples, fraction=0.25)
    for r in batched:
        assert r.counts().sum() == 0 and r.samples_iterated == 0
    shapes = b.shapes()
    assert len(shapes) >= 4
    for cache, kib, block, rounds in shapes:
        assert cache in (0, 1) and block in (128, 256, 512, 1024) and rounds in (1, 4)
    assert all(b.render(samples))
    for r, s in zip(alone, samples):
        assert r.render(s)
    for x, y in zip(batched, alone):
        assert_same(x, y)
    b.close()
    for r in batched + alone:
        r.close()


def test_bad_value_limit_per_renderer(ffr_mod):
    good = ffr_mod.BufferRenderer(TF.color_flame(3), seed=1, batch_size=1000)
    bad = ffr_mod.BufferRenderer(TF.expanding(), seed=2, batch_size=1000, bad_value_limit=3)
    b = ffr_mod.Batch([good, bad])
    ok = b.render([100_000, 100_000])
    assert ok == [True, False]
    assert good.samples_iterated == 100_000
    assert bad.last_stats.n_bad >= 4
    b.close()
    good.close()
    bad.close()


def test_batch_of_many_small_flames(ffr_mod):
    """64 seeded random flames from the batch set, at 64x64, FP32."""
    import os
    import re
    from conftest import FLAMES
    names = sorted(os.listdir(os.path.join(FLAMES, "batch")))[:64]
    texts = []
    for n in names:
        with open(os.path.join(FLAMES, "batch", n)) as fh:
            d = json.loads(re.sub(r"^\s*//.*$", "", fh.read(), flags=re.M))
        d["size"] = [64, 64]
        texts.append(json.dumps(d))
    batched = [ffr_mod.BufferRenderer(t, precision="fp32", seed=100 + i, batch_size=500)
               for i, t in enumerate(texts)]
    alone = [ffr_mod.BufferRenderer(t, precision="fp32", seed=100 + i, batch_size=500)
             for i, t in enumerate(texts)]
    samples = [20_000 + 1000 * i for i in range(len(texts))]
    b = ffr_mod.Batch(batched)
    assert all(b.render(samples))
    for r, s in zip(alone, samples):
        assert r.render(s)
    for x, y in zip(batched, alone):
        assert np.array_equal(x.counts(), y.counts())
        assert np.array_equal(x.xform_distribution(), y.xform_distribution())
    b.close()
    for r in batched + alone:
        r.close()


# ------------------------------------------------- precision policy

def test_histogram_l1_matches_host(ffr_mod):
    text = TF.color_flame(3)
    a = ffr_mod.BufferRenderer(text, seed=1, batch_size=1000)
    b = ffr_mod.BufferRenderer(text, precision="fp32", word=32, seed=2, batch_size=1000)
    twin = ffr_mod.BufferRenderer(text, seed=1, batch_size=1000)
    for r in (a, b, twin):
        assert r.render(200_000)
    assert ffr_mod.histogram_l1(a, twin) == 0.0
    ca = a.counts().astype(np.float64)
    cb = b.counts().astype(np.float64)
    want = float(np.abs(ca / ca.sum() - cb / cb.sum()).sum())
    got = ffr_mod.histogram_l1(a, b)
    assert want > 0.01 and abs(got - want) <= 1e-12 * max(1.0, want)
    for r in (a, b, twin):
        r.close()


def chaotic(size=64):
    """Swirl and sinusoidal after expanding rotations: nearby points separate,
    so FP32 and FP64 trajectories on the same streams drift apart (the CPU
    oracle puts the FP32/FP64 L1 of this flame at 0.06 for 2^20 samples)."""
    def rot(t, k):
        c, s = float(np.cos(t)) * k, float(np.sin(t)) * k
        return {"A": [[c, -s], [s, c]], "b": [0, 0]}
    return json.dumps({"dimensions": 2, "size": [size, size], "bounds": [[-2, 2]] * 2,
                       "xforms": [{"weight": 1, "variations": [{"name": "swirl", "weight": 1}],
                                   "pre_affine": rot(0.5, 1.3)},
                                  {"weight": 1, "variations": [{"name": "sinusoidal", "weight": 1}],
                                   "pre_affine": rot(1.0, 2.0)}]})


def test_precision_policy(ffr_mod):
    sier = read_flame("sierpinski_1024.json").replace("[1024, 1024]", "[64, 64]")
    # renders of 2^26 samples: the probe takes 2^20 of them, its bar is
    # 0.5 * 1e-3 * sqrt(64) = 4e-3
    prec, d = ffr_mod.choose_precision(sier, 1 << 26, seed=3)
    assert prec == "fp32" and d < 1e-3, d
    prec, d = ffr_mod.choose_precision(chaotic(), 1 << 26, seed=3)
    assert prec == "fp64" and d > 4e-3, d


def test_precision_policy_two_stages(ffr_mod):
    """the batched policy: clear cases decided by the first probe, the band
    between target and bar by the second (forced here by a target of half
    the chaotic flame's first-probe L1)"""
    sier = read_flame("sierpinski_1024.json").replace("[1024, 1024]", "[64, 64]")
    texts = [sier, chaotic()]
    n = [1 << 26, 1 << 26]
    prec, d = ffr_mod.choose_precisions(texts, n, [3, 3])
    assert prec[0] == "fp32" and d[0] < 1e-3
    assert prec[1] == "fp64" and d[1] > 4e-3
    # the same first probe with a target that puts the chaotic flame in
    # the band (target, 4 target]: the second stage decides it
    prec2, d2 = ffr_mod.choose_precisions(texts, n, [3, 3], target=d[1] / 2)
    # the first probe is the same render (its L1 is summed with atomics,
    # so only the last bits may differ)
    assert d2 == pytest.approx(d, rel=1e-12)
    assert prec2[0] == "fp32" and prec2[1] in ("fp32", "fp64")


def test_hot_cell_hints_leave_results_alone(ffr_mod):
    """40 flames of 256^2 hashed-cache cells, 16 tasks each: every flame has
    at most ~6 workgroups on it at a time, so its tasks seed their caches
    from the hints its earlier tasks left (written while other tasks of the
    same flame read them).  Counts equal the same renders done alone, which
    use no hints; colours agree to rounding."""
    flames = [TF.color_flame(3, size=256, seed=200 + i) for i in range(40)]
    n = 1 << 22
    batched = [ffr_mod.BufferRenderer(t, precision="fp32", seed=300 + i, batch_size=256)
               for i, t in enumerate(flames)]
    alone = [ffr_mod.BufferRenderer(t, precision="fp32", seed=300 + i, batch_size=256)
             for i, t in enumerate(flames)]
    with ffr_mod.Batch(batched) as b:
        for rep in range(2):          # the second pass starts from the first's hints
            assert all(b.render([n] * len(batched)))
            for r in alone:
                assert r.render(n)
            for x, y in zip(batched, alone):
                assert_same(x, y)
    for r in batched + alone:
        r.close()
