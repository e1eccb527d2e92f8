"""GPU checks of the measurement hooks: the walker kernel's count of global
atomic adds, and the atomic-rate probe behind the atomic roofline."""

import numpy as np
import pytest

import flames as TF

pytestmark = pytest.mark.gpu


def test_atomics_without_cache_are_one_per_word_per_plot(ffr_mod):
    """with the LDS cache off every plotted sample adds its count and each
    colour word with one global atomic each"""
    text = TF.color_flame(r=3, size=64)
    r = ffr_mod.BufferRenderer(text, seed=3, batch_size=512, cell_cache=0)
    assert r.render(1 << 20)
    assert r.last_stats.atomics == r.last_stats.s_plot * 4
    r.close()
    r = ffr_mod.BufferRenderer(TF.fern(), seed=3, batch_size=512, cell_cache=0)
    assert r.render(1 << 20)
    assert r.last_stats.atomics == r.last_stats.s_plot   # counts only
    r.close()


def test_atomics_with_direct_cache_are_the_flushes(ffr_mod):
    """a buffer small enough for a direct-indexed cache never misses: the
    atomics are the flushes, at most one cell update per task and cell"""
    text = TF.color_flame(r=3, size=32)
    r = ffr_mod.BufferRenderer(text, seed=3, batch_size=256, cell_cache=1)
    assert r.render(1 << 22)
    occupied = int(np.count_nonzero(r.counts()))
    a = r.last_stats.atomics
    assert a % 4 == 0 and a >= occupied * 4
    assert a < r.last_stats.s_plot   # far fewer than one per plot
    assert r.last_stats.flush_atomics == a   # nothing missed
    r.close()


def test_flush_atomics_without_cache_are_zero(ffr_mod):
    r = ffr_mod.BufferRenderer(TF.color_flame(r=3, size=64), seed=3, batch_size=512,
                               cell_cache=0)
    assert r.render(1 << 20)
    assert r.last_stats.flush_atomics == 0 and r.last_stats.atomics > 0
    r.close()


def test_atomic_rate_probe(ffr_mod):
    rate = ffr_mod.atomic_rate(1 << 20, 4, 8, 1 << 22)
    assert 1e8 < rate < 1e12
    rate32 = ffr_mod.atomic_rate(1 << 20, 1, 4, 1 << 22)
    assert 1e8 < rate32 < 1e12
    with pytest.raises(ffr_mod.FfrError, match="ffr_atomic_rate"):
        ffr_mod.atomic_rate(0, 4, 8, 1 << 20)
