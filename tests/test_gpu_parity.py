"""GPU parity: the HIP render path against the CPU oracle, same walker streams.

Bit-exact for integer histograms of linear flames in FP64 (and in FP32 against
the oracle's float build), exact counts plus colour sums within 1e-9 relative
for colour flames, and normalised-histogram L1 for transcendental flames.
"""

import os

import numpy as np
import pytest

from conftest import FLAMES, read_flame

pytestmark = pytest.mark.gpu

SEED = 0x5EED0001


def sierpinski(size=64):
    return read_flame("sierpinski_1024.json").replace("[1024, 1024]", f"[{size}, {size}]")


def gpu_render(ffr, text, samples, batch, precision="fp64", word=64, settle=20,
               seed=SEED, **kw):
    r = ffr.BufferRenderer(text, precision=precision, word=word, seed=seed,
                           batch_size=batch, settle=settle, **kw)
    ok = r.render(samples)
    return r, ok


def test_sierpinski_fp64_bit_exact(ffr_mod, oracle_mod):
    text = sierpinski(256)
    samples, batch = 1 << 20, 1 << 14
    r, ok = gpu_render(ffr_mod, text, samples, batch)
    print("sierpinski 256^2, 2^20 samples, batch 2^14", flush=True)
    ref = oracle_mod.OracleFlame(text, oracle_mod.F64_U64, settle=20).render(
        samples, batch, seed=SEED, threads=8)
    gpu = r.read_buffer()
    assert ok and ref.success
    assert np.array_equal(gpu, ref.buffer)
    assert np.array_equal(r.xform_distribution(), ref.xf_dist)
    assert r.samples_iterated == ref.s_iter == samples
    assert r.samples_plotted == ref.s_plot == samples
    ext = r.point_extremes
    for i in range(2):
        assert ext[i][0] == ref.pt_min[i] and ext[i][1] == ref.pt_max[i]
    r.close()
