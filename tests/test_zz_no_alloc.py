"""ffr_add_buffer_noalloc (through the C ABI) against ffr_add_buffer.

The no-allocation variant stages host words through the library's fixed
16 MiB device array instead of a per-call device buffer, with the same add
kernel, so its result must equal ffr_add_buffer's bit for bit: counts added
as integers, colours as floats, each word once.  Two calls in a row on one
GPU: the first adds a 265 MB colour buffer (17 chunks of the array), the
second a small one, so the second runs over what the first left in the
array.  The u32 float build's words and a plain count buffer are covered
too, as are the expected sums computed on the host.
"""

import numpy as np
import pytest

import flames as TF
from conftest import read_flame

pytestmark = pytest.mark.gpu


def words_for(r, rng):
    """synthetic buffer words of r's layout: small counts, colour sums"""
    cw = r.cell_words
    n = r.buffer_bytes // r.word_bytes
    cells = n // cw
    if r.word_bytes == 8:
        w = np.empty(n, dtype=np.uint64)
        w.reshape(cells, cw)[:, 0] = rng.integers(0, 1 << 20, cells, dtype=np.uint64)
        if cw > 1:
            w.reshape(cells, cw)[:, 1:] = rng.uniform(0, 50, (cells, cw - 1)).view(np.uint64)
    else:
        w = np.empty(n, dtype=np.uint32)
        w.reshape(cells, cw)[:, 0] = rng.integers(0, 1 << 20, cells, dtype=np.uint32)
        if cw > 1:
            w.reshape(cells, cw)[:, 1:] = rng.uniform(0, 50, (cells, cw - 1)).astype(
                np.float32).view(np.uint32)
    return w


def expected(r, base, add):
    cw = r.cell_words
    out = base.copy().reshape(-1, cw)
    b2, a2 = base.reshape(-1, cw), add.reshape(-1, cw)
    out[:, 0] = b2[:, 0] + a2[:, 0]
    if cw > 1:
        ft = np.float64 if r.word_bytes == 8 else np.float32
        out[:, 1:] = (b2[:, 1:].view(ft) + a2[:, 1:].view(ft)).view(out.dtype)
    return out.reshape(-1)


def check_pair(ffr, text, rng, **kw):
    a = ffr.BufferRenderer(text, seed=1, batch_size=1000, **kw)
    b = ffr.BufferRenderer(text, seed=1, batch_size=1000, **kw)
    base, add = words_for(a, rng), words_for(a, rng)
    a.write_buffer(base)
    b.write_buffer(base)
    a.add_buffer(add)           # the existing entry point (allocates its staging)
    b.add_buffer_noalloc(add)   # the new one (the library's fixed array)
    wa, wb = a.read_buffer(), b.read_buffer()
    assert np.array_equal(wa, wb)
    assert np.array_equal(wb, expected(b, base, add))
    a.close()
    b.close()
    return base.nbytes


def test_noalloc_add_equals_add_buffer(ffr_mod):
    rng = np.random.default_rng(7)
    big = read_flame("random_color_3840x2160.json")
    print("noalloc: 3840x2160 colour buffer (17 chunks)", flush=True)
    n1 = check_pair(ffr_mod, big, rng)
    assert n1 == 3840 * 2160 * 32 and n1 > 16 * (1 << 20)
    print("noalloc: second call, 64x64 colour buffer", flush=True)
    n2 = check_pair(ffr_mod, TF.color_flame(3, size=64), rng)
    assert n2 < n1
    print("noalloc: float build words and a count-only buffer", flush=True)
    check_pair(ffr_mod, TF.color_flame(2, size=96), rng, precision="fp32", word=32)
    check_pair(ffr_mod, TF.tetrahedron(size=40), rng, precision="fp32", word=32)


def test_noalloc_add_onto_a_render(ffr_mod):
    """ffr-buf -i semantics on a rendered buffer: a render, then the same
    buffer added through each entry point, gives twice the render"""
    text = read_flame("sierpinski_1024.json").replace("[1024, 1024]", "[256, 256]")
    r = ffr_mod.BufferRenderer(text, seed=3, batch_size=4096, settle=20)
    assert r.render(1 << 20)
    w = r.read_buffer()
    r.add_buffer_noalloc(w)
    assert np.array_equal(r.read_buffer(), 2 * w)
    with pytest.raises(ffr_mod.FfrError, match="sizes do not match"):
        r.add_buffer_noalloc(w[:-1])
    r.close()
