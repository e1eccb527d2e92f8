"""The oracle's image stage (ffr-img, src/ffr_img.cpp:199-309) against what
the mathematics fixes, on the CPU.

The oracle restates render_image in num_t with libm's log and pow.  These
tests check it against values computed independently in 50-digit decimal
arithmetic (Decimal.ln / exp, correctly rounded), pixel by pixel, except
where the exact value lies within a hair of a pixel boundary (there the
last bit of libm decides, and the reference does the same).  They also pin
the closed forms: count 0 -> 0, the brightest cell -> 2^bits - 1 (pix_scale
is 2^bits (1 - 2^-mantissa), so exactly 1.0 rounds down), colour = gray
value x colour sum / count, mono = 255 where the count is not 0, row order
and channel order, and the reference's rejections.
"""

from decimal import Decimal, getcontext

import numpy as np
import pytest

getcontext().prec = 50

W, H = 23, 17  # not square: a transposed image fails


def word_buffer(counts, colors=None, word=8):
    """cells of [count, colour...] words, counts (H, W), colors (H, W, r) floats"""
    it = np.uint64 if word == 8 else np.uint32
    ft = np.float64 if word == 8 else np.float32
    r = 0 if colors is None else colors.shape[-1]
    buf = np.zeros((counts.size, 1 + r), dtype=it)
    buf[:, 0] = counts.reshape(-1)
    if r:
        buf[:, 1:] = colors.reshape(-1, r).astype(ft).view(it)
    return buf.reshape(-1)


def exact_value(c, cmax, gamma):
    """pow(log(1 + c) / log(1 + cmax), 1 / gamma), exactly"""
    if c == 0:
        return Decimal(0)
    lv = Decimal(1 + c).ln() / Decimal(1 + cmax).ln()
    return (lv.ln() / Decimal(repr(gamma))).exp()


def scale_of(bits, word):
    m = 52 if word == 8 else 23
    return Decimal(2 ** bits) * (1 - Decimal(2) ** -m)


def expect_pixel(v: Decimal, bits, word, margin):
    """floor(v * pix_scale), or None within margin (relative) of a boundary"""
    x = v * scale_of(bits, word)
    margin = Decimal(repr(margin))
    f = int(x)
    frac = x - f
    if x > 0 and (frac < margin * x or (1 - frac) < margin * x):
        return None
    return f


def random_counts(rng, cmax=5000):
    c = rng.integers(0, cmax, size=(H, W), dtype=np.int64)
    c[rng.random((H, W)) < 0.3] = 0
    return c


@pytest.mark.parametrize("word", [8, 4])
@pytest.mark.parametrize("bits", [8, 16])
@pytest.mark.parametrize("gamma", [1.0, 2.2, 0.45])
def test_gray_matches_decimal(oracle_mod, word, bits, gamma):
    rng = np.random.default_rng(7 + bits + word)
    counts = random_counts(rng)
    st, pix, hist, scal = oracle_mod.tonemap(word_buffer(counts, word=word), W, H, 1, 1,
                                             gamma, bits)
    assert st == 0
    assert pix.dtype == (np.uint16 if bits == 16 else np.uint8)
    assert hist == (int(counts.min()), int(counts.max()))
    cmax = int(counts.max())
    margin = 1e-12 if word == 8 else 5e-7
    checked = 0
    for y in range(H):
        for x in range(W):
            c = int(counts[y, x])
            if c == cmax:
                assert pix[y, x] == 2 ** bits - 1
                continue
            e = expect_pixel(exact_value(c, cmax, gamma), bits, word, margin)
            if e is not None:
                assert pix[y, x] == e, (y, x, c, cmax, gamma)
                checked += 1
    # float arithmetic leaves a wider band around each boundary at 16 bits
    share = 0.9 if word == 8 or bits == 8 else 0.7
    assert checked > share * (W * H - np.count_nonzero(counts == cmax))
    # the scaler bounds: log(1 + min), log(1 + max)
    rel = 1e-15 if word == 8 else 1e-7
    assert abs(scal[1] - float(Decimal(1 + cmax).ln())) <= rel * scal[1]
    assert abs(scal[0] - float(Decimal(1 + int(counts.min())).ln())) <= rel * max(scal[0], 1)


@pytest.mark.parametrize("word", [8, 4])
@pytest.mark.parametrize("bits", [8, 16])
def test_color_matches_decimal(oracle_mod, word, bits):
    rng = np.random.default_rng(11 + bits + word)
    counts = random_counts(rng)
    ft = np.float64 if word == 8 else np.float32
    mean = rng.random((H, W, 3))
    sums = (mean * counts[..., None]).astype(ft)  # colour sums as the buffer holds them
    gamma = 2.2
    st, pix, _, _ = oracle_mod.tonemap(word_buffer(counts, sums, word=word), W, H, 4, 2,
                                       gamma, bits)
    assert st == 0 and pix.shape == (H, W, 3)
    cmax = int(counts.max())
    margin = 1e-12 if word == 8 else 5e-7
    for y in range(H):
        for x in range(W):
            c = int(counts[y, x])
            if c == 0:
                assert (pix[y, x] == 0).all()
                continue
            g = exact_value(c, cmax, gamma)
            for i in range(3):
                v = g * (Decimal(float(sums[y, x, i])) / Decimal(c))
                e = expect_pixel(v, bits, word, margin)
                if e is not None:
                    assert pix[y, x, i] == e, (y, x, i)


def test_closed_forms(oracle_mod):
    counts = np.zeros((H, W), dtype=np.int64)
    counts[0, 0] = 8   # the brightest cell
    counts[H - 1, 2] = 8
    counts[3, W - 1] = 1
    colors = np.zeros((H, W, 3))
    colors[0, 0] = [8.0, 4.0, 0.0]        # mean colour 1, 1/2, 0
    colors[H - 1, 2] = [0.0, 2.0, 6.0]    # 0, 1/4, 3/4
    buf = word_buffer(counts, colors)
    # mono: 255 exactly where the count is not 0, whatever bits says
    for bits in (8, 16):
        st, pix, _, _ = oracle_mod.tonemap(buf, W, H, 4, 0, 1.0, bits)
        assert st == 0 and pix.dtype == np.uint8
        assert np.array_equal(pix, np.where(counts != 0, 255, 0).astype(np.uint8))
    # gray: the brightest cells are 2^bits - 1, empty ones 0
    st, pix, hist, _ = oracle_mod.tonemap(buf, W, H, 4, 1, 1.0, 8)
    assert hist == (0, 8)
    assert pix[0, 0] == 255 and pix[H - 1, 2] == 255 and pix[5, 5] == 0
    # gamma 2: the value is the square root; log(2) / log(9) = 0.31546...
    st, pix, _, _ = oracle_mod.tonemap(buf, W, H, 4, 1, 2.0, 16)
    assert pix[3, W - 1] == int(Decimal(Decimal(2).ln() / Decimal(9).ln()).sqrt()
                                * 65536 * (1 - Decimal(2) ** -52))
    # colour: the gray value times the mean colour; 1/2 and 1/4 of 2^bits
    # land just below a boundary because pix_scale is below 2^bits
    st, pix, _, _ = oracle_mod.tonemap(buf, W, H, 4, 2, 1.0, 8)
    assert list(pix[0, 0]) == [255, 127, 0]
    assert list(pix[H - 1, 2]) == [0, 63, 191]
    st, pix, _, _ = oracle_mod.tonemap(buf, W, H, 4, 2, 1.0, 16)
    assert list(pix[0, 0]) == [65535, 32767, 0]


def test_rejections(oracle_mod):
    zero = word_buffer(np.zeros((H, W), dtype=np.int64), np.zeros((H, W, 3)))
    st, pix, hist, scal = oracle_mod.tonemap(zero, W, H, 4, 1, 1.0, 8)
    assert st == 1 and pix is None and hist == (0, 0) and scal == (0.0, 0.0)
    one = word_buffer(np.ones((H, W), dtype=np.int64), np.zeros((H, W, 3)))
    assert oracle_mod.tonemap(one, W, H, 4, 1, 1.0, 12)[0] == 2
    gray_only = word_buffer(np.ones((H, W), dtype=np.int64))
    assert oracle_mod.tonemap(gray_only, W, H, 1, 2, 1.0, 8)[0] == 2
    # a single non-zero count: log(2) > eps, so not empty; it maps to 1.0
    c = np.zeros((H, W), dtype=np.int64)
    c[1, 1] = 1
    st, pix, _, _ = oracle_mod.tonemap(word_buffer(c), W, H, 1, 1, 1.0, 8)
    assert st == 0 and pix[1, 1] == 255 and int(pix.sum()) == 255


def test_planted_mistake_is_caught(oracle_mod):
    """pow(l, gamma) in place of pow(l, 1 / gamma) fails the decimal check"""
    rng = np.random.default_rng(3)
    counts = random_counts(rng)
    cmax = int(counts.max())
    gamma = 2.2
    _, pix, _, _ = oracle_mod.tonemap(word_buffer(counts), W, H, 1, 1, gamma, 8)
    wrong = 0
    for y in range(H):
        for x in range(W):
            c = int(counts[y, x])
            if 0 < c < cmax:
                e = expect_pixel(exact_value(c, cmax, 1 / gamma), 8, 8, 1e-12)
                wrong += e is not None and pix[y, x] != e
    assert wrong > 0.5 * np.count_nonzero((counts > 0) & (counts < cmax))


def test_edge_values(oracle_mod):
    """gamma = inf (gp = 0: pow(l, 0) = 1 everywhere, empty cells too),
    gamma = NaN (passes the reference's check; pow(1, NaN) = 1, else NaN -> 0),
    colour values outside [0, 1]: the pixel keeps the low bits of a signed
    32-bit conversion, and values beyond the int32 range give 0"""
    counts = np.zeros((H, W), dtype=np.int64)
    counts[0, 0] = 4
    counts[1, 1] = 2
    colors = np.zeros((H, W, 3))
    colors[0, 0] = [4 * 300.0, -4 * 0.5, 4 * 1e10]   # mean 300, -0.5, 1e10
    colors[1, 1] = [1.0, 1.0, 1.0]
    buf = word_buffer(counts, colors)
    st, pix, _, _ = oracle_mod.tonemap(buf, W, H, 4, 1, float("inf"), 8)
    assert st == 0 and (pix == 255).all()
    st, pix, _, _ = oracle_mod.tonemap(buf, W, H, 4, 1, float("nan"), 8)
    assert st == 0 and pix[0, 0] == 255 and int(pix.sum()) == 255
    st, pix, _, _ = oracle_mod.tonemap(buf, W, H, 4, 2, 1.0, 8)
    # 300 * 255.99.. = 76799.99.. -> 76799 -> low byte 255; -0.5 * 255.99.. =
    # -127.99.. -> -127 -> 129; 2.56e12 is outside int32 -> 0
    assert list(pix[0, 0]) == [76799 & 255, (-127) & 255, 0]
    st, pix, _, _ = oracle_mod.tonemap(buf, W, H, 4, 2, 1.0, 16)
    assert list(pix[0, 0]) == [(300 * 65536 - 1) & 0xFFFF, (-32767) & 0xFFFF, 0]
