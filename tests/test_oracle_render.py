"""Pins of the oracle's render loop by arithmetic special cases, invariants and
brute force (the reference renderer cannot be built here: it needs
nlohmann/json and Boost headers the image lacks).

Each case is chosen so that a plausible restatement mistake fails it:
  - affine order b + ((0 + a0 x0) + a1 x1) (affine.hpp:104-110, point.hpp:228-234)
    vs (b + a0 x0) + a1 x1, a transposed A, a dropped b;
  - the variation sum in file order starting from zero (xform.hpp:214-221);
  - the identity post-affine applied in 2-D even when absent (xform.hpp:222);
  - cell index by truncation, dimension 0 fastest, mult_d = size / (hi - lo)
    * (1 - 2^-52) (buffer_renderer.hpp:114-140, 196-209);
  - Sierpinski invariants (removed triangle empty, sum of counts = samples);
  - bad-value record / re-init schedule and the limit (buffer_renderer.hpp:175-186).
"""

import json
import math

import numpy as np
import pytest

import flames as TF
from conftest import read_flame


def xf_flame(A, b, variations=None, post=None, dims=2):
    x = {"weight": 1, "variations": variations or [{"name": "linear", "weight": 1.0}],
         "pre_affine": {"A": A, "b": b}}
    if post is not None:
        x["post_affine"] = post
    return json.dumps({"dimensions": dims, "size": [4] * dims, "bounds": [[-1, 1]] * dims,
                       "xforms": [x]})


def apply(oracle_mod, text, point, mode=None):
    f = oracle_mod.OracleFlame(text, oracle_mod.F64_U64 if mode is None else mode)
    return f.apply_xform(0, point)[0]


def test_affine_adds_b_after_the_dot_product(oracle_mod):
    e = 2.0 ** -53
    out = apply(oracle_mod, xf_flame([[e, e], [0, 1]], [1.0, 0.0]), [1.0, 1.0])
    # b + ((0 + e) + e) = 1 + 2^-52; (b + e) + e would round to 1.0
    assert out[0] == 1.0 + 2.0 ** -52


def test_affine_row_major_and_offset(oracle_mod):
    out = apply(oracle_mod, xf_flame([[1, 2], [3, 4]], [0.5, -0.25]), [1.0, 10.0])
    assert out == [21.5, 42.75]


def test_affine_dot_starts_from_positive_zero(oracle_mod):
    # b = -0: -0 + ((+0 + (-0)) + (-0)) = +0; a dot starting at -0 gives -0
    out = apply(oracle_mod, xf_flame([[1, 0], [0, 1]], [-0.0, -0.0]), [-0.0, -0.0])
    assert out[0] == 0.0 and math.copysign(1.0, out[0]) == 1.0


def test_variation_sum_in_file_order(oracle_mod):
    e = 2.0 ** -53
    vs = [{"name": "linear", "weight": 1.0}, {"name": "linear", "weight": e},
          {"name": "linear", "weight": e}]
    out = apply(oracle_mod, xf_flame([[1, 0], [0, 1]], [0, 0], vs), [1.0, 1.0])
    assert out[0] == 1.0                       # ((0 + 1) + e) + e
    rev = apply(oracle_mod, xf_flame([[1, 0], [0, 1]], [0, 0], vs[::-1]), [1.0, 1.0])
    assert rev[0] == 1.0 + 2.0 ** -52          # ((0 + e) + e) + 1


def test_zero_weight_variation_removed(oracle_mod):
    # weight 0 * inf would be NaN if the variation were kept (xform.hpp:47-58)
    vs = [{"name": "linear", "weight": 1.0}, {"name": "spherical", "weight": 0.0}]
    out = apply(oracle_mod, xf_flame([[1, 0], [0, 1]], [0, 0], vs), [0.0, 0.0])
    assert out == [0.0, 0.0]


def test_identity_post_affine_always_applied_in_2d(oracle_mod):
    # v = (1e308 * 1, 1e308 * 10) = (1e308, inf); identity post-affine then
    # gives x = 0 + ((0 + 1 * 1e308) + 0 * inf) = NaN
    vs = [{"name": "linear", "weight": 1e308}]
    out2 = apply(oracle_mod, xf_flame([[1, 0], [0, 1]], [0, 0], vs), [1.0, 10.0])
    assert math.isnan(out2[0])
    # in 3-D an absent post-affine is skipped (xform.hpp:222)
    t3 = xf_flame([[1, 0, 0], [0, 1, 0], [0, 0, 1]], [0, 0, 0], vs, dims=3)
    out3 = apply(oracle_mod, t3, [1.0, 10.0, 0.5])
    assert out3[0] == 1e308 and math.isinf(out3[1])


def test_constant_map_index(oracle_mod):
    """A constant map sends every sample to one cell, found by truncation.
    y = 0.7 on 20 rows: (0.7 + 1) * (10 * (1 - 2^-52)) = 16.999999999999996,
    so the scale-down puts an exact cell edge in the cell below (16, not 17);
    x = hi lands in the last cell, x = lo in the first."""
    for (bx, by), size, cell in [((0.3, 0.7), (10, 20), (6, 16)),
                                 ((-1.0, -1.0), (7, 5), (0, 0)),
                                 ((1.0, 1.0), (7, 5), (6, 4)),
                                 ((0.0, 0.0), (2, 2), (0, 0))]:
        text = json.dumps({"dimensions": 2, "size": list(size), "bounds": [[-1, 1], [-1, 1]],
                           "xforms": [{"weight": 1, "variations": [{"name": "linear", "weight": 1}],
                                       "pre_affine": {"A": [[0, 0], [0, 0]], "b": [bx, by]}}]})
        f = oracle_mod.OracleFlame(text)
        r = f.render(1000, 100, seed=1)
        assert r.s_plot == 1000
        idx = cell[0] + cell[1] * size[0]
        assert r.counts[idx] == 1000 and r.counts.sum() == 1000


def test_mult_d_closed_form(oracle_mod):
    f = oracle_mod.OracleFlame(read_flame("sierpinski_1024.json"))
    assert f.mult_d() == [(1024 / 2.0) * (1.0 - 2.0 ** -52)] * 2


def test_index_brute_force_tiny_grid(oracle_mod):
    """Every sample's cell equals the brute-force cell of its final point:
    a constant-per-walker map (A = 0, b = the walker's first pre-image) makes
    pf = b exactly, so the histogram is predictable cell by cell."""
    rng = np.random.default_rng(3)
    pts = rng.uniform(-1, 1, (40, 2))
    xfs = [{"weight": 1, "variations": [{"name": "linear", "weight": 1}],
            "pre_affine": {"A": [[0, 0], [0, 0]], "b": [float(x), float(y)]}} for x, y in pts]
    size = (5, 3)
    text = json.dumps({"dimensions": 2, "size": list(size), "bounds": [[-1, 1], [-1, 1]],
                       "xforms": xfs})
    f = oracle_mod.OracleFlame(text)
    r = f.render(20000, 500, seed=2)
    counts = np.zeros(size[0] * size[1], dtype=np.int64)
    ids = np.array(f.xform_order)
    # per-xform selections times the brute-force cell of its constant point
    for xid, n in enumerate(r.xf_dist):
        x, y = pts[xid]
        ix = int((x - -1.0) * ((size[0] / 2.0) * (1 - 2.0 ** -52)))
        iy = int((y - -1.0) * ((size[1] / 2.0) * (1 - 2.0 ** -52)))
        counts[ix + iy * size[0]] += int(n)
    assert len(ids) == 40
    assert np.array_equal(counts, r.counts.astype(np.int64))


def test_sierpinski_invariants(oracle_mod):
    text = read_flame("sierpinski_1024.json").replace("[1024, 1024]", "[256, 256]")
    f = oracle_mod.OracleFlame(text, settle=20)
    n = 1 << 20
    r = f.render(n, 1 << 14, seed=0x5EED0001, threads=8)
    assert r.s_plot == n and int(r.counts.sum()) == n
    c = r.counts.reshape(256, 256)  # [y][x]
    # cell centres strictly inside the first-level removed triangle
    # (0.5,0), (0.75,0.5), (0.25,0.5), and outside the big triangle, hold
    # nothing: settle error is below 2^-20, far less than half a cell
    h = 2.0 / 256
    ys, xs = np.mgrid[0:256, 0:256]
    cx, cy = -1 + (xs + 0.5) * h, -1 + (ys + 0.5) * h
    margin = 2 * h
    # the hole: apex (0.5, 0), top edge y = 0.5 from x = 0.25 to 0.75
    inside_hole = (cy > margin) & (cy < 0.5 - margin) & (np.abs(cx - 0.5) < cy / 2 - margin)
    outside = (cy < -margin) | (cx < cy / 2 - margin) | (cx > 1 - cy / 2 + margin) | (cy > 1 + margin)
    assert inside_hole.sum() > 1000 and outside.sum() > 10000
    assert c[inside_hole].sum() == 0
    assert c[outside].sum() == 0
    # equal weights: each map taken about a third of the time (chi-square)
    exp = n / 3.0
    chi2 = float(((r.xf_dist - exp) ** 2 / exp).sum())
    assert chi2 < 20.0


def test_threads_give_identical_counts(oracle_mod):
    f = oracle_mod.OracleFlame(TF.color_flame(3))
    a = f.render(200_000, 1000, seed=4, threads=1)
    b = f.render(200_000, 1000, seed=4, threads=8)
    assert np.array_equal(a.counts, b.counts)
    assert np.allclose(a.colors, b.colors, rtol=1e-12)
    assert np.array_equal(a.xf_dist, b.xf_dist)


def test_bad_value_reinit_and_stale_final_point(oracle_mod):
    f = oracle_mod.OracleFlame(TF.occasional_bad())
    r = f.render(100_000, 1000, seed=5, bv_limit=10 ** 9)
    assert r.success and r.n_bad > 20
    # the final xform maps every point, bad or not, next to (0.25, 0.25), so
    # each iteration plots once, including the stale final point after a re-init
    assert r.s_plot == 100_000 and r.s_iter == 100_000
    # a record names the xform of the iteration on which the point passed
    # 1e20: the expanding map (file index 2) or a later contraction
    assert 2 in set(r.bad_xforms.tolist()) and set(r.bad_xforms.tolist()) <= {0, 1, 2}
    assert np.all(np.abs(r.bad_points).max(axis=1) > 1e20)


def test_bad_value_limit(oracle_mod):
    f = oracle_mod.OracleFlame(TF.expanding())
    r = f.render(50_000, 1000, seed=6, bv_limit=0)
    assert not r.success and r.n_bad == 1
    r = f.render(50_000, 1000, seed=6, bv_limit=3)
    assert not r.success and r.n_bad == 4


def test_walker_batches(oracle_mod):
    """Walkers [0,k) and [k,n) rendered apart and added equal one render."""
    text = read_flame("sierpinski_1024.json").replace("[1024, 1024]", "[64, 64]")
    f = oracle_mod.OracleFlame(text, settle=20)
    whole = f.render(64 * 512, 512, seed=8)
    a = f.render(24 * 512, 512, seed=8)
    b = f.render(40 * 512, 512, seed=8, walker_base=24)
    assert np.array_equal(whole.counts, a.counts + b.counts)
    # the last walker renders the remainder
    part = f.render(1000, 512, seed=8)
    assert part.s_iter == 1000
