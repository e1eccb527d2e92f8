"""The wave-per-walker kernel (few walkers of a linear flame, with up to four
colour dimensions, final xform or not) against the lane-per-walker kernel
and the oracle.  Flames with
identity post-affines take its constant-post build, the others (three_slots,
negzero_post) read the post-affine from the table; maps that share one
matrix and weight (sierpinski, tetrahedron) take the build that holds them
once, the others (uneven_weights, fern, ...) broadcast every coefficient.

A render with the cell cache left to the renderer (-1) and at most the
build's walkers per CU takes the wave-per-walker kernel; `cell_cache=0`
forces the lane-per-walker kernel on the same walkers and streams.
Buffers, samples iterated and plotted, xform counts, point extremes
(compared as numbers: the kernel leaves out operations that change only
the sign of a zero) and bad values must be equal, in every precision mode
and at walker shapes that leave partial blocks of 64 iterations (short
walkers, a short last walker).
"""

import json
import re

import numpy as np
import pytest

import flames as TF
from conftest import read_flame

pytestmark = pytest.mark.gpu

SEED = 0x5EED0007
MODES = {"fp64": 0, "fp32": 1, "fp32u64": 2}


def sierpinski(size=128):
    return read_flame("sierpinski_1024.json").replace("[1024, 1024]", f"[{size}, {size}]")


def load(text):
    """a flame document (the shipped files carry // comment lines)"""
    return json.loads(re.sub(r"^\s*//.*$", "", text, flags=re.M))


def three_slots(seed=11, size=96):
    """four maps, each the sum of three linear variations of random weight"""
    rng = np.random.default_rng(seed)
    xfs = []
    for k in range(4):
        A = (rng.uniform(-0.6, 0.6, (2, 2))).tolist()
        w = rng.dirichlet([1, 1, 1]).tolist()
        xfs.append({"weight": float(rng.uniform(0.5, 2)),
                    "variations": [{"name": "linear", "weight": x} for x in w],
                    "pre_affine": {"A": A, "b": rng.uniform(-1, 1, 2).tolist()},
                    "post_affine": {"A": [[1, 0.2], [-0.1, 0.9]],
                                    "b": rng.uniform(-0.2, 0.2, 2).tolist()}})
    return json.dumps({"dimensions": 2, "size": [size, size], "bounds": [[-3, 3], [-3, 3]],
                       "xforms": xfs})


def bad_no_final(size=64):
    """two halving maps and a rare one that throws the point beyond 1e20
    (a bad value: the walker starts again)"""
    return json.dumps({
        "dimensions": 2, "size": [size, size], "bounds": [[-2, 2], [-2, 2]],
        "xforms": [
            {"weight": 1.0, "variations": [{"name": "linear", "weight": 1.0}],
             "pre_affine": {"A": [[0.5, 0], [0, 0.5]], "b": [0.5, 0]}},
            {"weight": 1.0, "variations": [{"name": "linear", "weight": 1.0}],
             "pre_affine": {"A": [[0.5, 0], [0, 0.5]], "b": [-0.5, 0.3]}},
            {"weight": 0.004, "variations": [{"name": "linear", "weight": 1.0}],
             "pre_affine": {"A": [[1e25, 0], [0, 1e25]], "b": [0, 0]}},
        ]})


def negzero_post(size=96):
    """Sierpinski with post-affines whose offsets are -0.0: not the identity
    bit for bit (-0 + +0 is +0), so the kernel reads them from the table"""
    d = load(sierpinski(size))
    for x in d["xforms"]:
        x["post_affine"] = {"A": [[1, 0], [0, 1]], "b": [-0.0, -0.0]}
    return json.dumps(d)


def uneven_weights(size=96):
    """Sierpinski's shared matrix, but the three linear variations carry
    different weights (and one matrix a -0.0 entry): the kernel broadcasts
    every xform's coefficients instead of holding one copy"""
    d = load(sierpinski(size))
    for k, x in enumerate(d["xforms"]):
        x["variations"] = [{"name": "linear", "weight": [1.0, 0.9, 1.1][k]}]
    d["xforms"][2]["pre_affine"]["A"] = [[0.5, -0.0], [0.0, 0.5]]
    return json.dumps(d)


def shared_matrix(A, size=96):
    """Sierpinski's offsets under one shared pre-affine matrix A: the kernel
    holds the matrix once; a diagonal one (off-diagonal entries +0 or -0)
    takes the build that leaves the off-diagonal products out"""
    d = load(sierpinski(size))
    for x in d["xforms"]:
        x["pre_affine"]["A"] = A
    return json.dumps(d)


FLAMES = {
    "sierpinski": (sierpinski, 2),
    "shared_rotation": (lambda: shared_matrix([[0.4, -0.3], [0.3, 0.4]]), 2),
    "shared_negzero_diagonal": (lambda: shared_matrix([[0.5, -0.0], [-0.0, 0.45]]), 2),
    "uneven_weights": (uneven_weights, 2),
    "negzero_post": (negzero_post, 2),
    "fern": (TF.fern, 2),
    "tetrahedron": (TF.tetrahedron, 3),
    "tetrahedron_post": (lambda: TF.tetrahedron(post=True), 3),
    "one_d": (TF.one_d, 1),
    "tied_weights": (TF.tied_weights, 2),
    "three_slots": (three_slots, 2),
    "rotations_final": (TF.rotations, 2),
}


def render(ffr, text, samples, batch, mode, cache, settle=-1, seed=SEED, **kw):
    precision = "fp64" if mode == "fp64" else "fp32"
    word = 32 if mode == "fp32" else 64
    r = ffr.BufferRenderer(text, precision=precision, word=word, seed=seed,
                           batch_size=batch, settle=settle, cell_cache=cache, **kw)
    ok = r.render(samples)
    return r, ok


def records(ids, pts):
    """bad-value records as a sorted list; NaN coordinates (FP32 points that
    overflowed) compare equal to each other"""
    def coord(v):
        return (1, 0.0) if np.isnan(v) else (0, float(v))
    return sorted((int(i), tuple(coord(v) for v in row)) for i, row in zip(ids, pts))


def assert_same(a, b, dims):
    assert np.array_equal(a.read_buffer(), b.read_buffer())
    assert a.samples_iterated == b.samples_iterated
    assert a.samples_plotted == b.samples_plotted
    assert np.array_equal(a.xform_distribution(), b.xform_distribution())
    ea, eb = a.point_extremes, b.point_extremes
    for i in range(dims):
        assert ea[i] == eb[i], (i, ea, eb)
    assert a.last_stats.n_bad == b.last_stats.n_bad


@pytest.mark.parametrize("mode", ["fp64", "fp32", "fp32u64"])
@pytest.mark.parametrize("name", sorted(FLAMES))
def test_chain_equals_lane_kernel(ffr_mod, name, mode):
    maker, dims = FLAMES[name]
    text = maker()
    # 101 walkers, the last of 3 samples; then short walkers (37 samples,
    # under one block of 64), with settle iterations
    for samples, batch, settle in ((100_003, 1000, -1), (5_032, 37, 20)):
        print(f"{name} {mode} samples {samples} batch {batch}", flush=True)
        a, oka = render(ffr_mod, text, samples, batch, mode, -1, settle)
        b, okb = render(ffr_mod, text, samples, batch, mode, 0, settle)
        assert oka and okb
        assert a.last_stats.wave_per_walker == 1
        assert b.last_stats.wave_per_walker == 0
        assert_same(a, b, dims)
        a.close()
        b.close()


@pytest.mark.parametrize("flame", ["three_slots", "negzero_post", "sierpinski",
                                   "uneven_weights", "rotations_final", "shared_rotation",
                                   "shared_negzero_diagonal"])
@pytest.mark.parametrize("mode", ["fp64", "fp32"])
def test_chain_matches_oracle(ffr_mod, oracle_mod, mode, flame):
    text = FLAMES[flame][0]()
    samples, batch = 200_000, 2000
    r, ok = render(ffr_mod, text, samples, batch, mode, -1, settle=20)
    assert ok and r.last_stats.wave_per_walker == 1
    f = oracle_mod.OracleFlame(text, MODES[mode], settle=20)
    ref = f.render(samples, batch, seed=SEED, threads=8)
    assert ref.success
    assert np.array_equal(r.read_buffer(), ref.buffer)
    assert r.samples_iterated == ref.s_iter and r.samples_plotted == ref.s_plot
    assert np.array_equal(r.xform_distribution(), ref.xf_dist)
    ext = r.point_extremes
    for i in range(2):
        assert ext[i] == (ref.pt_min[i], ref.pt_max[i])
    r.close()


@pytest.mark.parametrize("flame", ["no_final", "final"])
@pytest.mark.parametrize("mode", ["fp64", "fp32u64"])
def test_chain_bad_values_restart_exact(ffr_mod, oracle_mod, mode, flame):
    """bad values end a block: the walker starts again from the next word;
    every bad point, its xform, the counts and extremes equal the oracle's
    (with a final xform, the plotted point of a bad step is the final
    xform's image of the bad point)"""
    text = bad_no_final() if flame == "no_final" else TF.occasional_bad()
    samples, batch = 300_000, 1500
    r, ok = render(ffr_mod, text, samples, batch, mode, -1, bad_value_limit=10 ** 9)
    assert ok and r.last_stats.wave_per_walker == 1
    f = oracle_mod.OracleFlame(text, MODES[mode], settle=-1)
    ref = f.render(samples, batch, seed=SEED, threads=1, bv_limit=10 ** 9, bv_cap=4096)
    assert ref.success and ref.n_bad > 200
    assert np.array_equal(r.read_buffer(), ref.buffer)
    assert r.samples_iterated == ref.s_iter and r.samples_plotted == ref.s_plot
    assert np.array_equal(r.xform_distribution(), ref.xf_dist)
    ids, pts = r.bad_values()
    assert len(ids) == ref.n_bad
    assert records(ids, pts) == records(ref.bad_xforms, ref.bad_points)
    lane, _ = render(ffr_mod, text, samples, batch, mode, 0, bad_value_limit=10 ** 9)
    assert_same(r, lane, 2)
    r.close()
    lane.close()


@pytest.mark.parametrize("r_dims,final_color", [(3, True), (3, False), (1, True), (4, True)])
@pytest.mark.parametrize("mode", ["fp64", "fp32", "fp32u64"])
def test_chain_colour_equals_lane_kernel(ffr_mod, oracle_mod, r_dims, final_color, mode):
    """linear colour flames (a final xform with or without its own colour):
    counts and statistics equal the lane kernel's and the oracle's, colour
    sums agree to the rounding of their atomic order"""
    text = TF.color_flame(r_dims, final_color)
    samples, batch = 400_000, 2000
    print(f"colour r={r_dims} final_color={final_color} {mode}", flush=True)
    a, oka = render(ffr_mod, text, samples, batch, mode, -1)
    b, okb = render(ffr_mod, text, samples, batch, mode, 0)
    assert oka and okb
    assert a.last_stats.wave_per_walker == 1 and b.last_stats.wave_per_walker == 0
    wa, wb = a.read_buffer(), b.read_buffer()
    assert np.array_equal(a.counts(wa), b.counts(wb))
    tol = 1e-12 if mode != "fp32" else 2e-5
    ca, cb = a.colors(wa), b.colors(wb)
    assert np.allclose(ca, cb, rtol=tol, atol=tol * float(np.abs(cb).max()))
    assert a.samples_iterated == b.samples_iterated
    assert a.samples_plotted == b.samples_plotted
    assert np.array_equal(a.xform_distribution(), b.xform_distribution())
    assert a.point_extremes == b.point_extremes
    ref = oracle_mod.OracleFlame(text, MODES[mode], settle=-1).render(
        samples, batch, seed=SEED, threads=1)
    assert ref.success and np.array_equal(a.counts(wa), ref.counts)
    assert np.allclose(ca, ref.colors, rtol=tol, atol=tol * float(np.abs(ref.colors).max()))
    a.close()
    b.close()


def with_colour(text, r=3, seed=21):
    """the flame with r colour dimensions: every xform but the last gets a
    colour, the second its own speed"""
    d = load(text)
    rng = np.random.default_rng(seed)
    d["color_dimensions"] = r
    d["color_speed"] = 0.45
    for k, x in enumerate(d["xforms"][:-1]):
        x["color"] = rng.uniform(0, 1, r).tolist()
        if k == 1:
            x["color_speed"] = 0.2
    return json.dumps(d)


# the kernel builds the colour flames above leave out: colour in 3-D (the
# has_pre / has_post flags beside the colour blend), colour with
# post-affines read from the table, colour with one shared matrix and weight
COLOUR_BUILDS = {
    "tetrahedron_3d": (lambda: with_colour(TF.tetrahedron()), 3),
    "tetrahedron_post_3d": (lambda: with_colour(TF.tetrahedron(post=True), r=2), 3),
    "three_slots_post": (lambda: with_colour(three_slots()), 2),
    "negzero_post": (lambda: with_colour(negzero_post(), r=4), 2),
    "sierpinski_shared": (lambda: with_colour(sierpinski(96)), 2),
}


@pytest.mark.parametrize("mode", ["fp64", "fp32", "fp32u64"])
@pytest.mark.parametrize("name", sorted(COLOUR_BUILDS))
def test_chain_colour_builds(ffr_mod, oracle_mod, name, mode):
    maker, dims = COLOUR_BUILDS[name]
    text = maker()
    samples, batch = 300_000, 1500
    print(f"colour build {name} {mode}", flush=True)
    a, oka = render(ffr_mod, text, samples, batch, mode, -1, settle=20)
    b, okb = render(ffr_mod, text, samples, batch, mode, 0, settle=20)
    assert oka and okb
    assert a.last_stats.wave_per_walker == 1 and b.last_stats.wave_per_walker == 0
    wa, wb = a.read_buffer(), b.read_buffer()
    assert np.array_equal(a.counts(wa), b.counts(wb))
    tol = 1e-12 if mode != "fp32" else 2e-5
    ca, cb = a.colors(wa), b.colors(wb)
    assert np.allclose(ca, cb, rtol=tol, atol=tol * float(np.abs(cb).max()))
    assert a.samples_iterated == b.samples_iterated
    assert a.samples_plotted == b.samples_plotted
    assert np.array_equal(a.xform_distribution(), b.xform_distribution())
    assert a.point_extremes == b.point_extremes
    ref = oracle_mod.OracleFlame(text, MODES[mode], settle=20).render(
        samples, batch, seed=SEED, threads=8)
    assert ref.success and np.array_equal(a.counts(wa), ref.counts)
    assert np.allclose(ca, ref.colors, rtol=tol, atol=tol * float(np.abs(ref.colors).max()))
    assert a.samples_plotted == ref.s_plot
    assert np.array_equal(a.xform_distribution(), ref.xf_dist)
    a.close()
    b.close()


def bad_colour(size=64):
    """bad_no_final with two colour dimensions: a restart draws new colour
    words, and the bad step's plotted colour is the one before it"""
    d = json.loads(bad_no_final(size))
    d["color_dimensions"] = 2
    for k, x in enumerate(d["xforms"]):
        x["color"] = [0.2 + 0.3 * k, 0.9 - 0.4 * k]
    return json.dumps(d)


@pytest.mark.parametrize("mode", ["fp64", "fp32"])
def test_chain_colour_bad_values(ffr_mod, oracle_mod, mode):
    text = bad_colour()
    samples, batch = 300_000, 1500
    a, oka = render(ffr_mod, text, samples, batch, mode, -1, bad_value_limit=10 ** 9)
    b, okb = render(ffr_mod, text, samples, batch, mode, 0, bad_value_limit=10 ** 9)
    assert oka and okb
    assert a.last_stats.wave_per_walker == 1 and b.last_stats.wave_per_walker == 0
    wa, wb = a.read_buffer(), b.read_buffer()
    assert np.array_equal(a.counts(wa), b.counts(wb))
    tol = 1e-12 if mode != "fp32" else 2e-5
    ca, cb = a.colors(wa), b.colors(wb)
    assert np.allclose(ca, cb, rtol=tol, atol=tol * float(np.abs(cb).max()))
    assert a.samples_iterated == b.samples_iterated
    assert a.samples_plotted == b.samples_plotted
    assert np.array_equal(a.xform_distribution(), b.xform_distribution())
    assert a.last_stats.n_bad == b.last_stats.n_bad > 200
    ia, pa = a.bad_values()
    ib, pb = b.bad_values()
    assert records(ia, pa) == records(ib, pb)
    ref = oracle_mod.OracleFlame(text, MODES[mode], settle=-1).render(
        samples, batch, seed=SEED, threads=1, bv_limit=10 ** 9, bv_cap=4096)
    assert ref.success and np.array_equal(a.counts(wa), ref.counts)
    assert np.allclose(ca, ref.colors, rtol=tol, atol=tol * float(np.abs(ref.colors).max()))
    a.close()
    b.close()


def test_chain_bad_value_limit_stops(ffr_mod):
    text = TF.expanding()
    r, ok = render(ffr_mod, text, 100_000, 1000, "fp64", -1, bad_value_limit=0)
    assert not ok and r.last_stats.n_bad >= 1
    assert r.last_stats.wave_per_walker == 1
    r.close()


def test_chain_at_walker_limit_and_cumulative(ffr_mod):
    """the most walkers the kernel takes (64 per CU), then a second render
    onto the same buffer: cumulative statistics equal the lane kernel's"""
    text = sierpinski(256)
    a = ffr_mod.BufferRenderer(text, seed=SEED, batch_size=300)
    b = ffr_mod.BufferRenderer(text, seed=SEED, batch_size=300, cell_cache=0)
    walkers = 64 * ffr_mod.DeviceMemory.cu_count()
    for rep in range(2):
        print(f"walker limit {walkers}, render {rep}", flush=True)
        assert a.render(walkers * 300) and b.render(walkers * 300)
        assert a.last_stats.walkers == walkers
        assert a.last_stats.wave_per_walker == 1
        assert_same(a, b, 2)
    assert a.render((walkers + 1) * 300)
    assert a.last_stats.wave_per_walker == 0
    a.close()
    b.close()


@pytest.mark.parametrize("name,per_cu", [("sierpinski", 64), ("uneven_weights", 64),
                                         ("three_slots", 16), ("colour", 16)])
def test_chain_cutoff_per_build(ffr_mod, name, per_cu):
    """the most walkers per CU the wave-per-walker kernel takes depends on
    the length of its serial step: 64 with identity post-affines and no
    colour, 16 with post-affines from the table or with colour"""
    text = TF.color_flame(3) if name == "colour" else FLAMES[name][0]()
    cu = ffr_mod.DeviceMemory.cu_count()
    for walkers, want in ((per_cu * cu, 1), (per_cu * cu + 1, 0)):
        r = ffr_mod.BufferRenderer(text, seed=1, batch_size=64)
        assert r.render(walkers * 64)
        assert r.last_stats.walkers == walkers
        assert r.last_stats.wave_per_walker == want, (name, walkers)
        r.close()


@pytest.mark.parametrize("case", ["colour5", "cache", "block", "five_slots", "sinusoidal"])
def test_chain_not_taken(ffr_mod, case):
    text, kw = sierpinski(64), {}
    if case == "colour5":
        d = load(text)
        d["color_dimensions"] = 5
        for x in d["xforms"]:
            x["color"] = [0.5] * 5
        text = json.dumps(d)
    elif case == "cache":
        kw["cell_cache"] = 1
    elif case == "block":
        kw["block"] = 64
    elif case == "five_slots":
        d = load(text)
        for x in d["xforms"]:
            x["variations"] = [{"name": "linear", "weight": 0.2}] * 5
        text = json.dumps(d)
    else:
        text = TF.one_d(nonlinear=True)
    r = ffr_mod.BufferRenderer(text, seed=1, batch_size=1000, **kw)
    assert r.render(50_000)
    assert r.last_stats.wave_per_walker == 0
    r.close()
