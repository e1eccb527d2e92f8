"""Pins of the oracle's variations at points where the mathematics fixes the
value (closed forms, symmetries, exact arithmetic), and for variations that
draw random words, at images fixed by the words the stream holds.  Every
one of the 98 variations has at least one such pin.  Values involving libm
functions are compared with Python's math module (the same C library) to
a few units in the last place.  The reference renderer itself cannot be
built in this image; formulas are from src/variations/variations.hpp.
"""

import json
import math

import pytest

S1, C1 = math.sin(1.0), math.cos(1.0)
PI = math.pi

CASES = [
    # name, params, point, expected
    ("linear", {}, (0.3, -0.7), (0.3, -0.7)),
    ("sinusoidal", {}, (PI / 2, 0.0), (1.0, 0.0)),
    ("spherical", {}, (3.0, 4.0), (3.0 / 25.0, 4.0 / 25.0)),
    ("spherical", {}, (2.0, 0.0), (0.5, 0.0)),
    ("swirl", {}, (math.sqrt(PI / 2), 0.0),
     (math.sqrt(PI / 2) * math.sin(PI / 2), math.sqrt(PI / 2) * math.cos(PI / 2))),
    ("horseshoe", {}, (1.0, 1.0), (0.0, math.sqrt(2.0))),
    ("polar", {}, (0.0, 2.0), (0.5, 1.0)),
    ("polar", {}, (-1.0, 0.0), (1.0, 0.0)),
    ("polar2", {}, (0.0, 1.0), (PI / 2, 0.0)),
    ("handkerchief", {}, (1.0, 0.0), (S1, C1)),
    ("heart", {}, (1.0, 0.0), (0.0, -1.0)),
    ("disc", {}, (1.0, 0.0), (0.0, 0.0)),
    ("spiral", {}, (1.0, 0.0), (1.0 + S1, -C1)),
    ("hyperbolic", {}, (1.0, 0.0), (0.0, 1.0)),
    ("diamond", {}, (1.0, 0.0), (0.0, S1)),
    ("ex", {}, (1.0, 0.0), (S1 ** 3 + C1 ** 3, S1 ** 3 - C1 ** 3)),
    ("exponential", {}, (1.0, 0.0), (1.0, 0.0)),
    ("power", {}, (0.0, 2.0), (0.0, 2.0)),
    ("power", {}, (1.0, 0.0), (1.0, 0.0)),
    ("cosine", {}, (0.0, 0.0), (1.0, 0.0)),
    ("cylinder", {}, (PI / 2, 3.0), (1.0, 3.0)),
    ("tangent", {}, (PI / 2, 0.0), (1.0, 0.0)),
    ("cross", {}, (2.0, 1.0), (2.0 / 3.0, 1.0 / 3.0)),
    ("exp", {}, (0.0, PI / 2), (math.cos(PI / 2), 1.0)),
    ("log", {}, (0.0, 1.0), (0.0, PI / 2)),
    ("sin", {}, (PI / 2, 0.0), (1.0, 0.0)),
    ("cos", {}, (0.0, 0.0), (1.0, 0.0)),
    ("tan", {}, (PI / 4, 0.0), (1.0, 0.0)),
    ("sec", {}, (0.0, 0.0), (0.5, 0.0)),
    ("csc", {}, (PI / 2, 0.0), (0.5, 0.0)),
    ("cot", {}, (PI / 4, 0.0), (1.0, 0.0)),
    ("sinh", {}, (0.0, PI / 2), (0.0, 1.0)),
    ("cosh", {}, (0.0, 0.0), (1.0, 0.0)),
    ("tanh", {}, (0.0, PI / 4), (0.0, 1.0)),
    ("sech", {}, (0.0, 0.0), (0.5, 0.0)),
    ("coth", {}, (1.0, 0.0), (1.0 / math.tanh(1.0), 0.0)),
    ("csch", {}, (1.0, 0.0), (1.0 / (2.0 * math.sinh(1.0)), 0.0)),
    ("butterfly", {}, (1.0, 1.0), (1 / math.sqrt(5.0), 2 / math.sqrt(5.0))),
    ("fisheye", {"addval": 1.0}, (3.0, 4.0), (0.5, 4.0 / 6.0)),
    ("bubble", {"addval": 4.0}, (1.0, 0.0), (0.2, 0.0)),
    ("bent", {"scales_neg": [2.0, 0.5], "scales_pos": [1.0, 1.2]}, (-1.0, 1.0), (-2.0, 1.2)),
    ("rectangles", {"params": [0.3, 0.0]}, (0.1, 0.7), (0.3 - 0.1, 0.7)),
    ("unit_sphere", {}, (3.0, 4.0), (0.6, 0.8)),
    ("unit_cube", {}, (3.0, -4.0), (0.75, -1.0)),
    ("spherical_p", {"norm": 2.0}, (3.0, 4.0), (3.0 / 25.0, 4.0 / 25.0)),
    ("unit_sphere_p", {"norm": 2.0}, (3.0, 4.0), (0.6, 0.8)),
    ("modulus", {"params": [0.7, 0.7]}, (1.0, 0.2), (1.0 - 1.4, 0.2)),
    ("splits", {"params": [0.2, 0.3]}, (0.5, -0.5), (0.7, -0.8)),
    ("stripes", {"space": 0.3, "warp": 0.8}, (0.3, 1.0), (0.3 * 0.7, 1.0 + 0.09 * 0.8)),
    ("waves", {"xfreq": 2.0, "xscale": 0.3, "yfreq": 1.5, "yscale": 0.2}, (0.0, 0.0), (0.0, 0.0)),
    ("popcorn", {"x": 0.3, "y": 0.2, "c": 0.9}, (0.0, 0.0), (0.0, 0.0)),
    ("curl", {"c1": 0.0, "c2": 0.0}, (0.3, -0.4), (0.3, -0.4)),
    ("mobius", {"a": [1, 0], "b": [0, 0], "c": [0, 0], "d": [1, 0]}, (0.3, -0.4), (0.3, -0.4)),
    ("perspective", {"distance": 2.0, "angle": 0.0}, (0.3, -0.4), (0.3, -0.4)),
    ("foci", {}, (0.0, PI), (0.0, math.sin(PI) / 2)),
    ("split", {"xsize": 0.0, "ysize": 0.0}, (0.3, -0.4), (0.3, -0.4)),
    ("secant", {"flam3_weight": 0.0}, (0.3, -0.4), (0.3, 1.0 - 1.0)),
    # s * (sqrt(x^2 + sep^2) - s * inside), s = sign(x): sqrt(0.5625 + 1) = 1.25
    ("separation", {"params": [1.0, 1.0], "inside": [0.25, 0.5]}, (0.75, -0.75),
     (1.0, -1.75)),
    # cell index floor(x / size); the cell moves to 2k (k >= 0) or -2k - 1
    ("celln", {"sizes": [0.5, 0.5]}, (1.25, -0.75), (2.25, 1.75)),
    ("cell", {"size": 0.5}, (1.25, -0.75), (2.25, -1.75)),
    # no rotation, no twist: angle(t) * (1, 0)
    ("disc2", {"rotation": 0.0, "twist": 0.0}, (0.0, 1.0), (PI / 2, 0.0)),
    # dx = pi, a = 0 + pi/2 (fmod(0, pi) = 0 < pi/2): |t| (cos, sin)(pi/2)
    ("fan", {"x": 1.0, "y": 0.0}, (1.0, 0.0), (math.cos(PI / 2), 1.0)),
    # dx = 1: r = fmod(3 + 1, 2) - 1 + 3 * 0 = -1
    ("rings", {"value": 1.0}, (3.0, 0.0), (-1.0, 0.0)),
    # r * (mid + amp sin(waves * pi/2)) = 2 * 1.5 along the y axis
    ("blob", {"low": 0.5, "high": 1.5, "waves": 1.0}, (0.0, 2.0), (0.0, 3.0)),
    # (sin(a y) - cos(b x), sin(c x) - cos(d y)) at the origin
    ("pdj", {"a": 1.1, "b": 0.9, "c": -0.7, "d": 1.3}, (0.0, 0.0), (-1.0, -1.0)),
    ("pdj", {"a": 1.0, "b": 1.0, "c": 1.0, "d": 1.0}, (PI / 2, 0.0),
     (-math.cos(PI / 2), 0.0)),
    # corners 0, circle 1, power 2: t / |t|^2
    ("ngon", {"sides": 5.0, "power": 2.0, "corners": 0.0, "circle": 1.0}, (3.0, 4.0),
     (3.0 / 25.0, 4.0 / 25.0)),
    # power 0, corners 1, square: 1/cos(pi/4) - 1 at 45 degrees, 0 on an axis
    ("ngon", {"sides": 4.0, "power": 0.0, "corners": 1.0, "circle": 0.0}, (1.0, 1.0),
     (math.sqrt(2.0) - 1.0, math.sqrt(2.0) - 1.0)),
    ("ngon", {"sides": 4.0, "power": 0.0, "corners": 1.0, "circle": 0.0}, (1.0, 0.0),
     (0.0, 0.0)),
    # x + sym (dx - x) with sin(pi/2) = 1: pi/2 + (1 + pi/2)
    ("auger", {"freq": 1.0, "flam3_weight": 1.0, "scale": 2.0, "sym": 1.0},
     (PI / 2, PI / 2), (1.0 + PI, 1.0 + PI)),
    # weight 0: ratio 1, half-angle 0 -> (sin 0, cos 0) * (2 + spread)
    ("flux", {"spread": 0.0, "flam3_weight": 0.0}, (0.3, 0.4), (0.0, 2.0)),
    # 1 / (|t| (|t|^2 + 1/w)) with w = 1: 1/(5 * 26)
    ("scry", {"flam3_weight": 1.0}, (3.0, 4.0), (3.0 / 130.0, 4.0 / 130.0)),
    # no wedges, no swirl: (|t| + hole) in the direction of t
    ("wedge", {"swirl": 0.0, "count": 3.0, "angle": 0.0, "hole": 0.5}, (0.6, 0.8),
     (0.9, 1.2)),
    ("wedge_sph", {"angle": 0.0, "count": 4.0, "swirl": 0.0, "hole": 0.0}, (3.0, 4.0),
     (0.12, 0.16)),
    # inside (r < w): angle + 0; outside: angle + pi/4 / (0.5 - 1) = angle - pi/2
    ("whorl", {"inside": 0.0, "outside": 1.0, "flam3_weight": 2.0}, (0.6, 0.8), (0.6, 0.8)),
    ("whorl", {"inside": 0.0, "outside": PI / 4, "flam3_weight": 0.5}, (0.6, 0.8),
     (0.8, -0.6)),
    # origin: (0 + 1) / (1 - 0), atan2(0, -1) / 2 = pi/2 wraps to -pi/2
    ("bipolar", {"shift": 0.0}, (0.0, 0.0), (0.0, -PI / 2)),
    # at the origin the offsets are the amplitudes
    ("curve", {"xamp": 0.4, "yamp": 0.6, "xlen": 1.2, "ylen": 0.8}, (0.0, 0.0), (0.4, 0.6)),
    # xmax = 1, a1 = 0, a2 = -pi/2: (cosh(pi/2), -0 * sinh(-pi/2))
    ("edisc", {}, (0.0, 0.0), (math.cosh(PI / 2), 0.0)),
    # xmax = sqrt(2), a = 0: (atan2(0, 1), log(sqrt(2) + sqrt(sqrt(2) - 1)))
    ("elliptic", {}, (0.0, 1.0),
     (0.0, math.log(math.sqrt(2.0) + math.sqrt(math.sqrt(2.0) - 1.0)))),
    # beta 0: exp(ln |t|) (cos, sin)(angle) = t
    ("escher", {"beta": 0.0}, (0.6, 0.8), (0.6, 0.8)),
    # outside the disc: (1 + space/|t|) t; inside with no spin or twist: t
    ("lazysusan", {"x": 0.0, "y": 0.0, "spin": 0.0, "twist": 0.0, "space": 5.0,
                   "flam3_weight": 0.0}, (3.0, 4.0), (6.0, 8.0)),
    ("lazysusan", {"x": 0.0, "y": 0.0, "spin": 0.0, "twist": 0.0, "space": 5.0,
                   "flam3_weight": 10.0}, (3.0, 4.0), (3.0, 4.0)),
    # |t|^2 = 0.25 < w^2 = 1: sqrt(1/0.25 - 1) t; outside: w t
    ("loonie", {"flam3_weight": 1.0}, (0.3, 0.4), (math.sqrt(3.0) * 0.3,
                                                   math.sqrt(3.0) * 0.4)),
    ("loonie", {"flam3_weight": 1.0}, (3.0, 4.0), (3.0, 4.0)),
    # x = 0: t = amplitude + separation = 0.8; |y| below it flips y
    ("oscope", {"frequency": 1.5, "amplitude": 0.6, "damping": 0.3, "separation": 0.2},
     (0.0, 0.5), (0.0, -0.5)),
    ("oscope", {"frequency": 1.5, "amplitude": 0.6, "damping": 0.3, "separation": 0.2},
     (0.0, 2.0), (0.0, 2.0)),
]


def flame_with(name, params, dims=2):
    v = {"name": name, "weight": 1.0}
    v.update(params)
    return json.dumps({"dimensions": dims, "size": [4] * dims, "bounds": [[-1, 1]] * dims,
                       "xforms": [{"weight": 1, "variations": [v]}]})


def close(a, b, tol=4e-16):
    return abs(a - b) <= tol * max(1.0, abs(b)) or (a == b)


@pytest.mark.parametrize("name,params,point,expected", CASES)
def test_special_point(oracle_mod, name, params, point, expected):
    f = oracle_mod.OracleFlame(flame_with(name, params))
    out, _ = f.eval_variation(0, 0, point)
    # a few libm calls in a row: within a few units in the last place
    assert all(close(o, e, 2e-15) for o, e in zip(out, expected)), (out, expected)


def test_every_variation_name_parses(oracle_mod):
    import flames as TF
    import oracle.model as M
    for name, _ in M.VARIATIONS:
        f = oracle_mod.OracleFlame(TF.single_variation_flame(name, 2))
        out, _ = f.eval_variation(0, 0, (0.4, -0.3))
        assert len(out) == 2


def _word_u(oracle_mod, seed, n):
    return (oracle_mod.stream_word(seed, 0, n) >> 11) / 2.0 ** 53


def test_blur_radius_is_the_first_draw(oracle_mod):
    f = oracle_mod.OracleFlame(flame_with("blur", {}))
    for d in range(5):
        out, used = f.eval_variation(0, 0, (0.3, 0.1), seed=11, draw=d)
        r = _word_u(oracle_mod, 11, d)
        a = 2.0 * PI * _word_u(oracle_mod, 11, d + 1)
        assert used == d + 2
        assert close(out[0], math.cos(a) * r) and close(out[1], math.sin(a) * r)


def test_square_noise_draws(oracle_mod):
    f = oracle_mod.OracleFlame(flame_with("square_noise", {}, dims=3))
    out, used = f.eval_variation(0, 0, (0.0, 0.0, 0.0), seed=12, draw=3)
    assert used == 6
    assert out == [_word_u(oracle_mod, 12, 3 + i) - 0.5 for i in range(3)]


def test_julia_branch_is_the_low_bit(oracle_mod):
    """The reference scales by |t| (variations.hpp Julia::calc2d uses
    tx.norm2()), with the half-angle shifted by pi when the word's low bit
    is set."""
    f = oracle_mod.OracleFlame(flame_with("julia", {}))
    for d in range(8):
        out, used = f.eval_variation(0, 0, (4.0, 0.0), seed=13, draw=d)
        bit = oracle_mod.stream_word(13, 0, d) & 1
        assert used == d + 1
        assert close(out[0], -4.0 if bit else 4.0) and abs(out[1]) < 1e-14


def test_draw_counts(oracle_mod):
    """Words one evaluation takes (randDirection<d>: 1 word for d <= 2, 2 for
    d = 3; Box-Muller: 2 words)."""
    import flames as TF
    want2 = {"noise": 2, "blur": 2, "gaussian_blur": 3, "pre_blur": 3, "square_noise": 2,
             "julia": 1, "julian": 1, "juliascope": 2, "radial_blur": 2, "pie": 3,
             "arch": 1, "rays": 1, "blade": 1, "twintrian": 1, "wedge_julia": 1,
             "supershape": 1, "flower": 1, "conic": 1, "parabola": 2, "boarders": 1,
             "cpow": 1}
    import oracle.model as M
    for name, _ in M.VARIATIONS:
        f = oracle_mod.OracleFlame(TF.single_variation_flame(name, 2))
        _, used = f.eval_variation(0, 0, (0.4, -0.3), draw=0)
        assert used == want2.get(name, 0), name
    for name, n3 in (("noise", 3), ("blur", 3), ("gaussian_blur", 4), ("square_noise", 3)):
        f = oracle_mod.OracleFlame(TF.single_variation_flame(name, 3))
        _, used = f.eval_variation(0, 0, (0.4, -0.3, 0.2), draw=0)
        assert used == n3, name


# ---------------------------------------------------------------- draws
# Variations that draw words: the expected image follows from the words the
# stream holds at the evaluation's draw position (flame_rng.hpp:60-206).


def _gauss(oracle_mod, seed, d):
    """randGaussian from words d, d+1 (Box-Muller, first output)."""
    u1 = _word_u(oracle_mod, seed, d)
    u2 = 2.0 * PI * _word_u(oracle_mod, seed, d + 1)
    return math.sqrt(-2.0 * math.log(u1)) * math.cos(u2)


def _dir2(oracle_mod, seed, d):
    a = 2.0 * PI * _word_u(oracle_mod, seed, d)
    return math.cos(a), math.sin(a)


def test_noise_scales_by_draws(oracle_mod):
    f = oracle_mod.OracleFlame(flame_with("noise", {}))
    for d in range(4):
        out, used = f.eval_variation(0, 0, (0.3, -0.2), seed=21, draw=d)
        r = _word_u(oracle_mod, 21, d)
        cx, sy = _dir2(oracle_mod, 21, d + 1)
        assert used == d + 2
        assert close(out[0], r * 0.3 * cx, 2e-15) and close(out[1], r * -0.2 * sy, 2e-15)


@pytest.mark.parametrize("name", ["gaussian_blur", "pre_blur"])
def test_gaussian_blur_draws(oracle_mod, name):
    f = oracle_mod.OracleFlame(flame_with(name, {}))
    for d in range(3):
        out, used = f.eval_variation(0, 0, (0.3, -0.2), seed=22, draw=d)
        g = _gauss(oracle_mod, 22, d)
        cx, sy = _dir2(oracle_mod, 22, d + 2)
        assert used == d + 3
        assert close(out[0], g * cx, 2e-15) and close(out[1], g * sy, 2e-15)


def test_julian_power_one_is_identity_and_two_halves(oracle_mod):
    f = oracle_mod.OracleFlame(flame_with("julian", {"power": 1.0, "dist": 1.0}))
    out, used = f.eval_variation(0, 0, (0.6, 0.8), seed=23, draw=0)
    assert used == 1 and close(out[0], 0.6, 2e-15) and close(out[1], 0.8, 2e-15)
    # power 2, dist 1 on the positive x axis: r = (|t|^2)^(1/4) = 4, the
    # angle 0 or pi by the draw
    f = oracle_mod.OracleFlame(flame_with("julian", {"power": 2.0, "dist": 1.0}))
    for d in range(6):
        out, _ = f.eval_variation(0, 0, (16.0, 0.0), seed=23, draw=d)
        t = int(2.0 * _word_u(oracle_mod, 23, d))
        assert close(out[0], -4.0 if t else 4.0, 2e-15) and abs(out[1]) < 1e-14


def test_juliascope_random_mirror(oracle_mod):
    f = oracle_mod.OracleFlame(flame_with("juliascope", {"power": 1.0, "dist": 1.0}))
    for d in range(6):
        out, used = f.eval_variation(0, 0, (0.6, 0.8), seed=24, draw=d)
        sign = math.copysign(1.0, _word_u(oracle_mod, 24, d + 1) - 0.5)
        assert used == d + 2
        assert close(out[0], 0.6, 2e-15) and close(out[1], sign * 0.8, 2e-15)


def test_radial_blur_is_gaussian_zoom(oracle_mod):
    # angle 0: spin 0, zoom 1, so the image is (w * g) t
    f = oracle_mod.OracleFlame(flame_with("radial_blur", {"angle": 0.0, "flam3_weight": 0.5}))
    for d in range(3):
        out, used = f.eval_variation(0, 0, (0.6, 0.8), seed=25, draw=d)
        g = 0.5 * _gauss(oracle_mod, 25, d)
        assert used == d + 2
        assert close(out[0], g * 0.6, 4e-15) and close(out[1], g * 0.8, 4e-15)


def test_pie_slices(oracle_mod):
    f = oracle_mod.OracleFlame(flame_with("pie", {"slices": 6.0, "rotation": 0.0,
                                                  "thickness": 0.0}))
    for d in range(4):
        out, used = f.eval_variation(0, 0, (0.6, 0.8), seed=26, draw=d)
        sl = int(_word_u(oracle_mod, 26, d) * 6.0 + 0.5)
        r = _word_u(oracle_mod, 26, d + 2)
        a = sl * (2.0 * PI / 6.0)
        assert used == d + 3
        assert close(out[0], r * math.cos(a), 2e-15) and close(out[1], r * math.sin(a), 2e-15)


def test_arch_rays_blade_twintrian(oracle_mod):
    seed = 27
    u = _word_u(oracle_mod, seed, 0)
    # arch: a = w u pi -> (sin a, sin^2 a / cos a)
    f = oracle_mod.OracleFlame(flame_with("arch", {"flam3_weight": 0.7}))
    out, used = f.eval_variation(0, 0, (0.3, 0.1), seed=seed, draw=0)
    a = 0.7 * u * PI
    assert used == 1 and close(out[0], math.sin(a), 2e-15)
    assert close(out[1], math.sin(a) ** 2 / math.cos(a), 1e-14)
    # rays on (1, 0): tan(w u pi) * w / (1 + eps) * (cos 1, sin 0)
    f = oracle_mod.OracleFlame(flame_with("rays", {"flam3_weight": 0.7}))
    out, _ = f.eval_variation(0, 0, (1.0, 0.0), seed=seed, draw=0)
    tr = math.tan(0.7 * u * PI) * 0.7
    assert close(out[0], tr * C1, 1e-14) and out[1] == 0.0
    # weight 0: blade gives x (1, 1); twintrian's log10(0) is a bad value, -30
    f = oracle_mod.OracleFlame(flame_with("blade", {"flam3_weight": 0.0}))
    out, used = f.eval_variation(0, 0, (0.7, 0.2), seed=seed, draw=0)
    assert used == 1 and out == [0.7, 0.7]
    f = oracle_mod.OracleFlame(flame_with("twintrian", {"flam3_weight": 0.0}))
    out, used = f.eval_variation(0, 0, (0.5, 0.1), seed=seed, draw=0)
    assert used == 1 and out == [-15.0, -15.0]


def test_wedge_julia_and_cpow_identities(oracle_mod):
    # power 1, dist 1, no wedges: the identity (one draw, unused)
    for name, params in (("wedge_julia", {"angle": 0.0, "count": 2.0, "power": 1.0,
                                          "dist": 1.0}),
                         ("cpow", {"r": 1.0, "i": 0.0, "power": 1.0})):
        f = oracle_mod.OracleFlame(flame_with(name, params))
        out, used = f.eval_variation(0, 0, (0.6, 0.8), seed=28, draw=0)
        assert used == 1 and close(out[0], 0.6, 2e-15) and close(out[1], 0.8, 2e-15), name


def test_supershape_flower_conic_parabola_boarders(oracle_mod):
    seed = 29
    u0 = _word_u(oracle_mod, seed, 0)
    u1 = _word_u(oracle_mod, seed, 1)
    # supershape m 0, n2 = n3 = 2, rnd 0, holes 0: cos^2 + sin^2 = 1, identity
    f = oracle_mod.OracleFlame(flame_with("supershape", {"n1": 2.0, "m": 0.0, "n2": 2.0,
                                                         "n3": 2.0, "rnd": 0.0,
                                                         "holes": 0.0}))
    out, used = f.eval_variation(0, 0, (0.3, -0.4), seed=seed, draw=0)
    assert used == 1 and close(out[0], 0.3, 2e-15) and close(out[1], -0.4, 2e-15)
    # flower with 0 petals: u t / |t|
    f = oracle_mod.OracleFlame(flame_with("flower", {"petals": 0.0, "holes": 0.0}))
    out, used = f.eval_variation(0, 0, (3.0, 4.0), seed=seed, draw=0)
    assert used == 1 and close(out[0], u0 * 0.6, 2e-15) and close(out[1], u0 * 0.8, 2e-15)
    # conic on the y axis: (0, (u - holes) eccen)
    f = oracle_mod.OracleFlame(flame_with("conic", {"eccen": 0.5, "holes": 0.0}))
    out, used = f.eval_variation(0, 0, (0.0, 2.0), seed=seed, draw=0)
    assert used == 1 and out[0] == 0.0 and close(out[1], 0.5 * u0, 2e-15)
    # parabola at the origin: (h sin^2 0 u0, w cos 0 u1) = (0, w u1)
    f = oracle_mod.OracleFlame(flame_with("parabola", {"height": 0.8, "width": 1.2}))
    out, used = f.eval_variation(0, 0, (0.0, 0.0), seed=seed, draw=0)
    assert used == 2 and out[0] == 0.0 and close(out[1], 1.2 * u1, 2e-15)
    # boarders with probability 0: always the half-way point to the lattice
    f = oracle_mod.OracleFlame(flame_with("boarders", {"prob": 0.0}))
    out, used = f.eval_variation(0, 0, (1.25, -0.75), seed=seed, draw=0)
    assert used == 1 and out == [1.125, -0.875]
