"""Pins of the oracle's variations at points where the mathematics fixes the
value (closed forms, symmetries, exact arithmetic).  Values involving libm
functions are compared with Python's math module (the same C library) to
a few units in the last place.  Variations not listed here are pinned only
through the GPU-vs-oracle agreement (tests/test_gpu_parity.py); the
reference renderer itself cannot be built in this image.
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
]


def flame_with(name, params, dims=2):
    v = {"name": name, "weight": 1.0}
    v.update(params)
    return json.dumps({"dimensions": dims, "size": [4] * dims, "bounds": [[-1, 1]] * dims,
                       "xforms": [{"weight": 1, "variations": [v]}]})


def close(a, b):
    return abs(a - b) <= 4e-16 * max(1.0, abs(b)) or (a == b)


@pytest.mark.parametrize("name,params,point,expected", CASES)
def test_special_point(oracle_mod, name, params, point, expected):
    f = oracle_mod.OracleFlame(flame_with(name, params))
    out, _ = f.eval_variation(0, 0, point)
    assert all(close(o, e) for o, e in zip(out, expected)), (out, expected)


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
