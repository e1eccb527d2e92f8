"""Closed-form pins of the parameterised variations at generic parameters.

The special-point pins (test_oracle_variations.py) fix most variations, but
for fourteen of them every pin there sits where the parameters make the map
the identity or zero, so a wrong parameter term would pass.  Here each of
them is evaluated at parameters and points where every term counts, and
compared with an independent closed form of the map, evaluated in 50-digit
arithmetic (mpmath) from the same double inputs.  The closed forms are the
maps' mathematical definitions, mostly as complex functions of z = x + iy,
not the reference's real-valued expressions:

  curl         z / (1 + c1 z + c2 z^2)
  mobius       (a z + b) / (c z + d)                  (a, b, c, d complex)
  perspective  the plane tilted by `angle` about the x axis, projected from
               `distance`: d (x, y cos t) / (d - y sin t)
  split        x takes the sign of cos(pi ysize y), y the sign of
               cos(pi xsize x)
  secant       (x, sec(w |z|) - sgn cos(w |z|))
  waves        z + (xscale sin(xfreq y), yscale sin(yfreq x))
  popcorn      z + (px sin tan(c y), py sin tan(c x))
  disc2        Arg z (exp(i pi rot (x + y)) + k (exp(i twist) - 1)),
               k = 1 + twist, less 2 pi beyond |twist| > 2 pi
  flux         q = (z - w) / (z + w): (2 + spread) |q|^(-1/2) times
               (Im, Re) of sqrt(q) / |sqrt(q)|  (the reference returns
               (sin, cos) of the half angle)
  escher       z ** ((1 + exp(i beta)) / 2), principal branch
  lazysusan    u = z - conj(p): inside |u| < w, |u| exp(i(Arg u + spin +
               twist (w - |u|))) + conj(p); outside u (1 + space/|u|) +
               conj(p)
  wedge_julia  the root |z|^(dist/power) exp(i (Arg z + 2 pi k) / power) of
               branch k = trunc(|power| u), turned by angle (floor((count a +
               pi) / 2 pi) - count a / 2 pi)
  cpow         z ** c exp(2 pi i k / power), c = (r + i i) / power,
               k = floor(power u)
  supershape   ((rnd u + (1 - rnd)|z|) - holes) rho(Arg z) z / |z| with the
               superformula rho = (|cos T|^n2 + |sin T|^n3)^(-1/n1),
               T = m Arg z / 4 + pi / 4

u is the walker stream's next number, (word >> 11) 2^-53.  Each pin also
checks that the closed form with one parameter term's sign flipped lands
far outside the tolerance, so a sign slip in the oracle fails it.
Reference formulas: src/variations/variations.hpp (Disc2 :624, Waves :660,
Perspective :943, Curl :1107, Secant :1218, Flux :1576, Mobius :1605,
Split :1655, WedgeJulia :1729, Supershape :1815, CPow :2038, Escher :2149,
LazySusan :2197, Popcorn :2286).
"""

import json

import mpmath as mp
import pytest

mp.mp.dps = 50

SEED = 31
F64_U64, F32_U32 = 0, 1


def word_u(oracle, n, seed=SEED):
    return mp.mpf(oracle.stream_word(seed, 0, n) >> 11) / mp.mpf(2) ** 53


def flame_with(name, params):
    v = {"name": name, "weight": 1.0}
    v.update(params)
    return json.dumps({"dimensions": 2, "size": [4, 4], "bounds": [[-1, 1], [-1, 1]],
                       "xforms": [{"weight": 1, "variations": [v]}]})


def Z(p):
    return mp.mpc(mp.mpf(p[0]), mp.mpf(p[1]))


def cplx(pair):
    return mp.mpc(mp.mpf(pair[0]), mp.mpf(pair[1]))


def sgn(v):
    return mp.mpf(1) if v >= 0 else mp.mpf(-1)


# ------------------------------------------------------------- closed forms
# f(params, z, u) -> complex image; `flip` names a parameter term whose sign
# the mutant turns over

def curl(p, z, u, flip=None):
    c1, c2 = mp.mpf(p["c1"]), mp.mpf(p["c2"])
    if flip == "c2":
        x, y = z.real, z.imag
        den = 1 + c1 * z + c2 * mp.mpc(x * x + y * y, 2 * x * y)   # x^2 + y^2 for x^2 - y^2
        return z / den
    return z / (1 + c1 * z + c2 * z * z)


def mobius(p, z, u, flip=None):
    a, b, c, d = (cplx(p[k]) for k in "abcd")
    if flip == "b":
        b = -b
    if flip == "a_im":
        a = mp.mpc(a.real, -a.imag)
    return (a * z + b) / (c * z + d)


def perspective(p, z, u, flip=None):
    d, t = mp.mpf(p["distance"]), mp.mpf(p["angle"])
    x, y = z.real, z.imag
    # rotate (x, y, 0) about the x axis by t, project from (0, 0, d)
    X, Y, H = x, y * mp.cos(t), y * mp.sin(t)
    if flip == "angle":
        H = -H
    s = d / (d - H)
    return mp.mpc(s * X, s * Y)


def split(p, z, u, flip=None):
    xs, ys = mp.mpf(p["xsize"]), mp.mpf(p["ysize"])
    x, y = z.real, z.imag
    sx = sgn(mp.cos(mp.pi * xs * x))
    sy = sgn(mp.cos(mp.pi * ys * y))
    if flip == "swap":
        return mp.mpc(x * sx, y * sy)
    return mp.mpc(x * sy, y * sx)


def secant(p, z, u, flip=None):
    w = mp.mpf(p["flam3_weight"])
    cr = mp.cos(w * abs(z))
    add = -sgn(cr) if flip != "sign" else sgn(cr)
    return mp.mpc(z.real, mp.sec(w * abs(z)) + add)


def waves(p, z, u, flip=None):
    x, y = z.real, z.imag
    dx = mp.mpf(p["xscale"]) * mp.sin(mp.mpf(p["xfreq"]) * y)
    dy = mp.mpf(p["yscale"]) * mp.sin(mp.mpf(p["yfreq"]) * x)
    if flip == "swap":
        dx = mp.mpf(p["xscale"]) * mp.sin(mp.mpf(p["xfreq"]) * x)
    return z + mp.mpc(dx, dy)


def popcorn(p, z, u, flip=None):
    x, y = z.real, z.imag
    c = mp.mpf(p["c"])
    dx = mp.mpf(p["x"]) * mp.sin(mp.tan(c * y))
    dy = mp.mpf(p["y"]) * mp.sin(mp.tan(c * x))
    if flip == "y":
        dy = -dy
    return z + mp.mpc(dx, dy)


def disc2(p, z, u, flip=None):
    rot, tw = mp.mpf(p["rotation"]), mp.mpf(p["twist"])
    k = 1 + tw
    if tw > 2 * mp.pi:
        k -= 2 * mp.pi
    if tw < -2 * mp.pi:
        k += 2 * mp.pi
    if flip == "twist":
        tw = -tw
    return mp.arg(z) * (mp.expj(mp.pi * rot * (z.real + z.imag)) + k * (mp.expj(tw) - 1))


def flux(p, z, u, flip=None):
    w = mp.mpf(p["flam3_weight"])
    if flip == "w":
        w = -w
    q = (z - w) / (z + w)
    e = mp.sqrt(q) / mp.sqrt(abs(q))
    return (2 + mp.mpf(p["spread"])) / mp.sqrt(abs(q)) * mp.mpc(e.imag, e.real)


def escher(p, z, u, flip=None):
    c = (1 + mp.expj(mp.mpf(p["beta"]))) / 2
    if flip == "beta":
        c = mp.conj(c)
    return mp.exp(c * mp.log(z))


def lazysusan(p, z, u, flip=None):
    pc = mp.mpc(mp.mpf(p["x"]), -mp.mpf(p["y"]))       # conj(px + i py)
    if flip == "py":
        pc = mp.conj(pc)
    w = mp.mpf(p["flam3_weight"])
    v = z - pc
    r = abs(v)
    if r < w:
        tw = mp.mpf(p["twist"])
        return r * mp.expj(mp.arg(v) + mp.mpf(p["spin"]) + tw * (w - r)) + pc
    return v * (1 + mp.mpf(p["space"]) / r) + pc


def wedge_julia(p, z, u, flip=None):
    power, dist = mp.mpf(p["power"]), mp.mpf(p["dist"])
    count, angle = mp.mpf(p["count"]), mp.mpf(p["angle"])
    k = int(abs(power) * u)
    a = (mp.arg(z) + 2 * mp.pi * k) / power
    root = mp.exp(dist / power * mp.log(abs(z)) + 1j * a)
    sector = mp.floor((count * a + mp.pi) / (2 * mp.pi))
    turn = angle * (sector - count * a / (2 * mp.pi))
    if flip == "angle":
        turn = -turn
    return root * mp.expj(turn)


def cpow(p, z, u, flip=None):
    power = mp.mpf(p["power"])
    c = mp.mpc(mp.mpf(p["r"]), mp.mpf(p["i"])) / power
    if flip == "i":
        c = mp.conj(c)
    k = mp.floor(power * u)
    return mp.exp(c * mp.log(z)) * mp.expj(2 * mp.pi * k / power)


def supershape(p, z, u, flip=None):
    n1, m, n2, n3 = (mp.mpf(p[k]) for k in ("n1", "m", "n2", "n3"))
    rnd, holes = mp.mpf(p["rnd"]), mp.mpf(p["holes"])
    T = m * mp.arg(z) / 4 + mp.pi / 4
    rho = (abs(mp.cos(T)) ** n2 + abs(mp.sin(T)) ** n3) ** (-1 / n1)
    if flip == "holes":
        holes = -holes
    return ((rnd * u + (1 - rnd) * abs(z)) - holes) * rho * z / abs(z)


# name, params, points, closed form, mutants, draws one word
CASES = [
    ("curl", {"c1": 0.3, "c2": -0.2}, [(0.7, 0.4), (-1.1, 0.6)], curl, ["c2"], False),
    ("mobius", {"a": [0.8, -0.3], "b": [0.2, 0.5], "c": [0.4, 0.1], "d": [1.1, -0.2]},
     [(0.7, 0.4), (-0.6, -0.9)], mobius, ["b", "a_im"], False),
    ("perspective", {"distance": 2.5, "angle": 0.6}, [(0.7, 0.4), (-0.3, -1.2)], perspective,
     ["angle"], False),
    ("split", {"xsize": 0.6, "ysize": 0.7}, [(0.9, 0.5), (0.2, -1.3)], split, ["swap"], False),
    ("secant", {"flam3_weight": 1.3}, [(0.7, 0.4), (1.5, 1.0)], secant, ["sign"], False),
    ("waves", {"xfreq": 2.0, "xscale": 0.3, "yfreq": 1.5, "yscale": 0.2},
     [(0.7, 0.4), (-0.5, 1.1)], waves, ["swap"], False),
    ("popcorn", {"x": 0.3, "y": 0.2, "c": 0.9}, [(0.7, 0.4), (-0.5, 0.8)], popcorn, ["y"],
     False),
    ("disc2", {"rotation": 0.7, "twist": 0.9}, [(0.7, 0.4), (-0.5, -0.2)], disc2, ["twist"],
     False),
    ("disc2", {"rotation": -0.4, "twist": 7.0}, [(0.3, 0.9)], disc2, ["twist"], False),
    ("flux", {"spread": 0.4, "flam3_weight": 0.6}, [(0.7, 0.4), (-0.3, 1.2)], flux, ["w"],
     False),
    ("escher", {"beta": 0.8}, [(0.7, 0.4), (-0.5, 0.3)], escher, ["beta"], False),
    # inside the disc (|u| < 2) and outside it
    ("lazysusan", {"x": 0.3, "y": -0.2, "spin": 0.5, "twist": 0.7, "space": 0.4,
                   "flam3_weight": 2.0}, [(0.7, 0.4), (3.0, 2.5)], lazysusan, ["py"], False),
    ("wedge_julia", {"angle": 0.4, "count": 3.0, "power": 3.0, "dist": 1.5},
     [(0.7, 0.4), (-0.5, -0.6)], wedge_julia, ["angle"], True),
    ("cpow", {"r": 1.3, "i": 0.4, "power": 3.0}, [(0.7, 0.4), (-0.5, 0.6)], cpow, ["i"], True),
    ("supershape", {"n1": 2.5, "m": 5.0, "n2": 1.7, "n3": 3.2, "rnd": 0.3, "holes": 0.1},
     [(0.7, 0.4), (-0.5, -0.6)], supershape, ["holes"], True),
]

IDS = [f"{c[0]}-{i}" for i, c in enumerate(CASES)]


@pytest.mark.parametrize("name,params,points,form,mutants,draws", CASES, ids=IDS)
def test_closed_form_fp64(oracle_mod, name, params, points, form, mutants, draws):
    f = oracle_mod.OracleFlame(flame_with(name, params), F64_U64)
    for draw in (0, 1, 2) if draws else (0,):
        u = word_u(oracle_mod, draw)
        for pt in points:
            out, used = f.eval_variation(0, 0, pt, seed=SEED, draw=draw)
            assert used == draw + (1 if draws else 0)
            want = form(params, Z(pt), u)
            scale = max(mp.mpf(1), abs(want))
            err = abs(mp.mpc(out[0], out[1]) - want) / scale
            # a few libm calls in double: well inside 1e-13 relative
            assert err <= 1e-13, (name, pt, out, complex(want), float(err))
            # the mutant (one parameter term's sign turned over) is far away
            for m in mutants:
                bad = form(params, Z(pt), u, flip=m)
                assert abs(bad - want) / scale > 1e-3, (name, pt, m)


@pytest.mark.parametrize("name,params,points,form,mutants,draws", CASES, ids=IDS)
def test_closed_form_float_build(oracle_mod, name, params, points, form, mutants, draws):
    """the float build (num_t = float): the same maps within float rounding
    (the oracle rounds the point and the precomputed parameters to float,
    so the closed form is taken at those rounded values)"""
    import numpy as np
    f = oracle_mod.OracleFlame(flame_with(name, params), F32_U32)
    for draw in (0, 1) if draws else (0,):
        # randNum in the float build: (w >> 40) 2^-24
        u = mp.mpf(oracle_mod.stream_word(SEED, 0, draw) >> 40) / mp.mpf(2) ** 24
        for pt in points:
            pt32 = tuple(float(np.float32(v)) for v in pt)
            out, used = f.eval_variation(0, 0, pt32, seed=SEED, draw=draw)
            want = form(params, Z(pt32), u)
            scale = max(mp.mpf(1), abs(want))
            err = abs(mp.mpc(out[0], out[1]) - want) / scale
            assert err <= 2e-5, (name, pt, out, complex(want), float(err))
