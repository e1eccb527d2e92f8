"""Flame JSON -> raw flame description for the CPU oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product.

This is an independent restatement of how the reference reads a flame file:
  - Json wrapper: nlohmann::json::parse with comments allowed
    (src/utils/json.cpp:15-23), typed accessors (json.cpp:60-185)
  - Flame(Json)   src/types/flame.hpp:91-210
  - XForm(Json)   src/types/xform.hpp:71-172
  - Affine(Json)  src/types/affine.hpp:45-92
  - Point(Json)   src/types/point.hpp:29-47
  - variation constructors and factory  src/variations/variations.hpp
It produces the raw (file-order, double precision) description; the C oracle
does the num_t conversion, zero-weight removal, sorting and cumulative
weights, exactly where the reference does them.

Error messages repeat the reference's exception texts, including its
wrapping ("Flame(): error parsing xforms[0]: XForm(): ...").
"""

from __future__ import annotations

import json
import math
from dataclasses import dataclass, field

# Catalogue: (name, [(json key, kind)]) in the reference's factory order,
# N-D variations first (ids 0-19), then the 2-D ones (ids 20-97).
# kind: "s" scalar, "p" point of flame dims, "p2" 2-D point.
VARIATIONS_ND = [
    ("linear", []),
    ("sinusoidal", []),
    ("spherical", []),
    ("bent", [("scales_neg", "p"), ("scales_pos", "p")]),
    ("rectangles", [("params", "p")]),
    ("fisheye", [("addval", "s")]),
    ("bubble", [("addval", "s")]),
    ("noise", []),
    ("blur", []),
    ("gaussian_blur", []),
    ("square_noise", []),
    ("separation", [("params", "p"), ("inside", "p")]),
    ("splits", [("params", "p")]),
    ("pre_blur", []),
    ("modulus", [("params", "p")]),
    ("celln", [("sizes", "p")]),
    ("spherical_p", [("norm", "s")]),
    ("unit_sphere", []),
    ("unit_sphere_p", [("norm", "s")]),
    ("unit_cube", []),
]
VARIATIONS_2D = [
    ("swirl", []),
    ("horseshoe", []),
    ("polar", []),
    ("polar2", []),
    ("handkerchief", []),
    ("heart", []),
    ("disc", []),
    ("disc2", [("rotation", "s"), ("twist", "s")]),
    ("waves", [("xfreq", "s"), ("xscale", "s"), ("yfreq", "s"), ("yscale", "s")]),
    ("fan", [("x", "s"), ("y", "s")]),
    ("rings", [("value", "s")]),
    ("spiral", []),
    ("hyperbolic", []),
    ("diamond", []),
    ("ex", []),
    ("julia", []),
    ("exponential", []),
    ("power", []),
    ("cosine", []),
    ("blob", [("low", "s"), ("high", "s"), ("waves", "s")]),
    ("pdj", [("a", "s"), ("b", "s"), ("c", "s"), ("d", "s")]),
    ("cylinder", []),
    ("perspective", [("distance", "s"), ("angle", "s")]),
    ("julian", [("power", "s"), ("dist", "s")]),
    ("juliascope", [("power", "s"), ("dist", "s")]),
    ("radial_blur", [("angle", "s"), ("flam3_weight", "s")]),
    ("pie", [("slices", "s"), ("rotation", "s"), ("thickness", "s")]),
    ("ngon", [("sides", "s"), ("power", "s"), ("corners", "s"), ("circle", "s")]),
    ("curl", [("c1", "s"), ("c2", "s")]),
    ("arch", [("flam3_weight", "s")]),
    ("tangent", []),
    ("rays", [("flam3_weight", "s")]),
    ("blade", [("flam3_weight", "s")]),
    ("secant", [("flam3_weight", "s")]),
    ("twintrian", [("flam3_weight", "s")]),
    ("cross", []),
    ("exp", []),
    ("log", []),
    ("sin", []),
    ("cos", []),
    ("tan", []),
    ("sec", []),
    ("csc", []),
    ("cot", []),
    ("sinh", []),
    ("cosh", []),
    ("tanh", []),
    ("sech", []),
    ("csch", []),
    ("coth", []),
    ("auger", [("freq", "s"), ("flam3_weight", "s"), ("scale", "s"), ("sym", "s")]),
    ("flux", [("spread", "s"), ("flam3_weight", "s")]),
    ("mobius", [("a", "p2"), ("b", "p2"), ("c", "p2"), ("d", "p2")]),
    ("scry", [("flam3_weight", "s")]),
    ("split", [("xsize", "s"), ("ysize", "s")]),
    ("stripes", [("space", "s"), ("warp", "s")]),
    ("wedge", [("swirl", "s"), ("count", "s"), ("angle", "s"), ("hole", "s")]),
    ("wedge_julia", [("angle", "s"), ("count", "s"), ("power", "s"), ("dist", "s")]),
    ("wedge_sph", [("angle", "s"), ("count", "s"), ("swirl", "s"), ("hole", "s")]),
    ("whorl", [("inside", "s"), ("outside", "s"), ("flam3_weight", "s")]),
    ("supershape", [("n1", "s"), ("m", "s"), ("n2", "s"), ("n3", "s"),
                    ("rnd", "s"), ("holes", "s")]),
    ("flower", [("petals", "s"), ("holes", "s")]),
    ("conic", [("eccen", "s"), ("holes", "s")]),
    ("parabola", [("height", "s"), ("width", "s")]),
    ("bipolar", [("shift", "s")]),
    ("boarders", [("prob", "s")]),
    ("butterfly", []),
    ("cell", [("size", "s")]),
    ("cpow", [("r", "s"), ("i", "s"), ("power", "s")]),
    ("curve", [("xamp", "s"), ("yamp", "s"), ("xlen", "s"), ("ylen", "s")]),
    ("edisc", []),
    ("elliptic", []),
    ("escher", [("beta", "s")]),
    ("foci", []),
    ("lazysusan", [("x", "s"), ("y", "s"), ("spin", "s"), ("twist", "s"),
                   ("space", "s"), ("flam3_weight", "s")]),
    ("loonie", [("flam3_weight", "s")]),
    ("oscope", [("frequency", "s"), ("amplitude", "s"), ("damping", "s"),
                ("separation", "s")]),
    ("popcorn", [("x", "s"), ("y", "s"), ("c", "s")]),
]
VARIATIONS = VARIATIONS_ND + VARIATIONS_2D
VAR_ID = {name: i for i, (name, _) in enumerate(VARIATIONS)}
N_ND = len(VARIATIONS_ND)

MAX_DIM = 65535
MAX_RECT_F64 = 1e10


class FlameError(Exception):
    """A flame file the reference would reject (its JsonError text)."""


# ---------------------------------------------------------------- JSON ---

def strip_comments(text: str) -> str:
    """Remove // and /* */ comments outside strings (nlohmann ignore_comments)."""
    out = []
    i, n = 0, len(text)
    in_str = False
    while i < n:
        ch = text[i]
        if in_str:
            out.append(ch)
            if ch == "\\" and i + 1 < n:
                out.append(text[i + 1])
                i += 2
                continue
            if ch == '"':
                in_str = False
            i += 1
            continue
        if ch == '"':
            in_str = True
            out.append(ch)
            i += 1
            continue
        if ch == "/" and i + 1 < n and text[i + 1] == "/":
            j = text.find("\n", i)
            i = n if j < 0 else j
            out.append(" ")
            continue
        if ch == "/" and i + 1 < n and text[i + 1] == "*":
            j = text.find("*/", i + 2)
            if j < 0:
                raise FlameError("parse error: unterminated comment")
            i = j + 2
            out.append(" ")
            continue
        out.append(ch)
        i += 1
    return "".join(out)


def _reject_constant(name):
    raise FlameError(f"parse error: invalid literal {name}")


def parse_json(text: str):
    try:
        return json.loads(strip_comments(text), parse_constant=_reject_constant)
    except json.JSONDecodeError as e:
        raise FlameError(f"parse error: {e}") from None


def _is_int(v) -> bool:
    return isinstance(v, int) and not isinstance(v, bool) and \
        -(1 << 63) <= v < (1 << 64)


def _is_num(v) -> bool:
    return (isinstance(v, (int, float)) and not isinstance(v, bool))


def _get(obj, key: str):
    if not isinstance(obj, dict):
        raise FlameError("Json::operator[](char*): not an object")
    if key not in obj:
        raise FlameError(f"Json::operator[](char*): key does not exist: {key}")
    return obj[key]


def _float(v) -> float:
    if not _is_num(v):
        raise FlameError("Json::floatValue(): type is not float")
    return float(v)


def _int(v) -> int:
    if not _is_int(v):
        raise FlameError("Json::intValue(): type is not int")
    if v >= (1 << 63):
        v -= 1 << 64
    return v


def _array(v) -> list:
    if not isinstance(v, list):
        raise FlameError("Json::arrayValue(): type is not array")
    return v


def _string(v) -> str:
    if not isinstance(v, str):
        raise FlameError("Json::stringValue(): type is not string")
    return v


def _point(v, n: int) -> list[float]:
    if not isinstance(v, list):
        raise FlameError("Point(Json&): not an array")
    if len(v) != n:
        raise FlameError(f"Point(Json&): incorrect array size: {len(v)}")
    out = []
    for e in v:
        if _is_int(e):
            out.append(float(_int(e)))
        elif _is_num(e):
            out.append(float(e))
        else:
            raise FlameError("Point(Json&): entry is not a number")
    return out


# --------------------------------------------------------------- model ---

@dataclass
class RawVar:
    type: int
    name: str
    weight: float
    params: list[float]
    axis_x: int = -1
    axis_y: int = -1


@dataclass
class RawXForm:
    weight: float
    pre: tuple | None          # (A rows flat, b) or None
    post: tuple | None
    color: list[float] | None
    speed: float | None
    vars: list[RawVar] = field(default_factory=list)


@dataclass
class RawFlame:
    dims: int
    size: list[int]
    bounds: list[tuple[float, float]]
    color_dims: int
    color_speed: float
    xforms: list[RawXForm]
    final: RawXForm | None


def _affine(j, n: int):
    if not isinstance(j, dict):
        raise FlameError("Affine(Json&): not an object")
    if "A" not in j:
        A = [1.0 if r == c else 0.0 for r in range(n) for c in range(n)]
    else:
        ja = j["A"]
        if not isinstance(ja, list):
            raise FlameError("Affine(Json&): A is not an array")
        if len(ja) != n:
            raise FlameError("Affine(Json&): A is wrong size")
        A = []
        for i in range(n):
            try:
                A.extend(_point(ja[i], n))
            except FlameError as e:
                raise FlameError(f"Affine(Json&): error parsing A[{i}]: {e}") from None
    if "b" not in j:
        b = [0.0] * n
    else:
        try:
            b = _point(j["b"], n)
        except FlameError as e:
            raise FlameError(f"Affine(Json&): error parsing b: {e}") from None
    return (A, b)


def _variation(j, dims: int) -> RawVar:
    name = _string(_get(j, "name"))
    vid = VAR_ID.get(name)
    if vid is None or (dims < 2 and vid >= N_ND):
        raise FlameError(f"unknown variation: {name}")
    weight = _float(_get(j, "weight"))
    ax = ay = -1
    if vid >= N_ND and dims > 2:
        ax = _int(_get(j, "axis_x"))
        ay = _int(_get(j, "axis_y"))
        if ax < 0 or ax >= dims:
            raise FlameError("axis_x index out of range")
        if ay < 0 or ay >= dims:
            raise FlameError("axis_y index out of range")
        if ax == ay:
            raise FlameError("axes are not distinct")
    params: list[float] = []
    for key, kind in VARIATIONS[vid][1]:
        v = _get(j, key)
        if kind == "s":
            params.append(_float(v))
        elif kind == "p":
            params.extend(_point(v, dims))
        else:
            params.extend(_point(v, 2))
    # constructor-time checks (variations.hpp: Boarders, SphericalP, UnitSphereP)
    if name == "boarders" and (params[0] < 0.0 or params[0] > 1.0):
        raise FlameError("boarders probability out of range")
    if name in ("spherical_p", "unit_sphere_p") and params[0] <= 0.0:
        raise FlameError("norm <= 0")
    return RawVar(vid, name, weight, params, ax, ay)


def _xform(j, dims: int, is_final: bool, color_dims: int, fp32: bool) -> RawXForm:
    if not is_final:
        try:
            weight = _float(_get(j, "weight"))
        except FlameError as e:
            raise FlameError(f"XForm(): cannot parse weight: {e}") from None
    else:
        weight = 1.0
    wn = _to_num(weight, fp32)
    if wn < 0.0:
        raise FlameError("XForm(): weight is negative")
    pre = post = None
    if isinstance(j, dict) and "pre_affine" in j:
        try:
            pre = _affine(j["pre_affine"], dims)
        except FlameError as e:
            raise FlameError(f"XForm(): cannot parse pre_affine{e}") from None
    if isinstance(j, dict) and "post_affine" in j:
        try:
            post = _affine(j["post_affine"], dims)
        except FlameError as e:
            raise FlameError(f"XForm(): cannot parse post_affine{e}") from None
    try:
        vars_ = [_variation(v, dims) for v in _array(_get(j, "variations"))]
    except FlameError as e:
        raise FlameError(f"XForm(): cannot parse variations: {e}") from None
    color = None
    if color_dims and isinstance(j, dict) and "color" in j:
        ca = _array(j["color"])
        try:
            if len(ca) != color_dims:
                raise FlameError("color length incorrect")
            color = []
            for v in ca:
                c = _float(v)
                cn = _to_num(c, fp32)
                if cn < 0.0 or cn > 1.0:
                    raise FlameError("color coordinate out of range")
                color.append(c)
        except FlameError as e:
            raise FlameError(f"XForm(): cannot parse color: {e}") from None
    speed = None
    if color_dims and isinstance(j, dict) and "color_speed" in j:
        try:
            speed = _float(j["color_speed"])
            sn = _to_num(speed, fp32)
            if sn < 0.0 or sn > 1.0:
                raise FlameError("color speed out of range")
        except FlameError as e:
            raise FlameError(f"XForm(): color speed issue: {e}") from None
    return RawXForm(weight, pre, post, color, speed, vars_)


def _to_num(v: float, fp32: bool) -> float:
    if not fp32:
        return v
    import numpy as np
    return float(np.float32(v))


def flame_dims(doc) -> int:
    """The CLI reads "dimensions" before building the flame (ffr_buf.cpp:131)."""
    return _int(_get(doc, "dimensions"))


def parse_flame(text_or_doc, fp32: bool = False) -> RawFlame:
    """Validate a flame document like Flame<dims>(Json) and return it raw."""
    doc = parse_json(text_or_doc) if isinstance(text_or_doc, str) else text_or_doc
    dims = flame_dims(doc)
    if dims not in (1, 2, 3):
        raise FlameError(f"ERROR: {dims}D not supported")
    maxrect = 1e5 if fp32 else MAX_RECT_F64
    size: list[int] = []
    bounds: list[tuple[float, float]] = []
    try:
        sizej = _array(_get(doc, "size"))
        boundsj = _array(_get(doc, "bounds"))
        if len(sizej) != dims:
            raise FlameError("incorrect size length")
        if len(boundsj) != dims:
            raise FlameError("incorrect bounds length")
        for i in range(dims):
            try:
                sv = _float(sizej[i])
            except FlameError as e:
                raise FlameError(f"error parsing size[{i}]: {e}") from None
            s = _size_t(sv)
            if s == 0 or s > MAX_DIM:
                raise FlameError(f"size[{i}] out of range")
            size.append(s)
            bj = _array(boundsj[i])
            if len(bj) != 2:
                raise FlameError(f"bounds[{i}] wrong format")
            try:
                lo = _float(bj[0])
                hi = _float(bj[1])
            except FlameError as e:
                raise FlameError(f"error parsing bounds[{i}]: {e}") from None
            lo_n, hi_n = _to_num(lo, fp32), _to_num(hi, fp32)
            if lo_n < -maxrect or lo_n > maxrect or hi_n < -maxrect or hi_n > maxrect:
                raise FlameError(f"bounds[{i}] out of range")
            if lo_n >= hi_n:
                raise FlameError(f"bounds[{i}] low >= high")
            bounds.append((lo, hi))
    except FlameError as e:
        raise FlameError(f"Flame(): {e}") from None
    has_final = isinstance(doc, dict) and "final_xform" in doc
    try:
        xfs = _array(_get(doc, "xforms"))
    except FlameError as e:
        raise FlameError(f"Flame(): cannot parse xforms: {e}") from None
    try:
        color_dims = _int(doc["color_dimensions"]) if "color_dimensions" in doc else 0
        if color_dims < 0 or color_dims > 127:
            raise FlameError("too many color dimensions")
        color_speed = _float(doc["color_speed"]) if "color_speed" in doc else 0.5
        cs = _to_num(color_speed, fp32)
        if cs < 0.0 or cs > 1.0:
            raise FlameError("color speed out of range")
    except FlameError as e:
        raise FlameError(f"Flame(): {e}") from None
    xforms = []
    for i, xj in enumerate(xfs):
        try:
            xforms.append(_xform(xj, dims, False, color_dims, fp32))
        except FlameError as e:
            raise FlameError(f"Flame(): error parsing xforms[{i}]: {e}") from None
    if not xforms:
        raise FlameError("Flame(): no xforms")
    final = None
    if has_final:
        try:
            final = _xform(doc["final_xform"], dims, True, color_dims, fp32)
        except FlameError as e:
            raise FlameError(f"Flame(): error parsing final xform: {e}") from None
    if all(_to_num(x.weight, fp32) == 0.0 for x in xforms):
        raise FlameError("Flame(): no xforms remaining after optimization")
    cells = 1
    for s in size:
        cells *= s
        if cells >= (1 << 48):
            raise FlameError("BufferRenderer(): histogram too big")
    return RawFlame(dims, size, bounds, color_dims, color_speed, xforms, final)


def _size_t(v: float) -> int:
    """double -> size_t as the reference's `size[i] = floatValue()` does."""
    if math.isnan(v) or v <= -1.0 or v >= 2.0 ** 64:
        return 0  # out of range for size_t; rejected as out of range
    return int(v) if v >= 0 else 0


def cells_of(raw: RawFlame) -> int:
    n = 1
    for s in raw.size:
        n *= s
    return n
