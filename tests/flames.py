"""Test flames written for this suite (small grids, seeded coefficients)."""

import json

import numpy as np

# parameters for every variation that takes some (catalogue keys)
VAR_PARAMS = {
    "bent": {"scales_neg": [2.0, 0.5, 1.5], "scales_pos": [1.0, 1.2, 0.9]},
    "rectangles": {"params": [0.3, 0.4, 0.5]},
    "fisheye": {"addval": 1.0},
    "bubble": {"addval": 4.0},
    "separation": {"params": [0.2, 0.3, 0.25], "inside": [0.1, 0.15, 0.05]},
    "splits": {"params": [0.2, 0.3, 0.1]},
    "modulus": {"params": [0.7, 0.9, 0.8]},
    "celln": {"sizes": [0.6, 0.8, 0.7]},
    "spherical_p": {"norm": 1.5},
    "unit_sphere_p": {"norm": 3.0},
    "disc2": {"rotation": 0.6, "twist": 0.4},
    "waves": {"xfreq": 2.0, "xscale": 0.3, "yfreq": 1.5, "yscale": 0.2},
    "fan": {"x": 0.4, "y": 0.3},
    "rings": {"value": 0.5},
    "blob": {"low": 0.2, "high": 1.1, "waves": 3.0},
    "pdj": {"a": 1.1, "b": 0.9, "c": -0.7, "d": 1.3},
    "perspective": {"distance": 2.5, "angle": 0.6},
    "julian": {"power": 3.0, "dist": 1.2},
    "juliascope": {"power": -2.5, "dist": 0.8},
    "radial_blur": {"angle": 0.7, "flam3_weight": 0.5},
    "pie": {"slices": 6.0, "rotation": 0.3, "thickness": 0.5},
    "ngon": {"sides": 5.0, "power": 2.0, "corners": 0.8, "circle": 1.0},
    "curl": {"c1": 0.6, "c2": 0.2},
    "arch": {"flam3_weight": 0.7},
    "rays": {"flam3_weight": 0.7},
    "blade": {"flam3_weight": 0.7},
    "secant": {"flam3_weight": 0.7},
    "twintrian": {"flam3_weight": 0.7},
    "auger": {"freq": 2.0, "flam3_weight": 0.5, "scale": 1.0, "sym": 0.3},
    "flux": {"spread": 0.6, "flam3_weight": 0.4},
    "mobius": {"a": [0.8, 0.1], "b": [0.2, -0.3], "c": [0.1, 0.4], "d": [1.0, 0.2]},
    "scry": {"flam3_weight": 0.7},
    "split": {"xsize": 0.4, "ysize": 0.6},
    "stripes": {"space": 0.3, "warp": 0.8},
    "wedge": {"swirl": 0.2, "count": 3.0, "angle": 0.5, "hole": 0.1},
    "wedge_julia": {"angle": 0.4, "count": 2.0, "power": 3.0, "dist": 1.0},
    "wedge_sph": {"angle": 0.5, "count": 4.0, "swirl": 0.3, "hole": 0.2},
    "whorl": {"inside": 0.5, "outside": 0.8, "flam3_weight": 1.2},
    "supershape": {"n1": 2.0, "m": 5.0, "n2": 1.5, "n3": 1.2, "rnd": 0.3, "holes": 0.1},
    "flower": {"petals": 5.0, "holes": 0.2},
    "conic": {"eccen": 0.7, "holes": 0.1},
    "parabola": {"height": 0.8, "width": 1.2},
    "bipolar": {"shift": 0.4},
    "boarders": {"prob": 0.6},
    "cell": {"size": 0.7},
    "cpow": {"r": 1.2, "i": 0.3, "power": 3.0},
    "curve": {"xamp": 0.4, "yamp": 0.6, "xlen": 1.2, "ylen": 0.8},
    "escher": {"beta": 0.6},
    "lazysusan": {"x": 0.2, "y": 0.1, "spin": 0.5, "twist": 0.3, "space": 0.4,
                  "flam3_weight": 1.5},
    "loonie": {"flam3_weight": 0.7},
    "oscope": {"frequency": 1.5, "amplitude": 0.6, "damping": 0.3, "separation": 0.2},
    "popcorn": {"x": 0.3, "y": 0.2, "c": 0.9},
}


def variation(name: str, weight: float, dims: int, axes=None) -> dict:
    v = {"name": name, "weight": weight}
    for k, val in VAR_PARAMS.get(name, {}).items():
        if isinstance(val, list) and name != "mobius":
            val = val[:dims]
        v[k] = val
    if axes is not None:
        v["axis_x"], v["axis_y"] = axes
    return v


def single_variation_flame(name: str, dims: int = 2, axes=None, scale=2.5) -> str:
    """One xform: pre-affine scale * I + offset, the variation, identity post."""
    A = (np.eye(dims) * scale).tolist()
    b = [0.3, -0.2, 0.15][:dims]
    doc = {
        "dimensions": dims,
        "size": [16] * dims,
        "bounds": [[-1e3, 1e3]] * dims,
        "xforms": [{"weight": 1.0, "variations": [variation(name, 1.0, dims, axes)],
                    "pre_affine": {"A": A, "b": b}}],
    }
    return json.dumps(doc)


def fern(size=128) -> str:
    # Barnsley's fern (Barnsley, Fractals Everywhere, 1988)
    maps = [
        (0.01, [[0, 0], [0, 0.16]], [0, 0]),
        (0.85, [[0.85, 0.04], [-0.04, 0.85]], [0, 1.6]),
        (0.07, [[0.2, -0.26], [0.23, 0.22]], [0, 1.6]),
        (0.07, [[-0.15, 0.28], [0.26, 0.24]], [0, 0.44]),
    ]
    doc = {"dimensions": 2, "size": [size, size], "bounds": [[-3, 3], [-0.5, 10.5]],
           "xforms": [{"weight": w, "variations": [{"name": "linear", "weight": 1.0}],
                       "pre_affine": {"A": A, "b": b}} for w, A, b in maps]}
    return json.dumps(doc)


def rotations(size=96) -> str:
    # four rotate-and-shrink maps with a linear final xform and post-affines
    xf = []
    for k, (bx, by) in enumerate([(-1, 0), (1, 0), (0, 1), (0, -1)]):
        a = 0.6 + 0.05 * k
        c, s = np.cos(0.3 * (k + 1)), np.sin(0.3 * (k + 1))
        xf.append({"weight": 1.0 + k,
                   "variations": [{"name": "linear", "weight": 0.7},
                                  {"name": "linear", "weight": 0.3}],
                   "pre_affine": {"A": [[a * c, -a * s], [a * s, a * c]], "b": [bx, by]},
                   "post_affine": {"A": [[1.0, 0.1], [-0.1, 1.0]], "b": [0.05, -0.05]}})
    doc = {"dimensions": 2, "size": [size, size], "bounds": [[-3, 3], [-3, 3]],
           "xforms": xf,
           "final_xform": {"variations": [{"name": "linear", "weight": 0.9}],
                           "pre_affine": {"A": [[1, 0.2], [0, 1]], "b": [0.1, 0]}}}
    return json.dumps(doc)


def tetrahedron(size=32, post=False) -> str:
    half = [[0.5, 0, 0], [0, 0.5, 0], [0, 0, 0.5]]
    xfs = []
    for b in ([0, 0, 0], [0.5, 0, 0], [0, 0.5, 0], [0, 0, 0.5]):
        x = {"weight": 1, "variations": [{"name": "linear", "weight": 1.0}],
             "pre_affine": {"A": half, "b": b}}
        if post:
            x["post_affine"] = {"A": [[1, 0, 0], [0, 1, 0], [0, 0, 1]], "b": [0, 0, 0]}
        xfs.append(x)
    return json.dumps({"dimensions": 3, "size": [size] * 3,
                       "bounds": [[0, 1]] * 3, "xforms": xfs})


def one_d(size=200, nonlinear=False) -> str:
    xf = [{"weight": 1, "variations": [{"name": "linear", "weight": 1.0}],
           "pre_affine": {"A": [[0.5]], "b": [0]}},
          {"weight": 2, "variations": [{"name": "linear", "weight": 1.0}],
           "pre_affine": {"A": [[0.4]], "b": [0.6]}}]
    if nonlinear:
        xf[1]["variations"].append({"name": "sinusoidal", "weight": 0.3})
    return json.dumps({"dimensions": 1, "size": [size], "bounds": [[-0.5, 1.5]],
                       "xforms": xf})


def color_flame(r=3, final_color=True, size=128, seed=4) -> str:
    rng = np.random.default_rng(seed)
    xfs = []
    for k in range(3):
        A = (rng.uniform(-1, 1, (2, 2)) * 0.45).tolist()
        b = rng.uniform(-1, 1, 2).tolist()
        x = {"weight": float(rng.uniform(0.5, 2.0)),
             "variations": [{"name": "linear", "weight": 1.0}],
             "pre_affine": {"A": A, "b": b},
             "color": rng.uniform(0, 1, r).tolist()}
        if k == 1:
            x["color_speed"] = 0.25
        if k == 2:
            del x["color"]   # this xform leaves the colour alone
        xfs.append(x)
    doc = {"dimensions": 2, "size": [size, size], "bounds": [[-3, 3], [-3, 3]],
           "color_dimensions": r, "color_speed": 0.6, "xforms": xfs,
           "final_xform": {"variations": [{"name": "linear", "weight": 1.0}],
                           "pre_affine": {"A": [[0.9, 0.1], [-0.1, 0.9]], "b": [0, 0]}}}
    if final_color:
        doc["final_xform"]["color"] = rng.uniform(0, 1, r).tolist()
        doc["final_xform"]["color_speed"] = 0.3
    return json.dumps(doc)


def tied_weights(n=24, seed=5, size=64) -> str:
    rng = np.random.default_rng(seed)
    xfs = []
    for k in range(n):
        A = (np.eye(2) * 0.3).tolist()
        b = rng.uniform(-1, 1, 2).tolist()
        xfs.append({"weight": float(rng.choice([0.5, 1.0, 1.0, 2.0, 0.0])),
                    "variations": [{"name": "linear", "weight": 1.0},
                                   {"name": "linear", "weight": 0.0}],
                    "pre_affine": {"A": A, "b": b}})
    return json.dumps({"dimensions": 2, "size": [size, size], "bounds": [[-2, 2], [-2, 2]],
                       "xforms": xfs})


def occasional_bad(size=64) -> str:
    # a rare map throws points far beyond the bad-value threshold
    return json.dumps({
        "dimensions": 2, "size": [size, size], "bounds": [[-2, 2], [-2, 2]],
        "xforms": [
            {"weight": 1.0, "variations": [{"name": "linear", "weight": 1.0}],
             "pre_affine": {"A": [[0.5, 0], [0, 0.5]], "b": [0.5, 0]}},
            {"weight": 1.0, "variations": [{"name": "linear", "weight": 1.0}],
             "pre_affine": {"A": [[0.5, 0], [0, 0.5]], "b": [-0.5, 0.3]}},
            {"weight": 0.002, "variations": [{"name": "linear", "weight": 1.0}],
             "pre_affine": {"A": [[1e25, 0], [0, 1e25]], "b": [0, 0]}},
        ],
        "final_xform": {"variations": [{"name": "linear", "weight": 1.0}],
                        "pre_affine": {"A": [[1e-30, 0], [0, 1e-30]], "b": [0.25, 0.25]}},
    })


def expanding(size=32) -> str:
    return json.dumps({
        "dimensions": 2, "size": [size, size], "bounds": [[-2, 2], [-2, 2]],
        "xforms": [{"weight": 1.0, "variations": [{"name": "linear", "weight": 1.0}],
                    "pre_affine": {"A": [[2, 0], [0, 2]], "b": [0.1, 0.1]}}]})
