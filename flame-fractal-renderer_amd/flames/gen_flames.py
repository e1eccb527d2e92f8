"""Seeded generators of the benchmark flames; writes the JSON files beside it.

  random_color_3840x2160.json   random 2-D colour flame (seed 2): 6 xforms,
      each mixing 3 distinct variations from {linear, sinusoidal, spherical,
      swirl, horseshoe, polar} with Dirichlet(1) weights; pre- and
      post-affine entries uniform in [-1, 1], rescaled to spectral norm
      <= 0.9; xform weights Dirichlet(1); RGB colour per xform uniform in
      [0, 1]^3 (color_dimensions 3, color_speed 0.5); 3840x2160 cells.
  tetrahedron_512.json          3-D Sierpinski tetrahedron: four maps
      A = 0.5 I toward (0,0,0), (1,0,0), (0,1,0), (0,0,1); 512^3 cells over
      [0, 1]^3.
  batch/flame_NNNN.json         256 flames from the same generator with 2-8
      xforms (seeds 1000-1255), square sizes drawn from 256..2048.
  serve3d/flame_NNNN.json       16 random 3-D linear flames (seeds 3000-3015):
      4 xforms, affine maps with entries uniform in [-1, 1] rescaled to
      spectral norm <= 0.6, Dirichlet(1) weights; 256^3 cells framed by the
      0.5/99.5 percentile box of a chaos game.

The bounds of a random flame frame its attractor: a short numpy chaos game
(the generator's own, double precision) gives the 0.5/99.5 percentile box,
widened by 10 % and padded to the image aspect ratio.

Run: python gen_flames.py   (deterministic; the committed files are its output)
"""

from __future__ import annotations

import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
VARS6 = ["linear", "sinusoidal", "spherical", "swirl", "horseshoe", "polar"]


def _affine(rng):
    A = rng.uniform(-1.0, 1.0, size=(2, 2))
    s = np.linalg.norm(A, 2)
    if s > 0.9:
        A *= 0.9 / s
    b = rng.uniform(-1.0, 1.0, size=2)
    return {"A": [[float(v) for v in row] for row in A], "b": [float(v) for v in b]}


def _var(name, x, y):
    if name == "linear":
        return x, y
    if name == "sinusoidal":
        return np.sin(x), np.sin(y)
    if name == "spherical":
        r = 1.0 / (x * x + y * y + 1e-20)
        return x * r, y * r
    if name == "swirl":
        r = x * x + y * y
        return x * np.sin(r) - y * np.cos(r), x * np.cos(r) + y * np.sin(r)
    if name == "horseshoe":
        r = 1.0 / (np.sqrt(x * x + y * y) + 1e-20)
        return (x - y) * (x + y) * r, 2.0 * x * y * r
    if name == "polar":
        return np.arctan2(y, x) / np.pi, np.sqrt(x * x + y * y) - 1.0
    raise ValueError(name)


def _frame(xforms, aspect, rng, walkers=4096, iters=256, settle=32):
    """Percentile box of a short chaos game over the xforms (no final xform)."""
    w = np.array([x["weight"] for x in xforms])
    cw = np.cumsum(w / w.sum())
    x = rng.uniform(-1, 1, walkers)
    y = rng.uniform(-1, 1, walkers)
    xs, ys = [], []
    for it in range(settle + iters):
        k = np.minimum(np.searchsorted(cw, rng.uniform(0, 1, walkers)), len(xforms) - 1)
        nx = np.zeros(walkers)
        ny = np.zeros(walkers)
        for j, xf in enumerate(xforms):
            m = k == j
            if not m.any():
                continue
            A = np.array(xf["pre_affine"]["A"])
            b = np.array(xf["pre_affine"]["b"])
            tx = b[0] + A[0, 0] * x[m] + A[0, 1] * y[m]
            ty = b[1] + A[1, 0] * x[m] + A[1, 1] * y[m]
            vx = np.zeros(m.sum())
            vy = np.zeros(m.sum())
            for v in xf["variations"]:
                a, c = _var(v["name"], tx, ty)
                vx += v["weight"] * a
                vy += v["weight"] * c
            A = np.array(xf["post_affine"]["A"])
            b = np.array(xf["post_affine"]["b"])
            nx[m] = b[0] + A[0, 0] * vx + A[0, 1] * vy
            ny[m] = b[1] + A[1, 0] * vx + A[1, 1] * vy
        x, y = nx, ny
        if it >= settle:
            ok = np.isfinite(x) & np.isfinite(y)
            xs.append(x[ok])
            ys.append(y[ok])
    xs = np.concatenate(xs)
    ys = np.concatenate(ys)
    x0, x1 = np.percentile(xs, [0.5, 99.5])
    y0, y1 = np.percentile(ys, [0.5, 99.5])
    cx, cy = (x0 + x1) / 2, (y0 + y1) / 2
    hw, hh = 1.1 * (x1 - x0) / 2, 1.1 * (y1 - y0) / 2
    if hw / hh < aspect:
        hw = hh * aspect
    else:
        hh = hw / aspect
    return [[float(cx - hw), float(cx + hw)], [float(cy - hh), float(cy + hh)]]


def random_color_flame(seed: int, n_xforms: int, size: tuple[int, int]) -> dict:
    rng = np.random.default_rng(seed)
    weights = rng.dirichlet(np.ones(n_xforms))
    xforms = []
    for i in range(n_xforms):
        names = rng.choice(VARS6, size=3, replace=False)
        vw = rng.dirichlet(np.ones(3))
        xforms.append({
            "weight": float(weights[i]),
            "color": [float(v) for v in rng.uniform(0.0, 1.0, 3)],
            "variations": [{"name": str(n), "weight": float(w)} for n, w in zip(names, vw)],
            "pre_affine": _affine(rng),
            "post_affine": _affine(rng),
        })
    bounds = _frame(xforms, size[0] / size[1], np.random.default_rng(seed + 7919))
    return {
        "dimensions": 2,
        "size": list(size),
        "bounds": bounds,
        "color_dimensions": 3,
        "color_speed": 0.5,
        "xforms": xforms,
    }


def tetrahedron(size: int = 512) -> dict:
    half = [[0.5, 0, 0], [0, 0.5, 0], [0, 0, 0.5]]
    bs = [[0, 0, 0], [0.5, 0, 0], [0, 0.5, 0], [0, 0, 0.5]]
    return {
        "dimensions": 3,
        "size": [size, size, size],
        "bounds": [[0, 1], [0, 1], [0, 1]],
        "xforms": [{"weight": 1, "variations": [{"name": "linear", "weight": 1.0}],
                    "pre_affine": {"A": half, "b": b}} for b in bs],
    }


def batch_flames(n: int = 256, seed0: int = 1000):
    sizes = [256, 512, 1024, 2048]
    out = []
    for i in range(n):
        s = seed0 + i
        rng = np.random.default_rng(s)
        nx = int(rng.integers(2, 9))
        side = int(rng.choice(sizes))
        out.append((s, random_color_flame(s, nx, (side, side))))
    return out


def random_3d_flame(seed: int, size: int = 256) -> dict:
    rng = np.random.default_rng(seed)
    n = 4
    weights = rng.dirichlet(np.ones(n))
    maps = []
    for _ in range(n):
        A = rng.uniform(-1.0, 1.0, size=(3, 3))
        s = np.linalg.norm(A, 2)
        if s > 0.6:
            A *= 0.6 / s
        maps.append((A, rng.uniform(-1.0, 1.0, size=3)))
    # frame: a short chaos game in double precision
    frng = np.random.default_rng(seed + 7919)
    cw = np.cumsum(weights)
    p = frng.uniform(-1, 1, (4096, 3))
    pts = []
    for it in range(32 + 128):
        k = np.minimum(np.searchsorted(cw, frng.uniform(0, 1, len(p))), n - 1)
        q = np.empty_like(p)
        for j, (A, b) in enumerate(maps):
            m = k == j
            q[m] = p[m] @ A.T + b
        p = q
        if it >= 32:
            pts.append(p.copy())
    pts = np.concatenate(pts)
    lo = np.percentile(pts, 0.5, axis=0)
    hi = np.percentile(pts, 99.5, axis=0)
    c, h = (lo + hi) / 2, 1.1 * (hi - lo) / 2
    return {
        "dimensions": 3,
        "size": [size, size, size],
        "bounds": [[float(c[i] - h[i]), float(c[i] + h[i])] for i in range(3)],
        "xforms": [{"weight": float(w), "variations": [{"name": "linear", "weight": 1.0}],
                    "pre_affine": {"A": [[float(v) for v in row] for row in A],
                                   "b": [float(v) for v in b]}}
                   for w, (A, b) in zip(weights, maps)],
    }


def _write(path, doc, header=None):
    text = json.dumps(doc, indent=1)
    if header:
        text = "// " + header + "\n" + text
    with open(path, "w") as f:
        f.write(text + "\n")


def main():
    _write(os.path.join(HERE, "random_color_3840x2160.json"),
           random_color_flame(2, 6, (3840, 2160)),
           "random 2-D colour flame, seed 2 (gen_flames.py)")
    _write(os.path.join(HERE, "tetrahedron_512.json"), tetrahedron(512),
           "3-D Sierpinski tetrahedron (gen_flames.py)")
    bdir = os.path.join(HERE, "batch")
    os.makedirs(bdir, exist_ok=True)
    for s, doc in batch_flames():
        _write(os.path.join(bdir, f"flame_{s}.json"), doc,
               f"batch flame, seed {s} (gen_flames.py)")
    sdir = os.path.join(HERE, "serve3d")
    os.makedirs(sdir, exist_ok=True)
    for s in range(3000, 3016):
        _write(os.path.join(sdir, f"flame_{s}.json"), random_3d_flame(s),
               f"3-D serving flame, seed {s} (gen_flames.py)")


if __name__ == "__main__":
    main()
