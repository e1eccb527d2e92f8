"""ctypes front end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, to check results against.  The product never imports it.
"""

from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import model

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

F64_U64, F32_U32, F32_U64 = 0, 1, 2
MAX_PARAMS = 16


class OrcVar(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("axis_x", C.c_int32),
        ("axis_y", C.c_int32),
        ("n_params", C.c_int32),
        ("weight", C.c_double),
        ("params", C.c_double * MAX_PARAMS),
    ]


class OrcXForm(C.Structure):
    _fields_ = [
        ("weight", C.c_double),
        ("has_pre", C.c_int32),
        ("has_post", C.c_int32),
        ("has_color", C.c_int32),
        ("has_speed", C.c_int32),
        ("pre_A", C.c_double * 9),
        ("pre_b", C.c_double * 3),
        ("post_A", C.c_double * 9),
        ("post_b", C.c_double * 3),
        ("speed", C.c_double),
        ("color", C.POINTER(C.c_double)),
        ("n_vars", C.c_int32),
        ("vars", C.POINTER(OrcVar)),
    ]


class OrcFlameDesc(C.Structure):
    _fields_ = [
        ("dims", C.c_int32),
        ("color_dims", C.c_int32),
        ("color_speed", C.c_double),
        ("size", C.c_uint64 * 3),
        ("lo", C.c_double * 3),
        ("hi", C.c_double * 3),
        ("n_xforms", C.c_int32),
        ("xforms", C.POINTER(OrcXForm)),
        ("has_final", C.c_int32),
        ("final_xform", OrcXForm),
    ]


class OrcStats(C.Structure):
    _fields_ = [
        ("s_iter", C.c_uint64),
        ("s_plot", C.c_uint64),
        ("n_bad", C.c_uint64),
        ("success", C.c_int32),
        ("pt_min", C.c_double * 3),
        ("pt_max", C.c_double * 3),
    ]


class OrcIsaac64(C.Structure):
    _fields_ = [("cnt", C.c_uint64), ("rsl", C.c_uint64 * 16),
                ("mem", C.c_uint64 * 16), ("a", C.c_uint64),
                ("b", C.c_uint64), ("c", C.c_uint64)]


class OrcIsaac32(C.Structure):
    _fields_ = [("cnt", C.c_uint64), ("rsl", C.c_uint32 * 16),
                ("mem", C.c_uint32 * 16), ("a", C.c_uint32),
                ("b", C.c_uint32), ("c", C.c_uint32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library not built: {_LIB_PATH} "
                               "(run `make -C oracle`)")
        L = C.CDLL(_LIB_PATH)
        L.orc_flame_new.restype = C.c_void_p
        L.orc_flame_new.argtypes = [C.POINTER(OrcFlameDesc), C.c_int32,
                                    C.c_int32, C.c_char_p, C.c_size_t]
        L.orc_flame_free.argtypes = [C.c_void_p]
        L.orc_flame_n_xforms.restype = C.c_int32
        L.orc_flame_n_xforms.argtypes = [C.c_void_p]
        L.orc_flame_xform_id.restype = C.c_int32
        L.orc_flame_xform_id.argtypes = [C.c_void_p, C.c_int32]
        L.orc_flame_cum_weight.restype = C.c_double
        L.orc_flame_cum_weight.argtypes = [C.c_void_p, C.c_int32]
        L.orc_flame_cells.restype = C.c_size_t
        L.orc_flame_cells.argtypes = [C.c_void_p]
        L.orc_flame_cell_words.restype = C.c_size_t
        L.orc_flame_cell_words.argtypes = [C.c_void_p]
        L.orc_flame_word_bytes.restype = C.c_size_t
        L.orc_flame_word_bytes.argtypes = [C.c_void_p]
        L.orc_flame_mult_d.restype = C.c_double
        L.orc_flame_mult_d.argtypes = [C.c_void_p, C.c_int32]
        L.orc_render.restype = C.c_int
        L.orc_render.argtypes = [
            C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64,
            C.c_uint64, C.c_uint64, C.c_int32, C.POINTER(C.c_uint64),
            C.POINTER(OrcStats), C.POINTER(C.c_uint64),
            C.POINTER(C.c_double), C.c_uint64]
        L.orc_stream_word.restype = C.c_uint64
        L.orc_stream_word.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32),
                                        C.POINTER(C.c_uint32),
                                        C.POINTER(C.c_uint32)]
        L.orc_isaac64_seed.argtypes = [C.POINTER(OrcIsaac64), C.c_uint64]
        L.orc_isaac64_next.restype = C.c_uint64
        L.orc_isaac64_next.argtypes = [C.POINTER(OrcIsaac64)]
        L.orc_isaac32_seed_u32.argtypes = [C.POINTER(OrcIsaac32), C.c_uint32]
        L.orc_isaac32_seed_u64.argtypes = [C.POINTER(OrcIsaac32), C.c_uint64]
        L.orc_isaac32_next.restype = C.c_uint32
        L.orc_isaac32_next.argtypes = [C.POINTER(OrcIsaac32)]
        L.orc_eval_variation.restype = C.c_int
        L.orc_eval_variation.argtypes = [
            C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_double),
            C.POINTER(C.c_double), C.c_uint64, C.c_uint64,
            C.POINTER(C.c_uint64)]
        L.orc_sort_order.argtypes = [C.POINTER(C.c_double), C.c_int32,
                                     C.c_int32, C.POINTER(C.c_int32)]
        L.orc_tonemap.restype = C.c_int
        L.orc_tonemap.argtypes = [
            C.c_void_p, C.c_uint64, C.c_int32, C.c_int32, C.c_int32,
            C.c_double, C.c_int32, C.c_void_p, C.POINTER(C.c_uint64),
            C.POINTER(C.c_double)]
        L.orc_apply_xform.restype = C.c_int
        L.orc_apply_xform.argtypes = [
            C.c_void_p, C.c_int32, C.POINTER(C.c_double),
            C.POINTER(C.c_double), C.c_uint64, C.c_uint64,
            C.POINTER(C.c_uint64)]
        _lib = L
    return _lib


def _fill_xform(dst: OrcXForm, xf: model.RawXForm, dims: int, keep: list):
    dst.weight = xf.weight
    if xf.pre is not None:
        dst.has_pre = 1
        A, b = xf.pre
    else:
        A = [1.0 if r == c else 0.0 for r in range(dims) for c in range(dims)]
        b = [0.0] * dims
    for i, v in enumerate(A):
        dst.pre_A[i] = v
    for i, v in enumerate(b):
        dst.pre_b[i] = v
    if xf.post is not None:
        dst.has_post = 1
        A, b = xf.post
    else:
        A = [1.0 if r == c else 0.0 for r in range(dims) for c in range(dims)]
        b = [0.0] * dims
    for i, v in enumerate(A):
        dst.post_A[i] = v
    for i, v in enumerate(b):
        dst.post_b[i] = v
    if xf.color is not None:
        arr = (C.c_double * len(xf.color))(*xf.color)
        keep.append(arr)
        dst.has_color = 1
        dst.color = C.cast(arr, C.POINTER(C.c_double))
    if xf.speed is not None:
        dst.has_speed = 1
        dst.speed = xf.speed
    vs = (OrcVar * max(1, len(xf.vars)))()
    keep.append(vs)
    for k, v in enumerate(xf.vars):
        vs[k].type = v.type
        vs[k].axis_x = v.axis_x
        vs[k].axis_y = v.axis_y
        vs[k].n_params = len(v.params)
        vs[k].weight = v.weight
        for i, q in enumerate(v.params):
            vs[k].params[i] = q
    dst.n_vars = len(xf.vars)
    dst.vars = C.cast(vs, C.POINTER(OrcVar))


@dataclass
class RenderResult:
    buffer: np.ndarray      # raw words, cells * (1 + r)
    counts: np.ndarray      # cells
    colors: np.ndarray      # cells x r
    xf_dist: np.ndarray
    s_iter: int
    s_plot: int
    n_bad: int
    success: bool
    pt_min: np.ndarray
    pt_max: np.ndarray
    bad_xforms: np.ndarray
    bad_points: np.ndarray


class OracleFlame:
    """The oracle's optimised flame in one precision mode."""

    def __init__(self, raw: model.RawFlame | str, mode: int = F64_U64,
                 settle: int = -1):
        if isinstance(raw, str):
            raw = model.parse_flame(raw, fp32=(mode != F64_U64))
        self.raw = raw
        self.mode = mode
        self._keep: list = []
        d = OrcFlameDesc()
        d.dims = raw.dims
        d.color_dims = raw.color_dims
        d.color_speed = raw.color_speed
        for i in range(raw.dims):
            d.size[i] = raw.size[i]
            d.lo[i] = raw.bounds[i][0]
            d.hi[i] = raw.bounds[i][1]
        xfs = (OrcXForm * len(raw.xforms))()
        self._keep.append(xfs)
        for i, xf in enumerate(raw.xforms):
            _fill_xform(xfs[i], xf, raw.dims, self._keep)
        d.n_xforms = len(raw.xforms)
        d.xforms = C.cast(xfs, C.POINTER(OrcXForm))
        if raw.final is not None:
            d.has_final = 1
            _fill_xform(d.final_xform, raw.final, raw.dims, self._keep)
        self._desc = d
        err = C.create_string_buffer(512)
        h = lib().orc_flame_new(C.byref(d), mode, settle, err, 512)
        if not h:
            raise model.FlameError(err.value.decode())
        self._h = h
        self.dims = raw.dims
        self.r = raw.color_dims
        self.cells = lib().orc_flame_cells(h)
        self.cell_words = lib().orc_flame_cell_words(h)
        self.word_bytes = lib().orc_flame_word_bytes(h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_flame_free(h)
            self._h = None

    @property
    def xform_order(self) -> list[int]:
        L = lib()
        return [L.orc_flame_xform_id(self._h, i)
                for i in range(L.orc_flame_n_xforms(self._h))]

    @property
    def cum_weights(self) -> list[float]:
        L = lib()
        return [L.orc_flame_cum_weight(self._h, i)
                for i in range(L.orc_flame_n_xforms(self._h))]

    def mult_d(self) -> list[float]:
        return [lib().orc_flame_mult_d(self._h, i) for i in range(self.dims)]

    def new_buffer(self) -> np.ndarray:
        dt = np.uint32 if self.word_bytes == 4 else np.uint64
        return np.zeros(self.cells * self.cell_words, dtype=dt)

    def render(self, samples: int, batch: int, seed: int, walker_base: int = 0,
               bv_limit: int = 256, threads: int = 1,
               buffer: np.ndarray | None = None, bv_cap: int = 1024
               ) -> RenderResult:
        buf = self.new_buffer() if buffer is None else buffer
        n_ids = len(self.raw.xforms)
        xf = (C.c_uint64 * max(1, n_ids))()
        st = OrcStats()
        for i in range(3):
            st.pt_min[i] = float("inf")
            st.pt_max[i] = float("-inf")
        bvx = (C.c_uint64 * max(1, bv_cap))()
        bvp = (C.c_double * (3 * max(1, bv_cap)))()
        lib().orc_render(self._h, buf.ctypes.data_as(C.c_void_p), seed,
                         walker_base, samples, batch, bv_limit, threads, xf,
                         C.byref(st), bvx, bvp, bv_cap)
        return self._result(buf, xf, st, bvx, bvp, bv_cap, n_ids)

    def _result(self, buf, xf, st, bvx, bvp, bv_cap, n_ids) -> RenderResult:
        cw = self.cell_words
        words = buf.reshape(self.cells, cw)
        counts = words[:, 0].copy()
        if self.mode == F32_U32:
            colors = words[:, 1:].copy().view(np.float32)
        else:
            colors = words[:, 1:].copy().view(np.float64)
        nb = min(st.n_bad, bv_cap)
        return RenderResult(
            buffer=buf, counts=counts, colors=colors,
            xf_dist=np.array([xf[i] for i in range(n_ids)], dtype=np.uint64),
            s_iter=st.s_iter, s_plot=st.s_plot, n_bad=st.n_bad,
            success=bool(st.success),
            pt_min=np.array(st.pt_min[:self.dims]),
            pt_max=np.array(st.pt_max[:self.dims]),
            bad_xforms=np.array([bvx[i] for i in range(nb)], dtype=np.uint64),
            bad_points=np.array([[bvp[3 * i + k] for k in range(self.dims)]
                                 for i in range(nb)]).reshape(nb, self.dims))

    def apply_xform(self, xform: int, point, seed: int = 0, walker: int = 0,
                    draw: int = 0):
        """XForm::applyIteration on one point (sorted xform index, -1 = final)."""
        inp = (C.c_double * 3)(*([float(v) for v in point] + [0.0] * 3)[:3])
        out = (C.c_double * 3)()
        d = C.c_uint64(draw)
        rc = lib().orc_apply_xform(self._h, xform, inp, out, seed, walker,
                                   C.byref(d))
        if rc != 0:
            raise IndexError("no such xform")
        return [out[i] for i in range(self.dims)], d.value

    def eval_variation(self, xform: int, var: int, point, seed: int = 0,
                       walker: int = 0, draw: int = 0):
        """Evaluate one variation (sorted xform index, -1 = final)."""
        inp = (C.c_double * 3)(*([float(v) for v in point] + [0.0] * 3)[:3])
        out = (C.c_double * 3)()
        d = C.c_uint64(draw)
        rc = lib().orc_eval_variation(self._h, xform, var, inp, out, seed,
                                      walker, C.byref(d))
        if rc != 0:
            raise IndexError("no such variation")
        return [out[i] for i in range(self.dims)], d.value


def tonemap(buf: np.ndarray, width: int, height: int, cell_words: int, mode: int,
            gamma: float = 1.0, bits: int = 8):
    """ffr-img's image of a 2-D buffer (ffr_img.cpp:199-309); buf holds
    width*height*cell_words words (uint64 or uint32).  Returns (status,
    pixels or None, (hist_min, hist_max), (scale_min, scale_max)); status 0
    ok, 1 empty histogram, 2 rejected argument."""
    buf = np.ascontiguousarray(buf)
    cells = width * height
    assert buf.size == cells * cell_words
    wide = mode != 0 and bits == 16
    ch = 3 if mode == 2 else 1
    out = np.zeros(cells * ch, dtype=np.uint16 if wide else np.uint8)
    hist = (C.c_uint64 * 2)()
    scaler = (C.c_double * 2)()
    st = lib().orc_tonemap(buf.ctypes.data, cells, cell_words, buf.itemsize, mode,
                           gamma, bits, out.ctypes.data, hist, scaler)
    pix = out.reshape((height, width) + ((3,) if ch == 3 else ())) if st == 0 else None
    return st, pix, (hist[0], hist[1]), (scaler[0], scaler[1])


def stream_word(seed: int, walker: int, n: int) -> int:
    return lib().orc_stream_word(seed, walker, n)


def philox4x32_10(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return [o[i] for i in range(4)]


def isaac64_words(seed: int, n: int) -> list[int]:
    s = OrcIsaac64()
    lib().orc_isaac64_seed(C.byref(s), seed)
    return [lib().orc_isaac64_next(C.byref(s)) for _ in range(n)]


def isaac32_words(seed: int, n: int, seed_bits: int = 32) -> list[int]:
    s = OrcIsaac32()
    if seed_bits == 32:
        lib().orc_isaac32_seed_u32(C.byref(s), seed)
    else:
        lib().orc_isaac32_seed_u64(C.byref(s), seed)
    return [lib().orc_isaac32_next(C.byref(s)) for _ in range(n)]


def sort_order(weights, mode: int = F64_U64) -> list[int]:
    n = len(weights)
    w = (C.c_double * max(1, n))(*weights)
    o = (C.c_int32 * max(1, n))()
    lib().orc_sort_order(w, n, mode, o)
    return [o[i] for i in range(n)]
