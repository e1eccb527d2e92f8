"""Python binding of libffr_hip (the MI355X flame-buffer renderer).

`BufferRenderer` mirrors the reference's C++ class of the same name
(src/renderers/buffer_renderer.hpp:38-585): construct from a flame JSON,
`render(samples)`, `add_buffer`, `read_buffer`/`write_buffer`, and the
statistics getters.  The work runs in the HIP kernels of libffr_hip.so; this
module only marshals arguments over the C ABI (include/ffr.h).  There is no
CPU fallback: if the library is missing, loading fails with an error.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libffr_hip.so")
CLI_PATH = os.path.join(_HERE, "ffr-buf")

FFR_OK, FFR_E_ARG, FFR_E_FLAME, FFR_E_DEVICE = 0, 1, 2, 3
FFR_E_BADVALUE_LIMIT, FFR_E_SIZE, FFR_E_UNSUPPORTED = 4, 5, 6
FFR_FP64, FFR_FP32 = 0, 1
FFR_WORD64, FFR_WORD32 = 0, 1


class FfrError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


class Options(C.Structure):
    _fields_ = [
        ("precision", C.c_uint32),
        ("word", C.c_uint32),
        ("settle_iters", C.c_int32),
        ("device", C.c_int32),
        ("seed", C.c_uint64),
        ("batch_size", C.c_uint64),
        ("bad_value_limit", C.c_uint64),
        ("bad_value_cap", C.c_uint64),
        ("cell_cache", C.c_int32),
        ("cache_kib", C.c_int32),
        ("block", C.c_int32),
        ("binned", C.c_int32),
        ("compiled", C.c_int32),
        ("reserved", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("s_iter", C.c_uint64),
        ("s_plot", C.c_uint64),
        ("n_bad", C.c_uint64),
        ("walkers", C.c_uint64),
        ("batch_size", C.c_uint64),
        ("pt_min", C.c_double * 3),
        ("pt_max", C.c_double * 3),
        ("kernel_ms", C.c_double),
        ("atomics", C.c_uint64),
        ("flush_atomics", C.c_uint64),
        ("wave_per_walker", C.c_uint32),
        ("cache_slots", C.c_uint32),
        ("binned", C.c_uint32),
        ("compiled", C.c_uint32),
    ]


class BatchStats(C.Structure):
    _fields_ = [
        ("kernel_ms", C.c_double),
        ("samples", C.c_uint64),
        ("tasks", C.c_uint64),
        ("launches", C.c_uint32),
        ("graph", C.c_uint32),
    ]


class ServerOptions(C.Structure):
    _fields_ = [
        ("hbm_budget", C.c_uint64),
        ("pass_ms", C.c_double),
        ("device", C.c_int32),
        ("reserved", C.c_int32),
    ]


class JobInfo(C.Structure):
    _fields_ = [
        ("state", C.c_uint32),
        ("chunks_done", C.c_uint32),
        ("samples_done", C.c_uint64),
        ("bytes", C.c_uint64),
        ("batch_size", C.c_uint64),
        ("submit_s", C.c_double),
        ("admit_s", C.c_double),
        ("first_chunk_s", C.c_double),
        ("done_s", C.c_double),
        ("status", C.c_int32),
        ("reserved", C.c_int32),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64),
        ("done", C.c_uint64),
        ("failed", C.c_uint64),
        ("passes", C.c_uint64),
        ("samples", C.c_uint64),
        ("bytes_held", C.c_uint64),
        ("peak_bytes", C.c_uint64),
        ("queued", C.c_uint32),
        ("running", C.c_uint32),
        ("busy_ms", C.c_double),
    ]


class ImageInfo(C.Structure):
    _fields_ = [
        ("hist_min", C.c_uint64),
        ("hist_max", C.c_uint64),
        ("scale_min", C.c_double),
        ("scale_max", C.c_double),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("channels", C.c_uint32),
        ("bytes_per_channel", C.c_uint32),
        ("kernel_ms", C.c_double),
    ]


JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_FAILED = 0, 1, 2, 3
FFR_E_TIMEOUT = 7
FFR_E_EMPTY = 8
FFR_E_IO = 9

# ffr_read_fn / ffr_write_fn: size_t (*)(void *ctx, void *buf, size_t bytes)
READ_FN = C.CFUNCTYPE(C.c_size_t, C.c_void_p, C.c_void_p, C.c_size_t)
WRITE_FN = C.CFUNCTYPE(C.c_size_t, C.c_void_p, C.c_void_p, C.c_size_t)
IMG_MONO, IMG_GRAY, IMG_COLOR = 0, 1, 2


# every symbol include/ffr.h declares
EXPORTS = [
    "ffr_default_options", "ffr_create", "ffr_destroy", "ffr_last_error",
    "ffr_render", "ffr_render_async", "ffr_sync", "ffr_set_stream",
    "ffr_device_buffer", "ffr_buffer_cells", "ffr_cell_words", "ffr_word_bytes",
    "ffr_buffer_bytes", "ffr_dims", "ffr_color_dims", "ffr_read_buffer",
    "ffr_add_buffer", "ffr_add_buffer_noalloc", "ffr_write_buffer", "ffr_clear_buffer",
    "ffr_add_buffer_stream", "ffr_read_buffer_stream",
    "ffr_histogram_stats", "ffr_xform_distribution", "ffr_xform_id_count",
    "ffr_bad_values", "ffr_xform_order", "ffr_auto_batch", "ffr_check_flame",
    "ffr_histogram_l1", "ffr_choose_precision", "ffr_atomic_rate",
    "ffr_image_bytes", "ffr_render_image",
    "ffr_render_image_device",
    "ffr_batch_create", "ffr_batch_destroy", "ffr_batch_render", "ffr_batch_tune",
    "ffr_batch_set_graph", "ffr_batch_kinds", "ffr_batch_shape",
    "ffr_server_default_options", "ffr_server_create", "ffr_server_destroy",
    "ffr_server_submit", "ffr_server_job", "ffr_server_job_error", "ffr_server_wait",
    "ffr_server_read", "ffr_server_release", "ffr_server_get_stats",
]

_lib = None


def lib():
    """Load libffr_hip.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libffr_hip.so is not built ({LIB_PATH}); run "
            "`make -C flame-fractal-renderer_amd` or __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    L.ffr_default_options.argtypes = [C.POINTER(Options)]
    L.ffr_create.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(Options), C.POINTER(P)]
    L.ffr_destroy.argtypes = [P]
    L.ffr_last_error.restype = C.c_char_p
    L.ffr_render.argtypes = [P, C.c_uint64, C.POINTER(Stats)]
    L.ffr_render_async.argtypes = [P, C.c_uint64]
    L.ffr_sync.argtypes = [P, C.POINTER(Stats)]
    L.ffr_set_stream.argtypes = [P, P]
    L.ffr_device_buffer.restype = P
    L.ffr_device_buffer.argtypes = [P]
    L.ffr_buffer_cells.restype = C.c_uint64
    L.ffr_buffer_cells.argtypes = [P]
    L.ffr_cell_words.restype = C.c_uint32
    L.ffr_cell_words.argtypes = [P]
    L.ffr_word_bytes.restype = C.c_uint32
    L.ffr_word_bytes.argtypes = [P]
    L.ffr_buffer_bytes.restype = C.c_size_t
    L.ffr_buffer_bytes.argtypes = [P]
    L.ffr_dims.restype = C.c_int32
    L.ffr_dims.argtypes = [P]
    L.ffr_color_dims.restype = C.c_int32
    L.ffr_color_dims.argtypes = [P]
    for name in ("ffr_read_buffer", "ffr_add_buffer", "ffr_add_buffer_noalloc",
                 "ffr_write_buffer"):
        getattr(L, name).argtypes = [P, P, C.c_size_t]
    L.ffr_clear_buffer.argtypes = [P]
    L.ffr_add_buffer_stream.argtypes = [P, READ_FN, P]
    L.ffr_read_buffer_stream.argtypes = [P, WRITE_FN, P]
    L.ffr_histogram_stats.argtypes = [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                      C.POINTER(C.c_uint64)]
    L.ffr_xform_distribution.argtypes = [P, C.POINTER(C.c_uint64), C.c_size_t]
    L.ffr_xform_id_count.restype = C.c_size_t
    L.ffr_xform_id_count.argtypes = [P]
    L.ffr_bad_values.argtypes = [P, C.POINTER(C.c_uint64), C.POINTER(C.c_double),
                                 C.c_size_t, C.POINTER(C.c_size_t)]
    L.ffr_xform_order.argtypes = [P, C.POINTER(C.c_int32), C.POINTER(C.c_double),
                                  C.c_size_t]
    L.ffr_auto_batch.restype = C.c_uint64
    L.ffr_auto_batch.argtypes = [P, C.c_uint64]
    L.ffr_check_flame.argtypes = [C.c_char_p, C.c_size_t, C.c_uint32,
                                  C.POINTER(C.c_int32), C.POINTER(C.c_double),
                                  C.c_size_t, C.POINTER(C.c_size_t)]
    L.ffr_histogram_l1.argtypes = [P, P, C.POINTER(C.c_double)]
    L.ffr_atomic_rate.argtypes = [C.c_int32, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64,
                                  C.POINTER(C.c_double)]
    L.ffr_image_bytes.restype = C.c_size_t
    L.ffr_image_bytes.argtypes = [P, C.c_int, C.c_int]
    L.ffr_render_image.argtypes = [P, C.c_int, C.c_double, C.c_int, P, C.c_size_t,
                                   C.POINTER(ImageInfo)]
    L.ffr_render_image_device.argtypes = [P, C.c_int, C.c_double, C.c_int, P, C.c_size_t,
                                          C.POINTER(ImageInfo)]
    L.ffr_choose_precision.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(Options), C.c_uint64,
                                       C.c_double, C.POINTER(C.c_uint32),
                                       C.POINTER(C.c_double)]
    L.ffr_batch_create.argtypes = [C.POINTER(P), C.c_size_t, C.POINTER(P)]
    L.ffr_batch_destroy.argtypes = [P]
    L.ffr_batch_destroy.restype = None
    L.ffr_batch_render.argtypes = [P, C.POINTER(C.c_uint64), C.POINTER(C.c_int),
                                   C.POINTER(BatchStats)]
    L.ffr_batch_tune.argtypes = [P, C.POINTER(C.c_uint64), C.c_double]
    L.ffr_batch_set_graph.argtypes = [P, C.c_int]
    L.ffr_batch_kinds.restype = C.c_size_t
    L.ffr_batch_kinds.argtypes = [P]
    L.ffr_batch_shape.argtypes = [P, C.c_size_t, C.POINTER(C.c_int32)]
    L.ffr_server_default_options.argtypes = [C.POINTER(ServerOptions)]
    L.ffr_server_default_options.restype = None
    L.ffr_server_create.argtypes = [C.POINTER(ServerOptions), C.POINTER(P)]
    L.ffr_server_destroy.argtypes = [P]
    L.ffr_server_destroy.restype = None
    L.ffr_server_submit.argtypes = [P, C.c_char_p, C.c_size_t, C.POINTER(Options), C.c_uint64,
                                    C.c_uint32, C.c_void_p, C.c_size_t,
                                    C.POINTER(C.c_uint64)]
    L.ffr_server_job.argtypes = [P, C.c_uint64, C.POINTER(JobInfo)]
    L.ffr_server_job_error.argtypes = [P, C.c_uint64, C.c_char_p, C.c_size_t]
    L.ffr_server_wait.argtypes = [P, C.c_uint64, C.c_double]
    L.ffr_server_read.argtypes = [P, C.c_uint64, C.c_void_p, C.c_size_t]
    L.ffr_server_release.argtypes = [P, C.c_uint64]
    L.ffr_server_get_stats.argtypes = [P, C.POINTER(ServerStats)]
    for name in EXPORTS:
        if not hasattr(L, name):
            raise RuntimeError(f"libffr_hip.so does not export {name}")
    # functions not given a restype above return int (the ctypes default)
    L.ffr_default_options.restype = None
    L.ffr_destroy.restype = None
    _lib = L
    return L


def _check(rc: int):
    if rc != FFR_OK:
        raise FfrError(rc, lib().ffr_last_error().decode(errors="replace"))


def check_flame(text: str, precision: str = "fp64"):
    """Validate a flame without a device; returns (xform ids, cum weights)."""
    L = lib()
    b = text.encode()
    n = C.c_size_t(0)
    cap = 4096
    ids = (C.c_int32 * cap)()
    cw = (C.c_double * cap)()
    prec = FFR_FP32 if precision == "fp32" else FFR_FP64
    _check(L.ffr_check_flame(b, len(b), prec, ids, cw, cap, C.byref(n)))
    return [ids[i] for i in range(n.value)], [cw[i] for i in range(n.value)]


def atomic_rate(cells: int, cell_words: int = 4, word_bytes: int = 8,
                updates: int = 1 << 28, device: int = 0) -> float:
    """Global atomic adds per second this device sustains with cell updates
    scattered uniformly over `cells` cells (the atomic roofline)."""
    out = C.c_double()
    _check(lib().ffr_atomic_rate(device, cells, cell_words, word_bytes, updates, C.byref(out)))
    return out.value


class DeviceMemory:
    """Device memory from the HIP runtime libffr_hip.so itself uses (one
    runtime per process: a second HIP runtime, such as the one a PyTorch
    wheel bundles, may not see the device once this one holds it)."""

    _hip = None

    @classmethod
    def _runtime(cls):
        if cls._hip is None:
            lib()
            h = C.CDLL("libamdhip64.so.7")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipFree.argtypes = [C.c_void_p]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
            cls._hip = h
        return cls._hip

    def __init__(self, nbytes: int, fill: int | None = None):
        self.nbytes = nbytes
        p = C.c_void_p()
        if self._runtime().hipMalloc(C.byref(p), max(nbytes, 1)) != 0:
            raise FfrError(3, f"hipMalloc of {nbytes} bytes failed")
        self.ptr = p.value
        if fill is not None and self._runtime().hipMemset(C.c_void_p(self.ptr), fill,
                                                          nbytes) != 0:
            raise FfrError(3, "hipMemset failed")

    def to_host(self, dtype=np.uint8) -> np.ndarray:
        out = np.empty(self.nbytes, dtype=np.uint8)
        if self._runtime().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr),
                                     self.nbytes, 2) != 0:  # device to host
            raise FfrError(3, "hipMemcpy failed")
        return out.view(dtype)

    @classmethod
    def cu_count(cls, device: int = 0) -> int:
        """compute units of a device (hipDeviceAttributeMultiprocessorCount)"""
        h = cls._runtime()
        v = C.c_int()
        h.hipDeviceGetAttribute.argtypes = [C.POINTER(C.c_int), C.c_int, C.c_int]
        # hipDeviceAttributeMultiprocessorCount
        if h.hipDeviceGetAttribute(C.byref(v), 63, device) != 0:
            raise FfrError(3, "hipDeviceGetAttribute failed")
        return v.value

    def close(self):
        if getattr(self, "ptr", None):
            self._runtime().hipFree(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        self.close()


class BufferRenderer:
    """A flame and its device buffer (reference BufferRenderer<dims>)."""

    def __init__(self, flame_json: str, precision: str = "fp64", word: int = 64,
                 seed: int = 0, batch_size: int = 0, bad_value_limit: int = 256,
                 settle: int = -1, device: int = 0, bad_value_cap: int = 4096,
                 cell_cache: int = -1, cache_kib: int = 0, block: int = 0,
                 binned: int = -1, compiled: int = -1):
        L = lib()
        o = Options()
        L.ffr_default_options(C.byref(o))
        o.precision = FFR_FP32 if precision == "fp32" else FFR_FP64
        o.word = FFR_WORD32 if word == 32 else FFR_WORD64
        o.settle_iters = settle
        o.device = device
        o.seed = seed
        o.batch_size = batch_size
        o.bad_value_limit = bad_value_limit
        o.bad_value_cap = bad_value_cap
        o.cell_cache = cell_cache
        o.cache_kib = cache_kib
        o.block = block
        o.binned = binned
        o.compiled = compiled
        h = C.c_void_p()
        b = flame_json.encode()
        _check(L.ffr_create(b, len(b), C.byref(o), C.byref(h)))
        self._h = h
        self.precision = precision
        self.dims = L.ffr_dims(h)
        self.color_dims = L.ffr_color_dims(h)
        self.cells = L.ffr_buffer_cells(h)
        self.cell_words = L.ffr_cell_words(h)
        self.word_bytes = L.ffr_word_bytes(h)
        self.last_stats = Stats()

    def close(self):
        if getattr(self, "_h", None):
            lib().ffr_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # --- rendering (buffer_renderer.hpp:269-372) ---
    def render(self, samples: int) -> bool:
        """Render; False when the bad-value limit stopped it."""
        rc = lib().ffr_render(self._h, samples, C.byref(self.last_stats))
        if rc == FFR_E_BADVALUE_LIMIT:
            return False
        _check(rc)
        return True

    def render_async(self, samples: int):
        _check(lib().ffr_render_async(self._h, samples))

    def sync(self) -> bool:
        rc = lib().ffr_sync(self._h, C.byref(self.last_stats))
        if rc == FFR_E_BADVALUE_LIMIT:
            return False
        _check(rc)
        return True

    def set_stream(self, stream_ptr: int | None):
        _check(lib().ffr_set_stream(self._h, C.c_void_p(stream_ptr or 0)))

    def device_buffer(self) -> int:
        return lib().ffr_device_buffer(self._h)

    def auto_batch(self, samples: int) -> int:
        return lib().ffr_auto_batch(self._h, samples)

    # --- buffer (buffer_renderer.hpp:375-480) ---
    @property
    def buffer_bytes(self) -> int:
        return lib().ffr_buffer_bytes(self._h)

    def _dtype(self):
        return np.uint32 if self.word_bytes == 4 else np.uint64

    def read_buffer(self) -> np.ndarray:
        out = np.empty(self.cells * self.cell_words, dtype=self._dtype())
        _check(lib().ffr_read_buffer(self._h, out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out

    def write_buffer(self, words: np.ndarray):
        w = np.ascontiguousarray(words, dtype=self._dtype())
        _check(lib().ffr_write_buffer(self._h, w.ctypes.data_as(C.c_void_p), w.nbytes))

    def add_buffer_stream(self, f):
        """addBuffer(std::istream&): add a buffer read from a binary file
        object in chunks, streamed to the device"""
        def rd(_ctx, dst, n):
            view = memoryview((C.c_char * n).from_address(dst)).cast("B")
            if hasattr(f, "readinto"):
                got = f.readinto(view)
                return got or 0
            data = f.read(n)
            view[: len(data)] = data
            return len(data)
        _check(lib().ffr_add_buffer_stream(self._h, READ_FN(rd), None))

    def write_buffer_stream(self, f):
        """writeBuffer(std::ostream&): write the buffer to a binary file
        object in chunks, streamed from the device"""
        def wr(_ctx, src, n):
            try:
                # a view of the pinned chunk, no copy on this side
                f.write(memoryview((C.c_char * n).from_address(src)).cast("B"))
            except OSError:
                return 0
            return n
        _check(lib().ffr_read_buffer_stream(self._h, WRITE_FN(wr), None))

    def add_buffer(self, words: np.ndarray):
        w = np.ascontiguousarray(words, dtype=self._dtype())
        _check(lib().ffr_add_buffer(self._h, w.ctypes.data_as(C.c_void_p), w.nbytes))

    def add_buffer_noalloc(self, words: np.ndarray):
        """add_buffer staged through the library's fixed device array (no
        device allocation; calls overlapping on one GPU race)"""
        w = np.ascontiguousarray(words, dtype=self._dtype())
        _check(lib().ffr_add_buffer_noalloc(self._h, w.ctypes.data_as(C.c_void_p), w.nbytes))

    def clear_buffer(self):
        _check(lib().ffr_clear_buffer(self._h))

    def counts(self, words: np.ndarray | None = None) -> np.ndarray:
        w = self.read_buffer() if words is None else words
        return w.reshape(self.cells, self.cell_words)[:, 0].copy()

    def colors(self, words: np.ndarray | None = None) -> np.ndarray:
        w = self.read_buffer() if words is None else words
        cols = w.reshape(self.cells, self.cell_words)[:, 1:].copy()
        return cols.view(np.float32 if self.word_bytes == 4 else np.float64)

    def histogram_stats(self):
        s, mn, mx = C.c_uint64(), C.c_uint64(), C.c_uint64()
        _check(lib().ffr_histogram_stats(self._h, C.byref(s), C.byref(mn), C.byref(mx)))
        return s.value, mn.value, mx.value

    def image_bytes(self, mode: int = IMG_GRAY, bits: int = 8) -> int:
        return int(lib().ffr_image_bytes(self._h, mode, bits))

    def render_image(self, mode: int = IMG_GRAY, gamma: float = 1.0, bits: int = 8):
        """The image ffr-img makes of this buffer (ffr_img.cpp:199-309), on the
        device: an array (height, width) or (height, width, 3) of uint8 or
        uint16 (mono is always uint8), and the bounds ffr-img prints."""
        info = ImageInfo()
        n = lib().ffr_image_bytes(self._h, mode, bits)
        wide = mode != IMG_MONO and bits == 16
        out = np.zeros(max(n // (2 if wide else 1), 1), dtype=np.uint16 if wide else np.uint8)
        _check(lib().ffr_render_image(self._h, mode, float(gamma), bits,
                                      out.ctypes.data if n else None, n, C.byref(info)))
        shape = (info.height, info.width) + ((3,) if info.channels == 3 else ())
        return out[: n // out.itemsize].reshape(shape), info

    def render_image_device(self, device_ptr: int, nbytes: int, mode: int = IMG_GRAY,
                            gamma: float = 1.0, bits: int = 8):
        """render_image into device memory (e.g. a torch tensor's data_ptr())"""
        info = ImageInfo()
        _check(lib().ffr_render_image_device(self._h, mode, float(gamma), bits,
                                             C.c_void_p(device_ptr), nbytes, C.byref(info)))
        return info

    # --- statistics getters (buffer_renderer.hpp:534-564) ---
    @property
    def samples_iterated(self) -> int:
        return self.last_stats.s_iter

    @property
    def samples_plotted(self) -> int:
        return self.last_stats.s_plot

    @property
    def point_extremes(self):
        return [(self.last_stats.pt_min[i], self.last_stats.pt_max[i])
                for i in range(self.dims)]

    @property
    def kernel_ms(self) -> float:
        return self.last_stats.kernel_ms

    def xform_distribution(self) -> np.ndarray:
        n = lib().ffr_xform_id_count(self._h)
        out = (C.c_uint64 * max(1, n))()
        _check(lib().ffr_xform_distribution(self._h, out, n))
        return np.array([out[i] for i in range(n)], dtype=np.uint64)

    def bad_values(self):
        n = C.c_size_t(0)
        lib().ffr_bad_values(self._h, None, None, 0, C.byref(n))
        k = n.value
        ids = (C.c_uint64 * max(1, k))()
        pts = (C.c_double * max(1, k * self.dims))()
        _check(lib().ffr_bad_values(self._h, ids, pts, k, C.byref(n)))
        return (np.array([ids[i] for i in range(k)], dtype=np.uint64),
                np.array([pts[i] for i in range(k * self.dims)]).reshape(k, self.dims))

    def xform_order(self):
        n = len(self.xform_distribution())
        ids = (C.c_int32 * max(1, n))()
        cw = (C.c_double * max(1, n))()
        # the number of xforms kept is at most the number in the file
        for m in range(n, 0, -1):
            if lib().ffr_xform_order(self._h, ids, cw, m) == FFR_OK:
                return [ids[i] for i in range(m)], [cw[i] for i in range(m)]
        raise FfrError(FFR_E_SIZE, "no xforms")


def histogram_l1(a: "BufferRenderer", b: "BufferRenderer") -> float:
    """L1 distance of two renderers' normalised count histograms (device)."""
    out = C.c_double()
    _check(lib().ffr_histogram_l1(a._h, b._h, C.byref(out)))
    return out.value


# the precision policy's bar: L1 between the FP32 and FP64 histograms of a
# full render; the share of its samples the probe renders; the margin the
# probe's estimate keeps below the bar (the estimate is itself a sample)
POLICY_L1_TARGET = 1e-3
POLICY_PROBE_FRACTION = 1.0 / 64
POLICY_MARGIN = 0.5
# the batched policy's second stage: the share of the target its estimate
# of the full render's L1 must stay below
POLICY_STAGE2_MARGIN = 0.9


def choose_precision(flame_json: str, samples: int, seed: int = 0, batch_size: int = 0,
                     settle: int = -1, target: float = POLICY_L1_TARGET,
                     fraction: float = POLICY_PROBE_FRACTION):
    """The precision policy for a render of `samples` samples: ("fp32" |
    "fp64", probe L1).  The probe renders `fraction` of the samples in FP32
    and in FP64 on the same streams, with the walker length of the full
    render (drift grows along a walker).  Where the trajectories drift
    apart, the drifted samples are independent draws, and the L1 they add
    scales as one over the square root of the sample count; so the probe's
    bar is margin * target / sqrt(fraction)."""
    L = lib()
    o = Options()
    L.ffr_default_options(C.byref(o))
    o.seed = seed
    o.batch_size = batch_size or L.ffr_auto_batch(None, int(samples))
    o.settle_iters = settle
    probe = max(int(samples * fraction), int(o.batch_size))
    threshold = POLICY_MARGIN * target / (min(1.0, probe / max(samples, 1)) ** 0.5)
    prec = C.c_uint32()
    d = C.c_double()
    b = flame_json.encode()
    _check(L.ffr_choose_precision(b, len(b), C.byref(o), probe, float(threshold),
                                  C.byref(prec), C.byref(d)))
    return ("fp32" if prec.value == FFR_FP32 else "fp64"), d.value


def choose_precisions(flame_jsons, samples, seeds, target: float = POLICY_L1_TARGET,
                      fraction: float = POLICY_PROBE_FRACTION, settle: int = 53):
    """choose_precision for many flames at once: every probe pair (FP32 and
    FP64 on the same streams) renders in one batched launch.  Returns
    ([("fp32" | "fp64")], [probe L1]).

    Two stages.  The first probe decides where it is clear: FP64 above the
    bar of choose_precision, FP32 at or below the target itself (the full
    render's L1 is at most about the probe's).  Between the two, the L1 may
    hold a part that shrinks with more samples (drifted walkers are
    independent draws) and a part that does not (FP32 moves the attractor).
    For those flames the same renderers render three times the probe again
    (their streams continue, so the sum is a probe four times larger), and
    L1(s) = b + k sqrt(s1 / s) is fitted through the two points:
    k = 2 (L1_1 - L1_2), b = L1_1 - k.  The full render's estimate
    b + k sqrt(s1 / n) must stay below POLICY_STAGE2_MARGIN * target."""
    L = lib()
    rs, bars = [], []
    try:
        for text, n, seed in zip(flame_jsons, samples, seeds):
            batch = int(L.ffr_auto_batch(None, int(n)))
            probe = max(int(n * fraction), batch)
            bars.append((probe, POLICY_MARGIN * target / (min(1.0, probe / max(n, 1)) ** 0.5)))
            for prec in ("fp32", "fp64"):
                rs.append(BufferRenderer(text, precision=prec, seed=seed, batch_size=batch,
                                         settle=settle))
        with Batch(rs) as b:
            ok = b.render([p for p, _ in bars for _ in range(2)])
            out, dist, second = [], [], []
            for i, (_, bar) in enumerate(bars):
                d = histogram_l1(rs[2 * i], rs[2 * i + 1])
                escapes = (not ok[2 * i]) and ok[2 * i + 1]
                if escapes or d > bar:
                    out.append("fp64")
                elif d <= target:
                    out.append("fp32")
                else:
                    out.append(None)
                    second.append(i)
                dist.append(d)
            if second:
                more = [0] * len(rs)
                for i in second:
                    more[2 * i] = more[2 * i + 1] = 3 * bars[i][0]
                ok2 = b.render(more)
                for i in second:
                    d2 = histogram_l1(rs[2 * i], rs[2 * i + 1])
                    k = 2.0 * (dist[i] - d2)
                    est = (dist[i] - k) + k * (bars[i][0] / max(samples[i], 1)) ** 0.5
                    escapes = (not ok2[2 * i]) and ok2[2 * i + 1]
                    out[i] = ("fp32" if (not escapes and est <= POLICY_STAGE2_MARGIN * target)
                              else "fp64")
        return out, dist
    finally:
        for r in rs:
            r.close()


class Batch:
    """Several BufferRenderers rendered together: one persistent launch per
    kernel kind, replayed as a HIP graph.  Each renderer renders exactly
    what its own render() would with the same samples.  No counterpart in
    the reference, which renders one flame per process."""

    def __init__(self, renderers):
        self.renderers = list(renderers)
        arr = (C.c_void_p * len(self.renderers))(*[r._h for r in self.renderers])
        h = C.c_void_p()
        _check(lib().ffr_batch_create(arr, len(self.renderers), C.byref(h)))
        self._h = h
        self.last_stats = BatchStats()

    def close(self):
        if getattr(self, "_h", None):
            lib().ffr_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _samples(self, samples):
        if len(samples) != len(self.renderers):
            raise ValueError("one sample count per renderer")
        return (C.c_uint64 * len(samples))(*[int(s) for s in samples])

    def render(self, samples) -> list[bool]:
        """Render samples[i] on renderer i; per renderer, False when its
        bad-value limit stopped it."""
        st = (C.c_int * len(self.renderers))()
        rc = lib().ffr_batch_render(self._h, self._samples(samples), st,
                                    C.byref(self.last_stats))
        if rc not in (FFR_OK, FFR_E_BADVALUE_LIMIT):
            _check(rc)
        for r in self.renderers:
            lib().ffr_sync(r._h, C.byref(r.last_stats))
        return [v == FFR_OK for v in st]

    def tune(self, samples, fraction: float = 1.0 / 32):
        _check(lib().ffr_batch_tune(self._h, self._samples(samples), float(fraction)))

    def set_graph(self, on: bool):
        _check(lib().ffr_batch_set_graph(self._h, 1 if on else 0))

    def shapes(self):
        out = []
        for k in range(lib().ffr_batch_kinds(self._h)):
            v = (C.c_int32 * 4)()
            _check(lib().ffr_batch_shape(self._h, k, v))
            out.append(tuple(v))
        return out


class Server:
    """A render server on one device (no counterpart in the reference):
    jobs are admitted within an HBM budget, keep their buffers on the device
    across `chunks` increments, and share batched passes run by a worker
    thread; each renders what BufferRenderer.render would, chunk by chunk."""

    def __init__(self, hbm_budget: int = 0, pass_ms: float = 0.0, device: int = 0):
        L = lib()
        o = ServerOptions()
        L.ffr_server_default_options(C.byref(o))
        if hbm_budget:
            o.hbm_budget = hbm_budget
        if pass_ms:
            o.pass_ms = pass_ms
        o.device = device
        h = C.c_void_p()
        _check(L.ffr_server_create(C.byref(o), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().ffr_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, flame_json: str, samples: int, chunks: int = 4, precision: str = "fp32",
               word: int = 64, seed: int = 0, batch_size: int = 0, settle: int = -1,
               bad_value_limit: int = 256, init: np.ndarray | None = None) -> int:
        L = lib()
        o = Options()
        L.ffr_default_options(C.byref(o))
        o.precision = FFR_FP32 if precision == "fp32" else FFR_FP64
        o.word = FFR_WORD32 if word == 32 else FFR_WORD64
        o.seed = seed
        o.batch_size = batch_size
        o.settle_iters = settle
        o.bad_value_limit = bad_value_limit
        b = flame_json.encode()
        job = C.c_uint64()
        if init is not None:
            init = np.ascontiguousarray(init)
            ptr, n = init.ctypes.data_as(C.c_void_p), init.nbytes
        else:
            ptr, n = None, 0
        _check(L.ffr_server_submit(self._h, b, len(b), C.byref(o), int(samples), int(chunks),
                                   ptr, n, C.byref(job)))
        return job.value

    def job(self, job: int) -> JobInfo:
        info = JobInfo()
        _check(lib().ffr_server_job(self._h, job, C.byref(info)))
        return info

    def job_error(self, job: int) -> str:
        buf = C.create_string_buffer(1024)
        _check(lib().ffr_server_job_error(self._h, job, buf, len(buf)))
        return buf.value.decode()

    def wait(self, job: int, timeout: float = 600.0) -> JobInfo:
        rc = lib().ffr_server_wait(self._h, job, float(timeout))
        if rc == FFR_E_TIMEOUT:
            raise TimeoutError(f"job {job} did not finish in {timeout} s")
        _check(rc)
        return self.job(job)

    def read(self, job: int, word: int = 64) -> np.ndarray:
        info = self.job(job)
        out = np.empty(info.bytes // (word // 8), dtype=np.uint64 if word == 64 else np.uint32)
        _check(lib().ffr_server_read(self._h, job, out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out

    def release(self, job: int):
        _check(lib().ffr_server_release(self._h, job))

    def stats(self) -> ServerStats:
        st = ServerStats()
        _check(lib().ffr_server_get_stats(self._h, C.byref(st)))
        return st
