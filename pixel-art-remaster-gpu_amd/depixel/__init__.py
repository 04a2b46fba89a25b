"""ctypes binding of libdepixel.so, the MI355X depixelization path.

The library is built in-tree (``make -C pixel-art-remaster-gpu_amd``); this
module refuses to import without it, so nothing silently falls back to a CPU
path.  Device buffers are plain pointers: the helpers below take torch HIP
tensors (``tensor.data_ptr()``) and torch streams (``stream.cuda_stream``).
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_bool, c_char_p, c_double, c_float, c_int, c_uint32, c_uint64, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "build", "libdepixel.so")
HEADER_PATH = os.path.join(os.path.dirname(os.path.dirname(_HERE)), "include", "depixel.h")

CELL_SIZE = 45
RGB, TOPDOWN, ALPHA, FLIP_OUT, INDEXED = 1, 2, 4, 8, 16
CLASS_SPRITE, CLASS_FRAME, CLASS_LARGE = 0, 1, 2
WALK_AUTO, WALK_ALWAYS, WALK_NEVER = 0, 1, 2
TILE_WIDTHS = (16, 32, 64)
VIEW_CELLS, VIEW_QUADS = 0, 1
VIEW_AA, VIEW_TOPDOWN = 1, 2

OK, EINVAL, ENOMEM, EHIP, ENODEV, EBUDGET, ETIMEOUT, ECANCELLED, ECLOSED, EINJECTED = (
    0, -1, -2, -3, -4, -5, -6, -7, -8, -9)


class Point(ctypes.Structure):
    _fields_ = [("x", c_float), ("y", c_float)]


class Params(ctypes.Structure):
    _fields_ = [("scale", c_int), ("subdivide", c_int), ("flags", c_uint32), ("reserved", c_int)]


class Item(ctypes.Structure):
    _fields_ = [("in_offset", c_uint64), ("out_offset", c_uint64), ("width", c_int), ("height", c_int),
                ("stride", c_int), ("channels", c_int), ("scale", c_int), ("subdivide", c_int),
                ("flags", c_uint32), ("out_stride", c_int), ("palette_offset", c_uint64)]


class View(ctypes.Structure):
    """dpx_view: the GLUT window a frame is presented in (defaults: the reference's)."""
    _fields_ = [("mode", c_int), ("scale_x100", c_int), ("translate_x", c_int), ("translate_y", c_int),
                ("width", c_int), ("height", c_int), ("flags", c_uint32), ("reserved", c_int)]


class BatchTimes(ctypes.Structure):
    _fields_ = [("graph_ms", c_double), ("raster_ms", c_double), ("batches", c_uint64)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", c_uint64), ("completed", c_uint64), ("failed", c_uint64), ("rejected", c_uint64),
                ("cancelled", c_uint64), ("batches", c_uint64), ("frames_per_s", c_double), ("p50_ms", c_double),
                ("p99_ms", c_double), ("mean_batch", c_double), ("hbm_in_use", c_uint64), ("hbm_peak", c_uint64),
                ("hbm_budget", c_uint64)]


class DepixelError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"depixel error {code}: {msg}")
        self.code = code


if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built; run `make -C pixel-art-remaster-gpu_amd` "
                      "(or __graft_entry__.build()) first")

lib = ctypes.CDLL(LIB_PATH)

_u8p = POINTER(ctypes.c_uint8)
_SIGS = {
    "dpx_last_error": (c_char_p, []),
    "dpx_version": (c_char_p, []),
    "launch_kernel": (c_void_p, [c_void_p, c_void_p, c_float, c_void_p, c_int, c_int, c_int, c_void_p, c_void_p,
                                 c_bool]),
    "dpx_frame": (c_int, [c_void_p, c_int, c_int, c_int, c_int, POINTER(Params), c_void_p, c_int, c_void_p,
                          c_void_p, c_void_p]),
    "dpx_stage_similarity_graph": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_uint32, c_void_p, c_void_p]),
    "dpx_stage_trivial_cross": (c_int, [c_void_p, c_int, c_int, c_void_p]),
    "dpx_stage_heuristic_cross": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p]),
    "dpx_stage_cells": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "dpx_stage_subdivide": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_uint32, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_void_p, c_void_p]),
    "dpx_stage_triangulate": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p]),
    "dpx_stage_export_vbo": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_uint32, c_void_p, c_void_p, c_void_p,
                                     c_void_p, c_void_p]),
    "dpx_stage_rasterize": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_uint32, c_void_p, c_void_p, c_int,
                                    c_void_p, c_int, c_void_p]),
    "dpx_stage_present": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_uint32, c_void_p, c_void_p, POINTER(View),
                                  c_void_p, c_int, c_void_p]),
    "dpx_stage_components": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "dpx_stage_border_chains": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_int, c_void_p, c_int,
                                        c_void_p, c_void_p]),
    "dpx_context_create": (c_int, [c_int, POINTER(c_void_p)]),
    "dpx_context_destroy": (None, [c_void_p]),
    "dpx_batch_run": (c_int, [c_void_p, c_void_p, c_void_p, POINTER(Item), c_int, c_void_p]),
    "dpx_batch_run_noalloc": (c_int, [c_void_p, c_void_p, POINTER(Item), c_int, c_void_p]),
    "dpx_batch_graph": (c_int, [c_void_p, POINTER(c_void_p), POINTER(POINTER(c_uint64))]),
    "dpx_context_keep_graph": (c_int, [c_void_p, c_int]),
    "dpx_context_set_tile": (c_int, [c_void_p, c_int, c_int, c_int, c_int]),
    "dpx_context_get_tile": (c_int, [c_void_p, c_int, c_int, c_int, POINTER(c_int)]),
    "dpx_context_set_walk": (c_int, [c_void_p, c_int, c_int, c_int, c_int]),
    "dpx_context_get_walk": (c_int, [c_void_p, c_int, c_int, c_int, POINTER(c_int)]),
    "dpx_context_reserve": (c_int, [c_void_p, c_int, ctypes.c_int64, c_int]),
    "dpx_batch_timing": (c_int, [c_void_p, c_int]),
    "dpx_plan_create": (c_int, [c_void_p, POINTER(Item), c_int, c_uint64, c_uint32, POINTER(c_void_p)]),
    "dpx_plan_run": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "dpx_plan_graph": (c_int, [c_void_p, POINTER(c_void_p), POINTER(POINTER(c_uint64))]),
    "dpx_plan_destroy": (None, [c_void_p]),
    "dpx_batch_timing_read": (c_int, [c_void_p, POINTER(BatchTimes)]),
    "dpx_server_create": (c_int, [c_int, c_uint64, c_int, POINTER(c_void_p)]),
    "dpx_server_destroy": (None, [c_void_p]),
    "dpx_server_open_stream": (c_int, [c_void_p, c_int, c_int, c_int, POINTER(Params), POINTER(c_int)]),
    "dpx_server_close_stream": (c_int, [c_void_p, c_int]),
    "dpx_server_set_palette": (c_int, [c_void_p, c_int, c_void_p]),
    "dpx_server_submit": (c_int, [c_void_p, c_int, c_void_p, c_int, c_void_p, c_int, POINTER(c_uint64)]),
    "dpx_server_wait": (c_int, [c_void_p, c_uint64, c_int]),
    "dpx_server_wait_timed": (c_int, [c_void_p, c_uint64, c_int, POINTER(c_double)]),
    "dpx_server_cancel": (c_int, [c_void_p, c_uint64]),
    "dpx_server_stats_get": (c_int, [c_void_p, POINTER(ServerStats)]),
    "dpx_server_inject_failure": (c_int, [c_void_p, c_uint64]),
    "dpx_server_stats_reset": (c_int, [c_void_p]),
    "dpx_host_alloc": (c_int, [ctypes.c_size_t, POINTER(c_void_p)]),
    "dpx_host_free": (c_int, [c_void_p]),
    "dpx_synth_frame": (c_int, [c_int, c_uint64, c_int, c_int, c_int, c_void_p, c_int]),
    "dpx_synth_sprite_size": (c_int, [c_uint64, c_int, POINTER(c_int), POINTER(c_int)]),
    "dpx_synth_stream_frame": (c_int, [c_uint64, c_int, c_int, c_int, c_void_p, c_int, c_void_p]),
}
for _name, (_res, _args) in _SIGS.items():
    if hasattr(lib, _name):
        _f = getattr(lib, _name)
        _f.restype = _res
        _f.argtypes = _args

libc = ctypes.CDLL(None)
libc.free.argtypes = [c_void_p]


def last_error() -> str:
    return (lib.dpx_last_error() or b"").decode(errors="replace")


def check(rc: int) -> int:
    if rc != OK:
        raise DepixelError(rc, last_error())
    return rc


def _ptr(a) -> int:
    """Address of a numpy array or a torch tensor (host or device)."""
    if a is None:
        return 0
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"]
        return a.ctypes.data
    return a.data_ptr()


def _stream(s) -> int:
    if s is None:
        return 0
    return getattr(s, "cuda_stream", s)


# --------------------------------------------------------------------------
# synthetic inputs
# --------------------------------------------------------------------------
def synth_frame(kind: int, seed: int, width: int, height: int, channels: int = 3,
                stride: int | None = None) -> np.ndarray:
    """Deterministic synthetic frame as a (height, stride) uint8 array."""
    if stride is None:
        stride = (width * channels + 3) // 4 * 4
    buf = np.zeros((height, stride), dtype=np.uint8)
    check(lib.dpx_synth_frame(kind, seed, width, height, channels, _ptr(buf), stride))
    return buf


def pinned_empty(shape, dtype=np.uint8) -> np.ndarray:
    """A numpy array in pinned host memory (dpx_host_alloc), freed with the
    last array that views it.  The server moves frames and results in such
    arrays by DMA alone."""
    import weakref
    dt = np.dtype(dtype)
    n = int(np.prod(shape)) * dt.itemsize
    p = c_void_p()
    check(lib.dpx_host_alloc(n, ctypes.byref(p)))
    buf = (ctypes.c_uint8 * n).from_address(p.value)
    fin = weakref.finalize(buf, lib.dpx_host_free, p.value)
    fin.atexit = False  # at interpreter exit the process releases it
    return np.frombuffer(buf, dtype=dt).reshape(shape)


def stream_frame(seed: int, t: int, width: int = 320, height: int = 224):
    """Frame t of synthetic game stream `seed`: (indices (height, width) uint8, palette (256, 4) uint8)."""
    idx = np.zeros((height, width), dtype=np.uint8)
    pal = np.zeros((256, 4), dtype=np.uint8)
    check(lib.dpx_synth_stream_frame(seed, t, width, height, _ptr(idx), width, _ptr(pal)))
    return idx, pal


def expand_indexed(idx: np.ndarray, pal: np.ndarray) -> np.ndarray:
    """The 4-channel frame of an indexed frame's palette entries, (H, 4*W) uint8."""
    return np.ascontiguousarray(pal[idx].reshape(idx.shape[0], idx.shape[1] * 4))


def sprite_size(seed: int, index: int) -> tuple[int, int]:
    w, h = c_int(), c_int()
    check(lib.dpx_synth_sprite_size(seed, index, ctypes.byref(w), ctypes.byref(h)))
    return w.value, h.value


# --------------------------------------------------------------------------
# plain per-frame path
# --------------------------------------------------------------------------
def frame(img: np.ndarray, width: int, height: int, channels: int = 3, scale: int = 4, subdivide: bool = False,
          flags: int = 0, want_diagram: bool = False):
    """Runs dpx_frame on a host frame; returns (rgba (sH, sW, 4), graph (H, W), edge_count (H, W)[, diagram])."""
    stride = img.shape[1]
    out = np.zeros((height * scale, width * scale * 4), dtype=np.uint8)
    graph = np.zeros((height, width), dtype=np.uint8)
    ec = np.zeros((height, width), dtype=np.int32)
    diag = np.zeros((height * width * CELL_SIZE, 2), dtype=np.float32) if want_diagram else None
    p = Params(scale, int(subdivide), flags, 0)
    check(lib.dpx_frame(_ptr(img), width, height, stride, channels, ctypes.byref(p), _ptr(out), out.shape[1],
                        _ptr(graph), _ptr(ec), _ptr(diag)))
    rgba = out.reshape(height * scale, width * scale, 4)
    if want_diagram:
        return rgba, graph, ec, diag
    return rgba, graph, ec


def launch_kernel(img: np.ndarray, width: int, height: int, subdivide: bool = False, pos=None, color=None):
    """The reference-compatible drop-in; returns (diagram (W*H*45, 2), edge_count, graph)."""
    stride = img.shape[1]
    ec = np.zeros(width * height, dtype=np.int32)
    graph = np.zeros(width * height, dtype=np.uint8)
    ptr = lib.launch_kernel(_ptr(pos), _ptr(color), 0.0, _ptr(img), width, height, stride, _ptr(ec), _ptr(graph),
                            bool(subdivide))
    if not ptr:
        raise DepixelError(-1, last_error())
    n = width * height * CELL_SIZE * 2
    diag = np.ctypeslib.as_array(ctypes.cast(ptr, POINTER(c_float)), shape=(n,)).copy().reshape(-1, 2)
    libc.free(ptr)
    return diag, ec, graph


# --------------------------------------------------------------------------
# batched path
# --------------------------------------------------------------------------
def items_array(items: list[Item]):
    """The C array of dpx_item that dpx_batch_run takes."""
    return (Item * len(items))(*items)


class Context:
    def __init__(self, device: int = 0):
        h = c_void_p()
        check(lib.dpx_context_create(device, ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib.dpx_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, d_in, d_out, items, stream=None):
        """items: a list of Item, or an `items_array(...)` built once for repeated batches."""
        arr = items if isinstance(items, ctypes.Array) else items_array(items)
        check(lib.dpx_batch_run(self.h, _ptr(d_in), _ptr(d_out), arr, len(arr), _stream(stream)))

    def reserve(self, max_items: int, max_item_pixels: int, scale: int = 0):
        """Sizes the context's buffers (and builds the tables of `scale`) ahead of the batches."""
        check(lib.dpx_context_reserve(self.h, max_items, max_item_pixels, scale))

    def set_tile(self, size_class: int, scale: int, subdivide: bool, tile_w: int):
        """Raster tile width (16, 32, 64; 0 = default) for items of a size class at a scale."""
        check(lib.dpx_context_set_tile(self.h, size_class, scale, int(subdivide), tile_w))

    def get_tile(self, size_class: int, scale: int, subdivide: bool) -> int:
        t = c_int()
        check(lib.dpx_context_get_tile(self.h, size_class, scale, int(subdivide), ctypes.byref(t)))
        return t.value

    def set_walk(self, size_class: int, scale: int, subdivide: bool, walk: int):
        """Walk window mode (WALK_AUTO, WALK_ALWAYS, WALK_NEVER) per key."""
        check(lib.dpx_context_set_walk(self.h, size_class, scale, int(subdivide), walk))

    def get_walk(self, size_class: int, scale: int, subdivide: bool) -> int:
        w = c_int()
        check(lib.dpx_context_get_walk(self.h, size_class, scale, int(subdivide), ctypes.byref(w)))
        return w.value

    def timing(self, enable: bool):
        check(lib.dpx_batch_timing(self.h, int(enable)))

    def keep_graph(self, enable: bool):
        """Whether the following batches also write the resolved graphs (off by default)."""
        check(lib.dpx_context_keep_graph(self.h, int(enable)))

    def timing_read(self) -> BatchTimes:
        t = BatchTimes()
        check(lib.dpx_batch_timing_read(self.h, ctypes.byref(t)))
        return t

    def graph(self, n_items: int):
        """(device pointer of the final graphs, per-item byte offsets) of the last batch."""
        p = c_void_p()
        offs = POINTER(c_uint64)()
        check(lib.dpx_batch_graph(self.h, ctypes.byref(p), ctypes.byref(offs)))
        return p.value, [offs[k] for k in range(n_items)]


PLAN_GRAPH, PLAN_KEEP_GRAPH = 1, 2


class Plan:
    """A batch layout fixed once (dpx_plan): its own descriptors and graph
    scratch, so runs can be captured into HIP graphs; with PLAN_GRAPH the
    library launches it as a HIP graph itself."""

    def __init__(self, ctx: Context, items, out_align: int = 0, flags: int = 0):
        arr = items if isinstance(items, ctypes.Array) else items_array(items)
        h = c_void_p()
        check(lib.dpx_plan_create(ctx.h, arr, len(arr), out_align, flags, ctypes.byref(h)))
        self.h = h
        self.ctx = ctx  # plans go before their context
        self.n = len(arr)

    def run(self, d_in, d_out, stream=None):
        check(lib.dpx_plan_run(self.h, _ptr(d_in), _ptr(d_out), _stream(stream)))

    def graph(self):
        """(device pointer of the resolved graphs, per-item byte offsets)."""
        p = c_void_p()
        offs = POINTER(c_uint64)()
        check(lib.dpx_plan_graph(self.h, ctypes.byref(p), ctypes.byref(offs)))
        return p.value, [offs[k] for k in range(self.n)]

    def close(self):
        if self.h:
            lib.dpx_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Serving loop: continuous batching of host frames under an HBM budget."""

    def __init__(self, budget_bytes: int, max_batch: int = 64, device: int = 0):
        h = c_void_p()
        check(lib.dpx_server_create(device, budget_bytes, max_batch, ctypes.byref(h)))
        self.h = h
        self._keep = {}

    def close(self):
        if self.h:
            lib.dpx_server_destroy(self.h)
            self.h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def open_stream(self, width: int, height: int, channels: int = 3, scale: int = 4, subdivide: bool = False,
                    flags: int = 0) -> int:
        sid = c_int()
        p = Params(scale, int(subdivide), flags, 0)
        check(lib.dpx_server_open_stream(self.h, width, height, channels, ctypes.byref(p), ctypes.byref(sid)))
        return sid.value

    def close_stream(self, sid: int):
        check(lib.dpx_server_close_stream(self.h, sid))

    def set_palette(self, sid: int, palette: np.ndarray):
        """The palette (256 x 4 bytes) of an INDEXED stream, for frames submitted from now on."""
        pal = np.ascontiguousarray(palette, dtype=np.uint8).reshape(-1)
        assert pal.size == 1024
        check(lib.dpx_server_set_palette(self.h, sid, _ptr(pal)))

    def submit(self, sid: int, img: np.ndarray, out: np.ndarray) -> int:
        """Queues a host frame; `out` (rows of >= 4*scale*W bytes) is filled on completion."""
        t = c_uint64()
        rc = lib.dpx_server_submit(self.h, sid, _ptr(img), img.shape[1], _ptr(out), out.shape[1], ctypes.byref(t))
        check(rc)
        self._keep[t.value] = (img, out)
        return t.value

    def wait(self, ticket: int, timeout_ms: int = -1) -> int:
        """Returns the request's status code (0 = done); raises nothing."""
        rc = lib.dpx_server_wait(self.h, ticket, timeout_ms)
        if rc != ETIMEOUT:
            self._keep.pop(ticket, None)
        return rc

    def wait_timed(self, ticket: int, timeout_ms: int = -1) -> tuple[int, float]:
        """(status, submit -> completion ms) of a request."""
        ms = c_double()
        rc = lib.dpx_server_wait_timed(self.h, ticket, timeout_ms, ctypes.byref(ms))
        if rc != ETIMEOUT:
            self._keep.pop(ticket, None)
        return rc, ms.value

    def cancel(self, ticket: int) -> int:
        return lib.dpx_server_cancel(self.h, ticket)

    def inject_failure(self, request_index: int):
        check(lib.dpx_server_inject_failure(self.h, request_index))

    def stats(self) -> ServerStats:
        s = ServerStats()
        check(lib.dpx_server_stats_get(self.h, ctypes.byref(s)))
        return s

    def stats_reset(self):
        check(lib.dpx_server_stats_reset(self.h))


def header_symbols() -> list[str]:
    """Function names declared in include/depixel.h."""
    import re
    src = open(HEADER_PATH).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", src, flags=re.M)
    return sorted(set(n for n in names if n not in ("defined",)))
