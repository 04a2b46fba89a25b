"""Python view of the CPU oracle (liboracle.so) and of the reference harness.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  Build with ``make -C oracle``
(and ``make -C oracle ref`` for the reference harness, which needs the
reference sources and runs on a GPU).
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_float, c_int, c_uint8, c_uint32, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libdepixel_ref.so")
CS = 45
RGB, TOPDOWN, ALPHA = 1, 2, 4

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise ImportError(f"{ORACLE_SO} not built; run `make -C oracle`")
        L = ctypes.CDLL(ORACLE_SO)
        L.orc_rgb_to_yuv.restype = c_uint32
        L.orc_rgb_to_yuv.argtypes = [c_uint32]
        L.orc_colors_similar.argtypes = [c_uint32, c_uint32]
        L.orc_cell_points.argtypes = [c_uint8, c_uint8, c_uint8, c_void_p]
        L.orc_convex_hull.argtypes = [c_void_p, c_int, c_void_p]
        L.orc_cell_from_pattern.argtypes = [c_uint8, c_uint8, c_uint8, c_void_p]
        L.orc_heuristic_decision.argtypes = [c_void_p, c_int, c_int, c_int, c_int]
        L.orc_triangulate_polygon.argtypes = [c_void_p, c_int, c_void_p]
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays passed to the oracle must be contiguous"
    return a.ctypes.data_as(c_void_p)


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with {rc}")


# ---- A2 ----
def rgb_to_yuv(c0: int, c1: int, c2: int) -> int:
    return lib().orc_rgb_to_yuv(c0 | (c1 << 8) | (c2 << 16))


def colors_similar(a: tuple, b: tuple) -> bool:
    pa = a[0] | (a[1] << 8) | (a[2] << 16)
    pb = b[0] | (b[1] << 8) | (b[2] << 16)
    return bool(lib().orc_colors_similar(pa, pb))


def similarity_graph(img: np.ndarray, W: int, H: int, channels: int = 3, flags: int = 0) -> np.ndarray:
    g = np.zeros((H, W), dtype=np.uint8)
    _ok(lib().orc_similarity_graph(_p(img), W, H, img.shape[1], channels, flags, _p(g)), "graph")
    return g


# ---- A3-A7 ----
def trivial_cross(g: np.ndarray) -> np.ndarray:
    g = np.ascontiguousarray(g, dtype=np.uint8).copy()
    H, W = g.shape
    _ok(lib().orc_trivial_cross(_p(g), W, H), "trivial")
    return g


def heuristic_cross(g: np.ndarray) -> np.ndarray:
    g = np.ascontiguousarray(g, dtype=np.uint8).copy()
    H, W = g.shape
    _ok(lib().orc_heuristic_cross(_p(g), W, H), "heuristic")
    return g


def heuristic_decision(g: np.ndarray, i: int, j: int) -> bool:
    g = np.ascontiguousarray(g, dtype=np.uint8)
    H, W = g.shape
    return bool(lib().orc_heuristic_decision(_p(g), W, H, i, j))


# ---- A8-A9 ----
def cell_points(node: int, left: int, right: int) -> np.ndarray:
    pts = np.zeros((16, 2), dtype=np.float32)
    n = lib().orc_cell_points(node, left, right, _p(pts))
    return pts[:n].copy()


def convex_hull(pts: np.ndarray) -> np.ndarray:
    pts = np.ascontiguousarray(pts, dtype=np.float32)
    out = np.zeros((CS, 2), dtype=np.float32)
    n = lib().orc_convex_hull(_p(pts), len(pts), _p(out))
    return out[:n].copy()


def cell_from_pattern(node: int, left: int, right: int) -> np.ndarray:
    out = np.zeros((CS, 2), dtype=np.float32)
    n = lib().orc_cell_from_pattern(node, left, right, _p(out))
    return out[:n].copy()


def cell_table() -> tuple[np.ndarray, np.ndarray]:
    """(points[4096, 45, 2], counts[4096]) for keys node | l2<<8 | l7<<9 | r0<<10 | r5<<11."""
    pts = np.zeros((4096, CS, 2), dtype=np.float32)
    cnt = np.zeros(4096, dtype=np.int32)
    for key in range(4096):
        node, left, right = key_bytes(key)
        cnt[key] = lib().orc_cell_from_pattern(node, left, right, _p(pts[key]))
    return pts, cnt


def key_bytes(key: int) -> tuple[int, int, int]:
    node = key & 0xFF
    left = (((key >> 8) & 1) << 2) | (((key >> 9) & 1) << 7)
    right = (((key >> 10) & 1) << 0) | (((key >> 11) & 1) << 5)
    return node, left, right


def cells(g: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    g = np.ascontiguousarray(g, dtype=np.uint8)
    H, W = g.shape
    d = np.zeros((H * W, CS, 2), dtype=np.float32)
    ec = np.zeros(H * W, dtype=np.int32)
    _ok(lib().orc_cells(_p(g), W, H, _p(d), _p(ec)), "cells")
    return d, ec


# ---- A10-A16 ----
def subdivide(img, W, H, g, diagram, ec, channels=3, flags=0):
    d = np.ascontiguousarray(diagram, dtype=np.float32).copy()
    e = np.ascontiguousarray(ec, dtype=np.int32).copy()
    es = np.zeros((W * H, CS), dtype=np.uint8)
    li = np.zeros((W * H, CS), dtype=np.int32)
    _ok(lib().orc_subdivide(_p(img), W, H, img.shape[1], channels, flags, _p(np.ascontiguousarray(g)), _p(d), _p(e),
                            _p(es), _p(li)), "subdivide")
    return d, e, es, li


def triangulate_polygon(poly: np.ndarray) -> tuple[bool, np.ndarray]:
    poly = np.ascontiguousarray(poly, dtype=np.float32)
    out = np.zeros((CS, 2), dtype=np.float32)
    ok = lib().orc_triangulate_polygon(_p(poly), len(poly), _p(out))
    return bool(ok), out


def triangulate(diagram, ec, W, H):
    d = np.ascontiguousarray(diagram, dtype=np.float32).copy()
    e = np.ascontiguousarray(ec, dtype=np.int32)
    nf = c_int(0)
    _ok(lib().orc_triangulate(_p(d), _p(e), W, H, ctypes.byref(nf)), "triangulate")
    return d, nf.value


def export_vbo(img, W, H, diagram, ec, channels=3, flags=0):
    pos = np.zeros((W * H * CS, 2), dtype=np.float32)
    col = np.zeros((W * H * CS, 4), dtype=np.uint8)
    _ok(lib().orc_export_vbo(_p(img), W, H, img.shape[1], channels, flags, _p(np.ascontiguousarray(diagram)),
                             _p(np.ascontiguousarray(ec)), _p(pos), _p(col)), "export")
    return pos, col


def rasterize(img, W, H, diagram, ec, scale, channels=3, flags=0):
    out = np.zeros((H * scale, W * scale * 4), dtype=np.uint8)
    _ok(lib().orc_rasterize(_p(img), W, H, img.shape[1], channels, flags, _p(np.ascontiguousarray(diagram)),
                            _p(np.ascontiguousarray(ec)), scale, _p(out), out.shape[1]), "rasterize")
    return out.reshape(H * scale, W * scale, 4)


def present(img, W, H, diagram, ec, mode=0, scale_x100=800, tx=1, ty=1, width=1200, height=800, view_flags=1,
            channels=3, flags=0):
    """The viewer window (defaults: the reference's 1200x800 window, scale 8, translate (1,1), AA on);
    returns (height, width, 4) RGBA in the requested row order."""
    out = np.zeros((height, width * 4), dtype=np.uint8)
    d = None if diagram is None else np.ascontiguousarray(diagram)
    e = None if ec is None else np.ascontiguousarray(ec)
    _ok(lib().orc_present(_p(img), W, H, img.shape[1], channels, flags, _p(d), _p(e), mode, scale_x100, tx, ty,
                          width, height, view_flags, _p(out), out.shape[1]), "present")
    return out.reshape(height, width, 4)


def depixelize(img, W, H, scale=4, subdivide=False, channels=3, flags=0, want_diagram=False):
    """Full per-frame sequence; returns dict(rgba, graph, edge_count[, diagram], tri_failed)."""
    out = np.zeros((H * scale, W * scale * 4), dtype=np.uint8)
    g = np.zeros((H, W), dtype=np.uint8)
    ec = np.zeros((H, W), dtype=np.int32)
    d = np.zeros((H * W * CS, 2), dtype=np.float32) if want_diagram else None
    nf = c_int(0)
    _ok(lib().orc_depixelize(_p(img), W, H, img.shape[1], channels, flags, scale, int(subdivide), _p(out),
                             out.shape[1], _p(g), _p(ec), _p(d), ctypes.byref(nf)), "depixelize")
    r = dict(rgba=out.reshape(H * scale, W * scale, 4), graph=g, edge_count=ec, tri_failed=nf.value)
    if want_diagram:
        r["diagram"] = d
    return r


# ---- A18 ----
def components(g: np.ndarray) -> np.ndarray:
    g = np.ascontiguousarray(g, dtype=np.uint8)
    H, W = g.shape
    lab = np.zeros((H, W), dtype=np.int32)
    _ok(lib().orc_components(_p(g), W, H, _p(lab)), "components")
    return lab


def border_chains(g: np.ndarray) -> list:
    """Border chains of a resolved graph in scan order, as lists of node
    indices j * W + i (buffer row order)."""
    g = np.ascontiguousarray(g, dtype=np.uint8)
    H, W = g.shape
    cap = 8 * W * H + 8
    nodes = np.zeros(cap, dtype=np.int32)
    sizes = np.zeros(W * H + 1, dtype=np.int32)
    nc, nn = c_int(0), c_int(0)
    _ok(lib().orc_border_chains(_p(g), W, H, _p(nodes), _p(sizes), cap, sizes.size, ctypes.byref(nc),
                                ctypes.byref(nn)), "border_chains")
    out, k = [], 0
    for s in sizes[:nc.value]:
        out.append(nodes[k:k + s].tolist())
        k += s
    return out


# --------------------------------------------------------------------------
# reference harness (the reference's own device functions, run on a GPU)
# --------------------------------------------------------------------------
_ref = None


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise ImportError(f"{REF_SO} not built (needs the reference sources; `make -C oracle ref`)")
        _ref = ctypes.CDLL(REF_SO)
    return _ref


def _rok(rc, what):
    if rc != 0:
        raise RuntimeError(f"reference harness {what} failed with {rc}")


def ref_graph(img, W, H):
    g = np.zeros((H, W), dtype=np.uint8)
    _rok(ref().ref_graph(_p(img), W, H, img.shape[1], _p(g)), "graph")
    return g


def ref_trivial(g):
    g = np.ascontiguousarray(g, dtype=np.uint8).copy()
    H, W = g.shape
    _rok(ref().ref_trivial(_p(g), W, H), "trivial")
    return g


def ref_heuristic(g):
    g = np.ascontiguousarray(g, dtype=np.uint8).copy()
    H, W = g.shape
    _rok(ref().ref_heuristic(_p(g), W, H), "heuristic")
    return g


def ref_decisions(g, blocks: np.ndarray) -> np.ndarray:
    g = np.ascontiguousarray(g, dtype=np.uint8)
    H, W = g.shape
    b = np.ascontiguousarray(blocks, dtype=np.int32)
    out = np.zeros(len(b), dtype=np.int32)
    _rok(ref().ref_decisions(_p(g), W, H, _p(b), len(b), _p(out)), "decisions")
    return out


def ref_cells(g):
    g = np.ascontiguousarray(g, dtype=np.uint8)
    H, W = g.shape
    d = np.zeros((H * W, CS, 2), dtype=np.float32)
    ec = np.zeros(H * W, dtype=np.int32)
    _rok(ref().ref_cells(_p(g), W, H, _p(d), _p(ec)), "cells")
    return d, ec


def ref_cell_table():
    pts = np.zeros((4096, CS, 2), dtype=np.float32)
    cnt = np.zeros(4096, dtype=np.int32)
    _rok(ref().ref_cell_table(_p(pts), _p(cnt)), "cell table")
    return pts, cnt


def ref_subdivide(img, W, H, g, diagram, ec):
    d = np.ascontiguousarray(diagram, dtype=np.float32).copy()
    e = np.ascontiguousarray(ec, dtype=np.int32).copy()
    _rok(ref().ref_subdivide(_p(img), W, H, img.shape[1], _p(np.ascontiguousarray(g)), _p(d), _p(e)), "subdivide")
    return d, e


def ref_triangulate(diagram, ec, W, H):
    d = np.ascontiguousarray(diagram, dtype=np.float32).copy()
    _rok(ref().ref_triangulate(_p(d), _p(np.ascontiguousarray(ec, dtype=np.int32)), W, H), "triangulate")
    return d
