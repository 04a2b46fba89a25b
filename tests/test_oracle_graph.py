"""Oracle pins for the similarity graph (A2) and the crossing resolution (A3-A7).

Known answers and closed forms, chosen so that a plausible slip fails:
a swapped channel order, FMA-contracted luma, a dropped borrow, a wrong
neighbour offset, the islands precedence, the walk cap, the tie rule.
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

BIT = lambda v, e: (int(v) >> e) & 1  # noqa: E731
EDX = [-1, 0, 1, -1, 1, -1, 0, 1]
EDY = [1, 1, 1, 0, 0, -1, -1, -1]


def unpack(packed):
    """(y, u, v) of a packed word y*65536 + u*256 + v (mod 2^32) with |u|,|v| < 128."""
    v = packed & 0xFF
    v = v - 256 if v >= 128 else v
    rest = (packed - (v & 0xFFFFFFFF)) & 0xFFFFFFFF
    u = (rest >> 8) & 0xFF
    u = u - 256 if u >= 128 else u
    rest = (rest - ((u << 8) & 0xFFFFFFFF)) & 0xFFFFFFFF
    return rest >> 16, u, v


def packed(y, u, v):
    return ((y << 16) + (u << 8) + v) & 0xFFFFFFFF


def test_pure_colours_follow_reference_byte_order(O):
    """Known answers, worked by hand from graph_functions.cu:84-97: byte 0 (blue
    in the reference's BGR buffer) is the one weighted 0.299."""
    # (255,0,0): y = int(76.245) = 76, u = int(-76*0.492) = -37, v = int(179*0.877) = 156
    assert O.rgb_to_yuv(255, 0, 0) == packed(76, -37, 156)
    # (0,255,0): y = int(149.685) = 149, u = int(-149*0.492) = -73, v = int(-149*0.877) = -130
    assert O.rgb_to_yuv(0, 255, 0) == packed(149, -73, -130)
    # (0,0,255): y = int(29.07) = 29, u = int(226*0.492) = 111, v = int(-29*0.877) = -25
    assert O.rgb_to_yuv(0, 0, 255) == packed(29, 111, -25)


def test_luma_is_uncontracted_double(O):
    """Luma over every grey and along the three primary ramps equals Python's
    IEEE double evaluation (0.299 c0 + 0.587 c1) + 0.114 c2 truncated, with
    no fused multiply-add; 65 of the 256 greys land below k."""
    below = 0
    for k in range(256):
        y, u, v = unpack(O.rgb_to_yuv(k, k, k))
        assert y == int((0.299 * k + 0.587 * k) + 0.114 * k)
        below += y != k
    assert below == 65
    for k in range(256):
        for c in ((k, 0, 0), (0, k, 0), (0, 0, k), (k, 255 - k, k // 2)):
            y = int((0.299 * c[0] + 0.587 * c[1]) + 0.114 * c[2])
            u = int(np.float32(c[2] - y) * np.float32(0.492))
            v = int(np.float32(c[0] - y) * np.float32(0.877))
            assert O.rgb_to_yuv(*c) == packed(y, u, v), c


def test_negative_uv_borrow_into_packed_word(O):
    # greys 1 and 2 sum to 0.9999999999999999 and 1.9999999999999998 in double: y = k-1, u = 0, v = 0
    assert O.rgb_to_yuv(1, 1, 1) == packed(0, 0, 0)
    assert O.rgb_to_yuv(2, 2, 2) == packed(1, 0, 0)
    assert O.rgb_to_yuv(255, 255, 255) == packed(255, 0, 0)
    # a colour with negative U and V: packed word wraps modulo 2^32 through the borrow
    p = O.rgb_to_yuv(10, 200, 10)
    y = int((0.299 * 10 + 0.587 * 200) + 0.114 * 10)
    u = int(np.float32(10 - y) * np.float32(0.492))
    v = int(np.float32(10 - y) * np.float32(0.877))
    assert u < 0 and v < 0
    assert p == ((y << 16) + (u << 8) + v) & 0xFFFFFFFF


def test_similarity_thresholds_on_greys(O):
    """Greys differ mainly in Y: |dY| <= 5 (0x50000) is similar, >= 7 never."""
    for k in range(0, 249):
        a = (k, k, k)
        assert O.colors_similar(a, a)
        for d in range(1, 7):
            b = (k + d, k + d, k + d)
            ya = unpack(O.rgb_to_yuv(*a))
            yb = unpack(O.rgb_to_yuv(*b))
            expect = abs(ya[0] - yb[0]) <= 5 and abs(ya[1] - yb[1]) <= 7 and abs(ya[2] - yb[2]) <= 6
            assert O.colors_similar(a, b) == expect == O.colors_similar(b, a)
        assert not O.colors_similar(a, (k + 7, k + 7, k + 7)) or k + 7 > 255


def test_similarity_is_symmetric_and_random_pairs(O):
    rng = np.random.default_rng(3)
    for _ in range(3000):
        a = tuple(int(x) for x in rng.integers(0, 256, 3))
        b = tuple(int(x) for x in np.clip(np.array(a) + rng.integers(-12, 13, 3), 0, 255))
        assert O.colors_similar(a, b) == O.colors_similar(b, a)


def symmetric(g):
    H, W = g.shape
    for j in range(H):
        for i in range(W):
            for e in range(8):
                ni, nj = i + EDX[e], j + EDY[e]
                inside = 0 <= ni < W and 0 <= nj < H
                if BIT(g[j, i], e) and (not inside or not BIT(g[nj, ni], 7 - e)):
                    return False
    return True


def crossings(g):
    return int((((g[:-1, :-1] >> 2) & 1) & ((g[1:, 1:] >> 5) & 1) & (g[:-1, 1:] & 1) & ((g[1:, :-1] >> 7) & 1)).sum())


def test_graph_invariants_on_synthetic_frames(D, O):
    for kind, seed, W, H in ((0, 1, 70, 50), (2, 2, 64, 40), (4, 3, 48, 48)):
        img = D.synth_frame(kind, seed, W, H, 3)
        g = O.similarity_graph(img, W, H)
        assert symmetric(g)
        assert not (g[:, 0] & 0b00101001).any() and not (g[:, -1] & 0b10010100).any()
        assert not (g[0] & 0b11100000).any() and not (g[-1] & 0b00000111).any()
        g3 = O.heuristic_cross(O.trivial_cross(g))
        assert symmetric(g3) and crossings(g3) == 0


def test_uniform_image_is_fully_connected(O):
    img = np.full((5, 7 * 3), 77, dtype=np.uint8)
    g = O.similarity_graph(img, 7, 5)
    assert g[2, 3] == 255 and g[0, 0] == 0b00010110 and g[4, 6] == 0b01101000


def test_layout_flags_are_pure_reindexing(D, O):
    img = D.synth_frame(2, 9, 40, 30, 3)
    g = O.similarity_graph(img, 40, 30)
    rgb = img.reshape(30, 40, 3)[:, :, ::-1].reshape(30, 120).copy()
    assert np.array_equal(O.similarity_graph(rgb, 40, 30, flags=O.RGB), g)
    flipped = img[::-1].copy()
    assert np.array_equal(O.similarity_graph(flipped, 40, 30, flags=O.TOPDOWN), g)


def test_alpha_extension(O):
    # 4x3 RGBA: left half transparent with random colours, right half opaque
    rng = np.random.default_rng(1)
    img = rng.integers(0, 256, (3, 16), dtype=np.uint8)
    px = img.reshape(3, 4, 4)
    px[:, :2, 3] = 0
    px[:, 2:, 3] = 255
    px[:, 2:, :3] = 40
    g = O.similarity_graph(img, 4, 3, channels=4, flags=O.ALPHA)
    assert BIT(g[1, 0], 4) and BIT(g[1, 0], 1) and BIT(g[1, 0], 6)  # transparent block connected
    assert not BIT(g[1, 1], 4)  # transparent | opaque
    assert BIT(g[1, 2], 4)      # opaque equal colours


# ---------------------------------------------------------------------------
# A3 trivial crossing
# ---------------------------------------------------------------------------
def random_symmetric_graph(rng, W, H, p):
    g = np.zeros((H, W), dtype=np.uint8)
    for j in range(H):
        for i in range(W):
            for e in (2, 4, 1, 7):  # each undirected edge once
                ni, nj = i + EDX[e], j + EDY[e]
                if 0 <= ni < W and 0 <= nj < H and rng.random() < p:
                    g[j, i] |= 1 << e
                    g[nj, ni] |= 1 << (7 - e)
    return g


def test_trivial_crossing_brute_force(O):
    rng = np.random.default_rng(5)
    for trial in range(20):
        g = random_symmetric_graph(rng, 9, 8, 0.8)
        t = O.trivial_cross(g)
        H, W = g.shape
        for j in range(H - 1):
            for i in range(W - 1):
                full = (BIT(g[j, i], 1) and BIT(g[j, i], 4) and BIT(g[j + 1, i + 1], 6) and BIT(g[j + 1, i + 1], 3))
                if full:  # both diagonals of a 4-connected block are gone, from both ends
                    assert not BIT(t[j, i], 2) and not BIT(t[j + 1, i + 1], 5)
                    assert not BIT(t[j, i + 1], 0) and not BIT(t[j + 1, i], 7)
        # only diagonal bits ever change, and only by clearing
        assert not ((t ^ g) & 0b01011010).any() and not (t & ~g).any()
        assert symmetric(t)


# ---------------------------------------------------------------------------
# A5-A7 heuristics: closed special cases
# ---------------------------------------------------------------------------
def checkerboard(W, H):
    img = np.zeros((H, W * 3), dtype=np.uint8)
    px = img.reshape(H, W, 3)
    px[(np.add.outer(np.arange(H), np.arange(W)) % 2) == 0] = (200, 30, 30)
    px[(np.add.outer(np.arange(H), np.arange(W)) % 2) == 1] = (20, 180, 60)
    return img


def test_checkerboard_interior_keeps_backslash(O):
    """Degrees are 4, both walks 0, the tie removes '/' (graph_functions.cu:830-839)."""
    W = H = 12
    g = O.heuristic_cross(O.trivial_cross(O.similarity_graph(checkerboard(W, H), W, H)))
    for j in range(2, H - 2):
        for i in range(2, W - 2):
            assert g[j, i] == 0b10000001, (i, j, g[j, i])  # bits 0 and 7: the "\" direction


def test_island_pair_stays_connected(O):
    W = H = 8
    img = np.full((H, W * 3), 90, dtype=np.uint8)
    px = img.reshape(H, W, 3)
    px[3, 3] = px[4, 4] = (250, 10, 10)  # a diagonal pair of one colour in a uniform field
    g1 = O.trivial_cross(O.similarity_graph(img, W, H))
    assert crossings(g1) == 1
    assert O.heuristic_decision(g1, 3, 3) is False  # islands rule: remove "\"
    g = O.heuristic_cross(g1)
    assert BIT(g[3, 3], 2) and BIT(g[4, 4], 5)
    assert not BIT(g[3, 4], 0) and not BIT(g[4, 3], 7)


def link(g, i, j, e):
    g[j, i] |= 1 << e
    g[j + EDY[e], i + EDX[e]] |= 1 << (7 - e)


def curve_graph(L1, L2, o=45):
    """Crossing at block (o,o): '/' from i1=(o,o) with a valence-2 chain of L1
    nodes going down-left, i4 of degree 3; '\\' from i3=(o+1,o) with a chain
    of L2 nodes going down-right, i2 of degree 3."""
    g = np.zeros((100, 100), dtype=np.uint8)
    link(g, o, o, 2)
    link(g, o + 1, o, 0)
    x, y = o, o
    for _ in range(L1):
        link(g, x, y, 5)
        x, y = x - 1, y - 1
    x, y = o + 1, o
    for _ in range(L2):
        link(g, x, y, 7)
        x, y = x + 1, y - 1
    link(g, o + 1, o + 1, 2)
    link(g, o + 1, o + 1, 1)
    link(g, o, o + 1, 0)
    link(g, o, o + 1, 1)
    return g


@pytest.mark.parametrize("L1,L2,remove_slash", [(9, 4, False), (4, 9, True), (6, 6, True), (40, 35, True),
                                                (31, 30, False), (30, 31, True), (28, 29, True)])
def test_curve_length_rule(O, L1, L2, remove_slash):
    """Walk sums are capped at 31 (graph_functions.cu:557); ties remove '/'."""
    g = curve_graph(L1, L2)
    assert crossings(g) == 1 and symmetric(g)
    assert O.heuristic_decision(g, 45, 45) is remove_slash
    h = O.heuristic_cross(g)
    assert crossings(h) == 0
    if remove_slash:
        assert not BIT(h[45, 45], 2) and BIT(h[45, 46], 0)
    else:
        assert BIT(h[45, 45], 2) and not BIT(h[45, 46], 0)


def test_islands_precedence_quirk(O):
    """checkValence2Vertex at 805-806 parses as A || (B && !C && !D): a
    degree-1 i3 removes '/' even when i1 is degree 1 too."""
    g = np.zeros((10, 10), dtype=np.uint8)
    link(g, 4, 4, 2)   # "/" i1=(4,4) - i4=(5,5)
    link(g, 5, 4, 0)   # "\" i3=(5,4) - i2=(4,5)
    link(g, 5, 5, 1)   # i4 degree 2, i2 degree 3, i1 and i3 degree 1
    link(g, 4, 5, 1)
    link(g, 4, 5, 0)
    # rule 3: (v1 || v4) && !v3 && !v2 -> v3 is true, so not taken; rule 4: v3 -> remove "/"
    assert O.heuristic_decision(g, 4, 4) is True


def test_reference_resolved_graphs_are_fixed_points(O):
    d = json.load(open(os.path.join(GOLDEN, "reference_graphs.json")))
    for name in ("c_pattern", "alex"):
        g = np.array(d[name]["graph"], dtype=np.uint8).reshape(d[name]["height"], d[name]["width"])
        assert symmetric(g) and crossings(g) == 0
        assert np.array_equal(O.heuristic_cross(O.trivial_cross(g)), g)
        assert int((g == 90).sum()) == {"c_pattern": 8, "alex": 178}[name]
