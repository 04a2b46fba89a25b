"""Oracle pins for cells, subdivision, triangulation, VBO packing and the
raster (A8-A17), from closed forms and exact rational arithmetic.
"""
import itertools
import random
from fractions import Fraction as F

import numpy as np
import pytest

CS = 45


def polygon_area2(p):
    return sum(p[k - 1][0] * p[k][1] - p[k][0] * p[k - 1][1] for k in range(len(p)))


def test_hull_table_properties(O):
    """All 4096 cells: 3..8 CCW strictly convex vertices on the quarter
    lattice inside [-0.25, 1.25]^2, containing the central square."""
    pts, cnt = O.cell_table()
    for key in range(4096):
        n = cnt[key]
        h = pts[key, :n].astype(np.float64)
        assert 3 <= n <= 8
        assert np.all(h * 4 == np.round(h * 4)) and h.min() >= -0.25 and h.max() <= 1.25
        assert tuple(pts[key, n]) == tuple(pts[key, 0])  # closing point
        for k in range(n):
            a, b, c = h[k - 2], h[k - 1], h[k]
            assert (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]) > 0, key
        for q in ((0.25, 0.25), (0.75, 0.25), (0.75, 0.75), (0.25, 0.75)):
            assert all((h[k][0] - h[k - 1][0]) * (q[1] - h[k - 1][1]) - (h[k][1] - h[k - 1][1]) * (q[0] - h[k - 1][0]) >= 0
                       for k in range(n)), key


def test_internal_node_is_unit_square(O):
    h = O.cell_from_pattern(90, 0, 0)
    assert sorted(map(tuple, h)) == [(0, 0), (0, 1), (1, 0), (1, 1)]


def test_triangulation_of_every_hull(O):
    """Ear clipping yields n-2 CCW triangles whose areas sum to the polygon's,
    starting with the ear (n-1, 0, 1) (triangulate_functions.cu:153-177)."""
    pts, cnt = O.cell_table()
    for key in range(4096):
        n = cnt[key]
        poly = pts[key, :n]
        ok, res = O.triangulate_polygon(poly)
        assert ok
        tris = res[:3 * (n - 2)].reshape(-1, 3, 2).astype(np.float64)
        areas = [polygon_area2(t) for t in tris]
        assert all(a > 0 for a in areas)
        assert sum(areas) == polygon_area2(poly.astype(np.float64))
        assert np.array_equal(tris[0], poly[[n - 1, 0, 1]])
        assert not res[3 * (n - 2):].any()


def test_subdivided_isolated_pixel_is_octagon(O):
    """An isolated pixel in a uniform field: the background's diagonals cut
    its four corners, so its hull is the square [1/4, 3/4]^2; every edge is a
    border, no vertex is a pixel corner, and each corner is cut at 1/4 and 3/4
    of the 1/2-long edges."""
    W = H = 5
    img = np.full((H, W * 3), 50, dtype=np.uint8)
    img.reshape(H, W, 3)[2, 2] = (250, 0, 0)
    g = O.heuristic_cross(O.trivial_cross(O.similarity_graph(img, W, H)))
    d, ec = O.cells(g)
    sd, sec, status, link = O.subdivide(img, W, H, g, d, ec)
    k = 2 * W + 2
    assert ec[k] == 4 and set(map(tuple, d[k, :4].astype(np.float64))) == {(0.25, 0.25), (0.75, 0.25), (0.75, 0.75),
                                                                             (0.25, 0.75)}
    assert sec[k] == 8 and not status[k].any()
    octagon = {(0.375, 0.25), (0.625, 0.25), (0.75, 0.375), (0.75, 0.625), (0.625, 0.75), (0.375, 0.75),
               (0.25, 0.625), (0.25, 0.375)}
    assert set(map(tuple, sd[k, :8].astype(np.float64))) == octagon


def test_subdivision_lattice_and_extent(D, O):
    for kind, seed in ((0, 1), (2, 2), (4, 3)):
        W, H = 48, 40
        img = D.synth_frame(kind, seed, W, H, 3)
        r = O.depixelize(img, W, H, scale=1, subdivide=True, want_diagram=True)
        g = r["graph"]
        d, ec = O.cells(g)
        sd, sec, _, _ = O.subdivide(img, W, H, g, d, ec)
        for k in range(W * H):
            p = sd[k, :sec[k]].astype(np.float64)
            assert np.all(p * 64 == np.round(p * 64))
            assert p.min() >= -0.5 and p.max() <= 1.5
            assert sec[k] <= 16
        assert r["tri_failed"] == 0


def test_vbo_packing(D, O):
    W, H = 6, 4
    img = D.synth_frame(0, 5, W, H, 3, 20)
    r = O.depixelize(img, W, H, scale=1, want_diagram=True)
    pos, col = O.export_vbo(img, W, H, r["diagram"], r["edge_count"].reshape(-1))
    for k in range(W * H):
        i, j = k % W, k // W
        n = r["edge_count"].reshape(-1)[k]
        b = img[j, 3 * i:3 * i + 3]
        assert (col[k * CS:(k + 1) * CS] == [b[2], b[1], b[0], 255]).all()
        assert (pos[k * CS + 3 * (n - 2):(k + 1) * CS] == -100).all()
        assert np.array_equal(pos[k * CS:k * CS + 3 * (n - 2)], r["diagram"][k * CS:k * CS + 3 * (n - 2)])


# ---------------------------------------------------------------------------
# raster
# ---------------------------------------------------------------------------
def stripes(W, H, seed):
    """Vertical 1-px stripes of alternating colour families: no diagonal
    similarity anywhere, so every cell is the unit square."""
    rng = np.random.default_rng(seed)
    img = np.zeros((H, W * 3), dtype=np.uint8)
    px = img.reshape(H, W, 3)
    a = rng.integers(0, 60, (W, 3))
    b = rng.integers(190, 256, (W, 3))
    for i in range(W):
        px[:, i] = a[i] if i % 2 == 0 else b[i]
    return img


@pytest.mark.parametrize("scale", [1, 2, 3, 4, 7])
def test_square_cells_raster_is_nearest_neighbour(O, scale):
    W, H = 9, 6
    img = stripes(W, H, scale)
    r = O.depixelize(img, W, H, scale=scale)
    assert (r["edge_count"] == 4).all()
    px = img.reshape(H, W, 3)
    nn = np.repeat(np.repeat(px, scale, axis=0), scale, axis=1)
    exp = np.concatenate([nn[..., ::-1], np.full(nn.shape[:2] + (1,), 255, np.uint8)], axis=-1)
    assert np.array_equal(r["rgba"], exp)


def test_raster_orientation_flags(D, O):
    W, H, s = 30, 20, 3
    img = D.synth_frame(0, 4, W, H, 3, W * 3)
    base = O.depixelize(img, W, H, scale=s, subdivide=True)["rgba"]
    top = O.depixelize(img[::-1].copy(), W, H, scale=s, subdivide=True, flags=O.TOPDOWN)["rgba"]
    assert np.array_equal(top[::-1], base)
    rgb = img.reshape(H, W, 3)[:, :, ::-1].reshape(H, W * 3).copy()
    assert np.array_equal(O.depixelize(rgb, W, H, scale=s, subdivide=True, flags=O.RGB)["rgba"], base)


def raster_triangles(O, tris, scale):
    """Rasterise a list of triangles (world coords inside [0,1]^2) as the one
    cell of a 1x1 white image; returns the coverage mask (scale x scale)."""
    img = np.full((1, 4), 255, dtype=np.uint8)
    d = np.zeros((CS, 2), dtype=np.float32)
    for t, tri in enumerate(tris):
        d[3 * t:3 * t + 3] = tri
    ec = np.array([len(tris) + 2], dtype=np.int32)
    out = O.rasterize(img, 1, 1, d, ec, scale)
    return out[..., 0] == 255


def perturbed_inside(tri, sx, sy):
    """Exact: is (sx, sy) + (e, e^2) strictly inside the triangle, e -> 0+ ?"""
    eps = F(1, 2 ** 40)
    px, py = sx + eps, sy + eps * eps
    (ax, ay), (bx, by), (cx, cy) = tri
    o = (bx - ax) * (cy - ay) - (by - ay) * (cx - ax)
    if o == 0:
        return False
    if o < 0:
        bx, by, cx, cy = cx, cy, bx, by
    for (p, q) in (((ax, ay), (bx, by)), ((bx, by), (cx, cy)), ((cx, cy), (ax, ay))):
        if (q[0] - p[0]) * (py - p[1]) - (q[1] - p[1]) * (px - p[0]) <= 0:
            return False
    return True


def test_edge_rule_equals_symbolic_perturbation(O):
    """Vertices on the 1/8 lattice put many samples exactly on edges and
    vertices; coverage must equal the exact perturbed point-in-triangle."""
    rnd = random.Random(7)
    scale = 4
    lattice = [F(k, 8) for k in range(9)]
    for trial in range(300):
        tri = [(rnd.choice(lattice), rnd.choice(lattice)) for _ in range(3)]
        got = raster_triangles(O, [[(float(x), float(y)) for x, y in tri]], scale)
        for Y in range(scale):
            for X in range(scale):
                sx, sy = F(2 * X + 1, 2 * scale), F(2 * Y + 1, 2 * scale)
                assert got[Y, X] == perturbed_inside(tri, sx, sy), (tri, X, Y)


def test_triangulations_of_a_polygon_cover_the_same_samples(O):
    """The fill rule partitions: two fans of one convex polygon cover the same
    samples, and each sample of a fan is covered by at most one triangle."""
    rnd = random.Random(11)
    scale = 8
    for trial in range(100):
        pts = {(F(rnd.randint(0, 16), 16), F(rnd.randint(0, 16), 16)) for _ in range(9)}
        P = sorted(pts)
        if len(P) < 3:
            continue
        cross = lambda o, a, b: (a[0] - o[0]) * (b[1] - o[1]) - (a[1] - o[1]) * (b[0] - o[0])  # noqa: E731
        lo, up = [], []
        for p in P:
            while len(lo) >= 2 and cross(lo[-2], lo[-1], p) <= 0:
                lo.pop()
            lo.append(p)
        for p in reversed(P):
            while len(up) >= 2 and cross(up[-2], up[-1], p) <= 0:
                up.pop()
            up.append(p)
        hull = lo[:-1] + up[:-1]
        if len(hull) < 3:
            continue
        n = len(hull)
        fans = [[(hull[a], hull[(a + k) % n], hull[(a + k + 1) % n]) for k in range(1, n - 1)] for a in (0, 1)]
        masks = []
        for fan in fans:
            per_tri = [raster_triangles(O, [[(float(x), float(y)) for x, y in t]], scale) for t in fan]
            assert max(int(sum(m[Y, X] for m in per_tri)) for Y in range(scale) for X in range(scale)) <= 1
            masks.append(raster_triangles(O, [[(float(x), float(y)) for x, y in t] for t in fan], scale))
        assert np.array_equal(masks[0], masks[1])
