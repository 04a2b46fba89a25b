"""Presentation in the viewer window (F4): the oracle's orc_present pinned
against the raster it generalises, the closed form of the GL_QUADS view
(simpleVBO.cpp:288-331) and a brute-force evaluation over exact rationals.

GL's own sampling and multisample resolve are closed: the window image is
the build's definition (samples at (1/2,1/2), or the rotated 4x grid
(3,1),(7,3),(1,5),(5,7)/8 averaged per channel), not a reference output."""
from fractions import Fraction

import numpy as np
import pytest

AA_OFFSETS = [(3, 1), (7, 3), (1, 5), (5, 7)]


def cell_colours(img, W, H):
    """(H, W, 4) RGBA of each pixel's cell: (b2, b1, b0, 255), kernel.cu:98-101."""
    px = img[:, :W * 3].reshape(H, W, 3)
    out = np.empty((H, W, 4), dtype=np.uint8)
    out[..., 0], out[..., 1], out[..., 2], out[..., 3] = px[..., 2], px[..., 1], px[..., 0], 255
    return out


@pytest.mark.parametrize("sub", [False, True])
def test_cells_view_is_the_raster(D, O, sub):
    """Integer scale, no translation, no antialiasing, window = the image:
    window pixel (X, Y) samples ((X + 1/2) / S, (Y + 1/2) / S), the raster's
    own sample, so the two agree bit for bit."""
    W, H = 37, 23
    img = D.synth_frame(0, 5, W, H, 3)
    for S in (1, 3, 4):
        r = O.depixelize(img, W, H, scale=S, subdivide=sub, want_diagram=True)
        win = O.present(img, W, H, r["diagram"], r["edge_count"], mode=0, scale_x100=100 * S, tx=0, ty=0,
                        width=W * S, height=H * S, view_flags=0)
        assert np.array_equal(win, r["rgba"]), f"scale {S}"


def test_quads_view_is_nearest_neighbour(D, O):
    W, H = 21, 17
    img = D.synth_frame(2, 3, W, H, 3)
    col = cell_colours(img, W, H)
    for S in (1, 2, 5):
        win = O.present(img, W, H, None, None, mode=1, scale_x100=100 * S, tx=0, ty=0, width=W * S,
                        height=H * S, view_flags=0)
        assert np.array_equal(win, np.repeat(np.repeat(col, S, axis=0), S, axis=1)), f"scale {S}"
        # the same with antialiasing: every sample of a window pixel lies in one square
        win_aa = O.present(img, W, H, None, None, mode=1, scale_x100=100 * S, tx=0, ty=0, width=W * S,
                           height=H * S, view_flags=1)
        assert np.array_equal(win_aa, win)


def test_quads_view_fractional_scale_and_translation(D, O):
    """Window pixel (X, Y) shows pixel floor((X + 1/2 - t) * 100 / sc), or the
    clear colour outside the image (squares include their left and bottom
    edges under the coverage rule)."""
    W, H = 9, 7
    img = D.synth_frame(0, 8, W, H, 3)
    col = cell_colours(img, W, H)
    for sc, tx, ty in ((850, 1, 1), (333, -4, 6), (25, 0, 0), (1250, 3, -2)):
        ww, wh = 40, 36
        win = O.present(img, W, H, None, None, mode=1, scale_x100=sc, tx=tx, ty=ty, width=ww, height=wh,
                        view_flags=0)
        want = np.zeros((wh, ww, 4), dtype=np.uint8)
        want[..., 3] = 255
        for Y in range(wh):
            j = (Fraction(2 * (Y - ty) + 1, 2) * 100 / sc).__floor__()
            for X in range(ww):
                i = (Fraction(2 * (X - tx) + 1, 2) * 100 / sc).__floor__()
                if 0 <= i < W and 0 <= j < H:
                    want[Y, X] = col[j, i]
        assert np.array_equal(win, want), f"scale {sc} translate {tx},{ty}"


def test_translation_shifts_the_window(D, O):
    W, H = 20, 14
    img = D.synth_frame(0, 2, W, H, 3)
    r = O.depixelize(img, W, H, scale=2, subdivide=True, want_diagram=True)
    kw = dict(mode=0, scale_x100=300, width=70, height=50, view_flags=1)
    base = O.present(img, W, H, r["diagram"], r["edge_count"], tx=0, ty=0, **kw)
    for tx, ty in ((1, 1), (5, -3), (-7, 2)):
        sh = O.present(img, W, H, r["diagram"], r["edge_count"], tx=tx, ty=ty, **kw)
        ys, xs = slice(max(ty, 0), 50 + min(ty, 0)), slice(max(tx, 0), 70 + min(tx, 0))
        yb, xb = slice(max(-ty, 0), 50 - max(ty, 0)), slice(max(-tx, 0), 70 - max(tx, 0))
        assert np.array_equal(sh[ys, xs], base[yb, xb]), (tx, ty)


def test_topdown_rows(D, O):
    W, H = 12, 10
    img = D.synth_frame(0, 4, W, H, 3)
    r = O.depixelize(img, W, H, scale=2, want_diagram=True)
    kw = dict(mode=0, scale_x100=250, tx=1, ty=1, width=33, height=27)
    up = O.present(img, W, H, r["diagram"], r["edge_count"], view_flags=1, **kw)
    down = O.present(img, W, H, r["diagram"], r["edge_count"], view_flags=3, **kw)
    assert np.array_equal(down, up[::-1])


def _inside(a, b, s):
    dx, dy = b[0] - a[0], b[1] - a[1]
    e = dx * (s[1] - a[1]) - dy * (s[0] - a[0])
    return e > 0 or (e == 0 and (dy < 0 or (dy == 0 and dx > 0)))


def _covers(tri, s):
    a, b, c = tri
    o = (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])
    if o == 0:
        return False
    if o < 0:
        b, c = c, b
    return _inside(a, b, s) and _inside(b, c, s) and _inside(c, a, s)


@pytest.mark.parametrize("sub", [False, True])
def test_antialiased_cells_brute_force(D, O, sub):
    """Every sample evaluated over exact rationals, all cells scanned from
    the last in draw order (no neighbourhood assumption)."""
    W, H = 6, 5
    img = D.synth_frame(0, 11, W, H, 3)
    r = O.depixelize(img, W, H, scale=2, subdivide=sub, want_diagram=True)
    diag, ec = r["diagram"].reshape(H * W, 45, 2), r["edge_count"].reshape(-1)
    col = cell_colours(img, W, H).reshape(-1, 4).astype(np.int64)
    sc, tx, ty, ww, wh = 275, 1, 1, 19, 16
    tris = []
    for k in range(W * H):
        cell = [tuple(Fraction(float(v)) for v in diag[k, t]) for t in range(45)]
        tris.append([cell[3 * t:3 * t + 3] for t in range(max(0, int(ec[k]) - 2))])
    want = np.zeros((wh, ww, 4), dtype=np.uint8)
    for Y in range(wh):
        for X in range(ww):
            acc = np.zeros(4, dtype=np.int64)
            for ox, oy in AA_OFFSETS:
                s = (Fraction(8 * (X - tx) + ox, 8) * 100 / sc, Fraction(8 * (Y - ty) + oy, 8) * 100 / sc)
                c = np.array([0, 0, 0, 255], dtype=np.int64)
                for k in range(W * H - 1, -1, -1):
                    if any(_covers(t, s) for t in tris[k]):
                        c = col[k]
                        break
                acc += c
            want[Y, X] = (acc + 2) // 4
    got = O.present(img, W, H, r["diagram"], r["edge_count"], mode=0, scale_x100=sc, tx=tx, ty=ty, width=ww,
                    height=wh, view_flags=1)
    assert np.array_equal(got, want)


def test_antialiasing_of_a_flat_frame(D, O):
    """A single-colour frame: interior pixels keep the colour, a pixel whose
    samples straddle the image edge mixes it with the clear colour."""
    W, H = 8, 6
    img = np.zeros((H, W * 3), dtype=np.uint8)
    img[:, 0::3], img[:, 1::3], img[:, 2::3] = 200, 100, 40
    r = O.depixelize(img, W, H, scale=2, want_diagram=True)
    win = O.present(img, W, H, r["diagram"], r["edge_count"], mode=0, scale_x100=250, tx=0, ty=0, width=24,
                    height=18, view_flags=1)
    colour = np.array([40, 100, 200, 255], dtype=np.uint8)
    assert (win[1:14, 1:19] == colour).all()
    assert (win[16:, :] == [0, 0, 0, 255]).all() and (win[:, 21:] == [0, 0, 0, 255]).all()
    # window column 20 covers image x in [8, 8.4): samples at x 20+1/8.. -> outside; column 19 holds
    # samples at (19 + k/8) / 2.5 = 7.65 .. 7.95 -> inside
    assert (win[1:14, 19] == colour).all()
