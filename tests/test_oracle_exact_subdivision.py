"""Subdivided vertices of one 640x480 frame of the large-output workload
(kind 4: 256 colours, 25% dither, a diagonal-crossing lattice) against an
exact-rational evaluation of the reference's formulas.

getQ_i / getR_i (subdivision_functions.cu:42-122) and midPoint (554-561)
are evaluated with Fractions on the cells' hulls: the corner cuts
3/4 P + 1/4 Pn and 1/4 P + 3/4 Pn when |Pn - P| <= 1, else 7/8 - 1/8; the
neighbour's cut moved into this cell's frame; midpoints.  Which case applies
at each vertex comes from the oracle's edge status and link direction, and a
two-border-edge vertex is either kept (a T-junction corner) or cut.  Every
vertex the oracle computed in its single/double-precision arithmetic must be
the exact rational: the vertex error of the FP64 evaluation is 0 (the GPU
equals the oracle bit for bit, tests/test_gpu_precision.py)."""
from fractions import Fraction as F

import numpy as np

W, H = 640, 480
EDX = [-1, 0, 1, -1, 1, -1, 0, 1]
EDY = [1, 1, 1, 0, 0, -1, -1, -1]


def q_cut(P, Pn):
    dx, dy = Pn[0] - P[0], Pn[1] - P[1]
    if dx * dx + dy * dy <= 1:
        return (P[0] * F(3, 4) + Pn[0] / 4, P[1] * F(3, 4) + Pn[1] / 4)
    return (P[0] * F(7, 8) + Pn[0] / 8, P[1] * F(7, 8) + Pn[1] / 8)


def r_cut(P, Pn):
    dx, dy = Pn[0] - P[0], Pn[1] - P[1]
    if dx * dx + dy * dy <= 1:
        return (P[0] / 4 + Pn[0] * F(3, 4), P[1] / 4 + Pn[1] * F(3, 4))
    return (P[0] / 8 + Pn[0] * F(7, 8), P[1] / 8 + Pn[1] * F(7, 8))


def hull(od, oec, k):
    n = int(oec[k])
    return [(F(float(od[k, t, 0])), F(float(od[k, t, 1]))) for t in range(n)]


def index_of(p, h):
    for i, q in enumerate(h):
        if q == p:
            return i
    return 0  # getPointIndex returns 0 when absent


def test_subdivided_vertices_are_exact_rationals(D, O):
    img = D.synth_frame(4, 0, W, H, 3)
    g = O.heuristic_cross(O.trivial_cross(O.similarity_graph(img, W, H)))
    od, oec = O.cells(g)
    sd, sec, es, li = O.subdivide(img, W, H, g, od, oec)
    sd = sd.reshape(W * H, 45, 2)
    gf = g.reshape(-1)
    n_vertices = n_cells = 0
    for k in range(W * H):
        if gf[k] == 90:
            continue
        h = hull(od, oec, k)
        n = len(h)
        i, j = k % W, k // W
        got = [(F(float(sd[k, t, 0])), F(float(sd[k, t, 1]))) for t in range(int(sec[k]))]
        pos = 0
        for v in range(n):
            pv = (v - 1) % n
            sk, sp = bool(es[k, v]), bool(es[k, pv])
            if not sk and not sp:
                corner = h[v]
                cut = [r_cut(h[pv], h[v]), q_cut(h[v], h[(v + 1) % n])]
                if got[pos] == corner and (pos + 1 >= len(got) or got[pos:pos + 2] != cut):
                    want = [corner]
                else:
                    want = cut
            elif not sk and sp:
                q = q_cut(h[v], h[(v + 1) % n])
                e = int(li[k, pv])
                nk = (j + EDY[e]) * W + i + EDX[e]
                nh = hull(od, oec, nk)
                op = index_of((h[v][0] - EDX[e], h[v][1] - EDY[e]), nh)
                r = r_cut(nh[(op - 1) % len(nh)], nh[op % len(nh)])
                r = (r[0] + EDX[e], r[1] + EDY[e])
                want = [((q[0] + r[0]) / 2, (q[1] + r[1]) / 2), q]
            elif sk and not sp:
                e = int(li[k, v])
                nk = (j + EDY[e]) * W + i + EDX[e]
                nh = hull(od, oec, nk)
                op = index_of((h[v][0] - EDX[e], h[v][1] - EDY[e]), nh)
                q = q_cut(nh[op % len(nh)], nh[(op + 1) % len(nh)])
                q = (q[0] + EDX[e], q[1] + EDY[e])
                r = r_cut(h[pv], h[v])
                want = [r, ((r[0] + q[0]) / 2, (r[1] + q[1]) / 2)]
            else:
                want = [h[v]]
            assert got[pos:pos + len(want)] == want, f"cell ({i},{j}) vertex {v}: {got[pos:pos + len(want)]} != {want}"
            pos += len(want)
            n_vertices += len(want)
        assert pos == len(got), f"cell ({i},{j}): {len(got)} vertices, {pos} evaluated"
        n_cells += 1
    print(f"640x480 kind-4 frame: {n_cells} smoothed cells, {n_vertices} vertices, all equal to the exact "
          f"rationals (vertex error 0)")
    assert n_vertices > 100000
