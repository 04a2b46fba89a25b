// geom.hpp -- exact fixed-point cell geometry for the depixelization path.
//
// Every coordinate the reference produces lies on a dyadic lattice: cell
// pattern points and hull vertices on the quarter lattice, corner cuts on the
// 1/32 lattice, their midpoints on the 1/64 lattice.  This header computes the
// same values in integers (quarter units for cells and hulls, 1/64 units for
// subdivided polygons), so results are exact on every device and equal to the
// reference's float/double arithmetic.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dpx {

constexpr int kCellSize = 45;   // CELL_SIZE, kernel.cu:8
constexpr int kMaxPts = 16;     // createCellFromPattern pushes at most 2 points per case
constexpr int kMaxPoly = 32;    // subdivision emits at most 2 points per hull vertex

// neighbour offsets of graph bit e (diff, graph_functions.cu:194-235)
__host__ __device__ inline int edx(int e) { return (e == 0 || e == 3 || e == 5) ? -1 : ((e == 1 || e == 6) ? 0 : 1); }
__host__ __device__ inline int edy(int e) { return e <= 2 ? 1 : (e <= 4 ? 0 : -1); }
__host__ __device__ inline uint32_t bit(uint32_t v, int e) { return (v >> e) & 1u; }

// ---------------------------------------------------------------------------
// Cell pattern: createCellFromPattern (diagram_functions.cu:319-535), points in
// quarter units, pushed in the reference's case order 0,1,2,4,7,6,5,3.
// ---------------------------------------------------------------------------
__host__ __device__ inline int cell_points(uint32_t node, uint32_t left, uint32_t right, int* px, int* py)
{
    int n = 0;
    auto push = [&](int x, int y) { px[n] = x; py[n] = y; n++; };
    // case 0, up-left
    if (bit(node, 0)) {
        if (bit(node, 1) && !bit(node, 3)) push(-1, 3);
        else if (bit(node, 3) && !bit(node, 1)) push(1, 5);
        else { push(-1, 3); push(1, 5); }
    } else {
        if (bit(left, 2)) push(1, 3); else push(0, 4);
    }
    // case 1, up
    if (bit(node, 1)) { push(0, 4); push(4, 4); } else push(2, 3);
    // case 2, up-right
    if (bit(node, 2)) {
        if (bit(node, 1) && !bit(node, 4)) push(5, 3);
        else if (bit(node, 4) && !bit(node, 1)) push(3, 5);
        else { push(3, 5); push(5, 3); }
    } else {
        if (bit(right, 0)) push(3, 3); else push(4, 4);
    }
    // case 4, right
    if (bit(node, 4)) { push(4, 4); push(4, 0); } else push(3, 3);
    // case 7, down-right
    if (bit(node, 7)) {
        if (bit(node, 4) && !bit(node, 6)) push(3, -1);
        else if (bit(node, 6) && !bit(node, 4)) push(5, 1);
        else { push(5, 1); push(3, -1); }
    } else {
        if (bit(right, 5)) push(3, 1); else push(4, 0);
    }
    // case 6, down
    if (bit(node, 6)) { push(4, 0); push(0, 0); } else push(1, 3);
    // case 5, down-left
    if (bit(node, 5)) {
        if (bit(node, 3) && !bit(node, 6)) push(1, -1);
        else if (bit(node, 6) && !bit(node, 3)) push(-1, 1);
        else { push(1, -1); push(-1, 1); }
    } else {
        if (bit(left, 7)) push(1, 1); else push(0, 0);
    }
    // case 3, left
    if (bit(node, 3)) { push(0, 0); push(0, 4); } else push(1, 2);
    return n;
}

// The cell depends on 12 bits: the node byte, the left neighbour's bits 2 and 7
// and the right neighbour's bits 0 and 5 (kernel.cu:204-206).
__host__ __device__ inline uint32_t cell_key(uint32_t node, uint32_t left, uint32_t right)
{
    return (node & 0xFFu) | (bit(left, 2) << 8) | (bit(left, 7) << 9) | (bit(right, 0) << 10) |
           (bit(right, 5) << 11);
}

__host__ __device__ inline void key_bytes(uint32_t key, uint32_t* node, uint32_t* left, uint32_t* right)
{
    *node = key & 0xFFu;
    *left = (bit(key, 8) << 2) | (bit(key, 9) << 7);
    *right = (bit(key, 10) << 0) | (bit(key, 11) << 5);
}

template <typename T>
__host__ __device__ inline void swap2(T* a, T* b)
{
    T t = *a;
    *a = *b;
    *b = t;
}

// The reference's lexicographic sort (sort diagram_functions.cu:82-129 with
// sort_y 36-80): repeated selection of the smallest-x group, whose members
// are swapped to the front in scan order, after which one selection pass
// brings the smallest y of the group first.  The order it leaves behind for
// groups of 3+ equal x is part of the hull's output and is kept as is.
__host__ __device__ inline void quirky_sort(int* px, int* py, int n)
{
    int tie_index[kMaxPts];
    int sorted = 0;
    while (sorted < n) {
        int minor = 1 << 20, tie = 0;
        tie_index[0] = sorted;
        for (int i = sorted; i < n; i++) {
            if (px[i] < minor) { minor = px[i]; tie = 0; tie_index[0] = i; }
            else if (px[i] == minor) { tie++; tie_index[tie] = i; }
        }
        if (tie > 0) {
            for (int t = 0; t <= tie; t++) {
                swap2(&px[t + sorted], &px[tie_index[t]]);
                swap2(&py[t + sorted], &py[tie_index[t]]);
            }
            // sort_y over [sorted, sorted + tie]
            int miny = 1 << 20, ytie = 0;
            int yidx[kMaxPts];
            yidx[0] = sorted;
            for (int i = sorted; i <= sorted + tie; i++) {
                if (py[i] < miny) { miny = py[i]; ytie = 0; yidx[0] = i; }
                else if (py[i] == miny) { ytie++; yidx[ytie] = i; }
            }
            if (ytie > 0) {
                for (int t = 0; t <= ytie; t++) {
                    swap2(&px[t + sorted], &px[yidx[t]]);
                    swap2(&py[t + sorted], &py[yidx[t]]);
                }
            } else {
                swap2(&px[sorted], &px[yidx[0]]);
                swap2(&py[sorted], &py[yidx[0]]);
            }
        } else {
            swap2(&px[sorted], &px[tie_index[0]]);
            swap2(&py[sorted], &py[tie_index[0]]);
        }
        sorted += tie + 1;
    }
}

// OA x OB (cross, diagram_functions.cu:225-230); the reference evaluates it on
// x100 float coordinates, which is exact and has the same sign.
__host__ __device__ inline int cross3(int ox, int oy, int ax, int ay, int bx, int by)
{
    return (ax - ox) * (by - oy) - (ay - oy) * (bx - ox);
}

// convex_hull (diagram_functions.cu:238-316): monotone chain on the sorted
// points, popping on cross <= 0.  Writes k points (the last repeats the
// first) and returns k-1, as the reference does.
__host__ __device__ inline int hull(int* px, int* py, int n, int* hx, int* hy)
{
    quirky_sort(px, py, n);
    int k = 0;
    for (int i = 0; i < n; i++) {
        while (k >= 2 && cross3(hx[k - 2], hy[k - 2], hx[k - 1], hy[k - 1], px[i], py[i]) <= 0) k--;
        hx[k] = px[i];
        hy[k] = py[i];
        k++;
    }
    for (int i = n - 2, t = k + 1; i >= 0; i--) {
        while (k >= t && cross3(hx[k - 2], hy[k - 2], hx[k - 1], hy[k - 1], px[i], py[i]) <= 0) k--;
        hx[k] = px[i];
        hy[k] = py[i];
        k++;
    }
    return k - 1;
}

// Hull of a 12-bit cell key in quarter units; returns the vertex count.
// hx/hy need 2*kMaxPts entries.
__host__ __device__ inline int cell_hull(uint32_t key, int* hx, int* hy)
{
    int px[kMaxPts], py[kMaxPts];
    uint32_t node, left, right;
    key_bytes(key, &node, &left, &right);
    int n = cell_points(node, left, right, px, py);
    return hull(px, py, n, hx, hy);
}

// ---------------------------------------------------------------------------
// Subdivision helpers (subdivision_functions.cu).  Inputs in quarter units,
// outputs in 1/64 units.
// ---------------------------------------------------------------------------
__host__ __device__ inline int mod_n(int i, int n) { return (n + (i % n)) % n; }

// The functions below take a polygon as any indexable type (a pointer to
// int, or an accessor whose operator[] reads packed storage).

// isLinkedEdge (245-424): link direction of hull edge idx, or -1 (border).
// slope == +-1 is dy == +-dx, dx != 0; mid_y > 0.5 is y_a + y_b > 2 quarter units * 2.
template <class HX, class HY>
__host__ __device__ inline int linked_edge(const HX& hx, const HY& hy, int n, uint32_t node, int idx)
{
    int b = mod_n(idx + 1, n);
    int ax = hx[idx], ay = hy[idx], bx = hx[b], by = hy[b];
    int dx = bx - ax, dy = by - ay;
    bool sp = dx != 0 && dy == dx;
    bool sm = dx != 0 && dy == -dx;
    int my2 = ay + by;  // 2 * mid_y in quarter units; mid_y > 0.5 <=> my2 > 4
    if (ay == by && ay > 2 && bit(node, 1)) return 1;
    if (ax == bx && ax > 2 && bit(node, 4)) return 4;
    if (ay == by && ay < 2 && bit(node, 6)) return 6;
    if (ax == bx && ax < 2 && bit(node, 3)) return 3;
    if (sp && my2 > 4 && bit(node, 0)) return 0;
    if (sm && my2 > 4 && bit(node, 2)) return 2;
    if (sp && my2 < 4 && bit(node, 7)) return 7;
    if (sm && my2 < 4 && bit(node, 5)) return 5;
    return -1;
}

// getQ_i / getR_i (42-122) on a quarter-unit polygon, result in 1/64 units.
// |P_next - P| <= 1  <=>  dx^2 + dy^2 <= 16 (quarter units).
template <class HX, class HY>
__host__ __device__ inline void q_point(const HX& hx, const HY& hy, int n, int i, int* qx, int* qy)
{
    int p = i % n;
    if (p < 0) p += n;
    int q = (i + 1) % n;
    int x0 = hx[p], y0 = hy[p], x1 = hx[q], y1 = hy[q];
    int dx = x1 - x0, dy = y1 - y0;
    if (dx * dx + dy * dy <= 16) { *qx = 4 * (3 * x0 + x1); *qy = 4 * (3 * y0 + y1); }
    else                         { *qx = 2 * (7 * x0 + x1); *qy = 2 * (7 * y0 + y1); }
}

template <class HX, class HY>
__host__ __device__ inline void r_point(const HX& hx, const HY& hy, int n, int i, int* rx, int* ry)
{
    int p = i % n;
    if (p < 0) p += n;
    int q = (i + 1) % n;
    int x0 = hx[p], y0 = hy[p], x1 = hx[q], y1 = hy[q];
    int dx = x1 - x0, dy = y1 - y0;
    if (dx * dx + dy * dy <= 16) { *rx = 4 * (x0 + 3 * x1); *ry = 4 * (y0 + 3 * y1); }
    else                         { *rx = 2 * (x0 + 7 * x1); *ry = 2 * (y0 + 7 * y1); }
}

// getPointIndex (527-538): index of (x,y) in the polygon, 0 if absent.
template <class HX, class HY>
__host__ __device__ inline int point_index(int x, int y, const HX& hx, const HY& hy, int n)
{
    for (int i = 0; i < n; i++)
        if (hx[i] == x && hy[i] == y) return i;
    return 0;
}

// subdivision (subdivision_functions.cu:564-669) of one cell, emitting the
// new polygon's points in order.
//   hx,hy,n   the cell's hull (quarter units, the snapshot of kernel.cu:434),
//             n <= 16
//   node      the cell's resolved graph byte
//   nbr(e, f) calls f(lx, ly, ln) with the hull of the neighbour across
//             graph edge e (snapshot) and returns what f returns
//   tjunction(x4, y4) -> bool   checkTJunction (170-242) for hull vertex (x4,y4)
//   emit(x, y)  receives each point of the new polygon, in 1/64 units
// Emits at most 2n points.
template <class HX, class HY, class NbrHull, class TJunction, class Emit>
__host__ __device__ inline void subdivide_cell_emit(const HX& hx, const HY& hy, int n, uint32_t node, NbrHull nbr,
                                                    TJunction tjunction, Emit emit)
{
    // link direction of every edge, 4 bits each (15 = border)
    uint64_t links = 0;
    for (int e = 0; e < n; e++) links |= (uint64_t)(linked_edge(hx, hy, n, node, e) & 15) << (4 * e);
    auto link = [&](int e) { return (int)((links >> (4 * e)) & 15u); };
    for (int k = 0; k < n; k++) {
        int prev = mod_n(k - 1, n);
        const int lk = link(k), lp = link(prev);
        bool sk = lk != 15, sp = lp != 15;
        if (!sk && !sp) {
            // two border edges: keep a T-junction corner, else cut it (R_{k-1}, Q_k)
            if (tjunction(hx[k], hy[k])) {
                emit(16 * hx[k], 16 * hy[k]);
            } else {
                int rx, ry, qx, qy;
                r_point(hx, hy, n, k - 1, &rx, &ry);
                q_point(hx, hy, n, k, &qx, &qy);
                emit(rx, ry);
                emit(qx, qy);
            }
        } else if (!sk && sp) {
            // edge k border, edge k-1 linked: blend with the neighbour's R
            int qx, qy;
            q_point(hx, hy, n, k, &qx, &qy);
            const int e = lp;
            int rx = 0, ry = 0;
            nbr(e, [&](const auto& lx, const auto& ly, int ln) {
                int op = point_index(hx[k] - 4 * edx(e), hy[k] - 4 * edy(e), lx, ly, ln);
                r_point(lx, ly, ln, op - 1, &rx, &ry);
                return 0;
            });
            rx += 64 * edx(e); ry += 64 * edy(e);
            emit((qx + rx) / 2, (qy + ry) / 2);
            emit(qx, qy);
        } else if (sk && !sp) {
            // edge k linked, edge k-1 border: blend with the neighbour's Q
            const int e = lk;
            int qx = 0, qy = 0;
            nbr(e, [&](const auto& lx, const auto& ly, int ln) {
                int op = point_index(hx[k] - 4 * edx(e), hy[k] - 4 * edy(e), lx, ly, ln);
                q_point(lx, ly, ln, op, &qx, &qy);
                return 0;
            });
            qx += 64 * edx(e); qy += 64 * edy(e);
            int rx, ry;
            r_point(hx, hy, n, k - 1, &rx, &ry);
            emit(rx, ry);
            emit((rx + qx) / 2, (ry + qy) / 2);
        } else {
            emit(16 * hx[k], 16 * hy[k]);
        }
    }
}

// subdivide_cell_emit into arrays: nbr(e, lx, ly) -> ln copies the
// neighbour's hull into lx, ly; the new polygon goes to ox, oy (2n entries);
// returns its size.
template <class NbrHull, class TJunction>
__host__ __device__ inline int subdivide_cell(const int* hx, const int* hy, int n, uint32_t node, NbrHull nbr,
                                              TJunction tjunction, int* ox, int* oy)
{
    int lx[2 * kMaxPts], ly[2 * kMaxPts];
    int out = 0;
    auto nbr_f = [&](int e, auto f) {
        int ln = nbr(e, lx, ly);
        return f((const int*)lx, (const int*)ly, ln);
    };
    subdivide_cell_emit(hx, hy, n, node, nbr_f, tjunction, [&](int x, int y) {
        ox[out] = x;
        oy[out] = y;
        out++;
    });
    return out;
}

// ---------------------------------------------------------------------------
// Ear clipping (process, triangulate_functions.cu:108-217) on a 1/64-unit
// polygon.  Writes triangle vertex indices into tri[3*(n-2)] and returns the
// number of triangles written; *ok = 0 when the reference's error counter
// runs out (the remaining slots stay degenerate).
// ---------------------------------------------------------------------------
__host__ __device__ inline bool inside_tri(int ax, int ay, int bx, int by, int cx, int cy, int px, int py)
{
    // insideTriangle (41-67): all three edge crosses >= 0
    int c1 = (cx - bx) * (py - by) - (cy - by) * (px - bx);
    int c2 = (bx - ax) * (py - ay) - (by - ay) * (px - ax);
    int c3 = (ax - cx) * (py - cy) - (ay - cy) * (px - cx);
    return c1 >= 0 && c2 >= 0 && c3 >= 0;
}

__host__ __device__ inline int ear_clip(const int* x, const int* y, int n, uint8_t* tri, int* ok)
{
    *ok = 1;
    if (n < 3 || n > kMaxPoly) { *ok = 0; return 0; }
    int V[kMaxPoly];
    // area (6-34): twice the signed area; > 0 keeps the order
    int a2 = 0;
    for (int p = n - 1, q = 0; q < n; p = q++) a2 += x[p] * y[q] - x[q] * y[p];
    if (a2 > 0) { for (int v = 0; v < n; v++) V[v] = v; }
    else        { for (int v = 0; v < n; v++) V[v] = n - 1 - v; }
    int nv = n, ntri = 0, count = 2 * nv;
    for (int v = nv - 1; nv > 2;) {
        if (0 >= (count--)) { *ok = 0; return ntri; }
        int u = v;
        if (nv <= u) u = 0;
        v = u + 1;
        if (nv <= v) v = 0;
        int w = v + 1;
        if (nv <= w) w = 0;
        // snip (70-105): reject if cross < EPSILON, i.e. cross <= 0 on the lattice
        int ax = x[V[u]], ay = y[V[u]], bx = x[V[v]], by = y[V[v]], cx = x[V[w]], cy = y[V[w]];
        bool ear = ((bx - ax) * (cy - ay) - (by - ay) * (cx - ax)) > 0;
        if (ear) {
            for (int p = 0; p < nv; p++) {
                if (p == u || p == v || p == w) continue;
                if (inside_tri(ax, ay, bx, by, cx, cy, x[V[p]], y[V[p]])) { ear = false; break; }
            }
        }
        if (ear) {
            tri[3 * ntri + 0] = (uint8_t)V[u];
            tri[3 * ntri + 1] = (uint8_t)V[v];
            tri[3 * ntri + 2] = (uint8_t)V[w];
            ntri++;
            for (int s = v, t = v + 1; t < nv; s++, t++) V[s] = V[t];
            nv--;
            count = 2 * nv;
        }
    }
    return ntri;
}

// ---------------------------------------------------------------------------
// Coverage rule of the raster (replaces the GL rasteriser).  A sample on an
// edge counts as inside iff it is inside after the symbolic shift (+e, +e^2):
// for a counter-clockwise edge (dx,dy) that is dy < 0, or dy == 0 and dx > 0.
// Coordinates are integers in units of 1/(128*scale) of a cell (int), or of
// 1/(512*scale_x100) of a cell for the presentation stage (long long).
// ---------------------------------------------------------------------------
template <class T>
__host__ __device__ inline bool edge_in(T px, T py, T qx, T qy, T sx, T sy)
{
    T dx = qx - px, dy = qy - py;
    T e = dx * (sy - py) - dy * (sx - px);
    return e > 0 || (e == 0 && (dy < 0 || (dy == 0 && dx > 0)));
}

// triangle (a,b,c) covers sample s; degenerate triangles cover nothing
template <class T>
__host__ __device__ inline bool tri_covers(T ax, T ay, T bx, T by, T cx, T cy, T sx, T sy)
{
    T o = (bx - ax) * (cy - ay) - (by - ay) * (cx - ax);
    if (o == 0) return false;
    if (o < 0) { T tx = bx, ty = by; bx = cx; by = cy; cx = tx; cy = ty; }
    return edge_in(ax, ay, bx, by, sx, sy) && edge_in(bx, by, cx, cy, sx, sy) &&
           edge_in(cx, cy, ax, ay, sx, sy);
}

// floor(a / d) for d > 0 and |a| < 2^26: the single-precision quotient is
// within 1 of it, and one exact step on each side fixes it.
__host__ __device__ inline int floor_div(int a, int d)
{
    int q = (int)floorf((float)a / (float)d);
    if (q * d > a) q--;
    else if ((q + 1) * d <= a) q++;
    return q;
}

// Bits b (0 <= b < mw <= 32) of the samples sx = 128 (kmin + b) + 64 on row
// sy covered by the counter-clockwise triangle (a, b, c), under edge_in's
// rule -- the same bits as edge_in on every sample.  Along a row an edge's
// function is e = K - dy * sx: with dy > 0 the covered columns are those
// with e > 0, b <= floor((N - 1) / D); with dy < 0 (ties cover) those with
// e >= 0, b >= -floor(N / -D); with dy = 0 all or none; N = e at b = 0, D =
// 128 dy.  Coordinates are bounded so that every product fits in int.
__host__ __device__ inline uint32_t tri_row_bits(int ax, int ay, int bx, int by, int cx, int cy, int sy, int kmin,
                                                 int mw)
{
    int lo = 0, hi = mw - 1;
    auto clip = [&](int px, int py, int qx, int qy) {
        const int dx = qx - px, dy = qy - py;
        const int K = dx * (sy - py) + dy * px;  // e at sx = 0
        if (dy == 0) {
            if (!(K > 0 || (K == 0 && dx > 0))) hi = -1;
            return;
        }
        const int N = K - dy * (128 * kmin + 64), D = 128 * dy;
        if (dy > 0) hi = min(hi, floor_div(N - 1, D));
        else lo = max(lo, -floor_div(N, -D));
    };
    clip(ax, ay, bx, by);
    clip(bx, by, cx, cy);
    clip(cx, cy, ax, ay);
    if (lo > hi) return 0u;
    return ((2u << hi) - 1u) & ~((1u << lo) - 1u);
}

// tri_row_bits of every row r = 0 .. MW - 1 at once (sample rows
// sy = 128 (kmin + r) + 64), ORed into cov[r].  Along the rows an edge's
// column bound moves by a fixed rational step, so each bound is kept as a
// quotient and remainder and advanced by one exact step per row instead of
// a division per row.  Same bits as tri_row_bits on every row.
template <int MW>
__host__ __device__ inline void tri_rows_or(int ax, int ay, int bx, int by, int cx, int cy, int kmin, uint32_t* cov)
{
    const int ymin = min(ay, min(by, cy)), ymax = max(ay, max(by, cy));
    const int sy0 = 128 * kmin + 64;
    // per edge: dy > 0 bounds hi, dy < 0 bounds lo, dy == 0 admits a range
    // of rows.  q, rem: floor of the bound's numerator over D = 128 |dy|.
    int q[3], rem[3], qs[3], rs[3], D[3], kind[3];  // kind: 1 hi, -1 lo, 0 horizontal
    int ylo = sy0, yhi = sy0 + 128 * (MW - 1);       // rows the horizontal edges admit (sample y)
    const int px_[3] = { ax, bx, cx }, py_[3] = { ay, by, cy };
#pragma unroll
    for (int e = 0; e < 3; e++) {
        const int px = px_[e], py = py_[e], qx = px_[(e + 1) % 3], qy = py_[(e + 1) % 3];
        const int dx = qx - px, dy = qy - py;
        kind[e] = dy > 0 ? 1 : (dy < 0 ? -1 : 0);
        D[e] = 128 * (dy < 0 ? -dy : dy);
        q[e] = rem[e] = qs[e] = rs[e] = 0;
        if (dy == 0) {
            // K = dx (sy - py): covered iff K > 0, or K == 0 and dx > 0
            if (dx > 0) ylo = max(ylo, py);          // sy >= py
            else if (dx < 0) yhi = min(yhi, py - 1);  // sy < py
            else yhi = ylo - 1;                      // (a point: nothing)
            continue;
        }
        // numerator at row 0: N = dx (sy0 - py) + dy (px - sy0); hi uses N - 1, lo uses N
        const int N0 = dx * (sy0 - py) + dy * (px - sy0) - (dy > 0 ? 1 : 0);
        q[e] = floor_div(N0, D[e]);
        rem[e] = N0 - q[e] * D[e];
        qs[e] = floor_div(128 * dx, D[e]);
        rs[e] = 128 * dx - qs[e] * D[e];
    }
#pragma unroll
    for (int r = 0; r < MW; r++) {
        const int sy = sy0 + 128 * r;
        int lo = 0, hi = MW - 1;
#pragma unroll
        for (int e = 0; e < 3; e++) {
            if (kind[e] > 0) hi = min(hi, q[e]);
            else if (kind[e] < 0) lo = max(lo, -q[e]);
            // advance to the next row
            q[e] += qs[e];
            rem[e] += rs[e];
            if (rem[e] >= D[e] && kind[e] != 0) {
                rem[e] -= D[e];
                q[e]++;
            }
        }
        if (sy >= ymin && sy <= ymax && sy >= ylo && sy <= yhi && lo <= hi)
            cov[r] |= ((2u << hi) - 1u) & ~((1u << lo) - 1u);
    }
}

// The coverage rows of a counter-clockwise polygon that ear clipping has
// cut completely into triangles, in one pass over its edges instead of one
// over each triangle.  Under edge_in's rule a sample is covered by a
// triangle iff the sample moved by (dx, dy), 0 < dy << dx infinitesimal, is
// inside it; the number of the cut (counter-clockwise) triangles covering a
// sample is then the polygon's winding number w >= 0 around the moved
// sample (the triangles' boundaries add up to the polygon's: each cut edge
// enters twice, opposite ways), and w = sum, over the edges that cross the
// moved row (min(py, qy) <= sy < max(py, qy)) with the sample left of the
// crossing, of +1 (upward edge) or -1 (downward).  The columns left of the
// crossing: an upward edge's covered side (tri_rows_or's hi bound), a
// downward edge's uncovered side (below its lo bound).  The function
// returns in cov the parity of w per sample (the exclusive or of those
// column masks) and in the result the sum of w over the window's samples
// (the masks' signed bit counts).  The union of the triangles (w >= 1) is
// the parity exactly when every w is 0 or 1, i.e. when that sum equals the
// number of set bits of cov: the caller's check.
// P: vertex k at P[stride * k], x | y << 16 in 1/64 units of a cell;
// coordinates are scaled by 2 S as for tri_rows_or.
template <int MW>
__host__ __device__ inline int poly_rows_xor(const uint32_t* P, int stride, int m, int S2, int kmin, uint32_t* cov)
{
    const int sy0 = 128 * kmin + 64;
    int wsum = 0;
    for (int k = 0; k < m; k++) {
        const uint32_t A = P[stride * k], B = P[stride * (k + 1 == m ? 0 : k + 1)];
        const int px = (int)(int16_t)(A & 0xFFFFu) * S2, py = ((int)A >> 16) * S2;
        const int qx = (int)(int16_t)(B & 0xFFFFu) * S2, qy = ((int)B >> 16) * S2;
        const int dx = qx - px, dy = qy - py;
        if (dy == 0) continue;
        const int ylo = min(py, qy), yhi = max(py, qy);
        const int D = 128 * (dy < 0 ? -dy : dy);
        // column bound at row 0 (as tri_rows_or): N = dx (sy0 - py) + dy (px - sy0)
        const int N0 = dx * (sy0 - py) + dy * (px - sy0) - (dy > 0 ? 1 : 0);
        int q = floor_div(N0, D), rem = N0 - q * D;
        const int qs = floor_div(128 * dx, D), rs = 128 * dx - qs * D;
#pragma unroll
        for (int r = 0; r < MW; r++) {
            const int sy = sy0 + 128 * r;
            const int b = dy > 0 ? q : -q - 1;  // last column left of the crossing
            const uint32_t msk = b < 0 ? 0u : (b >= MW - 1 ? (MW == 32 ? ~0u : (1u << MW) - 1u) : (2u << b) - 1u);
            if (sy >= ylo && sy < yhi) {
                cov[r] ^= msk;
                wsum += dy > 0 ? __builtin_popcount(msk) : -__builtin_popcount(msk);
            }
            q += qs;
            rem += rs;
            if (rem >= D) {
                rem -= D;
                q++;
            }
        }
    }
    return wsum;
}

}  // namespace dpx
