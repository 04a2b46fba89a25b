// stages.hip -- the per-stage kernels (one kernel per stage of kernel.cu's
// sequence, intermediates in HBM in the reference's formats), the stage entry
// points of the C ABI, the plain per-frame path and the launch_kernel drop-in.
//
// Launch shape: 16x16 threads, one thread per pixel (or per output pixel for
// the raster).  These kernels favour a direct mapping of the reference's data
// formats over speed; the batched path in fused.hip is the fast one.
#include <stdlib.h>
#include <string.h>

#include <utility>

#include "common.hpp"
#include "geom.hpp"
#include "graph.hpp"
#include "pixel.hpp"

namespace dpx {
namespace {

constexpr int kB = 16;

dim3 grid2(int W, int H) { return dim3((unsigned)((W + kB - 1) / kB), (unsigned)((H + kB - 1) / kB)); }

__device__ inline bool in_frame(int i, int j, int W, int H) { return i >= 0 && j >= 0 && i < W && j < H; }

// graph_Kernel (kernel.cu:140-159): 8 similarity bits per pixel
__global__ void k_graph(ImgView im, uint8_t* __restrict__ graph)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= im.W || j >= im.H) return;
    Yuv c = yuv_of(load_px(im, i, j), im.flags, colour_channels(im));
    uint32_t g = 0;
#pragma unroll
    for (int e = 0; e < 8; e++) {
        int ni = i + edx(e), nj = j + edy(e);
        if (!in_frame(ni, nj, im.W, im.H)) continue;
        if (similar(c, yuv_of(load_px(im, ni, nj), im.flags, colour_channels(im)))) g |= 1u << e;
    }
    graph[(size_t)j * im.W + i] = (uint8_t)g;
}

__device__ inline void gather_nb(const uint8_t* g, int W, int H, int i, int j, uint32_t* nb)
{
#pragma unroll
    for (int e = 0; e < 8; e++) {
        int ni = i + edx(e), nj = j + edy(e);
        nb[e] = in_frame(ni, nj, W, H) ? (uint32_t)g[(size_t)nj * W + ni] : 0u;
    }
}

// trivial_cross_Kernel (kernel.cu:162-177), in place: reads orthogonal bits,
// clears the pixel's own diagonal bits only.
__global__ void k_trivial(uint8_t* graph, int W, int H)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= W || j >= H) return;
    uint32_t nb[8];
    gather_nb(graph, W, H, i, j, nb);
    uint8_t* p = &graph[(size_t)j * W + i];
    uint32_t c = *p;
    *p = (uint8_t)(c & ~trivial_clear_mask(c, nb));
}

// ambiguous_cross_Kernel (kernel.cu:180-189): decisions on the snapshot aux
__global__ void k_heuristic(uint8_t* __restrict__ graph, const uint8_t* __restrict__ aux, int W, int H)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= W || j >= H) return;
    uint32_t nb[8];
    gather_nb(aux, W, H, i, j, nb);
    size_t idx = (size_t)j * W + i;
    uint32_t c = aux[idx];
    // a caller's graph may hold links that leave the frame (the similarity
    // stage never makes one): a walk reaching past the edge finds a node
    // without links there and stops
    auto get = [=](int x, int y) -> uint32_t {
        return (unsigned)x < (unsigned)W && (unsigned)y < (unsigned)H ? aux[(size_t)y * W + x] : 0u;
    };
    graph[idx] = (uint8_t)(c & ~heuristic_clear_mask(get, i, j, c, nb));
}

// cells_Kernel (kernel.cu:192-213): hull points in 45-slot cells.  Slots past
// the closing point are written as (0,0).
__global__ void k_cells(const uint8_t* __restrict__ graph, int W, int H, dpx_point* __restrict__ diagram,
                        int32_t* __restrict__ edge_count)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= W || j >= H) return;
    size_t N = (size_t)W * H, idx = (size_t)j * W + i;
    uint32_t left = idx > 0 ? graph[idx - 1] : 0u;
    uint32_t right = idx + 1 < N ? graph[idx + 1] : 0u;
    int hx[2 * kMaxPts], hy[2 * kMaxPts];
    int n = cell_hull(cell_key(graph[idx], left, right), hx, hy);
    dpx_point* c = &diagram[idx * kCellSize];
    for (int t = 0; t < kCellSize; t++) {
        dpx_point p;
        p.x = t <= n ? 0.25f * (float)hx[t] : 0.0f;
        p.y = t <= n ? 0.25f * (float)hy[t] : 0.0f;
        c[t] = p;
    }
    edge_count[idx] = n;
}

__device__ inline int load_quarter(const dpx_point* c, int n, int* x, int* y)
{
    for (int t = 0; t < n; t++) {
        x[t] = (int)(c[t].x * 4.0f);
        y[t] = (int)(c[t].y * 4.0f);
    }
    return n;
}

// subdivision_Kernel (kernel.cu:216-261) on the snapshots daux/eaux.
__global__ void k_subdivide(ImgView im, const uint8_t* __restrict__ graph, dpx_point* __restrict__ diagram,
                            const dpx_point* __restrict__ daux, int32_t* __restrict__ edge_count,
                            const int32_t* __restrict__ eaux)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int W = im.W;
    if (i >= W || j >= im.H) return;
    int idx = j * W + i;
    uint32_t node = graph[idx];
    if (node == 90) return;  // internal node (kernel.cu:231)
    int n = eaux[idx];
    // subdivide_cell emits up to 2n points into kMaxPoly slots; real hulls
    // have at most 8 vertices, larger caller-supplied counts leave the cell
    if (n < 3 || n > kMaxPoly / 2) return;
    int hx[2 * kMaxPts], hy[2 * kMaxPts];
    load_quarter(&daux[(size_t)idx * kCellSize], n, hx, hy);
    TjView tv = tj_view(im);
    auto nbr = [&](int e, int* lx, int* ly) -> int {
        int q = idx + edy(e) * W + edx(e);
        int ln = eaux[q];
        if (ln > 2 * kMaxPts - 1) ln = 2 * kMaxPts - 1;
        return load_quarter(&daux[(size_t)q * kCellSize], ln, lx, ly);
    };
    auto tj = [&](int x4, int y4) -> bool { return tjunction(tv, i, j, x4, y4); };
    int ox[kMaxPoly], oy[kMaxPoly];
    int m = subdivide_cell(hx, hy, n, node, nbr, tj, ox, oy);
    dpx_point* c = &diagram[(size_t)idx * kCellSize];
    for (int t = 0; t < m; t++) {
        dpx_point p;
        p.x = (float)ox[t] * (1.0f / 64.0f);
        p.y = (float)oy[t] * (1.0f / 64.0f);
        c[t] = p;
    }
    edge_count[idx] = m;
}

// triangulate_Kernel (kernel.cu:264-282): ear clipping, then every slot gets
// the pixel offset; unwritten slots stay at the pixel origin.
__global__ void k_triangulate(dpx_point* __restrict__ diagram, const int32_t* __restrict__ edge_count, int W,
                              int H)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= W || j >= H) return;
    size_t idx = (size_t)j * W + i;
    dpx_point* c = &diagram[idx * kCellSize];
    int n = edge_count[idx];
    int x[kMaxPoly], y[kMaxPoly];
    uint8_t tri[3 * kMaxPoly];
    int nt = 0, ok = 0;
    if (n >= 3 && n <= kMaxPoly) {
        for (int t = 0; t < n; t++) {
            x[t] = (int)(c[t].x * 64.0f);
            y[t] = (int)(c[t].y * 64.0f);
        }
        nt = ear_clip(x, y, n, tri, &ok);
    }
    float fi = (float)i, fj = (float)j;
    for (int t = 0; t < kCellSize; t++) {
        dpx_point p;
        if (t < 3 * nt) {
            p.x = (float)x[tri[t]] * (1.0f / 64.0f) + fi;
            p.y = (float)y[tri[t]] * (1.0f / 64.0f) + fj;
        } else {
            p.x = fi;
            p.y = fj;
        }
        c[t] = p;
    }
}

// color_kernel (kernel.cu:67-104) + position_kernel (kernel.cu:109-137)
__global__ void k_export_vbo(ImgView im, const dpx_point* __restrict__ diagram, const int32_t* __restrict__ edge_count,
                             dpx_point* __restrict__ pos, dpx_uchar4* __restrict__ color)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= im.W || j >= im.H) return;
    size_t idx = (size_t)j * im.W + i;
    uint32_t rgba = cell_rgba(load_px(im, i, j), colour_channels(im));
    dpx_uchar4 c4;
    c4.x = (uint8_t)(rgba & 0xFF);
    c4.y = (uint8_t)((rgba >> 8) & 0xFF);
    c4.z = (uint8_t)((rgba >> 16) & 0xFF);
    c4.w = (uint8_t)(rgba >> 24);
    int valid = (edge_count[idx] - 2) * 3;
    for (int t = 0; t < kCellSize; t++) {
        size_t s = idx * kCellSize + t;
        if (color) color[s] = c4;
        if (pos) {
            dpx_point p;
            if (t < valid) p = diagram[s];
            else { p.x = -100.0f; p.y = -100.0f; }
            pos[s] = p;
        }
    }
}

// glDrawArrays(GL_TRIANGLES) in VBO order (simpleVBO.cpp:281) as a gather:
// each output pixel takes the colour of the highest-index cell (the last one
// drawn) with a triangle covering its centre.  Cells reach at most 0.34 past
// their unit square, so the 3x3 cells around the pixel are the candidates.
__global__ void k_raster_tri(ImgView im, const dpx_point* __restrict__ diagram, const int32_t* __restrict__ edge_count,
                             int scale, uint8_t* __restrict__ out, int out_stride)
{
    int X = blockIdx.x * blockDim.x + threadIdx.x;
    int Y = blockIdx.y * blockDim.y + threadIdx.y;
    int OW = im.W * scale, OH = im.H * scale;
    if (X >= OW || Y >= OH) return;
    int cx = X / scale, cy = Y / scale;
    uint32_t color = 0xFF000000u;  // clear colour (0,0,0,1), main.cpp:260
    bool done = false;
    for (int dy = 1; dy >= -1 && !done; dy--)
        for (int dx = 1; dx >= -1 && !done; dx--) {
            int ci = cx + dx, cj = cy + dy;
            if (!in_frame(ci, cj, im.W, im.H)) continue;
            size_t k = (size_t)cj * im.W + ci;
            int nt = edge_count[k] - 2;
            if (nt > kCellSize / 3) nt = kCellSize / 3;
            int sx = (2 * X + 1) * 64 - ci * 128 * scale, sy = (2 * Y + 1) * 64 - cj * 128 * scale;
            const dpx_point* c = &diagram[k * kCellSize];
            float fi = (float)ci, fj = (float)cj;
            for (int t = 0; t < nt; t++) {
                int v[6];
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    dpx_point p = c[3 * t + q];
                    v[2 * q] = (int)((p.x - fi) * 64.0f) * 2 * scale;
                    v[2 * q + 1] = (int)((p.y - fj) * 64.0f) * 2 * scale;
                }
                if (tri_covers(v[0], v[1], v[2], v[3], v[4], v[5], sx, sy)) {
                    color = cell_rgba(load_px(im, ci, cj), colour_channels(im));
                    done = true;
                    break;
                }
            }
        }
    int row = out_topdown(im.flags) ? (OH - 1 - Y) : Y;
    *reinterpret_cast<uint32_t*>(out + (size_t)row * out_stride + (size_t)X * 4) = color;
}

// ---------------------------------------------------------------------------
// Presentation: the GLUT window (display() callbacks.cpp:39-72, renderCuda
// simpleVBO.cpp:185-440).  Window point (X + ox/8, Y + oy/8) maps to image
// point ((8 (X - tx) + ox) * 100) / (8 * sc), sc the scale in hundredths.  In
// units of 1/(512 sc) of a cell that is (8 (X - tx) + ox) * 6400 and a
// 1/64-lattice vertex v is 64 v * 8 sc: both integers, so coverage is exact.
// ---------------------------------------------------------------------------
struct ViewGeom {
    int mode, sc, tx, ty, W, H, aa, topdown;
};

__device__ inline long long floor_div(long long a, long long b)  // b > 0
{
    long long q = a / b;
    return (a % b != 0 && a < 0) ? q - 1 : q;
}

// colour of the sample at window position (8 X + ox, 8 Y + oy) / 8
__device__ inline uint32_t present_sample(const ImgView& im, const dpx_point* __restrict__ diagram,
                                          const int32_t* __restrict__ edge_count, const ViewGeom& v, int X, int Y,
                                          int ox, int oy)
{
    const long long unit = 512LL * v.sc;  // one cell
    const long long px = (8LL * (X - v.tx) + ox) * 6400, py = (8LL * (Y - v.ty) + oy) * 6400;
    const long long ci = floor_div(px, unit), cj = floor_div(py, unit);
    const int cc = colour_channels(im);
    if (v.mode == DPX_VIEW_QUADS) {
        // the two triangles of pixel (ci, cj) cover [ci, ci+1) x [cj, cj+1)
        // under the coverage rule (left and bottom edges in, right and top out)
        if (ci < 0 || cj < 0 || ci >= im.W || cj >= im.H) return 0xFF000000u;
        return cell_rgba(load_px(im, (int)ci, (int)cj), cc);
    }
    for (int dy = 1; dy >= -1; dy--)
        for (int dx = 1; dx >= -1; dx--) {  // cells in reverse draw order: the first hit wins
            long long i = ci + dx, j = cj + dy;
            if (i < 0 || j < 0 || i >= im.W || j >= im.H) continue;
            size_t k = (size_t)j * im.W + (size_t)i;
            int nt = edge_count[k] - 2;
            if (nt > kCellSize / 3) nt = kCellSize / 3;
            const long long sx = px - i * unit, sy = py - j * unit;
            const dpx_point* c = &diagram[k * kCellSize];
            const float fi = (float)i, fj = (float)j;
            for (int t = 0; t < nt; t++) {
                long long w[6];
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    dpx_point p = c[3 * t + q];
                    w[2 * q] = (long long)(int)((p.x - fi) * 64.0f) * 8 * v.sc;
                    w[2 * q + 1] = (long long)(int)((p.y - fj) * 64.0f) * 8 * v.sc;
                }
                if (tri_covers(w[0], w[1], w[2], w[3], w[4], w[5], sx, sy))
                    return cell_rgba(load_px(im, (int)i, (int)j), cc);
            }
        }
    return 0xFF000000u;  // clear colour (main.cpp:260)
}

// The four antialiasing samples of window pixel (X, Y) in one scan: the
// cells around all four, in reverse draw order, each triangle read once and
// tested against every sample not yet covered.  A cell reaches at most a
// quarter cell beyond its square, so the cells outside a sample's own 3x3
// never cover it and the first cover found is the sample's last in draw
// order, as in present_sample.
__device__ inline void present_samples4(const ImgView& im, const dpx_point* __restrict__ diagram,
                                        const int32_t* __restrict__ edge_count, const ViewGeom& v, int X, int Y,
                                        const int* sox, const int* soy, uint32_t* col)
{
    const long long unit = 512LL * v.sc;
    long long px[4], py[4];
    long long i0 = 0, i1 = 0, j0 = 0, j1 = 0;
#pragma unroll
    for (int s = 0; s < 4; s++) {
        px[s] = (8LL * (X - v.tx) + sox[s]) * 6400;
        py[s] = (8LL * (Y - v.ty) + soy[s]) * 6400;
        col[s] = 0xFF000000u;  // clear colour (main.cpp:260)
        const long long ci = floor_div(px[s], unit), cj = floor_div(py[s], unit);
        if (s == 0 || ci < i0) i0 = ci;
        if (s == 0 || ci > i1) i1 = ci;
        if (s == 0 || cj < j0) j0 = cj;
        if (s == 0 || cj > j1) j1 = cj;
    }
    // the window may lie far outside the image: clamp to the frame
    j0 = j0 - 1 < 0 ? 0 : j0 - 1;
    i0 = i0 - 1 < 0 ? 0 : i0 - 1;
    j1 = j1 + 1 > im.H - 1 ? im.H - 1 : j1 + 1;
    i1 = i1 + 1 > im.W - 1 ? im.W - 1 : i1 + 1;
    const int cc = colour_channels(im);
    uint32_t left = 0xFu;
    for (long long j = j1; j >= j0 && left; j--)
        for (long long i = i1; i >= i0 && left; i--) {
            size_t k = (size_t)j * im.W + (size_t)i;
            int nt = edge_count[k] - 2;
            if (nt > kCellSize / 3) nt = kCellSize / 3;
            const dpx_point* c = &diagram[k * kCellSize];
            const float fi = (float)i, fj = (float)j;
            long long sx[4], sy[4];
#pragma unroll
            for (int s = 0; s < 4; s++) { sx[s] = px[s] - i * unit; sy[s] = py[s] - j * unit; }
            uint32_t hit = 0;
            for (int t = 0; t < nt && (left & ~hit); t++) {
                long long w[6];
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    dpx_point p = c[3 * t + q];
                    w[2 * q] = (long long)(int)((p.x - fi) * 64.0f) * 8 * v.sc;
                    w[2 * q + 1] = (long long)(int)((p.y - fj) * 64.0f) * 8 * v.sc;
                }
#pragma unroll
                for (int s = 0; s < 4; s++)
                    if (((left & ~hit) >> s) & 1u)
                        if (tri_covers(w[0], w[1], w[2], w[3], w[4], w[5], sx[s], sy[s])) hit |= 1u << s;
            }
            if (hit) {
                const uint32_t rgba = cell_rgba(load_px(im, (int)i, (int)j), cc);
#pragma unroll
                for (int s = 0; s < 4; s++)
                    if ((hit >> s) & 1u) col[s] = rgba;
                left &= ~hit;
            }
        }
}

__global__ void k_present(ImgView im, const dpx_point* __restrict__ diagram, const int32_t* __restrict__ edge_count,
                          ViewGeom v, uint8_t* __restrict__ out, int out_stride)
{
    int X = blockIdx.x * blockDim.x + threadIdx.x;
    int Y = blockIdx.y * blockDim.y + threadIdx.y;
    if (X >= v.W || Y >= v.H) return;
    uint32_t color;
    if (!v.aa) {
        color = present_sample(im, diagram, edge_count, v, X, Y, 4, 4);
    } else {
        // rotated 4x grid, averaged per channel with halves rounded up
        const int sox[4] = { 3, 7, 1, 5 }, soy[4] = { 1, 3, 5, 7 };
        uint32_t sum[4] = { 0, 0, 0, 0 }, cs[4];
        if (v.mode == DPX_VIEW_QUADS) {
#pragma unroll
            for (int s = 0; s < 4; s++) cs[s] = present_sample(im, diagram, edge_count, v, X, Y, sox[s], soy[s]);
        } else {
            present_samples4(im, diagram, edge_count, v, X, Y, sox, soy, cs);
        }
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const uint32_t c = cs[s];
#pragma unroll
            for (int b = 0; b < 4; b++) sum[b] += (c >> (8 * b)) & 0xFFu;
        }
        color = 0;
#pragma unroll
        for (int b = 0; b < 4; b++) color |= ((sum[b] + 2) >> 2) << (8 * b);
    }
    int row = v.topdown ? (v.H - 1 - Y) : Y;
    *reinterpret_cast<uint32_t*>(out + (size_t)row * out_stride + (size_t)X * 4) = color;
}

// ---------------------------------------------------------------------------
// Components: lock-free union-find over the graph's edges.  Roots are hooked
// under smaller roots only, so each component's root is its smallest index.
// ---------------------------------------------------------------------------
__device__ inline int uf_find(int32_t* parent_, int v)
{
    // other threads hook roots and halve paths concurrently: read through volatile
    volatile int32_t* parent = parent_;
    int p = parent[v];
    while (p != v) {
        int gp = parent[p];
        if (gp != p) parent[v] = gp;  // path halving; any ancestor is a valid parent
        v = p;
        p = gp;
    }
    return v;
}

// Components in two levels: each 32x32 tile of nodes is labelled by one
// block with a union-find in LDS over the edges inside the tile (roots are
// the smallest node of each piece, so the tile's pieces come out with their
// smallest global index); the edges that cross tile borders are then united
// in global memory (k_cc_cross), and the final pass takes every node's root.
constexpr int kCcTile = 32;

__device__ inline int uf_find_lds(int* parent_, int v)
{
    volatile int* parent = parent_;
    int p = parent[v];
    while (p != v) {
        int gp = parent[p];
        if (gp != p) parent[v] = gp;  // path halving; any ancestor is a valid parent
        v = p;
        p = gp;
    }
    return v;
}

__global__ __launch_bounds__(256) void k_cc_tile(const uint8_t* __restrict__ graph, int W, int H, int32_t* parent)
{
    __shared__ int s_par[kCcTile * kCcTile];
    const int tx0 = blockIdx.x * kCcTile, ty0 = blockIdx.y * kCcTile;
    for (int k = threadIdx.x; k < kCcTile * kCcTile; k += 256) s_par[k] = k;
    __syncthreads();
    for (int k = threadIdx.x; k < kCcTile * kCcTile; k += 256) {
        const int lx = k % kCcTile, ly = k / kCcTile, i = tx0 + lx, j = ty0 + ly;
        if (i >= W || j >= H) continue;
        const uint32_t g = graph[(size_t)j * W + i];
        // edges towards smaller indices only (each edge once): bits 3, 5, 6, 7
        const int es[4] = { 3, 5, 6, 7 };
        for (int t = 0; t < 4; t++) {
            const int e = es[t];
            if (!bit(g, e)) continue;
            const int nlx = lx + edx(e), nly = ly + edy(e);
            if ((unsigned)nlx >= (unsigned)kCcTile || (unsigned)nly >= (unsigned)kCcTile) continue;  // k_cc_cross
            if (!in_frame(i + edx(e), j + edy(e), W, H)) continue;
            const int u = nly * kCcTile + nlx;
            for (;;) {
                const int ru = uf_find_lds(s_par, u), rv = uf_find_lds(s_par, k);
                if (ru == rv) break;
                const int hi = ru > rv ? ru : rv, lo = ru > rv ? rv : ru;
                if (atomicCAS(&s_par[hi], hi, lo) == hi) break;
            }
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kCcTile * kCcTile; k += 256) {
        const int lx = k % kCcTile, ly = k / kCcTile, i = tx0 + lx, j = ty0 + ly;
        if (i >= W || j >= H) continue;
        int r = k;
        for (int p = s_par[r]; p != r; p = s_par[r]) r = p;
        parent[(size_t)j * W + i] = (ty0 + r / kCcTile) * W + tx0 + r % kCcTile;
    }
}

// The edges between tiles, united in global memory (parents are tile roots
// after k_cc_tile).
__global__ void k_cc_cross(const uint8_t* __restrict__ graph, int W, int H, int32_t* parent)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= W || j >= H) return;
    const int lx = i % kCcTile, ly = j % kCcTile;
    if (lx != 0 && ly != 0 && lx != kCcTile - 1) return;  // no backward edge leaves the tile
    const int v = j * W + i;
    const uint32_t g = graph[v];
    const int es[4] = { 3, 5, 6, 7 };
    for (int t = 0; t < 4; t++) {
        const int e = es[t];
        if (!bit(g, e)) continue;
        const int ni = i + edx(e), nj = j + edy(e);
        if (!in_frame(ni, nj, W, H)) continue;
        if (ni / kCcTile == i / kCcTile && nj / kCcTile == j / kCcTile) continue;  // inside: k_cc_tile
        const int u = nj * W + ni;
        for (;;) {
            int ru = uf_find(parent, u), rv = uf_find(parent, v);
            if (ru == rv) break;
            int hi = ru > rv ? ru : rv, lo = ru > rv ? rv : ru;
            if (atomicCAS(&parent[hi], hi, lo) == hi) break;
        }
    }
}

// The root of v without shortening paths: the final pass writes each
// label over its own parent entry (the arrays alias), so it must write no
// other entry -- a halving write could replace a label another thread has
// already set to its root with a non-root ancestor.  Entries it reads are
// then an old parent or the root, both ancestors.
__device__ inline int uf_root(const int32_t* parent_, int v)
{
    const volatile int32_t* parent = parent_;
    int p = parent[v];
    while (p != v) {
        v = p;
        p = parent[v];
    }
    return v;
}

__global__ void k_cc_final(int32_t* parent, int32_t* labels, int N)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < N) labels[k] = uf_root(parent, k);
}

// Border chains (extractBorderPoints cc_functions.cu:348-503) of one frame.
// The scan: a walk starts at each node, in row-major order, that is neither
// internal (90) nor without links nor listed; it leaves by the node's first
// link clockwise from up-left and at each node turns clockwise from the link
// it came in by.  It is a chain when it comes back to its start through the
// link just counter-clockwise of the first one; it is abandoned on an
// internal, listed or discarded node, discarding the node it stepped from.
// Chain nodes are listed; an abandoned walk leaves its start unlisted (the
// behaviour that reproduces the chains recorded in alex_png.txt:1-21).
//
// Which walks become chains depends on every earlier walk, but where a walk
// goes does not: the turn rule gives each directed link (node s, link e) one
// successor (the neighbour t, its first link clockwise after 7 - e), and on
// a symmetric graph that is a permutation of the directed links whose
// cycles are the faces of the graph's plane embedding.  A walk from s runs
// along the cycle of s's first link and closes where that cycle comes back
// to it; a link into an internal node has no successor (the walk stops
// there), which cuts some cycles into paths.  So the cycles and paths are
// laid out first, all links in parallel: list ranking by pointer jumping
// (Wyllie) finds each link's distance to the end of its path and, on a
// cycle, the smallest link of the cycle, where the cycle is then cut; each
// path is written as an array of its links' nodes in walk order.  The scan
// itself (one wave) then needs, per walk, only the marks of the nodes of a
// contiguous range of that array, 64 at a time.
//
// Directed link id = node * 8 + link.  State per link (uint4): x = next
// link 2^k steps on (kEnd past the end, kNoLink for slots that are not
// links of a walkable node), y = steps to it, z = the furthest link reached
// (the path's last link once x = kEnd), w = the smallest link seen (cycle
// search) or flags (bit 0: on a cycle, bit 1: first link of its path).
constexpr int kEnd = -1, kNoLink = -2;

inline size_t round256(size_t v) { return (v + 255) / 256 * 256; }

// successor of directed link (s, e) of a walkable node, or kEnd when the
// neighbour is internal
__device__ inline int bc_succ(const uint8_t* g, int W, int s, int e)
{
    const int t = s + edy(e) * W + edx(e);
    const uint32_t vt = g[t];
    return vt == 90u ? kEnd : t * 8 + cw_turn(vt, 7 - e, +1);
}

__global__ void k_bc_clear(int* ctr)
{
    if (threadIdx.x < 4) ctr[threadIdx.x] = 0;
}

// ctr[0]: status (DPX_EINVAL: a link leaves the frame or has no reverse
// link), ctr[1]: entries of the path arrays handed out
__global__ __launch_bounds__(256) void k_bc_init(const uint8_t* __restrict__ g, int W, int H, uint4* __restrict__ S,
                                                 int* ctr)
{
    const uint32_t id = blockIdx.x * 256u + threadIdx.x, E = 8u * (uint32_t)W * (uint32_t)H;
    if (id >= E) return;
    const int s = (int)(id >> 3), e = (int)(id & 7u);
    const uint32_t v = g[s];
    if (!((v >> e) & 1u) || v == 90u) {
        S[id] = make_uint4((uint32_t)kNoLink, 0u, id, 0u);
        return;
    }
    const int ni = s % W + edx(e), nj = s / W + edy(e);
    if (!in_frame(ni, nj, W, H) || !((g[nj * W + ni] >> (7 - e)) & 1u)) {
        ctr[0] = DPX_EINVAL;
        S[id] = make_uint4((uint32_t)kNoLink, 0u, id, 0u);
        return;
    }
    const int nx = bc_succ(g, W, s, e);
    S[id] = make_uint4((uint32_t)nx, nx >= 0 ? 1u : 0u, nx >= 0 ? (uint32_t)nx : id, id);
}

// pointer jumping, in -> out; MIN: also the smallest link seen.  A link's
// state in `in` covers 2^k steps; composed with the state it points to
// HOPS times it covers (HOPS + 1) 2^k: HOPS = 1 is one doubling round,
// HOPS = 3 two rounds and HOPS = 7 three in one launch (dependent reads of
// `in` in place of launches that each read and write every state).
template <bool MIN, int HOPS>
__global__ __launch_bounds__(256) void k_bc_jump(const uint4* __restrict__ in, uint4* __restrict__ out, uint32_t E)
{
    const uint32_t id = blockIdx.x * 256u + threadIdx.x;
    if (id >= E) return;
    uint4 st = in[id];
#pragma unroll
    for (int h = 0; h < HOPS; h++) {
        if ((int)st.x >= 0) {
            const uint4 x = in[st.x];
            st = make_uint4(x.x, st.y + x.y, x.z, MIN ? (x.w < st.w ? x.w : st.w) : st.w);
        }
    }
    out[id] = st;
}

// after the cycle search: links still jumping lie on cycles; each cycle is
// cut before its smallest link, which becomes the first of its path.  A
// link on a path is the first of its path when its predecessor (the link
// that turns into it: from the neighbour across s's link just
// counter-clockwise of e) is no walkable node's link.
__global__ __launch_bounds__(256) void k_bc_cut(const uint8_t* __restrict__ g, int W, int H,
                                                const uint4* __restrict__ F, uint4* __restrict__ G, uint32_t E)
{
    const uint32_t id = blockIdx.x * 256u + threadIdx.x;
    if (id >= E) return;
    const uint4 st = F[id];
    if ((int)st.x == kNoLink) {
        G[id] = st;
        return;
    }
    const int s = (int)(id >> 3), e = (int)(id & 7u);
    if ((int)st.x >= 0) {
        const int nx = bc_succ(g, W, s, e);
        const uint32_t fl = 1u | (id == st.w ? 2u : 0u);
        G[id] = (uint32_t)nx == st.w ? make_uint4((uint32_t)kEnd, 0u, id, fl) : make_uint4((uint32_t)nx, 1u, (uint32_t)nx, fl);
        return;
    }
    const int a = cw_turn(g[s], e, -1);
    const int pi = s % W + edx(a), pj = s / W + edy(a);
    const bool first = !in_frame(pi, pj, W, H) || g[pj * W + pi] == 90u;
    G[id] = make_uint4((uint32_t)kEnd, st.y, st.z, first ? 2u : 0u);
}

// the first link of each path takes the path's range of the node array:
// info[last link] = (offset, length, on a cycle)
constexpr int kHeadsBlock = 1024;

__global__ __launch_bounds__(kHeadsBlock) void k_bc_heads(const uint4* __restrict__ F, uint4* __restrict__ info,
                                                          int* ctr, uint32_t E)
{
    const uint32_t id = blockIdx.x * (uint32_t)kHeadsBlock + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool ok = id < E && ctr[0] == 0;
    const uint4 st = ok ? F[id] : make_uint4((uint32_t)kNoLink, 0u, 0u, 0u);
    const bool link = (int)st.x != kNoLink;
    if (link && (int)st.x != kEnd) ctr[0] = DPX_EINVAL;  // not at the end: the links are no permutation
    const bool head = link && (int)st.x == kEnd && (st.w & 2u);
    // one atomic per block (one per wave made the counter the kernel's
    // bottleneck): the block's paths take consecutive ranges
    __shared__ uint64_t s_tot[kHeadsBlock / 64], s_base;
    const int wave = threadIdx.x >> 6;
    const uint64_t len = head ? (uint64_t)st.y + 1u : 0u;
    uint64_t incl = len;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t t = __shfl_up(incl, d);
        if (lane >= d) incl += t;
    }
    if (lane == 63) s_tot[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t total = 0;
        for (int w = 0; w < kHeadsBlock / 64; w++) total += s_tot[w];
        s_base = total == 0 ? 0u
                 : total > E ? (uint64_t)E + 1u
                             : (uint64_t)atomicAdd(reinterpret_cast<unsigned*>(&ctr[1]), (unsigned)total);
    }
    __syncthreads();
    if (!head) return;
    uint64_t base = s_base;
    for (int w = 0; w < wave; w++) base += s_tot[w];
    const uint64_t off = base + incl - len;
    if (off + len > E) {
        ctr[0] = DPX_EINVAL;
        return;
    }
    info[st.z] = make_uint4((uint32_t)off, (uint32_t)len, st.w & 1u, 0u);
}

// each link writes its node at its place in its path's array; the first
// link of a walkable node gives that node's walk (offset, length, place of
// the start, on a cycle)
__global__ __launch_bounds__(256) void k_bc_scatter(const uint8_t* __restrict__ g, const uint4* __restrict__ F,
                                                    const uint4* __restrict__ info, int32_t* __restrict__ A,
                                                    uint4* __restrict__ rec, const int* ctr, uint32_t E)
{
    const uint32_t id = blockIdx.x * 256u + threadIdx.x;
    if (id >= E || ctr[0] != 0) return;
    const uint4 st = F[id];
    if ((int)st.x == kNoLink) return;
    const uint4 inf = info[st.z];
    if (st.y >= inf.y || (uint64_t)inf.x + inf.y > E) return;  // (inconsistent only for graphs refused above)
    const uint32_t pos = inf.y - 1u - st.y;
    A[inf.x + pos] = (int32_t)(id >> 3);
    if ((int)(id & 7u) == cw_first_link(g[id >> 3])) rec[id >> 3] = make_uint4(inf.x, inf.y, pos, inf.z);
}

// The scan, one wave (wave 0 of the block).  Marks are two bitmaps of
// (N + 31) / 32 words each, in LDS when they fit, else in bits_g: listed
// (nodes of chains: a candidate that is listed starts no walk) and marked
// (listed or discarded: a walk stops at such a node), so that a walk's test
// reads one bit per node.  Which walks run depends on the marks, but what a walk
// reads does not, so the loads go out in groups and the wave waits once per
// group, not once per load:
// - the candidate flags and walk records of kBcBlk batches of 64 nodes are
//   loaded into registers while the block before is scanned, then staged
//   in LDS;
// - the first kBcChunk nodes of the next kBcAhead walks of a batch are
//   loaded together into registers (a walk whose start a chain lists
//   meanwhile is skipped);
// - the rest of a long walk, and a long chain's nodes, kBcChunk2 at a time.
// The loads of records and of a walk's nodes are unconditional (an index
// past an end is clamped, its value dropped): a load that the wave may skip
// leaves the compiler unsure how many are in flight, and it then waits for
// all of them at the next use.
// The other three waves read the path arrays, the graph and the records
// once, so that the scan's loads hit the L2 of its XCD.  The sizes are the
// fastest of a sweep on 256x224 and 640x480 frames (groups of 1-8 walks,
// first chunks of 64-256 nodes, later ones of 128-1024, blocks of 4-16
// batches): 85% of walks end within 64 nodes, and each load the wave
// issues for a slot it may not need costs more than the wait it saves.
constexpr int kBcU = 1, kBcChunk = 64 * kBcU, kBcAhead = 2, kBcBlk = 8, kBcU2 = 4, kBcChunk2 = 64 * kBcU2;

__device__ inline int lane_of(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

// place of the node before the first marked one in a chunk of U slots of 64
// (-1: the node before the chunk); -2 when no node of the chunk is marked
template <int U>
__device__ inline int bc_before_mark(const uint64_t (&bal)[U])
{
#pragma unroll
    for (int u = 0; u < U; u++)
        if (bal[u]) return 64 * u + __builtin_ctzll(bal[u]) - 1;
    return -2;
}

// the node at place p (uniform) of a chunk held U per lane
template <int U>
__device__ inline int bc_node_at(const int (&nd)[U], int p)
{
    int x = 0;
#pragma unroll
    for (int u = 0; u < U; u++)
        if ((p >> 6) == u) x = lane_of(nd[u], p & 63);
    return x;
}

template <bool LDS>
__global__ __launch_bounds__(256) void k_bc_scan(const uint8_t* __restrict__ g, int N, const uint4* __restrict__ rec,
                                                 const int32_t* __restrict__ A, uint32_t* bits_g, int32_t* nodes,
                                                 int32_t* sizes, int max_nodes, int max_chains, const int* ctr,
                                                 int32_t* counts)
{
    extern __shared__ uint32_t s_bits[];
    __shared__ uint4 s_rec[kBcBlk][64];
    __shared__ uint64_t s_cand[kBcBlk];
    uint32_t* bits = LDS ? s_bits : bits_g;
    const int NW = (N + 31) >> 5, tid = threadIdx.x, lane = tid & 63;
    uint32_t* listed = bits;
    uint32_t* marked = bits + NW;
    for (int k = tid; k < 2 * NW; k += 256) bits[k] = 0u;
    if constexpr (!LDS) __threadfence();
    __syncthreads();
    if (tid >= 64) {  // the helper waves: read the path arrays, the graph and the records, then leave
        const int total = ctr[0] == 0 ? ctr[1] : 0;
        uint32_t acc = 0u;
        for (int k = (tid - 64) * 16; k < total; k += 192 * 16) acc ^= (uint32_t)A[k];  // one read per 64 bytes
        for (int k = (tid - 64) * 64; k < N; k += 192 * 64) acc ^= g[k];
        for (int k = (tid - 64) * 4; k < N; k += 192 * 4) acc ^= rec[k].x;
        asm volatile("" ::"v"(acc));  // (keeps the reads)
        return;
    }
    auto get = [&](const uint32_t* b, int n) -> uint32_t {
        uint32_t w;
        if constexpr (LDS)
            w = b[n >> 5];
        else
            w = __hip_atomic_load(b + (n >> 5), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return (w >> (n & 31)) & 1u;
    };
    auto put = [&](uint32_t* b, int n) { atomicOr(b + (n >> 5), 1u << (n & 31)); };
    // the marks written so far are seen by the wave's next reads (one wave:
    // its LDS operations complete in order)
    auto settle = [&]() {
        if constexpr (!LDS) __threadfence();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
    };
    // node k (1-based) of a walk of n nodes from place pos of a path of L at
    // off; past the walk (k > n) the path arrays' first entry, unused
    auto walk_node = [&](int off, int L, int pos, int k, int n) -> int {
        int q = pos + k;
        if (q >= L) q -= L;
        return A[k <= n ? off + q : 0];
    };
    int status = ctr[0], nc = 0, nn = 0;
    // block k of kBcBlk batches: graph bytes and records into registers
    // (nodes past the frame read its last node), then flags and records
    // into LDS (the arrays stay in registers: both are inlined)
    const int nblk = (N + 64 * kBcBlk - 1) / (64 * kBcBlk);
    uint32_t vv[kBcBlk], rx[kBcBlk], ry[kBcBlk], rz[kBcBlk], rw[kBcBlk];  // (no uint4 array: it would go to scratch)
    auto issue = [&](int k) __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < kBcBlk; j++) {
            const int n = (k * kBcBlk + j) * 64 + lane, c = n < N ? n : N - 1;
            const uint4 t = rec[c];
            vv[j] = g[c];
            rx[j] = t.x;
            ry[j] = t.y;
            rz[j] = t.z;
            rw[j] = t.w;
        }
    };
    auto stage = [&](int k) __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < kBcBlk; j++) {
            const int n = (k * kBcBlk + j) * 64 + lane;
            const uint64_t cm = __ballot(n < N && vv[j] != 90u && vv[j] != 0u);
            if (lane == 0) s_cand[j] = cm;
            s_rec[j][lane] = make_uint4(rx[j], ry[j], rz[j], rw[j]);
        }
    };
    if (status == 0) issue(0);
    for (int k = 0; k < nblk && status == 0; k++) {
        stage(k);  // (the walks of block k - 1 are done with the LDS copy)
        if (k + 1 < nblk) issue(k + 1);
        for (int j = 0; j < kBcBlk && status == 0; j++) {
            const int base = (k * kBcBlk + j) * 64;
            if (base >= N) break;
            const int me = base + lane;
            const bool cand = (s_cand[j] >> lane) & 1u;
            uint64_t m = __ballot(cand && !get(listed, me));
            while (m && status == 0) {
                // the next walks of the batch: their first chunks, all loads at once
                int bs[kBcAhead], pv[kBcAhead][kBcU], ro[kBcAhead], rl[kBcAhead], rp[kBcAhead], rn[kBcAhead];
                bool rc[kBcAhead];
#pragma unroll
                for (int a = 0; a < kBcAhead; a++) {
                    bs[a] = m ? __builtin_ctzll(m) : -1;
                    m &= m ? m - 1 : 0;
                }
#pragma unroll
                for (int a = 0; a < kBcAhead; a++) {  // their records: path offset, length, place, on a cycle
                    const uint4 rb = s_rec[j][bs[a] < 0 ? 0 : bs[a]];
                    ro[a] = (int)rb.x;
                    rl[a] = (int)rb.y;
                    rp[a] = (int)rb.z;
                    rc[a] = rb.w != 0u;
                    rn[a] = bs[a] < 0 ? 0 : rc[a] ? rl[a] - 1 : rl[a] - 1 - rp[a];  // walk nodes after the start
                }
#pragma unroll
                for (int a = 0; a < kBcAhead; a++)
#pragma unroll
                    for (int u = 0; u < kBcU; u++)
                        pv[a][u] = walk_node(ro[a], rl[a], rp[a], 1 + lane + 64 * u, rn[a]);
                uint64_t alive = ~0ull;  // candidates not listed by a chain of this batch
#pragma unroll
                for (int a = 0; a < kBcAhead; a++) {  // (unrolled: pv[a] stays in registers)
                    if (status != 0) break;
                    const int b = bs[a];
                    if (b < 0) break;
                    if (!((alive >> b) & 1u)) continue;
                    const int s = base + b;
                    const int off = ro[a], L = rl[a], pos = rp[a], n = rn[a];
                    const bool cyc = rc[a];
                    int prev = s, nd[kBcU];
                    bool hit;
                    {  // the first chunk, loaded above
                        uint64_t bal[kBcU];
#pragma unroll
                        for (int u = 0; u < kBcU; u++) nd[u] = pv[a][u];
#pragma unroll
                        for (int u = 0; u < kBcU; u++) {
                            const int kk = 1 + lane + 64 * u;
                            bal[u] = 64 * u < n ? __ballot(kk <= n && get(marked, nd[u])) : 0ull;
                        }
                        const int f = bc_before_mark(bal);
                        hit = f != -2;
                        const int last = hit ? f : (n < kBcChunk ? n : kBcChunk) - 1;
                        if (last >= 0) prev = bc_node_at(nd, last);
                    }
                    for (int c = kBcChunk; !hit && c < n; c += kBcChunk2) {  // the rest of a long walk
                        int nd2[kBcU2];
                        uint64_t bal[kBcU2];
#pragma unroll
                        for (int u = 0; u < kBcU2; u++) nd2[u] = walk_node(off, L, pos, c + 1 + lane + 64 * u, n);
#pragma unroll
                        for (int u = 0; u < kBcU2; u++) {
                            const int kk = c + 1 + lane + 64 * u;
                            bal[u] = c + 64 * u < n ? __ballot(kk <= n && get(marked, nd2[u])) : 0ull;
                        }
                        const int f = bc_before_mark(bal);
                        hit = f != -2;
                        const int last = hit ? f : (n - c < kBcChunk2 ? n - c : kBcChunk2) - 1;
                        if (last >= 0) prev = bc_node_at(nd2, last);
                    }
                    if (hit || !cyc) {  // abandoned (a path ends at an internal node)
                        if (lane == 0) put(marked, prev);
                        settle();
                        continue;
                    }
                    if (nc >= max_chains || (int64_t)nn + L > max_nodes) {
                        status = DPX_ENOMEM;
                        break;
                    }
                    if (n < kBcChunk) {  // the whole walk is in nd
                        if (lane == 0) {
                            nodes[nn] = s;
                            put(listed, s);
                            put(marked, s);
                        }
#pragma unroll
                        for (int u = 0; u < kBcU; u++) {
                            if (64 * u >= n) break;  // (uniform)
                            const int kk = 1 + lane + 64 * u;
                            if (kk <= n) {
                                nodes[nn + kk] = nd[u];
                                put(listed, nd[u]);
                                put(marked, nd[u]);
                            }
                        }
                    } else {
                        for (int c = 0; c < L; c += kBcChunk2) {
                            int v[kBcU2];
#pragma unroll
                            for (int u = 0; u < kBcU2; u++) {
                                const int kk = c + lane + 64 * u;
                                v[u] = c + 64 * u >= L ? 0 : kk == 0 ? s : walk_node(off, L, pos, kk, n);  // (uniform)
                            }
#pragma unroll
                            for (int u = 0; u < kBcU2; u++) {
                                const int kk = c + lane + 64 * u;
                                if (kk < L) {
                                    nodes[nn + kk] = v[u];
                                    put(listed, v[u]);
                                    put(marked, v[u]);
                                }
                            }
                        }
                    }
                    if (lane == 0) sizes[nc] = L;
                    nc++;
                    nn += L;
                    settle();
                    const uint64_t still = __ballot(cand && !get(listed, me));
                    alive &= still;
                    m &= still;
                }
            }
        }
    }
    if (lane == 0) {
        counts[0] = nc;
        counts[1] = nn;
        counts[2] = status;
    }
}

int check_img(const void* p, int W, int H, int stride, int ch, uint32_t flags)
{
    DPX_CHECK(p != nullptr, DPX_EINVAL, "null image pointer");
    DPX_CHECK(W > 0 && H > 0 && W <= 32768 && H <= 32768, DPX_EINVAL, "bad frame size %dx%d", W, H);
    DPX_CHECK((flags & ~kKnownFlags) == 0, DPX_EINVAL, "unknown flags 0x%x", flags & ~kKnownFlags);
    DPX_CHECK(!(flags & DPX_INDEXED), DPX_EINVAL, "indexed frames: batched and serving paths only");
    DPX_CHECK(ch == 3 || ch == 4, DPX_EINVAL, "channels must be 3 or 4, got %d", ch);
    DPX_CHECK(stride >= W * ch, DPX_EINVAL, "stride %d < width*channels %d", stride, W * ch);
    return DPX_OK;
}

}  // namespace

// -------------------------------------------------------------------------
// stage entry points
// -------------------------------------------------------------------------
extern "C" int dpx_stage_similarity_graph(const uint8_t* d_img, int W, int H, int stride, int ch, uint32_t flags,
                                          uint8_t* d_graph, void* stream)
{
    if (int rc = check_img(d_img, W, H, stride, ch, flags)) return rc;
    DPX_CHECK(d_graph, DPX_EINVAL, "null graph pointer");
    ImgView im{ d_img, W, H, stride, ch, flags };
    hipLaunchKernelGGL(k_graph, grid2(W, H), dim3(kB, kB), 0, as_stream(stream), im, d_graph);
    DPX_HIP(hipGetLastError());
    return DPX_OK;
}

extern "C" int dpx_stage_trivial_cross(uint8_t* d_graph, int W, int H, void* stream)
{
    DPX_CHECK(d_graph && W > 0 && H > 0, DPX_EINVAL, "bad arguments");
    hipLaunchKernelGGL(k_trivial, grid2(W, H), dim3(kB, kB), 0, as_stream(stream), d_graph, W, H);
    DPX_HIP(hipGetLastError());
    return DPX_OK;
}

extern "C" int dpx_stage_heuristic_cross(uint8_t* d_graph, uint8_t* d_graph_aux, int W, int H, void* stream)
{
    DPX_CHECK(d_graph && d_graph_aux && W > 0 && H > 0, DPX_EINVAL, "bad arguments");
    DPX_HIP(hipMemcpyAsync(d_graph_aux, d_graph, (size_t)W * H, hipMemcpyDeviceToDevice, as_stream(stream)));
    hipLaunchKernelGGL(k_heuristic, grid2(W, H), dim3(kB, kB), 0, as_stream(stream), d_graph, d_graph_aux, W, H);
    DPX_HIP(hipGetLastError());
    return DPX_OK;
}

extern "C" int dpx_stage_cells(const uint8_t* d_graph, int W, int H, dpx_point* d_diagram, int32_t* d_edge_count,
                               void* stream)
{
    DPX_CHECK(d_graph && d_diagram && d_edge_count && W > 0 && H > 0, DPX_EINVAL, "bad arguments");
    hipLaunchKernelGGL(k_cells, grid2(W, H), dim3(kB, kB), 0, as_stream(stream), d_graph, W, H, d_diagram,
                       d_edge_count);
    DPX_HIP(hipGetLastError());
    return DPX_OK;
}

extern "C" int dpx_stage_subdivide(const uint8_t* d_img, int W, int H, int stride, int ch, uint32_t flags,
                                   const uint8_t* d_graph, dpx_point* d_diagram, dpx_point* d_diagram_aux,
                                   int32_t* d_edge_count, int32_t* d_edge_count_aux, void* stream)
{
    if (int rc = check_img(d_img, W, H, stride, ch, flags)) return rc;
    DPX_CHECK(d_graph && d_diagram && d_diagram_aux && d_edge_count && d_edge_count_aux, DPX_EINVAL,
              "bad arguments");
    size_t N = (size_t)W * H;
    hipStream_t s = as_stream(stream);
    DPX_HIP(hipMemcpyAsync(d_diagram_aux, d_diagram, N * kCellSize * sizeof(dpx_point), hipMemcpyDeviceToDevice, s));
    DPX_HIP(hipMemcpyAsync(d_edge_count_aux, d_edge_count, N * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
    ImgView im{ d_img, W, H, stride, ch, flags };
    hipLaunchKernelGGL(k_subdivide, grid2(W, H), dim3(kB, kB), 0, s, im, d_graph, d_diagram, d_diagram_aux,
                       d_edge_count, d_edge_count_aux);
    DPX_HIP(hipGetLastError());
    return DPX_OK;
}

extern "C" int dpx_stage_triangulate(dpx_point* d_diagram, const int32_t* d_edge_count, int W, int H, void* stream)
{
    DPX_CHECK(d_diagram && d_edge_count && W > 0 && H > 0, DPX_EINVAL, "bad arguments");
    hipLaunchKernelGGL(k_triangulate, grid2(W, H), dim3(kB, kB), 0, as_stream(stream), d_diagram, d_edge_count, W, H);
    DPX_HIP(hipGetLastError());
    return DPX_OK;
}

extern "C" int dpx_stage_export_vbo(const uint8_t* d_img, int W, int H, int stride, int ch, uint32_t flags,
                                    const dpx_point* d_diagram, const int32_t* d_edge_count, dpx_point* d_pos,
                                    dpx_uchar4* d_color, void* stream)
{
    if (int rc = check_img(d_img, W, H, stride, ch, flags)) return rc;
    DPX_CHECK(d_diagram && d_edge_count && (d_pos || d_color), DPX_EINVAL, "bad arguments");
    ImgView im{ d_img, W, H, stride, ch, flags };
    hipLaunchKernelGGL(k_export_vbo, grid2(W, H), dim3(kB, kB), 0, as_stream(stream), im, d_diagram, d_edge_count,
                       d_pos, d_color);
    DPX_HIP(hipGetLastError());
    return DPX_OK;
}

extern "C" int dpx_stage_rasterize(const uint8_t* d_img, int W, int H, int stride, int ch, uint32_t flags,
                                   const dpx_point* d_diagram, const int32_t* d_edge_count, int scale,
                                   uint8_t* d_out, int out_stride, void* stream)
{
    if (int rc = check_img(d_img, W, H, stride, ch, flags)) return rc;
    DPX_CHECK(d_diagram && d_edge_count && d_out, DPX_EINVAL, "bad arguments");
    DPX_CHECK(scale >= 1 && scale <= 32, DPX_EINVAL, "scale %d outside 1..32", scale);
    DPX_CHECK((long long)W * scale <= 65536 && (long long)H * scale <= 65536, DPX_EINVAL, "output too large");
    DPX_CHECK(out_stride >= 4 * W * scale && out_stride % 4 == 0, DPX_EINVAL, "bad out_stride %d", out_stride);
    ImgView im{ d_img, W, H, stride, ch, flags };
    hipLaunchKernelGGL(k_raster_tri, grid2(W * scale, H * scale), dim3(kB, kB), 0, as_stream(stream), im, d_diagram,
                       d_edge_count, scale, d_out, out_stride);
    DPX_HIP(hipGetLastError());
    return DPX_OK;
}

extern "C" int dpx_stage_present(const uint8_t* d_img, int W, int H, int stride, int ch, uint32_t flags,
                                 const dpx_point* d_diagram, const int32_t* d_edge_count, const dpx_view* view,
                                 uint8_t* d_out, int out_stride, void* stream)
{
    if (int rc = check_img(d_img, W, H, stride, ch, flags)) return rc;
    DPX_CHECK(view && d_out, DPX_EINVAL, "bad arguments");
    DPX_CHECK(view->mode == DPX_VIEW_CELLS || view->mode == DPX_VIEW_QUADS, DPX_EINVAL, "unknown view mode %d",
              view->mode);
    DPX_CHECK(view->mode == DPX_VIEW_QUADS || (d_diagram && d_edge_count), DPX_EINVAL,
              "the cells view needs the triangulated diagram");
    DPX_CHECK(view->scale_x100 >= 25 && view->scale_x100 <= 3200, DPX_EINVAL, "scale_x100 %d outside 25..3200",
              view->scale_x100);
    DPX_CHECK(view->translate_x >= -65536 && view->translate_x <= 65536 && view->translate_y >= -65536 &&
                  view->translate_y <= 65536,
              DPX_EINVAL, "translation outside -65536..65536");
    DPX_CHECK(view->width >= 1 && view->width <= 16384 && view->height >= 1 && view->height <= 16384, DPX_EINVAL,
              "window %dx%d outside 1..16384", view->width, view->height);
    DPX_CHECK((view->flags & ~(DPX_VIEW_AA | DPX_VIEW_TOPDOWN)) == 0 && view->reserved == 0, DPX_EINVAL,
              "unknown view flags");
    DPX_CHECK(out_stride >= 4 * view->width && out_stride % 4 == 0, DPX_EINVAL, "bad out_stride %d", out_stride);
    ImgView im{ d_img, W, H, stride, ch, flags };
    ViewGeom v{ view->mode, view->scale_x100, view->translate_x, view->translate_y, view->width, view->height,
                (view->flags & DPX_VIEW_AA) ? 1 : 0, (view->flags & DPX_VIEW_TOPDOWN) ? 1 : 0 };
    hipLaunchKernelGGL(k_present, grid2(v.W, v.H), dim3(kB, kB), 0, as_stream(stream), im, d_diagram, d_edge_count, v,
                       d_out, out_stride);
    DPX_HIP(hipGetLastError());
    return DPX_OK;
}

extern "C" int dpx_stage_components(const uint8_t* d_graph, int W, int H, int32_t* d_labels, void* stream)
{
    DPX_CHECK(d_graph && d_labels && W > 0 && H > 0, DPX_EINVAL, "bad arguments");
    int N = W * H;
    hipStream_t s = as_stream(stream);
    // d_labels doubles as the parent array, then receives the roots
    const dim3 tiles((unsigned)((W + kCcTile - 1) / kCcTile), (unsigned)((H + kCcTile - 1) / kCcTile));
    hipLaunchKernelGGL(k_cc_tile, tiles, dim3(256), 0, s, d_graph, W, H, d_labels);
    hipLaunchKernelGGL(k_cc_cross, grid2(W, H), dim3(kB, kB), 0, s, d_graph, W, H, d_labels);
    hipLaunchKernelGGL(k_cc_final, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, d_labels, d_labels, N);
    DPX_HIP(hipGetLastError());
    return DPX_OK;
}

extern "C" int dpx_stage_border_chains(const uint8_t* d_graph, int W, int H, uint8_t* d_scratch, int32_t* d_nodes,
                                       int max_nodes, int32_t* d_sizes, int max_chains, int32_t* d_counts,
                                       void* stream)
{
    DPX_CHECK(d_graph && d_scratch && d_nodes && d_sizes && d_counts && W > 0 && H > 0, DPX_EINVAL,
              "bad arguments");
    DPX_CHECK((int64_t)W * H <= (1 << 28), DPX_EINVAL, "frame too large (%dx%d)", W, H);
    DPX_CHECK(max_nodes >= 0 && max_chains >= 0, DPX_EINVAL, "negative capacity");
    const hipStream_t s = as_stream(stream);
    const uint32_t N = (uint32_t)W * (uint32_t)H, E = 8u * N;
    // the scan's marks in LDS when they fit a block, else in d_scratch
    int dev = 0, lds_max = 0;
    DPX_HIP(hipGetDevice(&dev));
    DPX_HIP(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    const size_t lds = 8 * (size_t)((N + 31) / 32);
    hipFuncAttributes fa{};
    DPX_HIP(hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(k_bc_scan<true>)));
    bool use_lds = lds + fa.sharedSizeBytes <= (size_t)lds_max &&
                   hipFuncSetAttribute(reinterpret_cast<const void*>(k_bc_scan<true>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) == hipSuccess;
    if (!use_lds) {
        (void)hipGetLastError();
        DPX_CHECK(reinterpret_cast<uintptr_t>(d_scratch) % 4 == 0, DPX_EINVAL,
                  "d_scratch must be 4-byte aligned for a frame of %dx%d", W, H);
    }
    int rounds = 0;
    while ((1ull << rounds) < E) rounds++;
    // working memory, stream-ordered: two link-state arrays, the path
    // arrays, the per-node walk records and the counters
    const size_t szS = round256((size_t)E * sizeof(uint4)), szA = round256((size_t)E * 4),
                 szR = round256((size_t)N * sizeof(uint4));
    uint8_t* work = nullptr;
    DPX_HIP(hipMallocAsync(reinterpret_cast<void**>(&work), 2 * szS + szA + szR + 256, s));
    struct FreeAfter {  // freed in stream order on every return below
        void* p;
        hipStream_t s;
        ~FreeAfter() { (void)hipFreeAsync(p, s); }
    } free_after{ work, s };
    uint4* SA = reinterpret_cast<uint4*>(work);
    uint4* SB = reinterpret_cast<uint4*>(work + szS);
    int32_t* A = reinterpret_cast<int32_t*>(work + 2 * szS);
    uint4* rec = reinterpret_cast<uint4*>(work + 2 * szS + szA);
    int* ctr = reinterpret_cast<int*>(work + 2 * szS + szA + szR);
    const dim3 grid((E + 255) / 256);
    hipLaunchKernelGGL(k_bc_clear, dim3(1), dim3(64), 0, s, ctr);
    hipLaunchKernelGGL(k_bc_init, grid, dim3(256), 0, s, d_graph, W, H, SA, ctr);
    uint4 *cur = SA, *oth = SB;
    for (int r = 0; r < rounds;) {  // (three rounds a launch while three are left)
        const int m = rounds - r >= 3 ? 3 : rounds - r;
        if (m == 3)
            hipLaunchKernelGGL((k_bc_jump<true, 7>), grid, dim3(256), 0, s, cur, oth, E);
        else if (m == 2)
            hipLaunchKernelGGL((k_bc_jump<true, 3>), grid, dim3(256), 0, s, cur, oth, E);
        else
            hipLaunchKernelGGL((k_bc_jump<true, 1>), grid, dim3(256), 0, s, cur, oth, E);
        r += m;
        std::swap(cur, oth);
    }
    hipLaunchKernelGGL(k_bc_cut, grid, dim3(256), 0, s, d_graph, W, H, cur, oth, E);
    std::swap(cur, oth);
    for (int r = 0; r < rounds;) {
        const int m = rounds - r >= 3 ? 3 : rounds - r;
        if (m == 3)
            hipLaunchKernelGGL((k_bc_jump<false, 7>), grid, dim3(256), 0, s, cur, oth, E);
        else if (m == 2)
            hipLaunchKernelGGL((k_bc_jump<false, 3>), grid, dim3(256), 0, s, cur, oth, E);
        else
            hipLaunchKernelGGL((k_bc_jump<false, 1>), grid, dim3(256), 0, s, cur, oth, E);
        r += m;
        std::swap(cur, oth);
    }
    hipLaunchKernelGGL(k_bc_heads, dim3((E + kHeadsBlock - 1) / kHeadsBlock), dim3(kHeadsBlock), 0, s, cur, oth, ctr, E);
    hipLaunchKernelGGL(k_bc_scatter, grid, dim3(256), 0, s, d_graph, cur, oth, A, rec, ctr, E);
    if (use_lds) {
        hipLaunchKernelGGL(k_bc_scan<true>, dim3(1), dim3(256), lds, s, d_graph, (int)N, rec, A, nullptr, d_nodes,
                           d_sizes, max_nodes, max_chains, ctr, d_counts);
    } else {
        hipLaunchKernelGGL(k_bc_scan<false>, dim3(1), dim3(256), 0, s, d_graph, (int)N, rec, A,
                           reinterpret_cast<uint32_t*>(d_scratch), d_nodes, d_sizes, max_nodes, max_chains, ctr,
                           d_counts);
    }
    DPX_HIP(hipGetLastError());
    return DPX_OK;
}

}  // namespace dpx

// -------------------------------------------------------------------------
// Plain per-frame path: kernel.cu's sequence with per-call allocation,
// blocking copies and one kernel per stage, plus the raster.
// -------------------------------------------------------------------------
namespace {

struct FrameBuffers {
    uint8_t *img = nullptr, *graph = nullptr, *graph_aux = nullptr, *out = nullptr;
    dpx_point *diagram = nullptr, *diagram_aux = nullptr, *pos = nullptr;
    int32_t *ec = nullptr, *ec_aux = nullptr;
    ~FrameBuffers()
    {
        (void)hipFree(img);
        (void)hipFree(graph);
        (void)hipFree(graph_aux);
        (void)hipFree(out);
        (void)hipFree(diagram);
        (void)hipFree(diagram_aux);
        (void)hipFree(ec);
        (void)hipFree(ec_aux);
    }
};

// Runs the stages up to triangulation on a frame already in fb.img.
int run_stages(FrameBuffers& fb, int W, int H, int stride, int ch, uint32_t flags, bool subdivide)
{
    using namespace dpx;
    size_t N = (size_t)W * H;
    if (int rc = dpx_stage_similarity_graph(fb.img, W, H, stride, ch, flags, fb.graph, nullptr)) return rc;
    if (int rc = dpx_stage_trivial_cross(fb.graph, W, H, nullptr)) return rc;
    if (int rc = dpx_stage_heuristic_cross(fb.graph, fb.graph_aux, W, H, nullptr)) return rc;
    if (int rc = dpx_stage_cells(fb.graph, W, H, fb.diagram, fb.ec, nullptr)) return rc;
    if (subdivide) {
        if (int rc = dpx_stage_subdivide(fb.img, W, H, stride, ch, flags, fb.graph, fb.diagram, fb.diagram_aux, fb.ec,
                                         fb.ec_aux, nullptr))
            return rc;
    }
    if (int rc = dpx_stage_triangulate(fb.diagram, fb.ec, W, H, nullptr)) return rc;
    (void)N;
    return DPX_OK;
}

int alloc_frame(FrameBuffers& fb, int W, int H, int stride, bool subdivide)
{
    size_t N = (size_t)W * H;
    DPX_HIP(hipMalloc(&fb.img, (size_t)stride * H));
    DPX_HIP(hipMalloc(&fb.graph, N));
    DPX_HIP(hipMalloc(&fb.graph_aux, N));
    DPX_HIP(hipMalloc(&fb.diagram, N * dpx::kCellSize * sizeof(dpx_point)));
    DPX_HIP(hipMalloc(&fb.ec, N * sizeof(int32_t)));
    if (subdivide) {
        DPX_HIP(hipMalloc(&fb.diagram_aux, N * dpx::kCellSize * sizeof(dpx_point)));
        DPX_HIP(hipMalloc(&fb.ec_aux, N * sizeof(int32_t)));
    }
    return DPX_OK;
}

}  // namespace

extern "C" int dpx_frame(const uint8_t* img, int W, int H, int stride, int ch, const dpx_params* params,
                         uint8_t* out_rgba, int out_stride, uint8_t* graph_out, int32_t* edge_count_out,
                         dpx_point* diagram_out)
{
    using namespace dpx;
    DPX_CHECK(params, DPX_EINVAL, "null params");
    if (int rc = check_img(img, W, H, stride, ch, params->flags)) return rc;
    DPX_CHECK(params->reserved == 0, DPX_EINVAL, "reserved must be 0");
    if (int rc = require_device()) return rc;
    int scale = params->scale;
    bool sub = params->subdivide != 0;
    if (out_rgba) {
        DPX_CHECK(scale >= 1 && scale <= 32, DPX_EINVAL, "scale %d outside 1..32", scale);
        DPX_CHECK(out_stride >= 4 * W * scale, DPX_EINVAL, "bad out_stride %d", out_stride);
    }
    size_t N = (size_t)W * H;
    FrameBuffers fb;
    if (int rc = alloc_frame(fb, W, H, stride, sub)) return rc;
    DPX_HIP(hipMemcpy(fb.img, img, (size_t)stride * H, hipMemcpyHostToDevice));
    if (int rc = run_stages(fb, W, H, stride, ch, params->flags, sub)) return rc;
    if (out_rgba) {
        DPX_HIP(hipMalloc(&fb.out, (size_t)out_stride * H * scale));
        if (int rc = dpx_stage_rasterize(fb.img, W, H, stride, ch, params->flags, fb.diagram, fb.ec, scale, fb.out,
                                         out_stride, nullptr))
            return rc;
    }
    DPX_HIP(hipDeviceSynchronize());
    if (out_rgba) DPX_HIP(hipMemcpy(out_rgba, fb.out, (size_t)out_stride * H * scale, hipMemcpyDeviceToHost));
    if (graph_out) DPX_HIP(hipMemcpy(graph_out, fb.graph, N, hipMemcpyDeviceToHost));
    if (edge_count_out) DPX_HIP(hipMemcpy(edge_count_out, fb.ec, N * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (diagram_out)
        DPX_HIP(hipMemcpy(diagram_out, fb.diagram, N * kCellSize * sizeof(dpx_point), hipMemcpyDeviceToHost));
    return DPX_OK;
}

extern "C" dpx_point* launch_kernel(dpx_point* pos, dpx_uchar4* colorPos, float time, char* img_data, int img_width,
                                    int img_height, int img_widthstep, int* edge_count_h, char* graph_h,
                                    bool subdivide)
{
    using namespace dpx;
    (void)time;
    int W = img_width, H = img_height;
    if (check_img(img_data, W, H, img_widthstep, 3, 0) || require_device()) return nullptr;
    size_t N = (size_t)W * H;
    FrameBuffers fb;
    if (alloc_frame(fb, W, H, img_widthstep, subdivide)) return nullptr;
    if (hipMemcpy(fb.img, img_data, (size_t)img_widthstep * H, hipMemcpyHostToDevice) != hipSuccess) {
        fail(DPX_EHIP, "image upload failed");
        return nullptr;
    }
    if (run_stages(fb, W, H, img_widthstep, 3, 0, subdivide)) return nullptr;
    if (pos || colorPos) {
        if (dpx_stage_export_vbo(fb.img, W, H, img_widthstep, 3, 0, fb.diagram, fb.ec, pos, colorPos, nullptr))
            return nullptr;
    }
    if (hipDeviceSynchronize() != hipSuccess) {
        fail(DPX_EHIP, "device error in the stage kernels");
        return nullptr;
    }
    dpx_point* diagram_h = (dpx_point*)malloc(N * kCellSize * sizeof(dpx_point));
    if (!diagram_h) {
        fail(DPX_ENOMEM, "host allocation of the diagram failed");
        return nullptr;
    }
    bool ok = true;
    if (edge_count_h)
        ok = ok && hipMemcpy(edge_count_h, fb.ec, N * sizeof(int32_t), hipMemcpyDeviceToHost) == hipSuccess;
    ok = ok && hipMemcpy(diagram_h, fb.diagram, N * kCellSize * sizeof(dpx_point), hipMemcpyDeviceToHost) == hipSuccess;
    if (graph_h) ok = ok && hipMemcpy(graph_h, fb.graph, N, hipMemcpyDeviceToHost) == hipSuccess;
    if (!ok) {
        free(diagram_h);
        fail(DPX_EHIP, "result download failed");
        return nullptr;
    }
    return diagram_h;
}
