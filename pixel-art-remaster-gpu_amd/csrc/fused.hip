// fused.hip -- the batched path: many frames or sprites of uneven size
// through two fused kernels.
//
//   kf_graph   tile of 64x32 pixels: lane words (YUV fields + class) once
//              per pixel, similarity in four directions (the other four by
//              symmetry) with block flags for the trivial crossing removal
//              -> snapshot graph (1 B/px)
//   kf_raster  tile of 32x8 cells: heuristic crossing resolution on the
//              snapshot (crossed blocks listed and decided once), cell keys
//              (and, with subdivision, hulls, smoothed polygons and their
//              coverage) in LDS; cells whose S*S pixels are all their own
//              colour are simple, the rest are listed and resolved row by
//              row (1x, 2x: cell by cell) against the 3x3 candidate cells
//              in draw order into an LDS pixel buffer; then every row of
//              the tile's output is stored once, simple and resolved pixels
//              by the same store instructions.
//
// Coverage of an unsmoothed cell depends only on its 12-bit key and the
// scale, so it is precomputed once per scale into a table of sample-row
// bitmasks (kt_mask): bit b of row r of key k says whether sample
// (kmin+b, kmin+r) of the cell's local sample grid is covered, kmin =
// -(S/2)-1, MW = 2S+2 rows and columns (local [-0.5, 1.5] of the cell); into
// a 16-bit summary per key (kt_info, coverage_info); and at 1x and 2x into
// each key's sample mask of each neighbour cell (kt_cell).
//
// With subdivision, kf_smooth runs between the two: it resolves the
// crossings of each tile, smooths and ear-clips its non-internal cells
// (polygons in LDS, each own cell once) and writes their coverage rows; the
// raster (kf_raster<S, true>) then reads the resolved graph and those rows.
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "geom.hpp"
#include "graph.hpp"
#include "pixel.hpp"

namespace dpx {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

[[maybe_unused]] constexpr int kMaxScale = 15;
[[maybe_unused]] constexpr int kKeys = 4096;
constexpr int kHullBytes = 32;  // n, x[8], y[8] (quarter units), padding

__host__ __device__ constexpr int mask_kmin(int S) { return -(S / 2) - 1; }
__host__ __device__ constexpr int mask_w(int S) { return 2 * S + 2; }

struct DevItem {
    uint64_t in_off, out_off, g_off;
    int32_t W, H, stride, ch;
    uint32_t flags;
    int32_t out_stride, scale, subdivide;
    int32_t aligned16;  // output rows 16-byte aligned
    int32_t graph_w4;   // graph rows 4-byte aligned (W % 4 == 0)
    uint64_t pal_off;   // DPX_INDEXED: the palette's byte offset in the input arena
    uint64_t rows_off;  // subdivision: byte offset of the item's cell rows (kf_smooth)
};

// ---------------------------------------------------------------------------
// tables
// ---------------------------------------------------------------------------
__global__ void kt_hull(int8_t* __restrict__ table)
{
    int key = blockIdx.x * blockDim.x + threadIdx.x;
    if (key >= kKeys) return;
    int hx[2 * kMaxPts], hy[2 * kMaxPts];
    int n = cell_hull((uint32_t)key, hx, hy);
    int8_t* e = table + (size_t)key * kHullBytes;
    e[0] = (int8_t)n;
    for (int t = 0; t < 8; t++) {
        e[1 + t] = (int8_t)(t < n ? hx[t] : 0);
        e[9 + t] = (int8_t)(t < n ? hy[t] : 0);
    }
}

// Coverage bits of one sample row of a polygon given in 1/64 units and
// already ear-clipped (tri, nt).  Units of the test: 1/(128 S) of a cell.
__device__ inline uint32_t coverage_row(const int* x, const int* y, const uint8_t* tri, int nt, int S, int r)
{
    const int kmin = mask_kmin(S), mw = mask_w(S);
    int sy = (2 * (kmin + r) + 1) * 64;
    uint32_t bits = 0;
    for (int t = 0; t < nt; t++) {
        int ax = x[tri[3 * t]] * 2 * S, ay = y[tri[3 * t]] * 2 * S;
        int bx = x[tri[3 * t + 1]] * 2 * S, by = y[tri[3 * t + 1]] * 2 * S;
        int cx = x[tri[3 * t + 2]] * 2 * S, cy = y[tri[3 * t + 2]] * 2 * S;
        int ymin = min(ay, min(by, cy)), ymax = max(ay, max(by, cy));
        if (sy < ymin || sy > ymax) continue;
        for (int b = 0; b < mw; b++) {
            int sx = (2 * (kmin + b) + 1) * 64;
            if (tri_covers(ax, ay, bx, by, cx, cy, sx, sy)) bits |= 1u << b;
        }
    }
    return bits;
}

__global__ void kt_mask(const int8_t* __restrict__ hull_table, int S, uint32_t* __restrict__ mask)
{
    int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int mw = mask_w(S);
    if (gid >= kKeys * mw) return;
    int key = gid / mw, r = gid % mw;
    const int8_t* e = hull_table + (size_t)key * kHullBytes;
    int n = e[0];
    int x[kMaxPoly], y[kMaxPoly];
    for (int t = 0; t < n; t++) { x[t] = 16 * e[1 + t]; y[t] = 16 * e[9 + t]; }
    uint8_t tri[3 * kMaxPoly];
    int ok;
    int nt = ear_clip(x, y, n, tri, &ok);
    mask[gid] = coverage_row(x, y, tri, nt, S, r);
}

// Summary of a cell's coverage rows at scale S: bit 0 = it covers every
// sample of its own square; bit 1 + (dy+1)*3 + (dx+1) = it covers a sample of
// the neighbouring square at (dx, dy).
__host__ __device__ inline uint32_t coverage_info(const uint32_t* rows, int S)
{
    const int kmin = mask_kmin(S), mw = mask_w(S);
    const int c0 = -kmin;  // bit of the own square's first column
    const uint32_t own = ((1u << S) - 1u) << c0;
    uint32_t info = 0;
    bool full = true;
    for (int r = 0; r < mw; r++) {
        int ky = kmin + r;
        int dy = ky < 0 ? -1 : (ky >= S ? 1 : 0);
        uint32_t row = rows[r];
        if (dy == 0 && (row & own) != own) full = false;
        if (row & ((1u << c0) - 1u)) info |= 1u << (1 + (dy + 1) * 3 + 0);
        if ((row & own) && dy != 0) info |= 1u << (1 + (dy + 1) * 3 + 1);
        if (row >> (c0 + S)) info |= 1u << (1 + (dy + 1) * 3 + 2);
    }
    return info | (full ? 1u : 0u);
}

// Cell masks (scales up to kMaxCellScale): entry j = (1 - dy) * 3 + (1 - dx)
// of key k holds, as bit ky * S + q, whether a cell of key k covers sample q
// of row ky of the cell at (-dx, -dy) from it, i.e. what k contributes when
// it is the candidate at (dx, dy) of that cell.  They follow the row table
// in the same allocation.
constexpr int kMaxCellScale = 2;
__global__ void kt_cell(const uint32_t* __restrict__ mask, int S, uint16_t* __restrict__ cell)
{
    int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= kKeys * 9) return;
    const int key = gid / 9, j = gid % 9, dy = 1 - j / 3, dx = 1 - j % 3;
    const int kmin = mask_kmin(S), mw = mask_w(S);
    const uint32_t full = (1u << S) - 1u;
    const int sh = -dx * S - kmin;
    uint32_t m = 0;
    for (int ky = 0; ky < S; ky++) {
        const int r = ky - dy * S - kmin;
        if (r < 0 || r >= mw) continue;
        const uint32_t row = mask[(size_t)key * mw + r];
        m |= ((sh >= 0 ? (row >> sh) : (row << -sh)) & full) << (ky * S);
    }
    cell[gid] = (uint16_t)m;
}

__global__ void kt_info(const uint32_t* __restrict__ mask, int S, uint16_t* __restrict__ info)
{
    int key = blockIdx.x * blockDim.x + threadIdx.x;
    if (key >= kKeys) return;
    info[key] = (uint16_t)coverage_info(mask + (size_t)key * mask_w(S), S);
}

// A cell whose own square it covers fully, and into which none of the four
// cells drawn after it (up-left, up, up-right, right) reaches, shows its own
// colour on all of its S*S output pixels.
__device__ inline bool simple_cell(uint32_t own, uint32_t upl, uint32_t up, uint32_t upr, uint32_t right)
{
    return (own & 1u) && !((upl >> 3) & 1u) && !((up >> 2) & 1u) && !((upr >> 1) & 1u) && !((right >> 4) & 1u);
}

// ---------------------------------------------------------------------------
// work distribution: block -> (item, tile column, tile row).  When every item
// of a launch has the same tile grid, the launch grid is (tile columns, tile
// rows, items) and the block index says it all; otherwise the grid is 1-D
// over all tiles and a map built on the host gives each block its item and
// tile: tile_map[b] = (item, tile row << 16 | tile column).
// ---------------------------------------------------------------------------
struct TileAt {
    int item, tx, ty;
};

// With RUN > 1 the blocks are permuted in groups of 8 * RUN: the hardware
// hands consecutive blocks to the 8 XCDs in turn, and within a group XCD x
// is given the RUN consecutive tiles x * RUN .. x * RUN + RUN - 1 (a row of
// horizontally adjacent tiles), so the input and graph lines that adjacent
// tiles share are fetched into one XCD's L2 once, not into several.  Blocks
// past the last whole group keep their place.
template <int RUN>
__device__ inline TileAt locate_tile(const uint2* __restrict__ tile_map, int grid3d)
{
    if constexpr (RUN == 1) {
        if (grid3d) return TileAt{ (int)blockIdx.z, (int)blockIdx.x, (int)blockIdx.y };
        const uint2 m = tile_map[blockIdx.x];
        return TileAt{ (int)m.x, (int)(m.y & 0xFFFFu), (int)(m.y >> 16) };
    } else {
        const uint32_t n = gridDim.x * gridDim.y * gridDim.z;
        uint32_t b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
        if (b < n / (8 * RUN) * (8 * RUN)) {
            const uint32_t w = b % (8 * RUN);
            b += (w % 8) * RUN + w / 8 - w;
        }
        if (grid3d) {
            const uint32_t plane = gridDim.x * gridDim.y, r = b % plane;
            return TileAt{ (int)(b / plane), (int)(r % gridDim.x), (int)(r / gridDim.x) };
        }
        const uint2 m = tile_map[b];
        return TileAt{ (int)m.x, (int)(m.y & 0xFFFFu), (int)(m.y >> 16) };
    }
}

// ---------------------------------------------------------------------------
// kf_graph: similarity graph + trivial crossing removal
// ---------------------------------------------------------------------------
// Three passes over LDS:
//   1. lane words (pixel class + YUV fields, see lane_ref) of the tile and a
//      halo: x0-1 .. x0+65, y0-2 .. y0+33;
//   2. per position of x0-1 .. x0+64, y0-1 .. y0+32, the four edges it owns
//      (up, up-right, right, down-right: bits 1, 2, 4, 7; similarity is
//      symmetric, so each edge is tested once);
//   3. per four pixels of a row at once (one byte each in a 32-bit word):
//      the other four edges from the left, down, down-left and up-left
//      neighbours' bytes, and the trivial removal: a 2x2 block is "full"
//      (all four orthogonal edges, crossCheck_4's condition for removing
//      both diagonals) iff its lower-left position has its up and right
//      edges, the position above it its right edge and the position right
//      of it its up edge; the flags of the four blocks around a pixel come
//      from those bits of the neighbouring bytes.
constexpr int GTX = 64, GTY = 32;
constexpr int GLX = GTX + 3, GLY = GTY + 4;  // lane words
// Lane-word rows in LDS: GLW words, column c + 3 holding x = x0 - 1 + c (so
// that x0 - 4 .. x0 + 67, 18 aligned groups of 4 pixels, fit a row).
constexpr int GLW = GTX + 8;
constexpr int GHX = GTX + 2, GHY = GTY + 2;  // positions of pass 2
constexpr int GHP = GTX + 8;                 // byte pitch of pass 2's rows: column c is x = x0 - 4 + c

__global__ __launch_bounds__(256) void kf_graph(const DevItem* __restrict__ items, const uint2* __restrict__ tile_map,
                                                int grid3d, const uint8_t* __restrict__ in,
                                                uint8_t* __restrict__ gaux)
{
    __shared__ __align__(16) uint32_t s_lane[GLW * GLY];
    __shared__ uint32_t s_h32[GHY * GHP / 4];
    uint8_t* s_h = reinterpret_cast<uint8_t*>(s_h32);
    const int t = threadIdx.x;
    const TileAt at = locate_tile<1>(tile_map, grid3d);
    const DevItem it = items[at.item];
    const int x0 = at.tx * GTX, y0 = at.ty * GTY;

    // pass 1.  Byte offsets within an item fit 31 bits (checked on the host).
    const uint8_t* base = in + it.in_off;
    const bool rgb = it.flags & DPX_RGB, topdown = it.flags & DPX_TOPDOWN;
    const bool indexed = it.flags & DPX_INDEXED;
    const bool alpha = (it.flags & DPX_ALPHA) && it.ch == 4;
    const uint32_t o0 = rgb ? 2 : 0, o2 = 2 - o0;
    if (indexed) {
        // the lane words of the 256 palette entries, in pass 2's rows until
        // pass 2 writes them
        const uint32_t e = reinterpret_cast<const uint32_t*>(in + it.pal_off)[t];
        const int c0 = (e >> (8 * o0)) & 0xFF, c1 = (e >> 8) & 0xFF, c2 = (e >> (8 * o2)) & 0xFF;
        uint32_t lw = yuv_lanes(yuv_pack(c0, c1, c2)) | kLaneOpaque;
        if ((it.flags & DPX_ALPHA) && (e >> 24) == 0) lw = kLaneTransparent;
        s_h32[t] = lw;
        __syncthreads();
    }
    // Lane word of pixel (x, y) at byte offset `off` of the item's buffer.
    auto lane_word = [&](uint32_t off) -> uint32_t {
        if (indexed) return s_h32[base[off]];
        int c0 = base[off + o0], c1 = base[off + 1], c2 = base[off + o2];
        if (alpha && base[off + 3] == 0) return kLaneTransparent;
        return yuv_lanes(yuv_pack(c0, c1, c2)) | kLaneOpaque;
    };
    // Rows of 3 or 4 channels that start on 4-byte boundaries are read as
    // groups of four pixels (3 or 4 dwords per group): x0 - 4 .. x0 + 67,
    // GLY rows, 18 groups a row; groups that straddle the frame's left or
    // right edge are read pixel by pixel.
    const bool grouped = !indexed && ((it.in_off | (uint64_t)it.stride) & 3) == 0;
    if (grouped) {
        for (int k = t; k < GLY * 18; k += 256) {
            const int r = k / 18, g = k % 18;
            const int x = x0 - 4 + 4 * g, y = y0 - 2 + r;
            uint32_t lw[4] = { 0, 0, 0, 0 };
            if ((unsigned)y < (unsigned)it.H && x < it.W && x + 4 > 0) {
                const uint32_t rowoff = (uint32_t)(topdown ? it.H - 1 - y : y) * (uint32_t)it.stride;
                if (x >= 0 && x + 3 < it.W) {
                    const uint32_t* p = reinterpret_cast<const uint32_t*>(base + rowoff + (uint32_t)x * it.ch);
                    uint32_t d[4];
                    d[0] = p[0];
                    d[1] = p[1];
                    d[2] = p[2];
                    d[3] = it.ch == 4 ? p[3] : 0u;
                    auto byte = [&](int j) { return (d[j >> 2] >> (8 * (j & 3))) & 0xFFu; };
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        int c0, c1, c2;
                        bool clear = false;
                        if (it.ch == 3) {  // (constant byte indices; the order is a select)
                            const int bA = (int)byte(3 * q), bC = (int)byte(3 * q + 2);
                            c0 = rgb ? bC : bA;
                            c1 = (int)byte(3 * q + 1);
                            c2 = rgb ? bA : bC;
                        } else {
                            c0 = (int)((d[q] >> (8 * o0)) & 0xFFu);
                            c1 = (int)((d[q] >> 8) & 0xFFu);
                            c2 = (int)((d[q] >> (8 * o2)) & 0xFFu);
                            clear = alpha && (d[q] >> 24) == 0;
                        }
                        lw[q] = clear ? kLaneTransparent : (yuv_lanes(yuv_pack(c0, c1, c2)) | kLaneOpaque);
                    }
                } else {
#pragma unroll
                    for (int q = 0; q < 4; q++)
                        if ((unsigned)(x + q) < (unsigned)it.W) lw[q] = lane_word(rowoff + (uint32_t)(x + q) * it.ch);
                }
            }
            u32x4 v = { lw[0], lw[1], lw[2], lw[3] };
            *reinterpret_cast<u32x4*>(&s_lane[r * GLW + 4 * g]) = v;
        }
    }
    // Otherwise pixel by pixel: window column c (x = x0 - 1 + c), row r
    // (y = y0 - 2 + r).  Columns 0..63: one per thread (t % 64), every fourth
    // row from t / 64, the byte offset stepping by four rows; columns
    // 64..GLX-1: one word per thread.
    if (!grouped) {
        const int c = t & 63, x = x0 - 1 + c;
        const bool xin = (unsigned)x < (unsigned)it.W;
        int y = y0 - 2 + (t >> 6);
        // modulo 2^32 (rows outside the frame may lie outside int); used only inside the frame
        uint32_t off = (uint32_t)(topdown ? it.H - 1 - y : y) * (uint32_t)it.stride + (uint32_t)(x * it.ch);
        const uint32_t step = (uint32_t)((topdown ? -4 : 4) * it.stride);
        for (int r = t >> 6; r < GLY; r += 4, y += 4, off += step) {
            uint32_t lw = 0;
            if (xin && (unsigned)y < (unsigned)it.H) lw = lane_word(off);
            s_lane[r * GLW + c + 3] = lw;
        }
    }
    if (!grouped && t < (GLX - 64) * GLY) {
        const int c = 64 + t % (GLX - 64), r = t / (GLX - 64);
        const int x = x0 - 1 + c, y = y0 - 2 + r;
        uint32_t lw = 0;
        if ((unsigned)x < (unsigned)it.W && (unsigned)y < (unsigned)it.H)
            lw = lane_word((uint32_t)((topdown ? it.H - 1 - y : y) * it.stride + x * it.ch));
        s_lane[r * GLW + c + 3] = lw;
    }
    __syncthreads();
    // pass 2: position (x0 - 1 + cx, y0 - 1 + cy), same column split
    auto edges = [&](int cx, int cy) {
        const int c = (cy + 1) * GLW + cx + 3;  // its lane word
        uint32_t up = s_lane[c + GLW], ur = s_lane[c + GLW + 1], rt = s_lane[c + 1], dr = s_lane[c - GLW + 1];
        LaneRef a = lane_ref(s_lane[c]);
        uint32_t h = lane_similar(a, up) << 1 | lane_similar(a, ur) << 2 | lane_similar(a, rt) << 4 |
                     lane_similar(a, dr) << 7;
        s_h[cy * GHP + cx + 3] = (uint8_t)h;
    };
    for (int cy = t >> 6; cy < GHY; cy += 4) edges(t & 63, cy);
    if (t < (GHX - 64) * GHY) edges(64 + t % (GHX - 64), t / (GHX - 64));
    __syncthreads();
    // pass 3
    uint8_t* g = gaux + it.g_off;
    for (int k = t; k < GTX / 4 * GTY; k += 256) {
        int q = k % (GTX / 4), ly = k / (GTX / 4);
        int x = x0 + 4 * q, y = y0 + ly;
        if (x >= it.W || y >= it.H) continue;
        const uint32_t* dn = s_h32 + ly * (GHP / 4) + q;  // row y-1; word q+1 holds x .. x+3
        const uint32_t* md = dn + GHP / 4;
        const uint32_t* up = md + GHP / 4;
        uint32_t own = md[1], left = __builtin_amdgcn_alignbyte(md[1], md[0], 3);
        uint32_t down = dn[1], dl = __builtin_amdgcn_alignbyte(dn[1], dn[0], 3);
        uint32_t ul = __builtin_amdgcn_alignbyte(up[1], up[0], 3);
        uint32_t edges = (own & 0x96969696u) | ((left & 0x10101010u) >> 1) | ((down & 0x02020202u) << 5) |
                         ((dl & 0x04040404u) << 3) | ((ul & 0x80808080u) >> 7);
        // full flags (bit 0 of each byte) of the blocks whose lower-left
        // positions are the bytes of h: its up (bit 1) and right (bit 4)
        // edges, the right edge of the byte above (in hu) and the up edge of
        // the byte to the right (in hr, h shifted by one position)
        auto full = [](uint32_t h, uint32_t hr, uint32_t hu) {
            return (h >> 1) & (h >> 4) & (hu >> 4) & (hr >> 1) & 0x01010101u;
        };
        const uint32_t fm1 = full(md[1], __builtin_amdgcn_alignbyte(md[2], md[1], 1), up[1]);  // x .. x+3, row y
        const uint32_t fm0 = full(md[0], __builtin_amdgcn_alignbyte(md[1], md[0], 1), up[0]);  // byte 3: x-1
        const uint32_t fd1 = full(dn[1], __builtin_amdgcn_alignbyte(dn[2], dn[1], 1), md[1]);  // row y-1
        const uint32_t fd0 = full(dn[0], __builtin_amdgcn_alignbyte(dn[1], dn[0], 1), md[0]);
        uint32_t clear = (fm1 << 2) | __builtin_amdgcn_alignbyte(fm1, fm0, 3) |
                         (__builtin_amdgcn_alignbyte(fd1, fd0, 3) << 5) | (fd1 << 7);
        uint32_t v = edges & ~clear;
        uint8_t* dst = g + (uint32_t)(y * it.W + x);
        if (it.graph_w4) {
            *reinterpret_cast<uint32_t*>(dst) = v;
        } else {
            int n = min(4, it.W - x);
            for (int b = 0; b < n; b++) dst[b] = (uint8_t)(v >> (8 * b));
        }
    }
}

// ---------------------------------------------------------------------------
// kf_raster: heuristic crossing resolution, cells (+ subdivision) and the
// gather raster of a 32x8-cell tile
// ---------------------------------------------------------------------------
constexpr int RTX = 32, RTY = 8;
constexpr int RCX = RTX + 2, RCY = RTY + 2;  // candidate cells, halo 1
constexpr uint16_t kNoCell = 0xFFFF;

// LDS windows of the raster tile hold graph bytes in rows of RWW 32-bit
// words; word w of a row holds x = x0 - 8 + 4w .. x0 - 5 + 4w.  Words 1..10
// (x0 - 4 .. x0 + 35) carry data, words 0 and 11 stay zero.  Candidate cells
// (halo 1) need the resolved bytes of their left and right neighbours for
// their keys: resolved rows at halo 1; with subdivision the hulls at halo 2
// are needed as well: resolved rows at halo 2.  Resolving a byte reads the
// snapshot rows above and below it.
constexpr int RWW = 12, RWB = 4 * RWW;
static_assert(RTX * RTY == 256, "one own cell per thread of a 256-thread block");
// Resolved rows y0 - 1 .. y0 + RTY (candidate cells need their left and
// right neighbours' bytes for their keys); resolving them reads one more
// snapshot row on each side.
constexpr int GY = RTY + 2, AY = GY + 2;

// Per-cell coverage rows of smoothed cells (kf_smooth -> kf_raster<S, true>):
// MW words of sample rows and one word of coverage summary per cell, cell
// (x, y) of an item at word (y * W + x) * (MW + 1) of its range.
__host__ __device__ constexpr int rows_words(int S) { return mask_w(S) + 1; }

// The snapshot rows ay0 .. ay0 + AY - 1 of a tile's window into LDS.  With
// W % 4 == 0 the graph rows are word aligned and a word is either wholly
// inside the frame or wholly outside.
__device__ inline void tile_snapshot_load(const DevItem& it, const uint8_t* __restrict__ aux, int x0, int ay0,
                                          uint32_t* s_aux32, int t)
{
    if (it.graph_w4) {
        if (t < AY * RWW) {
            int r = t / RWW, w = t % RWW;
            int x = x0 - 8 + 4 * w, y = ay0 + r;
            uint32_t v = 0;
            if (w >= 1 && w <= 10 && (unsigned)x < (unsigned)it.W && (unsigned)y < (unsigned)it.H)
                v = *reinterpret_cast<const uint32_t*>(aux + (uint32_t)(y * it.W + x));
            s_aux32[t] = v;
        }
    } else {
        uint8_t* s8 = reinterpret_cast<uint8_t*>(s_aux32);
        for (int k = t; k < AY * RWB; k += 256) {
            int r = k / RWB, c = k % RWB;
            int x = x0 - 8 + c, y = ay0 + r;
            bool ok = c >= 4 && c < 44 && (unsigned)x < (unsigned)it.W && (unsigned)y < (unsigned)it.H;
            s8[k] = ok ? aux[(uint32_t)(y * it.W + x)] : 0;
        }
    }
}

// Heuristic crossing resolution (ambiguous_cross_Kernel) of a tile's window,
// after tile_snapshot_load and a barrier.  Only the pixels of crossed 2x2
// blocks (both diagonals present) change, and in the symmetric snapshot the
// block with lower-left pixel p is crossed iff p has bit 2 and its right
// neighbour bit 0.  1: list the crossed blocks of the window, four origins
// of a row per thread; 2: decide each listed block once
// (processHeuristics2), the walks reading the snapshot in LDS where they can
// and in HBM beyond the window; 3: remove the losing diagonals, four
// positions per thread, into s_g32 (resolved rows y0 - 1 .. y0 + RTY), and
// the tile's own resolved bytes into gfin when it is given.  s_p32 must
// have been cleared and *s_nblk zeroed before the barrier.
__device__ inline void tile_resolve(const DevItem& it, const uint8_t* __restrict__ aux, uint8_t* __restrict__ gfin,
                                    int x0, int y0, int ay0, const uint32_t* s_aux32, uint32_t* s_p32,
                                    uint16_t* s_blk, int* s_nblk, uint32_t* s_g32, int t)
{
    for (int k = t; k < (GY + 1) * 10; k += 256) {
        const int r = k / 10, w = k % 10 + 1;  // word w of snapshot row r: origins of blocks on rows r, r + 1
        const uint32_t* a = s_aux32 + r * RWW + w;
        uint32_t D = a[0];
        uint32_t cross = (D >> 2) & __builtin_amdgcn_alignbyte(a[1], D, 1) & 0x01010101u;
        while (cross) {
            int q = __builtin_ctz(cross) >> 3;
            cross &= cross - 1;
            s_blk[atomicAdd(s_nblk, 1)] = (uint16_t)(r * RWB + 4 * w + q);
        }
    }
    __syncthreads();
    {
        // (the window is read by words, so that the compiler keeps the LDS
        // and HBM reads apart instead of merging them into one generic load)
        auto get = [&](int gx, int gy) -> uint32_t {
            int wx = gx - (x0 - 4), wy = gy - ay0;
            bool win = (unsigned)wx < 40u && (unsigned)wy < (unsigned)AY;
            int c = win ? wy * RWB + wx + 4 : 0;
            uint32_t v = (s_aux32[c >> 2] >> (8 * (c & 3))) & 0xFFu;
            if (!win) v = aux[(uint32_t)(gy * it.W + gx)];
            return v;
        };
        const int nblk = *s_nblk;
        uint8_t* p8 = reinterpret_cast<uint8_t*>(s_p32);
        for (int k = t; k < nblk; k += 256) {
            int e = s_blk[k];
            int r = e / RWB, c = e % RWB;
            // 1: "/" loses (bits 2 and 5), 2: "\" loses (bits 0 and 7)
            p8[e] = remove_slash(get, x0 - 8 + c, ay0 + r) ? 1 : 2;
        }
    }
    __syncthreads();
    for (int k = t; k < GY * 10; k += 256) {
        const int r = k / 10, w = k % 10 + 1;  // word w of resolved row r (snapshot row r + 1)
        const int y = ay0 + 1 + r, x = x0 - 8 + 4 * w;
        const uint32_t* pa = s_p32 + (r + 1) * RWW + w;  // decisions of the blocks with origin on row y
        const uint32_t* pb = pa - RWW;                   // ... on row y - 1
        uint32_t A = pa[0], Al = __builtin_amdgcn_alignbyte(A, pa[-1], 3);
        uint32_t B = pb[0], Bl = __builtin_amdgcn_alignbyte(B, pb[-1], 3);
        uint32_t clear = ((A & 0x01010101u) << 2) |        // up-right block: "/"
                         ((Al >> 1) & 0x01010101u) |        // up-left block: "\"
                         ((Bl & 0x01010101u) << 5) |        // down-left block: "/"
                         (((B >> 1) & 0x01010101u) << 7);  // down-right block: "\"
        uint32_t res = s_aux32[(r + 1) * RWW + w] & ~clear;
        s_g32[r * RWW + w] = res;
        // the resolved graph of the tile's own cells, when the caller keeps it
        if (gfin && y >= y0 && y < y0 + RTY && w >= 2 && w <= 9 && y < it.H && x < it.W) {
            uint8_t* dst = gfin + it.g_off + (uint32_t)(y * it.W + x);
            if (it.graph_w4) {
                *reinterpret_cast<uint32_t*>(dst) = res;
            } else {
                int n = min(4, it.W - x);
                for (int q = 0; q < n; q++) dst[q] = (uint8_t)(res >> (8 * q));
            }
        }
    }
}

// The resolved rows y0 - 1 .. y0 + RTY of a tile's window, read from the
// resolved graph written by kf_smooth (words laid out as in s_g32).
__device__ inline void tile_resolved_load(const DevItem& it, const uint8_t* __restrict__ gfin, int x0, int y0,
                                          uint32_t* s_g32, int t)
{
    const uint8_t* g = gfin + it.g_off;
    if (it.graph_w4) {
        if (t < GY * RWW) {
            int r = t / RWW, w = t % RWW;
            int x = x0 - 8 + 4 * w, y = y0 - 1 + r;
            uint32_t v = 0;
            if (w >= 1 && w <= 10 && (unsigned)x < (unsigned)it.W && (unsigned)y < (unsigned)it.H)
                v = *reinterpret_cast<const uint32_t*>(g + (uint32_t)(y * it.W + x));
            s_g32[t] = v;
        }
    } else {
        uint8_t* s8 = reinterpret_cast<uint8_t*>(s_g32);
        for (int k = t; k < GY * RWB; k += 256) {
            int r = k / RWB, c = k % RWB;
            int x = x0 - 8 + c, y = y0 - 1 + r;
            bool ok = c >= 4 && c < 44 && (unsigned)x < (unsigned)it.W && (unsigned)y < (unsigned)it.H;
            s8[k] = ok ? g[(uint32_t)(y * it.W + x)] : 0;
        }
    }
}

// ---------------------------------------------------------------------------
// kf_smooth: crossing resolution, cells and corner-cutting subdivision of a
// 32x8-cell tile (subdivision_Kernel kernel.cu:216-261 on the cells_Kernel
// hulls, then triangulate_Kernel's ear clipping, kernel.cu:264-282), each
// own cell once.  Writes the tile's resolved graph bytes to gfin and, for
// every smoothed (non-internal) cell, the coverage rows of its triangles at
// scale S with their summary to `rows` (rows_words(S) words per cell).
// Internal cells (node 90) keep their hull (kernel.cu:231): the raster takes
// their rows from the table.  The cells that need smoothing are listed first
// so that every lane of a wave works on one; each lane's polygon lives in
// LDS (a column of s_poly), where ear clipping scans and shifts it.
// ---------------------------------------------------------------------------
template <int S>
__global__ __launch_bounds__(256) void kf_smooth(const DevItem* __restrict__ items, const uint2* __restrict__ tile_map,
                                                 int grid3d, const uint8_t* __restrict__ in,
                                                 const uint8_t* __restrict__ gaux, uint8_t* __restrict__ gfin,
                                                 const int8_t* __restrict__ hull_table, uint32_t* __restrict__ rows)
{
    constexpr int MW = mask_w(S), KMIN = mask_kmin(S), RW = rows_words(S);
    constexpr int HB = 17;  // LDS bytes per hull: n, x[8], y[8] (quarter units)
    __shared__ uint32_t s_aux32[AY * RWW];
    __shared__ uint32_t s_p32[(GY + 1) * RWW];
    __shared__ uint16_t s_blk[(GY + 1) * 40];
    __shared__ int s_nblk;
    __shared__ uint32_t s_g32[GY * RWW];
    __shared__ int8_t s_hull[RCX * RCY * HB];
    __shared__ uint16_t s_list[RTX * RTY];
    __shared__ int s_nlist;
    __shared__ uint32_t s_poly[2 * kMaxPts * 256];  // point k of lane t at [k * 256 + t]: x | y << 16 (1/64 units)
    const uint8_t* s_g = reinterpret_cast<const uint8_t*>(s_g32);

    const int t = threadIdx.x;
    const TileAt at = locate_tile<8>(tile_map, grid3d);
    const DevItem it = items[at.item];
    const int x0 = at.tx * RTX, y0 = at.ty * RTY;
    const uint8_t* aux = gaux + it.g_off;
    const int ay0 = y0 - 2;
    if (t == 0) {
        s_nlist = 0;
        s_nblk = 0;
    }
    if (t < (GY + 1) * RWW) s_p32[t] = 0;
    tile_snapshot_load(it, aux, x0, ay0, s_aux32, t);
    __syncthreads();
    tile_resolve(it, aux, gfin, x0, y0, ay0, s_aux32, s_p32, s_blk, &s_nblk, s_g32, t);
    __syncthreads();
    // hulls of the own cells and of their neighbours: cell (x0 - 1 + lx,
    // y0 - 1 + ly) at resolved row ly, byte lx + 7
    for (int k = t; k < RCX * RCY; k += 256) {
        const int lx = k % RCX, ly = k / RCX;
        const int c = ly * RWB + lx + 7;
        const uint32_t key = cell_key(s_g[c], s_g[c - 1], s_g[c + 1]);
        const int8_t* e = hull_table + (size_t)key * kHullBytes;
#pragma unroll
        for (int q = 0; q < HB; q++) s_hull[k * HB + q] = e[q];
    }
    // own cells to smooth: in the frame and not internal
    {
        const int cx = t % RTX, cy = t / RTX;
        const uint32_t node = s_g[(cy + 1) * RWB + cx + 8];
        if (x0 + cx < it.W && y0 + cy < it.H && node != 90) s_list[atomicAdd(&s_nlist, 1)] = (uint16_t)t;
    }
    __syncthreads();
    if (t >= s_nlist) return;
    const int own = s_list[t];
    const int cx = own % RTX, cy = own / RTX;
    const int x = x0 + cx, y = y0 + cy;
    const uint32_t node = s_g[(cy + 1) * RWB + cx + 8];
    const int hk = (cy + 1) * RCX + cx + 1;
    const int8_t* h = s_hull + hk * HB;
    const int n = h[0];
    struct HullX {
        const int8_t* p;
        __device__ int operator[](int k) const { return p[1 + k]; }
    };
    struct HullY {
        const int8_t* p;
        __device__ int operator[](int k) const { return p[9 + k]; }
    };
    ImgView im{ in + it.in_off, it.W, it.H, it.stride, it.ch, it.flags,
                (it.flags & DPX_INDEXED) ? in + it.pal_off : nullptr };
    const TjView tv = tj_view(im);
    auto nbr = [&](int e, auto f) {
        const int8_t* q = s_hull + (hk + edy(e) * RCX + edx(e)) * HB;
        return f(HullX{ q }, HullY{ q }, (int)q[0]);
    };
    auto tj = [&](int x4, int y4) -> bool { return tjunction(tv, x, y, x4, y4); };
    uint32_t* P = s_poly + t;  // this lane's column
    int m = 0;
    subdivide_cell_emit(HullX{ h }, HullY{ h }, n, node, nbr, tj, [&](int px, int py) {
        P[256 * m] = ((uint32_t)px & 0xFFFFu) | ((uint32_t)py << 16);
        m++;
    });
    auto px_of = [](uint32_t p) { return (int)(int16_t)(p & 0xFFFFu); };
    auto py_of = [](uint32_t p) { return (int)p >> 16; };
    // Ear clipping (process, triangulate_functions.cu:108-217) as ear_clip
    // does it: vertices in counter-clockwise order (reversed when the
    // signed area is not positive), first ear tried at (m-1, 0, 1), an ear
    // needs a positive cross and no other vertex inside or on it, a failure
    // counter of 2 nv.  Each ear's coverage rows are accumulated as it is cut.
    uint32_t cov[MW];
#pragma unroll
    for (int r = 0; r < MW; r++) cov[r] = 0;
    if (m >= 3) {
        int a2 = 0;
        for (int p = m - 1, q = 0; q < m; p = q++) {
            const uint32_t A = P[256 * p], B = P[256 * q];
            a2 += px_of(A) * py_of(B) - px_of(B) * py_of(A);
        }
        if (a2 <= 0)
            for (int lo = 0, hi = m - 1; lo < hi; lo++, hi--) {
                const uint32_t tmp = P[256 * lo];
                P[256 * lo] = P[256 * hi];
                P[256 * hi] = tmp;
            }
        int nv = m, count = 2 * nv;
        for (int v = nv - 1; nv > 2;) {
            if (0 >= (count--)) break;  // the reference gives up; the triangles cut so far stand
            int u = v;
            if (nv <= u) u = 0;
            v = u + 1;
            if (nv <= v) v = 0;
            int w = v + 1;
            if (nv <= w) w = 0;
            const uint32_t A = P[256 * u], B = P[256 * v], Cc = P[256 * w];
            const int ax = px_of(A), ay = py_of(A), bx = px_of(B), by = py_of(B), cx_ = px_of(Cc), cy_ = py_of(Cc);
            bool ear = ((bx - ax) * (cy_ - ay) - (by - ay) * (cx_ - ax)) > 0;
            if (ear) {
                for (int p = 0; p < nv; p++) {
                    if (p == u || p == v || p == w) continue;
                    const uint32_t Q = P[256 * p];
                    if (inside_tri(ax, ay, bx, by, cx_, cy_, px_of(Q), py_of(Q))) {
                        ear = false;
                        break;
                    }
                }
            }
            if (ear) {
                // coverage of triangle (A, B, C) in units of 1/(128 S) of a
                // cell; it is counter-clockwise (positive cross)
                const int tax = ax * 2 * S, tay = ay * 2 * S, tbx = bx * 2 * S, tby = by * 2 * S;
                const int tcx = cx_ * 2 * S, tcy = cy_ * 2 * S;
                const int ymin = min(tay, min(tby, tcy)), ymax = max(tay, max(tby, tcy));
#pragma unroll
                for (int r = 0; r < MW; r++) {
                    const int sy = (2 * (KMIN + r) + 1) * 64;
                    if (sy < ymin || sy > ymax) continue;
                    cov[r] |= tri_row_bits(tax, tay, tbx, tby, tcx, tcy, sy, KMIN, MW);
                }
                for (int s2 = v, t2 = v + 1; t2 < nv; s2++, t2++) P[256 * s2] = P[256 * t2];
                nv--;
                count = 2 * nv;
            }
        }
    }
    uint32_t* dst = rows + it.rows_off / 4 + (size_t)((uint32_t)y * (uint32_t)it.W + (uint32_t)x) * RW;
#pragma unroll
    for (int r = 0; r < MW; r++) dst[r] = cov[r];
    dst[MW] = coverage_info(cov, S);
}

// `items` are the items of one (scale, subdivision) group, in launch order.
// With SUB the resolved graph comes from gfin and the coverage rows of the
// smoothed cells from `cell_rows` (kf_smooth ran before); without, the kernel
// resolves the crossings itself from the snapshot.
template <int S, bool SUB>
__global__ __launch_bounds__(256) void kf_raster(const DevItem* __restrict__ items, const uint2* __restrict__ tile_map,
                                                 int grid3d,
                                                 const uint8_t* __restrict__ in, const uint8_t* __restrict__ gaux,
                                                 uint8_t* __restrict__ gfin, const uint32_t* __restrict__ cell_rows,
                                                 const uint32_t* __restrict__ mask_table,
                                                 const uint16_t* __restrict__ info_table, uint8_t* __restrict__ out)
{
    constexpr int MW = mask_w(S), KMIN = mask_kmin(S), RW = rows_words(S);
    __shared__ uint32_t s_aux32[SUB ? 1 : AY * RWW];
    __shared__ uint32_t s_p32[SUB ? 1 : (GY + 1) * RWW];  // decisions of crossed blocks, by origin
    __shared__ uint16_t s_blk[SUB ? 1 : (GY + 1) * 40];   // crossed blocks: origin byte index in the window
    __shared__ int s_nblk;
    __shared__ uint32_t s_g32[GY * RWW];
    __shared__ uint16_t s_key[RCX * RCY];
    __shared__ uint8_t s_info[RCX * RCY];  // coverage summary bits 0..4 (simple_cell)
    __shared__ uint32_t s_col[RCX * RCY];
    __shared__ uint32_t s_src[SUB ? RCX * RCY : 1];  // smoothed candidates: word offset of their rows
    __shared__ uint16_t s_list[RTX * RTY];  // own cells that need the candidate search
    __shared__ int s_nlist;
    // resolved pixels of the first listed cells, so that every own
    // cell's rows are stored together (a line is then written by one store
    // instruction instead of being split between the direct and searched
    // cells).  A pixel is kept as its colour; from S = 6 a tile whose listed
    // cells would overflow that keeps the code of each pixel's winning
    // candidate instead (1 byte: (1 - dy) * 3 + (1 - dx), 9 = clear colour),
    // four times as many cells.  Measured: codes cost a decode per pixel, so
    // tiles that fit keep colours; at S = 5 codes lose to the overflow stores.
    constexpr bool CODES = S >= 6;
    constexpr int PX_WORDS = 2048, PX_CELLS = PX_WORDS / (S * S), PX_CODE_CELLS = 4 * PX_WORDS / (S * S);
    static_assert(PX_CELLS >= 1, "at least one buffered cell");
    __shared__ uint32_t s_px[PX_WORDS];
    uint8_t* s_px8 = reinterpret_cast<uint8_t*>(s_px);
    __shared__ uint16_t s_idx[RTX * RTY];    // list index of a listed own cell
    __shared__ uint8_t s_simple[RTX * RTY];  // own cell shows its own colour only
    const uint8_t* s_g = reinterpret_cast<const uint8_t*>(s_g32);

    const int t = threadIdx.x;
    const TileAt at = locate_tile<8>(tile_map, grid3d);
    const DevItem it = items[at.item];
    const int x0 = at.tx * RTX, y0 = at.ty * RTY;
    const uint8_t* aux = gaux + it.g_off;
    const int ay0 = y0 - 2;  // first snapshot row

    if (t == 0) {
        s_nlist = 0;
        s_nblk = 0;
    }
    // The snapshot bytes (or the resolved ones) and the candidates' colours
    // are independent loads: issue them together.
    if constexpr (SUB) {
        tile_resolved_load(it, gfin, x0, y0, s_g32, t);
    } else {
        if (t < (GY + 1) * RWW) s_p32[t] = 0;
        tile_snapshot_load(it, aux, x0, ay0, s_aux32, t);
    }
    {
        const uint8_t* base = in + it.in_off;
        const uint32_t o0 = (it.flags & DPX_RGB) ? 2 : 0, o2 = 2 - o0;
        const bool topdown = it.flags & DPX_TOPDOWN, four = it.ch == 4;
        const bool indexed = it.flags & DPX_INDEXED;
        const uint32_t* pal = reinterpret_cast<const uint32_t*>(in + it.pal_off);
        for (int k = t; k < RCX * RCY; k += 256) {
            int x = x0 - 1 + k % RCX, y = y0 - 1 + k / RCX;
            if ((unsigned)x < (unsigned)it.W && (unsigned)y < (unsigned)it.H) {
                int row = topdown ? it.H - 1 - y : y;
                // row * stride < 2^31 (checked on the host); the stride itself may exceed 2^24
                uint32_t off = (uint32_t)row * (uint32_t)it.stride + __umul24(x, it.ch);
                if (indexed) {
                    const uint32_t e = pal[base[off]];
                    s_col[k] = ((e >> (8 * o2)) & 0xFFu) | (e & 0xFF00u) | ((e >> (8 * o0)) & 0xFFu) << 16 |
                               (e & 0xFF000000u);
                } else {
                    uint32_t a = four ? base[off + 3] : 255u;
                    s_col[k] = (uint32_t)base[off + o2] | (uint32_t)base[off + 1] << 8 |
                               (uint32_t)base[off + o0] << 16 | a << 24;  // colorPos (kernel.cu:98-101)
                }
            }
        }
    }
    __syncthreads();
    if constexpr (!SUB) {
        tile_resolve(it, aux, gfin, x0, y0, ay0, s_aux32, s_p32, s_blk, &s_nblk, s_g32, t);
        __syncthreads();
    }
    // candidate cells: key and coverage summary.  Candidate (lx, ly) is the
    // cell at (x0 - 1 + lx, y0 - 1 + ly): resolved row ly, byte lx + 7.
    for (int k = t; k < RCX * RCY; k += 256) {
        int lx = k % RCX, ly = k / RCX;
        int x = x0 - 1 + lx, y = y0 - 1 + ly;
        uint16_t key = kNoCell;
        uint8_t info = 0;
        if ((unsigned)x < (unsigned)it.W && (unsigned)y < (unsigned)it.H) {
            int c = ly * RWB + lx + 7;
            key = (uint16_t)cell_key(s_g[c], s_g[c - 1], s_g[c + 1]);
            if constexpr (SUB) {
                // a smoothed cell's rows and summary come from kf_smooth;
                // an internal one (node 90) keeps its hull: the table's
                uint32_t src = 0xFFFFFFFFu;
                if ((key & 0xFFu) != 90) {
                    src = (uint32_t)(it.rows_off / 4) + ((uint32_t)y * (uint32_t)it.W + (uint32_t)x) * RW;
                    info = (uint8_t)(cell_rows[src + MW] & 0x1Fu);
                } else {
                    info = (uint8_t)(info_table[key] & 0x1Fu);
                }
                s_src[k] = src;
            } else {
                info = (uint8_t)(info_table[key] & 0x1Fu);
            }
        }
        s_key[k] = key;
        s_info[k] = info;
    }
    __syncthreads();

    // Output pixels are written once and not read back by the kernels:
    // non-temporal stores keep them from displacing the graph and table lines
    // the other tiles' loads need in the caches.
    const int OH = it.H * S;
    auto store_row = [&](int gx, int Y, const uint32_t* col) {
        int row = out_topdown(it.flags) ? (OH - 1 - Y) : Y;
        uint8_t* dst = out + it.out_off + (size_t)row * it.out_stride + (size_t)gx * S * 4;
        if constexpr (S % 4 == 0) {
            if (it.aligned16) {
#pragma unroll
                for (int q = 0; q < S; q += 4) {
                    u32x4 v = { col[q], col[q + 1], col[q + 2], col[q + 3] };
                    __builtin_nontemporal_store(v, reinterpret_cast<u32x4*>(dst + 4 * q));
                }
                return;
            }
        }
        if constexpr (S % 2 == 0) {
            if ((reinterpret_cast<uintptr_t>(dst) & 7) == 0) {
#pragma unroll
                for (int q = 0; q < S; q += 2) {
                    u32x2 v = { col[q], col[q + 1] };
                    __builtin_nontemporal_store(v, reinterpret_cast<u32x2*>(dst + 4 * q));
                }
                return;
            }
        }
#pragma unroll
        for (int q = 0; q < S; q++) __builtin_nontemporal_store(col[q], reinterpret_cast<uint32_t*>(dst + 4 * q));
    };

    // Own cells, one per thread: a cell whose S*S pixels all show its own
    // colour is written directly, the others are listed for the candidate
    // search, which resolves their pixels into LDS first.
    const int cx = t % RTX, cy = t / RTX;
    const int kown = (cy + 1) * RCX + cx + 1;
    const bool mine = x0 + cx < it.W && y0 + cy < it.H;
    const bool simple = mine && simple_cell(s_info[kown], s_info[kown + RCX - 1], s_info[kown + RCX],
                                            s_info[kown + RCX + 1], s_info[kown + 1]);
    if (mine && !simple) {
        const int i = atomicAdd(&s_nlist, 1);
        s_list[i] = (uint16_t)t;
        s_idx[t] = (uint16_t)i;
    }
    s_simple[t] = simple ? 1 : 0;
    __syncthreads();

    // One unit = one output row (S pixels) of a listed cell, resolved
    // against the 3x3 candidates in draw order (highest cell index first).
    constexpr uint32_t FULL = (S >= 32) ? 0xFFFFFFFFu : ((1u << S) - 1u);
    const int nunits = s_nlist * S;
    // uniform over the block; at 6x and 7x only tiles far beyond the colour
    // buffer gain from codes (measured: sprites lose 5% at a threshold of 1x)
    const bool codes = CODES && s_nlist > (S >= 8 ? 1 : 3) * PX_CELLS;
    const int px_cells = codes ? PX_CODE_CELLS : PX_CELLS;
    auto fetch_rows = [&](int u, uint32_t* rows) {
        int own = s_list[u / S], ky = u % S;
        int ux = own % RTX, uy = own / RTX;
        // an own cell that covers its whole square takes every pixel the
        // cells drawn after it leave: the candidates drawn before it (left,
        // down-left, down, down-right) can win none, and are not fetched
        const bool own_full = s_info[(uy + 1) * RCX + ux + 1] & 1u;
#pragma unroll
        for (int dy = 1; dy >= -1; dy--) {
#pragma unroll
            for (int dx = 1; dx >= -1; dx--) {
                int k = (uy + 1 + dy) * RCX + (ux + 1 + dx);
                int r = ky - dy * S - KMIN;
                uint32_t row = 0;
                const bool before_own = dy < 0 || (dy == 0 && dx < 0);
                if (!(own_full && before_own) && r >= 0 && r < MW && s_key[k] != kNoCell) {
                    if constexpr (SUB) {
                        const uint32_t src = s_src[k];
                        row = src != 0xFFFFFFFFu ? cell_rows[src + r] : mask_table[(uint32_t)s_key[k] * MW + r];
                    } else {
                        row = mask_table[(uint32_t)s_key[k] * MW + r];
                    }
                }
                rows[(1 - dy) * 3 + (1 - dx)] = row;
            }
        }
    };
    auto resolve_unit = [&](int u, const uint32_t* rows) {
        int own = s_list[u / S], ky = u % S;
        int ux = own % RTX, uy = own / RTX;
        uint32_t col[S];
        if constexpr (CODES) {
            if (codes) {  // the same search, keeping candidate codes
                uint8_t code[S];
#pragma unroll
                for (int q = 0; q < S; q++) code[q] = 9;
                uint32_t assigned = 0;
#pragma unroll
                for (int dy = 1; dy >= -1; dy--) {
#pragma unroll
                    for (int dx = 1; dx >= -1; dx--) {
                        uint32_t row = rows[(1 - dy) * 3 + (1 - dx)];
                        int sh = -dx * S - KMIN;
                        uint32_t nb = (sh >= 0 ? (row >> sh) : (row << -sh)) & FULL & ~assigned;
                        if (nb) {
#pragma unroll
                            for (int q = 0; q < S; q++)
                                if ((nb >> q) & 1u) code[q] = (uint8_t)((1 - dy) * 3 + (1 - dx));
                            assigned |= nb;
                        }
                    }
                }
                if (u / S < PX_CODE_CELLS) {
#pragma unroll
                    for (int q = 0; q < S; q++) s_px8[u * S + q] = code[q];
                } else {
#pragma unroll
                    for (int q = 0; q < S; q++)
                        col[q] = code[q] == 9 ? 0xFF000000u
                                              : s_col[(uy + 2 - code[q] / 3) * RCX + (ux + 2 - code[q] % 3)];
                    store_row(x0 + ux, (y0 + uy) * S + ky, col);  // beyond the buffer: written directly
                }
                return;
            }
        }
#pragma unroll
        for (int q = 0; q < S; q++) col[q] = 0xFF000000u;
        uint32_t assigned = 0;
#pragma unroll
        for (int dy = 1; dy >= -1; dy--) {
#pragma unroll
            for (int dx = 1; dx >= -1; dx--) {
                uint32_t row = rows[(1 - dy) * 3 + (1 - dx)];
                int sh = -dx * S - KMIN;
                uint32_t bits = (sh >= 0 ? (row >> sh) : (row << -sh)) & FULL;
                uint32_t nb = bits & ~assigned;
                if (nb) {
                    uint32_t c = s_col[(uy + 1 + dy) * RCX + (ux + 1 + dx)];
#pragma unroll
                    for (int q = 0; q < S; q++)
                        if ((nb >> q) & 1u) col[q] = c;
                    assigned |= nb;
                }
            }
        }
        if (u / S < PX_CELLS) {
#pragma unroll
            for (int q = 0; q < S; q++) s_px[u * S + q] = col[q];
        } else {
            store_row(x0 + ux, (y0 + uy) * S + ky, col);  // beyond the buffer: written directly
        }
    };

    if constexpr (!SUB && S <= kMaxCellScale) {
        // 1x and 2x: one thread per listed cell, the candidates' samples of
        // it from the cell-mask table (nine independent loads), resolved in
        // the same draw order as the rows of the other scales.  Measured: 2x
        // frames 8-11% faster than with row units; at 4x a thread resolving
        // 16 pixels alone lengthens sparse tiles (-3.5% on 16-colour frames).
        const uint16_t* cell_table = reinterpret_cast<const uint16_t*>(mask_table + kKeys * MW);
        const int nlist = s_nlist;
        for (int i = t; i < nlist; i += 256) {
            const int own = s_list[i];
            const int ux = own % RTX, uy = own / RTX;
            uint32_t m[9], c[9];
#pragma unroll
            for (int j = 0; j < 9; j++) {
                const int k = (uy + 2 - j / 3) * RCX + (ux + 2 - j % 3);
                const uint32_t key = s_key[k];
                m[j] = key != kNoCell ? cell_table[key * 9 + j] : 0u;
                c[j] = s_col[k];
            }
            uint32_t col[S * S];
#pragma unroll
            for (int p = 0; p < S * S; p++) col[p] = 0xFF000000u;
            uint32_t assigned = 0;
#pragma unroll
            for (int j = 0; j < 9; j++) {
                const uint32_t nb = m[j] & ~assigned;
#pragma unroll
                for (int p = 0; p < S * S; p++) col[p] = ((nb >> p) & 1u) ? c[j] : col[p];
                assigned |= nb;
            }
            if (i < PX_CELLS) {
#pragma unroll
                for (int p = 0; p < S * S; p++) s_px[i * S * S + p] = col[p];
            } else {
#pragma unroll
                for (int ky = 0; ky < S; ky++) store_row(x0 + ux, (y0 + uy) * S + ky, col + ky * S);
            }
        }
        __syncthreads();
    } else if (nunits > 0) {  // uniform: read from LDS after the barrier
        for (int u = t; u < nunits; u += 256) {
            uint32_t rows[9];
            fetch_rows(u, rows);
            resolve_unit(u, rows);
        }
        __syncthreads();
    }

    // At S = 1, 2 and 4 one output row of a cell is one store and the lanes
    // of a wave store adjacent cells, so each thread stores its own cell.  At
    // the other scales that would make every store instruction strided (S * 4
    // bytes between lanes, each line written in pieces by several
    // instructions), so the tile's output is stored in chunks of CW pixels
    // inside one cell, chunk c by thread c % 256: adjacent lanes store
    // adjacent chunks of a row.  Searched cells beyond the pixel buffer were
    // stored by the search itself; their chunks are skipped.  Measured: at 5x
    // to 7x, with chunks of fewer than 4 pixels, the chunk loop's own work
    // outweighs the coalescing once a tile's cells overflow the colour buffer
    // (dense frames, sprites), so such tiles keep the per-cell stores; at 8x
    // and from 9x up (rows of 36 bytes and more) chunks win on every tile.
    constexpr int CW = (S % 4 == 0) ? 4 : ((S % 2 == 0) ? 2 : 1);
    // pixel (kx, ky) of listed cell i, own cell (lx, ly), from the buffer;
    // use_codes (std::true_type: the tile keeps candidate codes) is fixed per
    // tile, outside the store loops
    auto buffered = [&](auto use_codes, int i, int kx, int ky, int lx, int ly) -> uint32_t {
        if constexpr (decltype(use_codes)::value) {
            const int code = s_px8[(i * S + ky) * S + kx];
            return code == 9 ? 0xFF000000u : s_col[(ly + 2 - code / 3) * RCX + (lx + 2 - code % 3)];
        } else {
            return s_px[(i * S + ky) * S + kx];
        }
    };
    bool chunked = false, wide = false;
    if constexpr (S != 1 && S != 2 && S != 4) {
        const uintptr_t base = reinterpret_cast<uintptr_t>(out + it.out_off);
        chunked = ((base | (uintptr_t)it.out_stride) & (4 * CW - 1)) == 0 &&
                  (CW == 4 || S >= 9 || s_nlist <= PX_CELLS);
        // odd scales: chunks of 4 pixels across cell boundaries when the rows
        // are 16-byte aligned and every listed cell is buffered (no pixel of
        // the tile was stored by the search)
        if constexpr (CW == 1) wide = ((base | (uintptr_t)it.out_stride) & 15) == 0 && s_nlist <= px_cells;
    }
    // Chunks of 4 pixels of a tile row (16-byte stores) that may span two
    // cells (S >= 3): each pixel takes its colour from its own cell, simple
    // or buffered.  Measured against one-pixel chunks: 3x 14% and 5x 24%
    // faster (7x: 8% before the path for two simple cells); against the
    // two-pixel chunks of 6x, 6-12% slower (the per-pixel lookups cost more
    // than the wider stores save).
    auto store_wide = [&](auto use_codes) {
        constexpr int CPR = RTX * S / 4;   // chunks per output row of the tile
        constexpr int NCH = RTY * S * CPR;
        static_assert(RTX * S % 4 == 0, "whole chunks per tile row");
        const int ncx = min(RTX, it.W - x0), rowpx = ncx * S;  // cells and pixels of a tile row in the frame
        for (int c = t; c < NCH; c += 256) {
            const int ry = c / CPR, px0 = 4 * (c % CPR);
            const int ly = ry / S, ky = ry % S;
            if (px0 >= rowpx || y0 + ly >= it.H) continue;
            // the chunk's cells; past the frame's last column the pixels are
            // not stored and take that column's cell
            const int la = px0 / S, lb = min((px0 + 3) / S, ncx - 1);
            const int oa = ly * RTX + la, ob = ly * RTX + lb;
            const bool sa = s_simple[oa], sb = s_simple[ob];
            const int ia = sa ? 0 : s_idx[oa], ib = sb ? 0 : s_idx[ob];  // simple: entry 0, ignored
            const uint32_t va = s_col[(ly + 1) * RCX + la + 1], vb = s_col[(ly + 1) * RCX + lb + 1];
            uint32_t col[4];
            if (sa && sb) {  // both cells simple (most chunks): no buffer reads
#pragma unroll
                for (int q = 0; q < 4; q++) col[q] = px0 + q >= (la + 1) * S ? vb : va;
            } else {
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int p = px0 + q;
                    const bool inb = p >= (la + 1) * S;  // pixel in the second cell
                    const int lx = inb ? lb : la, i = inb ? ib : ia;
                    const int kx = min(p - lx * S, S - 1);  // (past the frame: not stored)
                    const uint32_t bv = buffered(use_codes, i, kx, ky, lx, ly);
                    col[q] = (inb ? sb : sa) ? (inb ? vb : va) : bv;
                }
            }
            const int Y = (y0 + ly) * S + ky;
            const int row = out_topdown(it.flags) ? (OH - 1 - Y) : Y;
            uint8_t* dst = out + it.out_off + (size_t)row * it.out_stride + ((size_t)x0 * S + px0) * 4;
            if (px0 + 3 < rowpx) {
                u32x4 v = { col[0], col[1], col[2], col[3] };
                __builtin_nontemporal_store(v, reinterpret_cast<u32x4*>(dst));
            } else {
#pragma unroll
                for (int q = 0; q < 4; q++)
                    if (px0 + q < rowpx) __builtin_nontemporal_store(col[q], reinterpret_cast<uint32_t*>(dst + 4 * q));
            }
        }
    };
    auto store_chunks = [&](auto use_codes) {
        constexpr int CPC = S / CW;          // chunks per cell row
        constexpr int CPR = RTX * CPC;       // chunks per output row of the tile
        constexpr int NCH = RTY * S * CPR;   // chunks of the tile
        for (int c = t; c < NCH; c += 256) {
            const int ry = c / CPR, rc = c % CPR;
            const int lx = rc / CPC, kx = (rc % CPC) * CW;
            const int ly = ry / S, ky = ry % S;
            if (x0 + lx >= it.W || y0 + ly >= it.H) continue;
            const int own = ly * RTX + lx;
            uint32_t col[CW];
            if (s_simple[own]) {
                const uint32_t v = s_col[(ly + 1) * RCX + lx + 1];
#pragma unroll
                for (int q = 0; q < CW; q++) col[q] = v;
            } else {
                const int i = s_idx[own];
                if (i >= px_cells) continue;  // stored by the search
#pragma unroll
                for (int q = 0; q < CW; q++) col[q] = buffered(use_codes, i, kx + q, ky, lx, ly);
            }
            const int Y = (y0 + ly) * S + ky;
            const int row = out_topdown(it.flags) ? (OH - 1 - Y) : Y;
            uint8_t* dst = out + it.out_off + (size_t)row * it.out_stride + ((size_t)(x0 + lx) * S + kx) * 4;
            if constexpr (CW == 4) {
                u32x4 v = { col[0], col[1], col[2], col[3] };
                __builtin_nontemporal_store(v, reinterpret_cast<u32x4*>(dst));
            } else if constexpr (CW == 2) {
                u32x2 v = { col[0], col[1] };
                __builtin_nontemporal_store(v, reinterpret_cast<u32x2*>(dst));
            } else {
                __builtin_nontemporal_store(col[0], reinterpret_cast<uint32_t*>(dst));
            }
        }
    };
    // A thread's own cell, simple or buffered, through one sequence of store
    // instructions for the whole wave: with separate branches for the two
    // kinds, a wave holding both issues every row store twice, each writing
    // part of the same lines.
    auto store_cell = [&](auto use_codes, int i) {
        const uint32_t c = s_col[kown];
#pragma unroll
        for (int ky = 0; ky < S; ky++) {
            uint32_t col[S];
#pragma unroll
            for (int q = 0; q < S; q++) {
                const uint32_t b = buffered(use_codes, i, q, ky, cx, cy);
                col[q] = simple ? c : b;
            }
            store_row(x0 + cx, (y0 + cy) * S + ky, col);
        }
    };
    if (wide) {
        if constexpr (CODES) {
            if (codes) {
                store_wide(std::true_type{});
                return;
            }
        }
        store_wide(std::false_type{});
        return;
    }
    if (chunked) {
        if constexpr (CODES) {
            if (codes) {
                store_chunks(std::true_type{});
                return;
            }
        }
        store_chunks(std::false_type{});
        return;
    }
    if (mine) {
        const int i = simple ? 0 : s_idx[t];  // a simple cell reads entry 0 and ignores it
        if (simple || i < px_cells) {         // listed cells beyond the buffer: stored by the search
            if constexpr (CODES) {
                if (codes) {
                    store_cell(std::true_type{}, i);
                    return;
                }
            }
            store_cell(std::false_type{}, i);
        }
    }
}

using RasterFn = void (*)(const DevItem*, const uint2*, int, const uint8_t*, const uint8_t*, uint8_t*,
                          const uint32_t*, const uint32_t*, const uint16_t*, uint8_t*);
using SmoothFn = void (*)(const DevItem*, const uint2*, int, const uint8_t*, const uint8_t*, uint8_t*,
                          const int8_t*, uint32_t*);

// The raster kernel of scale S (1..15) with (SUB) or without subdivision.
template <bool SUB>
RasterFn pick_raster_of(int S)
{
    switch (S) {
        case 1: return kf_raster<1, SUB>;
        case 2: return kf_raster<2, SUB>;
        case 3: return kf_raster<3, SUB>;
        case 4: return kf_raster<4, SUB>;
        case 5: return kf_raster<5, SUB>;
        case 6: return kf_raster<6, SUB>;
        case 7: return kf_raster<7, SUB>;
        case 8: return kf_raster<8, SUB>;
        case 9: return kf_raster<9, SUB>;
        case 10: return kf_raster<10, SUB>;
        case 11: return kf_raster<11, SUB>;
        case 12: return kf_raster<12, SUB>;
        case 13: return kf_raster<13, SUB>;
        case 14: return kf_raster<14, SUB>;
        case 15: return kf_raster<15, SUB>;
    }
    return nullptr;
}

RasterFn pick_raster(int S, bool sub) { return sub ? pick_raster_of<true>(S) : pick_raster_of<false>(S); }

// The smoothing kernel of scale S (1..15).
SmoothFn pick_smooth(int S)
{
    switch (S) {
        case 1: return kf_smooth<1>;
        case 2: return kf_smooth<2>;
        case 3: return kf_smooth<3>;
        case 4: return kf_smooth<4>;
        case 5: return kf_smooth<5>;
        case 6: return kf_smooth<6>;
        case 7: return kf_smooth<7>;
        case 8: return kf_smooth<8>;
        case 9: return kf_smooth<9>;
        case 10: return kf_smooth<10>;
        case 11: return kf_smooth<11>;
        case 12: return kf_smooth<12>;
        case 13: return kf_smooth<13>;
        case 14: return kf_smooth<14>;
        case 15: return kf_smooth<15>;
    }
    return nullptr;
}

}  // namespace dpx


// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
// The launch plan of the last batch.  A batch whose items and output
// alignment equal the last one's reuses it (no validation, grouping, tile
// maps or descriptor comparison): for thousands of sprites that host work
// takes longer than the kernels.
struct PlanGroup {
    int scale;
    bool sub;
    dim3 grid;
    int grid3d;
    size_t off_items, off_map;  // in the descriptor buffer
};
struct BatchPlan {
    bool valid = false;
    std::vector<dpx_item> items;
    uintptr_t out_mod = 0;  // d_out % 16
    dim3 ggrid;
    int g3d = 0;
    size_t off_items = 0, off_gmap = 0;
    std::vector<PlanGroup> groups;
};

struct dpx_context {
    BatchPlan plan;
    int device = 0;
    int8_t* d_hull = nullptr;
    uint32_t* d_mask[dpx::kMaxScale + 1] = {};
    uint16_t* d_info[dpx::kMaxScale + 1] = {};
    uint8_t *d_gaux = nullptr, *d_gfin = nullptr;
    size_t g_cap = 0;
    uint32_t* d_rows = nullptr;  // cell rows of smoothed cells (kf_smooth)
    size_t rows_cap = 0;
    uint8_t* d_desc = nullptr;
    size_t desc_cap = 0;
    uint8_t* h_pinned = nullptr;
    size_t pinned_cap = 0;
    hipEvent_t upload_done = nullptr;
    bool upload_pending = false;
    std::vector<uint8_t> last_desc;
    std::vector<uint64_t> g_offsets;
    bool keep_graph = false;  // write the resolved graphs (dpx_context_keep_graph)
    bool last_kept = false;   // the last batch wrote them
    std::mutex mu;
    // kernel timing: 3 events per recorded batch (start, after graph, end)
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    dpx_batch_times acc{};
};

namespace {

using namespace dpx;

int ensure_mask(dpx_context* ctx, int S)
{
    if (ctx->d_mask[S] && ctx->d_info[S]) return DPX_OK;
    size_t n = (size_t)kKeys * mask_w(S);
    const size_t cell_bytes = S <= kMaxCellScale ? (size_t)kKeys * 9 * sizeof(uint16_t) : 0;
    if (!ctx->d_mask[S]) DPX_HIP(hipMalloc(&ctx->d_mask[S], n * sizeof(uint32_t) + cell_bytes));
    if (!ctx->d_info[S]) DPX_HIP(hipMalloc(&ctx->d_info[S], kKeys * sizeof(uint16_t)));
    hipLaunchKernelGGL(kt_mask, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, ctx->d_hull, S, ctx->d_mask[S]);
    if (cell_bytes)
        hipLaunchKernelGGL(kt_cell, dim3((unsigned)((kKeys * 9 + 255) / 256)), dim3(256), 0, 0, ctx->d_mask[S], S,
                           reinterpret_cast<uint16_t*>(ctx->d_mask[S] + n));
    hipLaunchKernelGGL(kt_info, dim3(kKeys / 256), dim3(256), 0, 0, ctx->d_mask[S], S, ctx->d_info[S]);
    DPX_HIP(hipGetLastError());
    DPX_HIP(hipDeviceSynchronize());
    return DPX_OK;
}

struct Group {
    int scale;
    bool sub;
    std::vector<DevItem> dev;  // the group's items, in launch order
    std::vector<uint2> map;    // tile map of a 1-D launch
    dim3 grid;
    int grid3d = 0;
};

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Growth of the context's device buffers.  A buffer the device may still be
// reading is only replaced after the device is idle.
int grow_graph(dpx_context* ctx, size_t need)
{
    if (need <= ctx->g_cap) return DPX_OK;
    DPX_HIP(hipDeviceSynchronize());
    (void)hipFree(ctx->d_gaux);
    (void)hipFree(ctx->d_gfin);
    ctx->d_gaux = ctx->d_gfin = nullptr;
    ctx->g_cap = 0;
    size_t cap = align_up(need + need / 4, 1 << 20);
    DPX_HIP(hipMalloc(&ctx->d_gaux, cap));
    DPX_HIP(hipMalloc(&ctx->d_gfin, cap));
    ctx->g_cap = cap;
    return DPX_OK;
}

int grow_rows(dpx_context* ctx, size_t need)
{
    if (need <= ctx->rows_cap) return DPX_OK;
    DPX_HIP(hipDeviceSynchronize());
    (void)hipFree(ctx->d_rows);
    ctx->d_rows = nullptr;
    ctx->rows_cap = 0;
    size_t cap = align_up(need + need / 4, 1 << 20);
    DPX_HIP(hipMalloc(&ctx->d_rows, cap));
    ctx->rows_cap = cap;
    return DPX_OK;
}

int grow_desc(dpx_context* ctx, size_t need)
{
    if (need <= ctx->desc_cap) return DPX_OK;
    if (ctx->upload_pending) {
        DPX_HIP(hipEventSynchronize(ctx->upload_done));
        ctx->upload_pending = false;
    }
    DPX_HIP(hipDeviceSynchronize());
    (void)hipFree(ctx->d_desc);
    ctx->d_desc = nullptr;
    if (ctx->h_pinned) (void)hipHostFree(ctx->h_pinned);
    ctx->h_pinned = nullptr;
    ctx->desc_cap = ctx->pinned_cap = 0;
    ctx->last_desc.clear();
    ctx->plan.valid = false;  // its descriptors are gone with the buffer
    size_t cap = align_up(need * 2, 1 << 16);
    DPX_HIP(hipMalloc(&ctx->d_desc, cap));
    DPX_HIP(hipHostMalloc(&ctx->h_pinned, cap, hipHostMallocDefault));
    ctx->desc_cap = cap;
    ctx->pinned_cap = cap;
    return DPX_OK;
}

// Descriptor bytes of a batch of n items whose items share a tile grid: the
// items twice (all, and per group), each array 64-byte aligned, with room for
// one group per (scale, subdivision) pair.  (Batches of mixed sizes add tile
// maps and may grow the buffer.)
size_t desc_bytes(size_t n)
{
    return 2 * n * sizeof(DevItem) + 64 * (2 * 2 * (kMaxScale + 1) + 4);
}

// Tiles per item when all items of a prefix have the same count, else 0.
// Launch grid of a kernel with TX x TY tiles over `dev`: (tile columns, tile
// rows, items) when all items have the same tile grid, else 1-D over all
// tiles with a map from block to (item, tile row << 16 | tile column).
// *grid3d tells the kernel which.  Tile columns and rows fit 16 bits: items
// are at most 32768 pixels wide and high.
template <int TX, int TY>
dim3 tile_grid(const std::vector<DevItem>& dev, int* grid3d, std::vector<uint2>* map)
{
    const unsigned tx0 = (unsigned)(dev[0].W + TX - 1) / TX, ty0 = (unsigned)(dev[0].H + TY - 1) / TY;
    bool same = dev.size() <= 65535;
    for (const DevItem& d : dev)
        same = same && (unsigned)(d.W + TX - 1) / TX == tx0 && (unsigned)(d.H + TY - 1) / TY == ty0;
    map->clear();
    *grid3d = same ? 1 : 0;
    if (same) return dim3(tx0, ty0, (unsigned)dev.size());
    for (size_t k = 0; k < dev.size(); k++) {
        const unsigned tx = (unsigned)(dev[k].W + TX - 1) / TX, ty = (unsigned)(dev[k].H + TY - 1) / TY;
        for (unsigned y = 0; y < ty; y++)
            for (unsigned x = 0; x < tx; x++) map->push_back(make_uint2((unsigned)k, y << 16 | x));
    }
    return dim3((unsigned)map->size(), 1, 1);
}

}  // namespace

extern "C" int dpx_context_create(int device, dpx_context** out)
{
    DPX_CHECK(out, DPX_EINVAL, "null output pointer");
    *out = nullptr;
    if (int rc = require_device()) return rc;
    DPX_HIP(hipSetDevice(device));
    dpx_context* ctx = new dpx_context();
    ctx->device = device;
    auto cleanup = [&](int rc) {
        dpx_context_destroy(ctx);
        return rc;
    };
    if (hipMalloc(&ctx->d_hull, (size_t)kKeys * kHullBytes) != hipSuccess)
        return cleanup(fail(DPX_ENOMEM, "table allocation failed"));
    hipLaunchKernelGGL(kt_hull, dim3(kKeys / 256), dim3(256), 0, 0, ctx->d_hull);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
        return cleanup(fail(DPX_EHIP, "cell table build failed"));
    if (hipEventCreateWithFlags(&ctx->upload_done, hipEventDisableTiming) != hipSuccess)
        return cleanup(fail(DPX_EHIP, "event creation failed"));
    *out = ctx;
    return DPX_OK;
}

extern "C" void dpx_context_destroy(dpx_context* ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    (void)hipFree(ctx->d_hull);
    for (auto* m : ctx->d_mask) (void)hipFree(m);
    for (auto* m : ctx->d_info) (void)hipFree(m);
    (void)hipFree(ctx->d_gaux);
    (void)hipFree(ctx->d_gfin);
    (void)hipFree(ctx->d_rows);
    (void)hipFree(ctx->d_desc);
    if (ctx->h_pinned) (void)hipHostFree(ctx->h_pinned);
    if (ctx->upload_done) (void)hipEventDestroy(ctx->upload_done);
    for (auto e : ctx->ev_pool) (void)hipEventDestroy(e);
    delete ctx;
}

extern "C" int dpx_context_reserve(dpx_context* ctx, int max_items, int64_t max_item_pixels, int scale)
{
    DPX_CHECK(ctx && max_items > 0 && max_item_pixels > 0, DPX_EINVAL, "bad arguments");
    DPX_CHECK(scale >= 0 && scale <= kMaxScale, DPX_EINVAL, "scale %d outside 0..%d", scale, kMaxScale);
    std::lock_guard<std::mutex> lock(ctx->mu);
    DPX_HIP(hipSetDevice(ctx->device));
    if (int rc = grow_graph(ctx, (size_t)max_items * align_up((size_t)max_item_pixels, 256))) return rc;
    if (int rc = grow_desc(ctx, desc_bytes((size_t)max_items))) return rc;
    if (scale > 0)
        if (int rc = ensure_mask(ctx, scale)) return rc;
    return DPX_OK;
}

namespace dpx {
int context_reserve_rows(dpx_context* ctx, size_t bytes)
{
    std::lock_guard<std::mutex> lock(ctx->mu);
    DPX_HIP(hipSetDevice(ctx->device));
    return grow_rows(ctx, bytes);
}
}  // namespace dpx

extern "C" int dpx_context_keep_graph(dpx_context* ctx, int enable)
{
    DPX_CHECK(ctx, DPX_EINVAL, "null context");
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->keep_graph = enable != 0;
    return DPX_OK;
}

extern "C" int dpx_batch_graph(dpx_context* ctx, const uint8_t** d_graph, const uint64_t** offsets_host)
{
    DPX_CHECK(ctx && d_graph && offsets_host, DPX_EINVAL, "bad arguments");
    DPX_CHECK(ctx->last_kept, DPX_EINVAL, "the last batch wrote no graphs (see dpx_context_keep_graph)");
    *d_graph = ctx->d_gfin;
    *offsets_host = ctx->g_offsets.data();
    return DPX_OK;
}

// Validates item k and fills its device descriptor; its graph bytes start
// at g_off of the graph scratch.
static int make_dev_item(const dpx_item& a, int k, const uint8_t* d_out, uint64_t g_off, DevItem& d)
{
    using namespace dpx;
    DPX_CHECK(a.width > 0 && a.height > 0 && a.width <= 32768 && a.height <= 32768, DPX_EINVAL,
              "item %d: bad size %dx%d", k, a.width, a.height);
    DPX_CHECK((a.flags & ~kKnownFlags) == 0, DPX_EINVAL, "item %d: unknown flags 0x%x", k, a.flags & ~kKnownFlags);
    const bool indexed = a.flags & DPX_INDEXED;
    DPX_CHECK(indexed ? a.channels == 1 : (a.channels == 3 || a.channels == 4), DPX_EINVAL, "item %d: channels %d%s",
              k, a.channels, indexed ? " (indexed frames have 1)" : "");
    DPX_CHECK(!indexed || a.palette_offset % 4 == 0, DPX_EINVAL, "item %d: palette offset %llu not a multiple of 4", k,
              (unsigned long long)a.palette_offset);
    DPX_CHECK(a.stride >= a.width * a.channels, DPX_EINVAL, "item %d: stride %d", k, a.stride);
    DPX_CHECK((int64_t)a.stride * a.height < (int64_t)1 << 31, DPX_EINVAL, "item %d: input larger than 2 GiB", k);
    DPX_CHECK(a.scale >= 1 && a.scale <= kMaxScale, DPX_EINVAL, "item %d: scale %d outside 1..%d", k, a.scale,
              kMaxScale);
    DPX_CHECK(a.out_stride >= 4 * a.width * a.scale && a.out_stride % 4 == 0 && a.out_offset % 4 == 0, DPX_EINVAL,
              "item %d: output layout (stride %d, offset %llu)", k, a.out_stride, (unsigned long long)a.out_offset);
    DPX_CHECK(!a.subdivide || (uint64_t)a.width * a.height * rows_words(a.scale) < ((uint64_t)1 << 32), DPX_EINVAL,
              "item %d: %dx%d at scale %d with subdivision exceeds 2^32 cell-row words", k, a.width, a.height, a.scale);
    memset(&d, 0, sizeof(d));
    d.in_off = a.in_offset;
    d.out_off = a.out_offset;
    d.g_off = g_off;
    d.W = a.width;
    d.H = a.height;
    d.stride = a.stride;
    d.ch = a.channels;
    d.flags = a.flags;
    d.pal_off = indexed ? a.palette_offset : 0;
    d.out_stride = a.out_stride;
    d.scale = a.scale;
    d.subdivide = a.subdivide ? 1 : 0;
    d.aligned16 = (a.out_stride % 16 == 0) && (a.out_offset % 16 == 0) && (reinterpret_cast<uintptr_t>(d_out) % 16 == 0);
    d.graph_w4 = a.width % 4 == 0;
    return DPX_OK;
}

// Validates the items and builds a launch plan: device items, graph tiles,
// raster groups by (scale, subdivision), the descriptor blob (offsets in the
// plan refer to it) and the graph scratch bytes the batch needs.
static int build_plan(dpx_context* ctx, const uint8_t* d_out, const dpx_item* items, int n_items, BatchPlan& plan,
                      std::vector<uint8_t>& blob, std::vector<uint64_t>* g_offsets, size_t* graph_bytes,
                      size_t* rows_bytes)
{
    using namespace dpx;
    plan.valid = false;
    std::vector<DevItem> dev(n_items);
    if (g_offsets) g_offsets->assign(n_items, 0);
    std::map<int, Group> groups;
    uint64_t goff = 0, roff = 0;
    for (int k = 0; k < n_items; k++) {
        const dpx_item& a = items[k];
        DevItem& d = dev[k];
        if (int rc = make_dev_item(a, k, d_out, goff, d)) return rc;
        if (g_offsets) (*g_offsets)[k] = goff;
        goff += align_up((size_t)a.width * a.height, 256);
        if (d.subdivide) {
            d.rows_off = roff;
            roff += align_up((size_t)a.width * a.height * rows_words(a.scale) * 4, 256);
        }
        int gkey = a.scale * 2 + d.subdivide;
        Group& grp = groups[gkey];
        grp.scale = a.scale;
        grp.sub = d.subdivide;
        grp.dev.push_back(d);
    }
    for (auto& kv : groups)
        if (int rc = ensure_mask(ctx, kv.second.scale)) return rc;

    // launch grids
    int g3d = 0;
    std::vector<uint2> gmap;
    const dim3 ggrid = tile_grid<GTX, GTY>(dev, &g3d, &gmap);
    for (auto& kv : groups) kv.second.grid = tile_grid<RTX, RTY>(kv.second.dev, &kv.second.grid3d, &kv.second.map);

    // descriptor blob: items | graph tile map | per group (items, tile map)
    blob.clear();
    auto put = [&](const void* p, size_t n) {
        size_t off = align_up(blob.size(), 64);
        blob.resize(off + n);
        if (n) memcpy(blob.data() + off, p, n);
        return off;
    };
    plan.off_items = put(dev.data(), dev.size() * sizeof(DevItem));
    plan.off_gmap = put(gmap.data(), gmap.size() * sizeof(uint2));
    plan.ggrid = ggrid;
    plan.g3d = g3d;
    plan.groups.clear();
    for (auto& kv : groups) {
        const Group& grp = kv.second;
        size_t a = put(grp.dev.data(), grp.dev.size() * sizeof(DevItem));
        size_t b = put(grp.map.data(), grp.map.size() * sizeof(uint2));
        plan.groups.push_back(PlanGroup{ grp.scale, grp.sub, grp.grid, grp.grid3d, a, b });
    }
    plan.items.assign(items, items + n_items);
    plan.out_mod = reinterpret_cast<uintptr_t>(d_out) % 16;
    *graph_bytes = goff;
    *rows_bytes = roff;
    plan.valid = true;
    return DPX_OK;
}

// The context's own plan for dpx_batch_run: built into the context's
// descriptor buffer (uploaded on `stream` when it differs from the one
// already on the device) and its graph scratch.
static int make_plan(dpx_context* ctx, const uint8_t* d_out, const dpx_item* items, int n_items, hipStream_t stream)
{
    BatchPlan& plan = ctx->plan;
    std::vector<uint8_t> blob;
    size_t gbytes = 0, rbytes = 0;
    if (int rc = build_plan(ctx, d_out, items, n_items, plan, blob, &ctx->g_offsets, &gbytes, &rbytes)) return rc;
    plan.valid = false;
    // graph and cell-row arenas (grown only when a batch needs more)
    if (int rc = grow_graph(ctx, gbytes)) return rc;
    if (int rc = grow_rows(ctx, rbytes)) return rc;
    if (blob != ctx->last_desc) {
        if (ctx->upload_pending) {
            DPX_HIP(hipEventSynchronize(ctx->upload_done));
            ctx->upload_pending = false;
        }
        if (int rc = grow_desc(ctx, blob.size())) return rc;
        memcpy(ctx->h_pinned, blob.data(), blob.size());
        DPX_HIP(hipMemcpyAsync(ctx->d_desc, ctx->h_pinned, blob.size(), hipMemcpyHostToDevice, stream));
        DPX_HIP(hipEventRecord(ctx->upload_done, stream));
        ctx->upload_pending = true;
        ctx->last_desc.swap(blob);
    }
    plan.valid = true;
    return DPX_OK;
}

// The kernels of one batch: kf_graph, then one kf_raster per (scale,
// subdivision) group.  `desc` holds the plan's descriptor blob.
struct LaunchArgs {
    const uint8_t* desc;
    uint8_t* gaux;
    uint8_t* gfin;  // the resolved graph (written by kf_smooth for subdivided items)
    bool keep;      // the raster of the other items writes the resolved graph too
    uint32_t* rows;  // cell rows of smoothed cells
};

static int launch_plan(dpx_context* ctx, const BatchPlan& plan, const LaunchArgs& la, const uint8_t* d_in,
                       uint8_t* d_out, hipStream_t stream, hipEvent_t ev_mid)
{
    const DevItem* d_items = reinterpret_cast<const DevItem*>(la.desc + plan.off_items);
    const uint2* d_gmap = reinterpret_cast<const uint2*>(la.desc + plan.off_gmap);
    hipLaunchKernelGGL(kf_graph, plan.ggrid, dim3(256), 0, stream, d_items, d_gmap, plan.g3d, d_in, la.gaux);
    if (ev_mid) DPX_HIP(hipEventRecord(ev_mid, stream));
    for (const PlanGroup& pg : plan.groups) {
        RasterFn fn = pick_raster(pg.scale, pg.sub);
        DPX_CHECK(fn, DPX_EINVAL, "no raster kernel for scale %d", pg.scale);
        const DevItem* d_gitems = reinterpret_cast<const DevItem*>(la.desc + pg.off_items);
        const uint2* d_rmap = reinterpret_cast<const uint2*>(la.desc + pg.off_map);
        if (pg.sub) {
            SmoothFn sf = pick_smooth(pg.scale);
            DPX_CHECK(sf && la.gfin && la.rows, DPX_EINVAL, "no smoothing buffers for scale %d", pg.scale);
            hipLaunchKernelGGL(sf, pg.grid, dim3(256), 0, stream, d_gitems, d_rmap, pg.grid3d, d_in, la.gaux, la.gfin,
                               ctx->d_hull, la.rows);
        }
        hipLaunchKernelGGL(fn, pg.grid, dim3(256), 0, stream, d_gitems, d_rmap, pg.grid3d, d_in, la.gaux,
                           (pg.sub || la.keep) ? la.gfin : nullptr, la.rows, ctx->d_mask[pg.scale],
                           ctx->d_info[pg.scale], d_out);
    }
    DPX_HIP(hipGetLastError());
    return DPX_OK;
}

extern "C" int dpx_batch_run(dpx_context* ctx, const uint8_t* d_in, uint8_t* d_out, const dpx_item* items,
                             int n_items, void* stream_)
{
    DPX_CHECK(ctx && d_in && d_out && items && n_items > 0, DPX_EINVAL, "bad arguments");
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipStream_t stream = as_stream(stream_);
    DPX_HIP(hipSetDevice(ctx->device));
    const BatchPlan& plan = ctx->plan;
    if (!(plan.valid && plan.items.size() == (size_t)n_items &&
          plan.out_mod == reinterpret_cast<uintptr_t>(d_out) % 16 &&
          memcmp(plan.items.data(), items, (size_t)n_items * sizeof(dpx_item)) == 0))
        if (int rc = make_plan(ctx, d_out, items, n_items, stream)) return rc;

    hipEvent_t ev[3] = {};
    if (ctx->timing) {
        while (ctx->ev_pool.size() < ctx->ev_used + 3) {
            hipEvent_t e;
            DPX_HIP(hipEventCreate(&e));
            ctx->ev_pool.push_back(e);
        }
        for (int q = 0; q < 3; q++) ev[q] = ctx->ev_pool[ctx->ev_used + q];
        ctx->ev_used += 3;
        DPX_HIP(hipEventRecord(ev[0], stream));
    }
    LaunchArgs la{ ctx->d_desc, ctx->d_gaux, ctx->d_gfin, ctx->keep_graph, ctx->d_rows };
    if (int rc = launch_plan(ctx, plan, la, d_in, d_out, stream, ev[1])) return rc;
    ctx->last_kept = ctx->keep_graph;
    if (ctx->timing) DPX_HIP(hipEventRecord(ev[2], stream));
    return DPX_OK;
}

// ---------------------------------------------------------------------------
// plans: a batch layout fixed once, with descriptors and graph scratch of its
// own, launched directly or as a HIP graph the library builds
// ---------------------------------------------------------------------------
struct dpx_plan {
    dpx_context* ctx = nullptr;
    BatchPlan plan;
    uint8_t* d_desc = nullptr;
    uint8_t* d_gaux = nullptr;
    uint8_t* d_gfin = nullptr;
    uint32_t* d_rows = nullptr;
    uint32_t flags = 0;
    std::mutex mu;
    // DPX_PLAN_GRAPH: the kernels as graph nodes (kf_graph, then per group
    // kf_smooth when it subdivides and kf_raster); the input and output
    // pointers of the instantiated graph are updated in place when they change
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipGraphNode_t> nodes;
    std::vector<std::pair<int, int>> node_kind;  // (0 graph / 1 smooth / 2 raster, group)
    const uint8_t* g_in = nullptr;
    uint8_t* g_out = nullptr;
    std::vector<uint64_t> g_offsets;  // per item, in d_gfin
};

namespace {

void plan_free(dpx_plan* p)
{
    if (!p) return;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    (void)hipFree(p->d_desc);
    (void)hipFree(p->d_gaux);
    (void)hipFree(p->d_gfin);
    (void)hipFree(p->d_rows);
    delete p;
}

// Kernel node parameters of one of the plan's kernels for (d_in, d_out);
// `a` keeps the argument values the parameter pointers refer to.
struct NodeArgs {
    const DevItem* items;
    const uint2* map;
    int g3d;
    const uint8_t* in;
    const uint8_t* gaux_c;
    uint8_t* gaux;
    uint8_t* gfin;
    const int8_t* hull;
    const uint32_t* rows_c;
    uint32_t* rows;
    const uint32_t* mask;
    const uint16_t* info;
    uint8_t* out;
    void* ptrs[10];
};

hipKernelNodeParams node_params(dpx_plan* p, int kind, int group, const uint8_t* d_in, uint8_t* d_out, NodeArgs& a)
{
    hipKernelNodeParams np;
    memset(&np, 0, sizeof(np));
    np.blockDim = dim3(256);
    a.in = d_in;
    a.out = d_out;
    a.gaux = p->d_gaux;
    a.gaux_c = p->d_gaux;
    if (kind == 0) {
        a.items = reinterpret_cast<const DevItem*>(p->d_desc + p->plan.off_items);
        a.map = reinterpret_cast<const uint2*>(p->d_desc + p->plan.off_gmap);
        a.g3d = p->plan.g3d;
        void* v[5] = { &a.items, &a.map, &a.g3d, &a.in, &a.gaux };
        memcpy(a.ptrs, v, sizeof(v));
        np.func = reinterpret_cast<void*>(kf_graph);
        np.gridDim = p->plan.ggrid;
    } else {
        const PlanGroup& pg = p->plan.groups[group];
        a.items = reinterpret_cast<const DevItem*>(p->d_desc + pg.off_items);
        a.map = reinterpret_cast<const uint2*>(p->d_desc + pg.off_map);
        a.g3d = pg.grid3d;
        a.gfin = (pg.sub || (p->flags & DPX_PLAN_KEEP_GRAPH)) ? p->d_gfin : nullptr;
        a.rows = p->d_rows;
        a.rows_c = p->d_rows;
        if (kind == 1) {
            a.hull = p->ctx->d_hull;
            void* v[8] = { &a.items, &a.map, &a.g3d, &a.in, &a.gaux_c, &a.gfin, &a.hull, &a.rows };
            memcpy(a.ptrs, v, sizeof(v));
            np.func = reinterpret_cast<void*>(pick_smooth(pg.scale));
        } else {
            a.mask = p->ctx->d_mask[pg.scale];
            a.info = p->ctx->d_info[pg.scale];
            void* v[10] = { &a.items, &a.map, &a.g3d, &a.in, &a.gaux_c, &a.gfin, &a.rows_c, &a.mask, &a.info, &a.out };
            memcpy(a.ptrs, v, sizeof(v));
            np.func = reinterpret_cast<void*>(pick_raster(pg.scale, pg.sub));
        }
        np.gridDim = pg.grid;
    }
    np.kernelParams = a.ptrs;
    return np;
}

int plan_build_graph(dpx_plan* p, const uint8_t* d_in, uint8_t* d_out)
{
    DPX_HIP(hipGraphCreate(&p->graph, 0));
    p->node_kind.assign(1, { 0, 0 });
    for (int g = 0; g < (int)p->plan.groups.size(); g++) {
        if (p->plan.groups[g].sub) p->node_kind.push_back({ 1, g });
        p->node_kind.push_back({ 2, g });
    }
    p->nodes.assign(p->node_kind.size(), nullptr);
    for (size_t k = 0; k < p->nodes.size(); k++) {
        NodeArgs a;
        memset(&a, 0, sizeof(a));
        const int kind = p->node_kind[k].first;
        hipKernelNodeParams np = node_params(p, kind, p->node_kind[k].second, d_in, d_out, a);
        DPX_CHECK(np.func, DPX_EINVAL, "no kernel for a group of the plan");
        // kf_smooth and the rasters of groups without subdivision follow the
        // graph kernel; a group's raster after subdivision follows its kf_smooth
        hipGraphNode_t dep = (kind == 2 && k > 0 && p->node_kind[k - 1].first == 1) ? p->nodes[k - 1] : p->nodes[0];
        DPX_HIP(hipGraphAddKernelNode(&p->nodes[k], p->graph, k ? &dep : nullptr, k ? 1 : 0, &np));
    }
    DPX_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
    p->g_in = d_in;
    p->g_out = d_out;
    return DPX_OK;
}

}  // namespace

extern "C" int dpx_plan_create(dpx_context* ctx, const dpx_item* items, int n_items, uint64_t out_align,
                               uint32_t flags, dpx_plan** out)
{
    DPX_CHECK(out, DPX_EINVAL, "null output pointer");
    *out = nullptr;
    DPX_CHECK(ctx && items && n_items > 0, DPX_EINVAL, "bad arguments");
    DPX_CHECK((flags & ~(uint32_t)(DPX_PLAN_GRAPH | DPX_PLAN_KEEP_GRAPH)) == 0, DPX_EINVAL, "unknown plan flags 0x%x",
              flags);
    std::lock_guard<std::mutex> lock(ctx->mu);
    DPX_HIP(hipSetDevice(ctx->device));
    auto* p = new dpx_plan();
    p->ctx = ctx;
    p->flags = flags;
    std::vector<uint8_t> blob;
    size_t gbytes = 0, rbytes = 0;
    // out_align stands for d_out modulo 16 (only the alignment of the output
    // arena enters the plan)
    const uint8_t* fake_out = reinterpret_cast<const uint8_t*>((uintptr_t)(out_align % 16) + 16);
    int rc = build_plan(ctx, fake_out, items, n_items, p->plan, blob, nullptr, &gbytes, &rbytes);
    const bool want_gfin = (flags & DPX_PLAN_KEEP_GRAPH) || rbytes > 0;
    if (rc == DPX_OK && (hipMalloc(&p->d_desc, blob.size()) != hipSuccess ||
                         hipMalloc(&p->d_gaux, std::max<size_t>(gbytes, 256)) != hipSuccess ||
                         (want_gfin && hipMalloc(&p->d_gfin, std::max<size_t>(gbytes, 256)) != hipSuccess) ||
                         (rbytes > 0 && hipMalloc(&p->d_rows, rbytes) != hipSuccess)))
        rc = fail(DPX_ENOMEM, "plan allocation failed (%zu descriptor, %zu graph and %zu cell-row bytes)", blob.size(),
                  gbytes, rbytes);
    if (rc == DPX_OK && hipMemcpy(p->d_desc, blob.data(), blob.size(), hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(DPX_EHIP, "descriptor upload failed");
    if (rc != DPX_OK) {
        plan_free(p);
        return rc;
    }
    *out = p;
    return DPX_OK;
}

extern "C" int dpx_plan_run(dpx_plan* p, const uint8_t* d_in, uint8_t* d_out, void* stream_)
{
    DPX_CHECK(p && d_in && d_out, DPX_EINVAL, "bad arguments");
    DPX_CHECK(reinterpret_cast<uintptr_t>(d_out) % 16 == p->plan.out_mod, DPX_EINVAL,
              "d_out %% 16 = %u, the plan was made for %u", (unsigned)(reinterpret_cast<uintptr_t>(d_out) % 16),
              (unsigned)p->plan.out_mod);
    std::lock_guard<std::mutex> lock(p->mu);
    hipStream_t stream = as_stream(stream_);
    DPX_HIP(hipSetDevice(p->ctx->device));
    if (!(p->flags & DPX_PLAN_GRAPH)) {
        LaunchArgs la{ p->d_desc, p->d_gaux, p->d_gfin, (p->flags & DPX_PLAN_KEEP_GRAPH) != 0, p->d_rows };
        return launch_plan(p->ctx, p->plan, la, d_in, d_out, stream, nullptr);
    }
    if (!p->exec) {
        if (int rc = plan_build_graph(p, d_in, d_out)) return rc;
    } else if (p->g_in != d_in || p->g_out != d_out) {
        for (size_t k = 0; k < p->nodes.size(); k++) {
            NodeArgs a;
            memset(&a, 0, sizeof(a));
            hipKernelNodeParams np = node_params(p, p->node_kind[k].first, p->node_kind[k].second, d_in, d_out, a);
            DPX_HIP(hipGraphExecKernelNodeSetParams(p->exec, p->nodes[k], &np));
        }
        p->g_in = d_in;
        p->g_out = d_out;
    }
    DPX_HIP(hipGraphLaunch(p->exec, stream));
    return DPX_OK;
}

extern "C" int dpx_plan_graph(dpx_plan* p, const uint8_t** d_graph, const uint64_t** offsets_host)
{
    DPX_CHECK(p && d_graph && offsets_host, DPX_EINVAL, "bad arguments");
    DPX_CHECK(p->flags & DPX_PLAN_KEEP_GRAPH, DPX_EINVAL, "the plan was made without DPX_PLAN_KEEP_GRAPH");
    std::lock_guard<std::mutex> lock(p->mu);
    if (p->g_offsets.empty()) {
        uint64_t off = 0;
        for (const dpx_item& a : p->plan.items) {
            p->g_offsets.push_back(off);
            off += align_up((size_t)a.width * a.height, 256);
        }
    }
    *d_graph = p->d_gfin;
    *offsets_host = p->g_offsets.data();
    return DPX_OK;
}

extern "C" void dpx_plan_destroy(dpx_plan* p)
{
    if (!p) return;
    (void)hipSetDevice(p->ctx->device);
    (void)hipDeviceSynchronize();
    plan_free(p);
}

extern "C" int dpx_batch_timing(dpx_context* ctx, int enable)
{
    DPX_CHECK(ctx, DPX_EINVAL, "null context");
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->timing = enable != 0;
    return DPX_OK;
}

extern "C" int dpx_batch_timing_read(dpx_context* ctx, dpx_batch_times* out)
{
    DPX_CHECK(ctx && out, DPX_EINVAL, "bad arguments");
    std::lock_guard<std::mutex> lock(ctx->mu);
    for (size_t b = 0; b + 3 <= ctx->ev_used; b += 3) {
        hipEvent_t* ev = &ctx->ev_pool[b];
        DPX_HIP(hipEventSynchronize(ev[2]));
        float t01, t12;
        DPX_HIP(hipEventElapsedTime(&t01, ev[0], ev[1]));
        DPX_HIP(hipEventElapsedTime(&t12, ev[1], ev[2]));
        ctx->acc.graph_ms += t01;
        ctx->acc.raster_ms += t12;
        ctx->acc.batches++;
    }
    ctx->ev_used = 0;
    *out = ctx->acc;
    ctx->acc = dpx_batch_times{};
    return DPX_OK;
}
