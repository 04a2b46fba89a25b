// fused.hip -- the batched path: many frames or sprites of uneven size
// through two fused kernels.
//
//   kf_graph   tile of 64x32 pixels (64x8 in small launches): lane words (YUV fields + class) once
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
#include <dlfcn.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "batch.hpp"

namespace dpx {

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


// Cell masks (scales up to kMaxCellScale): entry j = (1 - dy) * 3 + (1 - dx)
// of key k holds, as bit ky * S + q, whether a cell of key k covers sample q
// of row ky of the cell at (-dx, -dy) from it, i.e. what k contributes when
// it is the candidate at (dx, dy) of that cell.  They follow the row table
// in the same allocation.
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


// ---------------------------------------------------------------------------
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
// Tiles are GTX x TY pixels: TY = GTY or kGraphTallTY, whichever converts
// fewer rows (tiles x (TY + 4): 56 for 224-row frames, 4 tiles of 60 rows
// against 7 of 36, 6% faster), or kGraphSmallTY for launches that would have
// fewer than kGraphSmallTiles blocks of GTY rows (a launch of a few frames
// then spreads over more CUs; the extra halo rows cost throughput only in
// launches that fill the machine).
constexpr int GTX = 64, GTY = 32, kGraphSmallTY = 8, kGraphTallTY = 56;
constexpr uint64_t kGraphSmallTiles = 512;
constexpr int GLX = GTX + 3;  // lane words per row
constexpr int GHX = GTX + 2;  // positions of pass 2 per row
constexpr int GHP = GTX + 8;  // byte pitch of pass 2's rows: column c is x = x0 - 4 + c

template <int TY>
__global__ __launch_bounds__(256) void kf_graph(const DevItem* __restrict__ items, const uint2* __restrict__ tile_map,
                                                int grid3d, const uint8_t* __restrict__ in,
                                                uint8_t* __restrict__ gaux)
{
    constexpr int GLY = TY + 4, GHY = TY + 2;  // rows of lane words and of pass 2's positions
    static_assert(GLY * (GLX - 64) <= 256 && GHY * (GHX - 64) <= 256, "the columns past 64 take one pass");
    __shared__ uint32_t s_lane[GLX * GLY];
    // pass 2's rows; first, for indexed items, the 256 palette lane words
    __shared__ uint32_t s_h32[GHY * GHP / 4 > 256 ? GHY * GHP / 4 : 256];
    uint8_t* s_h = reinterpret_cast<uint8_t*>(s_h32);
    const int t = threadIdx.x;
    // runs of 8 tiles per XCD: the lines that neighbouring tiles' rows share
    // (a 64-pixel row of 3-byte pixels and its halo span 2-3 lines of 128
    // bytes) are fetched once into one L2 (HBM reads 365 -> 183 MB for 1024
    // frames of 256x224, 1.04x the input; the kernel's time is its VALU work)
    const TileAt at = locate_tile<8>(tile_map, grid3d);
    const DevItem it = items[at.item];
    const int x0 = at.tx * GTX, y0 = at.ty * TY;

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
    // Lane word of pixel (x, y) at byte offset `off` of the item's buffer:
    // one 4-byte load (any alignment) where the 4 bytes lie inside the
    // item's rows, else byte loads (the last pixel of the last row in
    // memory of a 3-channel frame without row padding).
    const uint32_t in_end = (uint32_t)(it.H - 1) * (uint32_t)it.stride + (uint32_t)(it.W * it.ch);
    auto lane_word = [&](uint32_t off) -> uint32_t {
        if (indexed) return s_h32[base[off]];
        uint32_t w;
        if (off + 4u <= in_end)
            __builtin_memcpy(&w, base + off, 4);
        else
            w = base[off] | (uint32_t)base[off + 1] << 8 | (uint32_t)base[off + 2] << 16 |
                (it.ch == 4 ? (uint32_t)base[off + 3] << 24 : 0u);
        const int c0 = (w >> (8 * o0)) & 0xFF, c1 = (w >> 8) & 0xFF, c2 = (w >> (8 * o2)) & 0xFF;
        if (alpha && (w >> 24) == 0) return kLaneTransparent;
        return yuv_lanes(yuv_pack(c0, c1, c2)) | kLaneOpaque;
    };
    // Window column c (x = x0 - 1 + c), row r (y = y0 - 2 + r).  Columns
    // 0..63: one per thread (t % 64), every fourth row from t / 64, the byte
    // offset stepping by four rows; columns 64..GLX-1: one word per thread.
    // (Reading aligned rows as groups of four pixels, dwordx3 per thread and
    // four conversions each, measured slower: 0.153 -> 0.160-0.173 ms.)
    {
        const int c = t & 63, x = x0 - 1 + c;
        const bool xin = (unsigned)x < (unsigned)it.W;
        int y = y0 - 2 + (t >> 6);
        // modulo 2^32 (rows outside the frame may lie outside int); used only inside the frame
        uint32_t off = (uint32_t)(topdown ? it.H - 1 - y : y) * (uint32_t)it.stride + (uint32_t)(x * it.ch);
        const uint32_t step = (uint32_t)((topdown ? -4 : 4) * it.stride);
        for (int r = t >> 6; r < GLY; r += 4, y += 4, off += step) {
            uint32_t lw = 0;
            if (xin && (unsigned)y < (unsigned)it.H) lw = lane_word(off);
            s_lane[r * GLX + c] = lw;
        }
    }
    if (t < (GLX - 64) * GLY) {
        const int c = 64 + t % (GLX - 64), r = t / (GLX - 64);
        const int x = x0 - 1 + c, y = y0 - 2 + r;
        uint32_t lw = 0;
        if ((unsigned)x < (unsigned)it.W && (unsigned)y < (unsigned)it.H)
            lw = lane_word((uint32_t)((topdown ? it.H - 1 - y : y) * it.stride + x * it.ch));
        s_lane[r * GLX + c] = lw;
    }
    __syncthreads();
    // pass 2: position (x0 - 1 + cx, y0 - 1 + cy), same column split
    auto edges = [&](int cx, int cy) {
        const int c = (cy + 1) * GLX + cx;  // its lane word
        uint32_t up = s_lane[c + GLX], ur = s_lane[c + GLX + 1], rt = s_lane[c + 1], dr = s_lane[c - GLX + 1];
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
    for (int k = t; k < GTX / 4 * TY; k += 256) {
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

using GraphFn = void (*)(const DevItem*, const uint2*, int, const uint8_t*, uint8_t*);
inline GraphFn graph_kernel(int ty)
{
    return ty == kGraphSmallTY ? kf_graph<kGraphSmallTY> : ty == kGraphTallTY ? kf_graph<kGraphTallTY> : kf_graph<GTY>;
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
    int tx;  // raster tile width (cells): 16, 32 or 64
    dim3 grid;
    int grid3d;
    size_t off_items, off_map;  // in the descriptor buffer
    bool walk;  // the crossing walks read a larger LDS window (tile_resolve's WALK)
};
struct BatchPlan {
    bool valid = false;
    std::vector<dpx_item> items;
    uintptr_t out_mod = 0;  // d_out % 16
    dim3 ggrid;
    int g3d = 0;
    int gty = dpx::GTY;  // graph tile rows (kf_graph<gty>)
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
    // raster tile width per (size class, scale, subdivision); 0: the default
    int8_t tile_pick[3][dpx::kMaxScale + 1][2] = {};
    // walk window per (size class, scale, subdivision): DPX_WALK_AUTO (the
    // tuned choice: DPX_WALK_ALWAYS where tuning.json asks for the window,
    // else none), DPX_WALK_ALWAYS or DPX_WALK_NEVER
    int8_t walk_pick[3][dpx::kMaxScale + 1][2] = {};
    std::mutex mu;
    // kernel timing: 3 events per recorded batch (start, after graph, end)
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    dpx_batch_times acc{};
};

namespace {

using namespace dpx;

// Size class of an item for the tile picks.
int size_class(int W, int H)
{
    if (W <= 64 && H <= 64) return DPX_CLASS_SPRITE;
    if (W <= 512 && H <= 512) return DPX_CLASS_FRAME;
    return DPX_CLASS_LARGE;
}

int item_tile(const dpx_context* ctx, const dpx_item& a)
{
    const int t = ctx->tile_pick[size_class(a.width, a.height)][a.scale][a.subdivide ? 1 : 0];
    return t ? t : kDefaultTileX;
}

int item_walk(const dpx_context* ctx, const dpx_item& a)
{
    return ctx->walk_pick[size_class(a.width, a.height)][a.scale][a.subdivide ? 1 : 0];
}

// The tile picks of a tuning file: a JSON object whose "picks" array holds
// objects {"class": "sprite" | "frame" | "large", "scale": 1..15,
// "subdivide": 0 | 1, "tile": 16 | 32 | 64, "walk": 0 | 1} (other keys are
// ignored; "walk": 1 picks the walk window for every launch of the key).
// Returns the number of picks applied.
int load_tuning(dpx_context* ctx, const char* path)
{
    FILE* f = fopen(path, "rb");
    if (!f) return 0;
    std::string text;
    char buf[4096];
    size_t n;
    while ((n = fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, n);
    fclose(f);
    const size_t picks = text.find("\"picks\"");
    if (picks == std::string::npos) return 0;
    int applied = 0;
    for (size_t a = text.find('{', picks); a != std::string::npos; a = text.find('{', a + 1)) {
        const size_t b = text.find('}', a);
        if (b == std::string::npos) break;
        const std::string obj = text.substr(a, b - a);
        auto num = [&](const char* key, int* v) {
            const size_t k = obj.find(key);
            if (k == std::string::npos) return false;
            const size_t c = obj.find(':', k);
            if (c == std::string::npos) return false;
            *v = atoi(obj.c_str() + c + 1);
            return true;
        };
        int cls = -1, scale = 0, sub = 0, tile = 0, walk = 0;
        if (obj.find("\"sprite\"") != std::string::npos) cls = DPX_CLASS_SPRITE;
        else if (obj.find("\"frame\"") != std::string::npos) cls = DPX_CLASS_FRAME;
        else if (obj.find("\"large\"") != std::string::npos) cls = DPX_CLASS_LARGE;
        if (cls < 0 || !num("\"scale\"", &scale) || !num("\"tile\"", &tile)) continue;
        num("\"subdivide\"", &sub);
        num("\"walk\"", &walk);
        if (scale < 1 || scale > kMaxScale || (tile != 16 && tile != 32 && tile != 64)) continue;
        ctx->tile_pick[cls][scale][sub ? 1 : 0] = (int8_t)tile;
        ctx->walk_pick[cls][scale][sub ? 1 : 0] = (int8_t)(walk == 1 ? DPX_WALK_ALWAYS : DPX_WALK_AUTO);
        applied++;
    }
    return applied;
}

// The tuning file: $DPX_TUNING, else tuning.json in the package directory
// (one level above the directory of this library).
std::string tuning_path()
{
    if (const char* e = getenv("DPX_TUNING")) return e;
    Dl_info info;
    if (dladdr(reinterpret_cast<const void*>(&tuning_path), &info) && info.dli_fname) {
        std::string p = info.dli_fname;
        const size_t slash = p.rfind('/');
        p = slash == std::string::npos ? std::string(".") : p.substr(0, slash);
        return p + "/../tuning.json";
    }
    return "tuning.json";
}

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
    int tx = kDefaultTileX;
    int walk = DPX_WALK_AUTO;
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
dim3 tile_grid(const std::vector<DevItem>& dev, int TX, int TY, int* grid3d, std::vector<uint2>* map)
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
    load_tuning(ctx, tuning_path().c_str());
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

extern "C" int dpx_context_set_tile(dpx_context* ctx, int size_class, int scale, int subdivide, int tile_w)
{
    DPX_CHECK(ctx, DPX_EINVAL, "null context");
    DPX_CHECK(size_class >= 0 && size_class <= 2, DPX_EINVAL, "size class %d", size_class);
    DPX_CHECK(scale >= 1 && scale <= kMaxScale, DPX_EINVAL, "scale %d outside 1..%d", scale, kMaxScale);
    DPX_CHECK(tile_w == 0 || tile_w == 16 || tile_w == 32 || tile_w == 64, DPX_EINVAL, "tile width %d", tile_w);
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->tile_pick[size_class][scale][subdivide ? 1 : 0] = (int8_t)tile_w;
    ctx->plan.valid = false;  // the next batch plans its groups again
    return DPX_OK;
}

extern "C" int dpx_context_get_tile(dpx_context* ctx, int size_class, int scale, int subdivide, int* tile_w)
{
    DPX_CHECK(ctx && tile_w, DPX_EINVAL, "bad arguments");
    DPX_CHECK(size_class >= 0 && size_class <= 2, DPX_EINVAL, "size class %d", size_class);
    DPX_CHECK(scale >= 1 && scale <= kMaxScale, DPX_EINVAL, "scale %d outside 1..%d", scale, kMaxScale);
    std::lock_guard<std::mutex> lock(ctx->mu);
    const int t = ctx->tile_pick[size_class][scale][subdivide ? 1 : 0];
    *tile_w = t ? t : kDefaultTileX;
    return DPX_OK;
}

extern "C" int dpx_context_set_walk(dpx_context* ctx, int size_class, int scale, int subdivide, int walk)
{
    DPX_CHECK(ctx, DPX_EINVAL, "null context");
    DPX_CHECK(size_class >= 0 && size_class <= 2, DPX_EINVAL, "size class %d", size_class);
    DPX_CHECK(scale >= 1 && scale <= kMaxScale, DPX_EINVAL, "scale %d outside 1..%d", scale, kMaxScale);
    DPX_CHECK(walk == DPX_WALK_AUTO || walk == DPX_WALK_ALWAYS || walk == DPX_WALK_NEVER, DPX_EINVAL,
              "walk window mode %d", walk);
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->walk_pick[size_class][scale][subdivide ? 1 : 0] = (int8_t)walk;
    ctx->plan.valid = false;  // the next batch plans its groups again
    return DPX_OK;
}

extern "C" int dpx_context_get_walk(dpx_context* ctx, int size_class, int scale, int subdivide, int* walk)
{
    DPX_CHECK(ctx && walk, DPX_EINVAL, "bad arguments");
    DPX_CHECK(size_class >= 0 && size_class <= 2, DPX_EINVAL, "size class %d", size_class);
    DPX_CHECK(scale >= 1 && scale <= kMaxScale, DPX_EINVAL, "scale %d outside 1..%d", scale, kMaxScale);
    std::lock_guard<std::mutex> lock(ctx->mu);
    *walk = ctx->walk_pick[size_class][scale][subdivide ? 1 : 0];
    return DPX_OK;
}

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
        const int tx = item_tile(ctx, a), wk = item_walk(ctx, a);
        const int gkey = wk << 12 | tx << 8 | a.scale << 1 | d.subdivide;
        Group& grp = groups[gkey];
        grp.tx = tx;
        grp.walk = wk;
        grp.scale = a.scale;
        grp.sub = d.subdivide;
        grp.dev.push_back(d);
    }
    for (auto& kv : groups)
        if (int rc = ensure_mask(ctx, kv.second.scale)) return rc;

    // launch grids
    int g3d = 0;
    std::vector<uint2> gmap;
    uint64_t gtiles = 0, gtall = 0;
    for (const DevItem& d : dev) {
        gtiles += (uint64_t)((d.W + GTX - 1) / GTX) * ((d.H + GTY - 1) / GTY);
        gtall += (uint64_t)((d.W + GTX - 1) / GTX) * ((d.H + kGraphTallTY - 1) / kGraphTallTY);
    }
    plan.gty = gtiles < kGraphSmallTiles                                ? kGraphSmallTY
               : gtall * (kGraphTallTY + 4) < gtiles * (GTY + 4) ? kGraphTallTY
                                                                 : GTY;
    const dim3 ggrid = tile_grid(dev, GTX, plan.gty, &g3d, &gmap);
    for (auto& kv : groups)
        kv.second.grid = tile_grid(kv.second.dev, kv.second.tx, 256 / kv.second.tx, &kv.second.grid3d, &kv.second.map);

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
        const bool walk = grp.walk == DPX_WALK_ALWAYS;
        plan.groups.push_back(PlanGroup{ grp.scale, grp.sub, grp.tx, grp.grid, grp.grid3d, a, b, walk });
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
    hipLaunchKernelGGL(graph_kernel(plan.gty), plan.ggrid, dim3(256), 0, stream, d_items, d_gmap, plan.g3d, d_in,
                       la.gaux);
    if (ev_mid) DPX_HIP(hipEventRecord(ev_mid, stream));
    for (const PlanGroup& pg : plan.groups) {
        RasterFn fn = pick_raster(pg.scale, pg.sub, pg.tx, pg.walk);
        DPX_CHECK(fn, DPX_EINVAL, "no raster kernel for scale %d", pg.scale);
        const DevItem* d_gitems = reinterpret_cast<const DevItem*>(la.desc + pg.off_items);
        const uint2* d_rmap = reinterpret_cast<const uint2*>(la.desc + pg.off_map);
        if (pg.sub) {
            SmoothFn sf = pick_smooth(pg.scale, pg.tx, pg.walk);
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
        np.func = reinterpret_cast<void*>(graph_kernel(p->plan.gty));
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
            np.func = reinterpret_cast<void*>(pick_smooth(pg.scale, pg.tx, pg.walk));
        } else {
            a.mask = p->ctx->d_mask[pg.scale];
            a.info = p->ctx->d_info[pg.scale];
            void* v[10] = { &a.items, &a.map, &a.g3d, &a.in, &a.gaux_c, &a.gfin, &a.rows_c, &a.mask, &a.info, &a.out };
            memcpy(a.ptrs, v, sizeof(v));
            np.func = reinterpret_cast<void*>(pick_raster(pg.scale, pg.sub, pg.tx, pg.walk));
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

// ---------------------------------------------------------------------------
// dpx_batch_run_noalloc: the batched path with every buffer beyond the
// caller's input and output in one static device array, `scratch`, so that a
// call allocates nothing.  Layout (all offsets 256-byte aligned):
//   tables     the hull table, and per scale 1..15 the coverage rows (with
//              the cell masks at 1x and 2x) and the coverage summaries,
//              built on the first call on each device;
//   items      the batch's device items, in item order and grouped by
//              (scale, subdivision);
//   tile maps  block -> (item, tile) of the graph launch and of the raster
//              launches;
//   graph      the post-trivial snapshot graph, 1 byte per pixel.
// A call's descriptors reach the device as kernel arguments (k_na_put,
// kNaPutChunk items per launch), not as a copy from host memory.
// ---------------------------------------------------------------------------
namespace {

constexpr size_t na_align(size_t v) { return (v + 255) / 256 * 256; }
constexpr size_t na_mask_bytes(int S)
{
    return na_align((size_t)kKeys * mask_w(S) * 4 + (S <= kMaxCellScale ? (size_t)kKeys * 9 * 2 : 0));
}
constexpr size_t na_info_bytes() { return na_align((size_t)kKeys * 2); }
constexpr size_t na_mask_off(int S)
{
    size_t o = na_align((size_t)kKeys * kHullBytes);
    for (int s = 1; s < S; s++) o += na_mask_bytes(s) + na_info_bytes();
    return o;
}
constexpr size_t na_info_off(int S) { return na_mask_off(S) + na_mask_bytes(S); }

constexpr int kNaMaxItems = 4096;
constexpr int kNaTX = kDefaultTileX, kNaTY = 256 / kDefaultTileX;  // raster tiles (the default shape)
constexpr size_t kNaGraphTiles = 65536, kNaRasterTiles = 524288;
constexpr size_t kNaGraphBytes = 48u << 20;
constexpr size_t kNaOffItems = na_mask_off(kMaxScale + 1);
constexpr size_t kNaOffGroupItems = kNaOffItems + na_align(kNaMaxItems * sizeof(DevItem));
constexpr size_t kNaOffGmap = kNaOffGroupItems + na_align(kNaMaxItems * sizeof(DevItem));
constexpr size_t kNaOffRmap = kNaOffGmap + na_align(kNaGraphTiles * sizeof(uint2));
constexpr size_t kNaOffGraph = kNaOffRmap + na_align(kNaRasterTiles * sizeof(uint2));
constexpr size_t kNaScratchBytes = kNaOffGraph + kNaGraphBytes;
static_assert(kNaScratchBytes < (64u << 20), "scratch stays under 64 MiB");

static __device__ __align__(256) uint8_t scratch[kNaScratchBytes];

// One item of a k_na_put launch: its descriptor, its slots in the item
// arrays and the first entries of its tiles in the two tile maps.
struct NaPut {
    DevItem d;
    uint32_t index;  // in the item array (the graph launch's items)
    uint32_t gslot;  // in the grouped item array
    uint32_t gpos;   // gslot - the group's first slot (the raster launch's items)
    uint32_t gtile, rtile;
};
constexpr int kNaPutChunk = 32;
struct NaPutArgs {
    uint32_t n;
    NaPut it[kNaPutChunk];
};
static_assert(sizeof(NaPutArgs) <= 4000, "kernel arguments stay under 4 KB");

__global__ __launch_bounds__(256) void k_na_put(NaPutArgs a, DevItem* __restrict__ items,
                                                DevItem* __restrict__ group_items, uint2* __restrict__ gmap,
                                                uint2* __restrict__ rmap)
{
    for (uint32_t i = 0; i < a.n; i++) {
        const NaPut& p = a.it[i];
        if (threadIdx.x == 0) {
            items[p.index] = p.d;
            group_items[p.gslot] = p.d;
        }
        const uint32_t gx = (uint32_t)(p.d.W + GTX - 1) / GTX, gy = (uint32_t)(p.d.H + GTY - 1) / GTY;
        for (uint32_t t = threadIdx.x; t < gx * gy; t += blockDim.x)
            gmap[p.gtile + t] = make_uint2(p.index, (t / gx) << 16 | (t % gx));
        const uint32_t rx = (uint32_t)(p.d.W + kNaTX - 1) / kNaTX, ry = (uint32_t)(p.d.H + kNaTY - 1) / kNaTY;
        for (uint32_t t = threadIdx.x; t < rx * ry; t += blockDim.x)
            rmap[p.rtile + t] = make_uint2(p.gpos, (t / rx) << 16 | (t % rx));
    }
}

constexpr int kNaMaxDevices = 64;
std::mutex g_na_mu;
uint8_t* g_na_ptr[kNaMaxDevices] = {};  // each device's address of `scratch`

// Set-up on the first call on a device: the array's address on this device
// and its tables, built on `stream` and waited for (later calls on any
// stream read them).
int na_setup(int dev, hipStream_t stream, uint8_t** out)
{
    std::lock_guard<std::mutex> lock(g_na_mu);
    if (!g_na_ptr[dev]) {
        uint8_t* ptr = nullptr;
        DPX_HIP(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)));
        DPX_CHECK(ptr, DPX_EHIP, "no device address for the scratch array");
        int8_t* hull = reinterpret_cast<int8_t*>(ptr);
        hipLaunchKernelGGL(kt_hull, dim3(kKeys / 256), dim3(256), 0, stream, hull);
        for (int S = 1; S <= kMaxScale; S++) {
            uint32_t* mask = reinterpret_cast<uint32_t*>(ptr + na_mask_off(S));
            const size_t n = (size_t)kKeys * mask_w(S);
            hipLaunchKernelGGL(kt_mask, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, hull, S, mask);
            if (S <= kMaxCellScale)
                hipLaunchKernelGGL(kt_cell, dim3((unsigned)((kKeys * 9 + 255) / 256)), dim3(256), 0, stream, mask, S,
                                   reinterpret_cast<uint16_t*>(mask + n));
            hipLaunchKernelGGL(kt_info, dim3(kKeys / 256), dim3(256), 0, stream, mask, S,
                               reinterpret_cast<uint16_t*>(ptr + na_info_off(S)));
        }
        DPX_HIP(hipGetLastError());
        DPX_HIP(hipStreamSynchronize(stream));
        g_na_ptr[dev] = ptr;
    }
    *out = g_na_ptr[dev];
    return DPX_OK;
}

}  // namespace

extern "C" int dpx_batch_run_noalloc(const uint8_t* d_in, uint8_t* d_out, const dpx_item* items, int n_items,
                                     void* stream_)
{
    DPX_CHECK(d_in && d_out && items && n_items > 0, DPX_EINVAL, "bad arguments");
    DPX_CHECK(n_items <= kNaMaxItems, DPX_EINVAL, "%d items: dpx_batch_run_noalloc takes at most %d", n_items,
              kNaMaxItems);
    hipStream_t stream = as_stream(stream_);
    int dev = 0;
    DPX_HIP(hipGetDevice(&dev));
    DPX_CHECK(dev >= 0 && dev < kNaMaxDevices, DPX_ENODEV, "device %d", dev);
    // pass 1: validate, graph bytes and tiles, per group (2 * scale + sub)
    // its item and tile counts
    uint32_t gcount[2 * (kMaxScale + 1)] = {}, gtiles[2 * (kMaxScale + 1)] = {};
    size_t gbytes = 0, graph_tiles = 0, raster_tiles = 0;
    for (int k = 0; k < n_items; k++) {
        DevItem d;
        if (int rc = make_dev_item(items[k], k, d_out, 0, d)) return rc;
        DPX_CHECK(!d.subdivide, DPX_EINVAL,
                  "item %d: subdivision needs per-cell row buffers; dpx_batch_run_noalloc takes items without it", k);
        gbytes += align_up((size_t)d.W * d.H, 256);
        graph_tiles += (size_t)((d.W + GTX - 1) / GTX) * ((d.H + GTY - 1) / GTY);
        const size_t rt = (size_t)((d.W + kNaTX - 1) / kNaTX) * ((d.H + kNaTY - 1) / kNaTY);
        raster_tiles += rt;
        const int g = 2 * d.scale + d.subdivide;
        gcount[g]++;
        gtiles[g] += (uint32_t)std::min<size_t>(rt, 0xFFFFFFFFu);
    }
    DPX_CHECK(gbytes <= kNaGraphBytes, DPX_EINVAL,
              "batch needs %zu graph bytes: dpx_batch_run_noalloc holds at most %zu (one byte per pixel, each item "
              "rounded up to 256)", gbytes, kNaGraphBytes);
    DPX_CHECK(graph_tiles <= kNaGraphTiles && raster_tiles <= kNaRasterTiles, DPX_EINVAL,
              "batch has %zu graph and %zu raster tiles: dpx_batch_run_noalloc holds at most %zu and %zu", graph_tiles,
              raster_tiles, kNaGraphTiles, kNaRasterTiles);
    uint8_t* base = nullptr;
    if (int rc = na_setup(dev, stream, &base)) return rc;
    DevItem* d_items = reinterpret_cast<DevItem*>(base + kNaOffItems);
    DevItem* d_gitems = reinterpret_cast<DevItem*>(base + kNaOffGroupItems);
    uint2* d_gmap = reinterpret_cast<uint2*>(base + kNaOffGmap);
    uint2* d_rmap = reinterpret_cast<uint2*>(base + kNaOffRmap);
    uint8_t* d_graph = base + kNaOffGraph;
    // groups in ascending (scale, subdivision): first item slot and tile
    uint32_t gslot0[2 * (kMaxScale + 1)], gtile0[2 * (kMaxScale + 1)];
    uint32_t slot = 0, tile = 0;
    for (int g = 0; g < 2 * (kMaxScale + 1); g++) {
        gslot0[g] = slot;
        gtile0[g] = tile;
        slot += gcount[g];
        tile += gtiles[g];
    }
    // pass 2: descriptors to the device, kNaPutChunk items per launch
    uint32_t gnext[2 * (kMaxScale + 1)] = {}, gtnext[2 * (kMaxScale + 1)] = {};
    NaPutArgs args;
    memset(&args, 0, sizeof(args));
    uint64_t goff = 0;
    uint32_t gt = 0;
    for (int k = 0; k < n_items; k++) {
        NaPut& p = args.it[args.n++];
        (void)make_dev_item(items[k], k, d_out, goff, p.d);  // (validated in pass 1)
        const int g = 2 * p.d.scale + p.d.subdivide;
        p.index = (uint32_t)k;
        p.gpos = gnext[g]++;
        p.gslot = gslot0[g] + p.gpos;
        p.gtile = gt;
        p.rtile = gtile0[g] + gtnext[g];
        gt += (uint32_t)(((p.d.W + GTX - 1) / GTX) * ((p.d.H + GTY - 1) / GTY));
        gtnext[g] += (uint32_t)(((p.d.W + kNaTX - 1) / kNaTX) * ((p.d.H + kNaTY - 1) / kNaTY));
        goff += align_up((size_t)p.d.W * p.d.H, 256);
        if (args.n == kNaPutChunk || k == n_items - 1) {
            hipLaunchKernelGGL(k_na_put, dim3(1), dim3(256), 0, stream, args, d_items, d_gitems, d_gmap, d_rmap);
            args.n = 0;
        }
    }
    hipLaunchKernelGGL(kf_graph<GTY>, dim3((unsigned)graph_tiles), dim3(256), 0, stream, d_items, d_gmap, 0, d_in,
                       d_graph);
    for (int g = 0; g < 2 * (kMaxScale + 1); g++) {
        if (!gcount[g]) continue;
        const int S = g / 2;
        RasterFn fn = pick_raster(S, g % 2, kNaTX);
        DPX_CHECK(fn, DPX_EINVAL, "no raster kernel for scale %d", S);
        hipLaunchKernelGGL(fn, dim3(gtiles[g]), dim3(256), 0, stream, d_gitems + gslot0[g], d_rmap + gtile0[g], 0,
                           d_in, d_graph, nullptr, nullptr,
                           reinterpret_cast<const uint32_t*>(base + na_mask_off(S)),
                           reinterpret_cast<const uint16_t*>(base + na_info_off(S)), d_out);
    }
    DPX_HIP(hipGetLastError());
    return DPX_OK;
}

