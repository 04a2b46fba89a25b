// batch.hpp -- device code of the batched path shared by its translation
// units: the item descriptor, tile placement, and the smoothing and raster
// kernels as templates on the scale and the tile shape.  fused.hip holds the
// tables, the graph kernel and the host side; raster_tiles.hip is built once
// per tile shape and instantiates the kernels of that shape.
#pragma once

#include <type_traits>

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

// Cell masks at 1x and 2x follow the row table (kt_cell).
constexpr int kMaxCellScale = 2;

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
// kf_raster: heuristic crossing resolution, cells (+ subdivision) and the
// gather raster of a tile of cells
// ---------------------------------------------------------------------------
// Raster tiles of TX x (256 / TX) cells, one own cell per thread of a
// 256-thread block (32x8 by default; 16x16 and 64x4 are the other shapes an
// item may be given).
template <int TX>
struct TileShape {
    static constexpr int X = TX, Y = 256 / TX;
    static_assert(X * Y == 256 && X % 4 == 0, "one own cell per thread, whole words per row");
    static constexpr int CX = X + 2, CY = Y + 2;       // candidate cells, halo 1
    static constexpr int DW = X / 4 + 2, DB = 4 * DW;  // data words / bytes of a window row
    static constexpr int WW = DW + 2, WB = 4 * WW;     // words / bytes of a window row
    static constexpr int GY = Y + 2, AY = GY + 2;      // resolved rows, snapshot rows
    // the window loads and the decision clear take one word per thread
    static_assert(AY * WW <= 256 && (GY + 1) * WW <= 256, "a window fits one pass of 256 threads");
};
constexpr int kDefaultTileX = 32;
constexpr uint16_t kNoCell = 0xFFFF;

// LDS windows of the raster tile hold graph bytes in rows of WW 32-bit
// words; word w of a row holds x = x0 - 8 + 4w .. x0 - 5 + 4w.  Words
// 1..DW (x0 - 4 .. x0 + X + 3) carry data, words 0 and DW + 1 stay zero.  Candidate cells
// (halo 1) need the resolved bytes of their left and right neighbours for
// their keys: resolved rows at halo 1; with subdivision the hulls at halo 2
// are needed as well: resolved rows at halo 2.  Resolving a byte reads the
// snapshot rows above and below it.
// Resolved rows y0 - 1 .. y0 + Y (GY rows: candidate cells need their left
// and right neighbours' bytes for their keys); resolving them reads one more
// snapshot row on each side (AY rows).
#define DPX_TILE_NAMES(TX)                                                                                 \
    using TS = TileShape<TX>;                                                                              \
    constexpr int RTX = TS::X, RTY = TS::Y, RCX = TS::CX, RCY = TS::CY, RWW = TS::WW, RWB = TS::WB;        \
    constexpr int GY = TS::GY, AY = TS::AY, RDW = TS::DW, RDB = TS::DB;                                    \
    (void)RTX; (void)RTY; (void)RCX; (void)RCY; (void)RWW; (void)RWB; (void)GY; (void)AY; (void)RDW; (void)RDB

// Per-cell coverage rows of smoothed cells (kf_smooth -> kf_raster<S, true>):
// MW words of sample rows and one word of coverage summary per cell, cell
// (x, y) of an item at word (y * W + x) * (MW + 1) of its range.
__host__ __device__ constexpr int rows_words(int S) { return mask_w(S) + 1; }

// The snapshot rows ay0 .. ay0 + AY - 1 of a tile's window into LDS.  With
// W % 4 == 0 the graph rows are word aligned and a word is either wholly
// inside the frame or wholly outside.
template <int TX>
__device__ inline void tile_snapshot_load(const DevItem& it, const uint8_t* __restrict__ aux, int x0, int ay0,
                                          uint32_t* s_aux32, int t)
{
    DPX_TILE_NAMES(TX);
    if (it.graph_w4) {
        if (t < AY * RWW) {
            int r = t / RWW, w = t % RWW;
            int x = x0 - 8 + 4 * w, y = ay0 + r;
            uint32_t v = 0;
            if (w >= 1 && w <= RDW && (unsigned)x < (unsigned)it.W && (unsigned)y < (unsigned)it.H)
                v = *reinterpret_cast<const uint32_t*>(aux + (uint32_t)(y * it.W + x));
            s_aux32[t] = v;
        }
    } else {
        uint8_t* s8 = reinterpret_cast<uint8_t*>(s_aux32);
        for (int k = t; k < AY * RWB; k += 256) {
            int r = k / RWB, c = k % RWB;
            int x = x0 - 8 + c, y = ay0 + r;
            bool ok = c >= 4 && c < 4 + RDB && (unsigned)x < (unsigned)it.W && (unsigned)y < (unsigned)it.H;
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
//
// The curve walks of the decisions (up to 31 steps) leave the small window
// on line-heavy frames, and each step outside it is a dependent HBM read.
// With WALK, a tile that has crossed blocks first loads a larger walk window
// of the snapshot into s_walk (an LDS buffer of WalkWindow<TX>::BYTES that
// the kernel uses only later): kWalkEX more bytes on each side and kWalkEY
// more rows above and below, so that no walk (at most 31 steps from a node
// next to a block of the small window) leaves it.  It costs every such tile
// 7-10 KB more reads, which a full batch pays in HBM time (20% of the
// 256x224 batch); it pays where dense crossing lattices make long walks the
// tile's critical path (640x480 frames at 8x), and the autotuner picks it
// per workload class.
constexpr int kWalkEX = 32, kWalkEY = 32;
template <int TX>
struct WalkWindow {
    static constexpr int B = TileShape<TX>::DB + 2 * kWalkEX;    // bytes per row (a multiple of 4)
    static constexpr int R = TileShape<TX>::AY + 2 * kWalkEY;    // rows
    static constexpr int BYTES = B * R;
};

template <int TX, bool WALK>
__device__ inline void tile_resolve(const DevItem& it, const uint8_t* __restrict__ aux, uint8_t* __restrict__ gfin,
                                    int x0, int y0, int ay0, const uint32_t* s_aux32, uint32_t* s_p32,
                                    uint16_t* s_blk, int* s_nblk, uint32_t* s_g32, uint32_t* s_walk, int t)
{
    DPX_TILE_NAMES(TX);
    using WW_ = WalkWindow<TX>;
    for (int k = t; k < (GY + 1) * RDW; k += 256) {
        const int r = k / RDW, w = k % RDW + 1;  // word w of snapshot row r: origins of blocks on rows r, r + 1
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
    const int nblk = *s_nblk;
    const bool walk_win = WALK && nblk > 0;  // uniform over the block
    const int ex0 = x0 - 4 - kWalkEX, ey0 = ay0 - kWalkEY;  // the walk window's first byte and row
    if (walk_win) {
        // all of a thread's loads are issued before the first is waited for
        // (a loop that stores each into LDS before the next load waits a
        // memory latency per iteration: 17 us instead of 2 for a tile)
        if (it.graph_w4) {  // x0 and kWalkEX are multiples of 4: whole words inside or outside the frame
            constexpr int NW = WW_::R * (WW_::B / 4), PER = (NW + 255) / 256;
            uint32_t v[PER];
#pragma unroll
            for (int q = 0; q < PER; q++) {
                const int k = t + 256 * q;
                const int r = k / (WW_::B / 4), w = k % (WW_::B / 4);
                const int x = ex0 + 4 * w, y = ey0 + r;
                v[q] = 0;
                if (k < NW && (unsigned)x < (unsigned)it.W && (unsigned)y < (unsigned)it.H)
                    v[q] = *reinterpret_cast<const uint32_t*>(aux + (uint32_t)(y * it.W + x));
            }
#pragma unroll
            for (int q = 0; q < PER; q++)
                if (t + 256 * q < NW) s_walk[t + 256 * q] = v[q];
        } else {
            constexpr int PER = (WW_::BYTES + 255) / 256;
            uint8_t* s8 = reinterpret_cast<uint8_t*>(s_walk);
            uint8_t v[PER];
#pragma unroll
            for (int q = 0; q < PER; q++) {
                const int k = t + 256 * q;
                const int x = ex0 + k % WW_::B, y = ey0 + k / WW_::B;
                v[q] = 0;
                if (k < WW_::BYTES && (unsigned)x < (unsigned)it.W && (unsigned)y < (unsigned)it.H)
                    v[q] = aux[(uint32_t)(y * it.W + x)];
            }
#pragma unroll
            for (int q = 0; q < PER; q++)
                if (t + 256 * q < WW_::BYTES) s8[t + 256 * q] = v[q];
        }
        __syncthreads();
    }
    {
        // (the window is read by words, so that the compiler keeps the LDS
        // and HBM reads apart instead of merging them into one generic load)
        auto get = [&](int gx, int gy) -> uint32_t {
            bool win;
            uint32_t v;
            if (walk_win) {
                const int wx = gx - ex0, wy = gy - ey0;
                win = (unsigned)wx < (unsigned)WW_::B && (unsigned)wy < (unsigned)WW_::R;
                const int c = win ? wy * WW_::B + wx : 0;
                v = (s_walk[c >> 2] >> (8 * (c & 3))) & 0xFFu;
            } else {
                const int wx = gx - (x0 - 4), wy = gy - ay0;
                win = (unsigned)wx < (unsigned)RDB && (unsigned)wy < (unsigned)AY;
                const int c = win ? wy * RWB + wx + 4 : 0;
                v = (s_aux32[c >> 2] >> (8 * (c & 3))) & 0xFFu;
            }
            if (!win) v = aux[(uint32_t)(gy * it.W + gx)];
            return v;
        };
        uint8_t* p8 = reinterpret_cast<uint8_t*>(s_p32);
        for (int k = t; k < nblk; k += 256) {
            int e = s_blk[k];
            int r = e / RWB, c = e % RWB;
            // 1: "/" loses (bits 2 and 5), 2: "\" loses (bits 0 and 7)
            p8[e] = remove_slash(get, x0 - 8 + c, ay0 + r) ? 1 : 2;
        }
    }
    __syncthreads();
    for (int k = t; k < GY * RDW; k += 256) {
        const int r = k / RDW, w = k % RDW + 1;  // word w of resolved row r (snapshot row r + 1)
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
        if (gfin && y >= y0 && y < y0 + RTY && w >= 2 && w <= RDW - 1 && y < it.H && x < it.W) {
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
template <int TX>
__device__ inline void tile_resolved_load(const DevItem& it, const uint8_t* __restrict__ gfin, int x0, int y0,
                                          uint32_t* s_g32, int t)
{
    DPX_TILE_NAMES(TX);
    const uint8_t* g = gfin + it.g_off;
    if (it.graph_w4) {
        if (t < GY * RWW) {
            int r = t / RWW, w = t % RWW;
            int x = x0 - 8 + 4 * w, y = y0 - 1 + r;
            uint32_t v = 0;
            if (w >= 1 && w <= RDW && (unsigned)x < (unsigned)it.W && (unsigned)y < (unsigned)it.H)
                v = *reinterpret_cast<const uint32_t*>(g + (uint32_t)(y * it.W + x));
            s_g32[t] = v;
        }
    } else {
        uint8_t* s8 = reinterpret_cast<uint8_t*>(s_g32);
        for (int k = t; k < GY * RWB; k += 256) {
            int r = k / RWB, c = k % RWB;
            int x = x0 - 8 + c, y = y0 - 1 + r;
            bool ok = c >= 4 && c < 4 + RDB && (unsigned)x < (unsigned)it.W && (unsigned)y < (unsigned)it.H;
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
template <int S, int TX, bool WALK>
__global__ __launch_bounds__(256, 4) void kf_smooth(const DevItem* __restrict__ items, const uint2* __restrict__ tile_map,
                                                 int grid3d, const uint8_t* __restrict__ in,
                                                 const uint8_t* __restrict__ gaux, uint8_t* __restrict__ gfin,
                                                 const int8_t* __restrict__ hull_table, uint32_t* __restrict__ rows)
{
    DPX_TILE_NAMES(TX);
    constexpr int MW = mask_w(S), KMIN = mask_kmin(S), RW = rows_words(S);
    constexpr int HB = 20;  // LDS bytes per hull: n, x[8], y[8] (quarter units), 3 bytes of padding
    __shared__ uint32_t s_aux32[AY * RWW];
    __shared__ uint32_t s_p32[(GY + 1) * RWW];
    __shared__ uint16_t s_blk[(GY + 1) * RDB];
    __shared__ int s_nblk;
    __shared__ uint32_t s_g32[GY * RWW];
    __shared__ uint32_t s_hull32[RCX * RCY * HB / 4];
    const int8_t* s_hull = reinterpret_cast<const int8_t*>(s_hull32);
    __shared__ uint16_t s_list[RTX * RTY];
    __shared__ int s_nlist;
    __shared__ int s_bin[16];  // own cells to smooth per hull size, then the first list place of each size
    // point k of lane t at [k * 256 + t]: x | y << 16 (1/64 units); a hull has
    // at most 8 vertices (the table's), its smoothed polygon at most 16
    __shared__ uint32_t s_poly[16 * 256];
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
    if (t < 16) s_bin[t] = 0;
    if (t < (GY + 1) * RWW) s_p32[t] = 0;
    tile_snapshot_load<TX>(it, aux, x0, ay0, s_aux32, t);
    __syncthreads();
    static_assert(16 * 256 * 4 >= WalkWindow<TX>::BYTES, "the polygon buffer holds the walk window first");
    tile_resolve<TX, WALK>(it, aux, gfin, x0, y0, ay0, s_aux32, s_p32, s_blk, &s_nblk, s_g32, s_poly, t);
    __syncthreads();
    // hulls of the own cells and of their neighbours: cell (x0 - 1 + lx,
    // y0 - 1 + ly) at resolved row ly, byte lx + 7
    for (int k = t; k < RCX * RCY; k += 256) {
        const int lx = k % RCX, ly = k / RCX;
        const int c = ly * RWB + lx + 7;
        const uint32_t key = cell_key(s_g[c], s_g[c - 1], s_g[c + 1]);
        // the first HB bytes of the table's 32-byte entry, as words
        const uint32_t* e = reinterpret_cast<const uint32_t*>(hull_table + (size_t)key * kHullBytes);
#pragma unroll
        for (int q = 0; q < HB / 4; q++) s_hull32[k * (HB / 4) + q] = e[q];
    }
    // own cells to smooth: in the frame and not internal, listed by the size
    // of their hull (a cell's work grows with it), largest first, so that
    // the lanes of a wave smooth polygons of about the same size
    int my_n = -1, my_rank = 0;
    {
        const int cx = t % RTX, cy = t / RTX;
        const int c = (cy + 1) * RWB + cx + 8;
        const uint32_t node = s_g[c];
        if (x0 + cx < it.W && y0 + cy < it.H && node != 90) {
            my_n = hull_table[(size_t)cell_key(node, s_g[c - 1], s_g[c + 1]) * kHullBytes];
            my_rank = atomicAdd(&s_bin[my_n], 1);
        }
    }
    __syncthreads();
    if (t == 0) {
        int acc = 0;
        for (int b = 15; b >= 0; b--) {
            const int v = s_bin[b];
            s_bin[b] = acc;
            acc += v;
        }
        s_nlist = acc;
    }
    __syncthreads();
    if (my_n >= 0) s_list[s_bin[my_n] + my_rank] = (uint16_t)t;
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
    auto px_of = [](uint32_t p) { return (int)(int16_t)(p & 0xFFFFu); };
    auto py_of = [](uint32_t p) { return (int)p >> 16; };
    // Ear clipping (process, triangulate_functions.cu:108-217) as ear_clip
    // does it: vertices in counter-clockwise order (reversed when the
    // signed area is not positive), first ear tried at (m-1, 0, 1), an ear
    // needs a positive cross and no other vertex inside or on it, a failure
    // counter of 2 nv.  The coverage is the union of the ears' rows.  Pass 0
    // takes it from the polygon's edges (poly_rows_xor: exact when the
    // clipping cuts the whole polygon and the winding sum certifies it) and
    // clips only to learn that; a lane where either fails (none in the
    // synthetic workloads) takes pass 1, which subdivides again and
    // accumulates each ear's rows as it is cut.
    uint32_t cov[MW];
    bool need = true;
#pragma clang loop unroll(disable)
    for (int pass = 0; pass < 2; pass++) {
        if (!need) continue;
        int m = 0;
        subdivide_cell_emit(HullX{ h }, HullY{ h }, n, node, nbr, tj, [&](int px, int py) {
            P[256 * m] = ((uint32_t)px & 0xFFFFu) | ((uint32_t)py << 16);
            m++;
        });
#pragma unroll
        for (int r = 0; r < MW; r++) cov[r] = 0;
        need = false;
        if (m < 3) continue;
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
        int wsum = 0;
        if (pass == 0) wsum = poly_rows_xor<MW>(P, 256, m, 2 * S, KMIN, cov);
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
                if (pass == 1)
                    tri_rows_or<MW>(ax * 2 * S, ay * 2 * S, bx * 2 * S, by * 2 * S, cx_ * 2 * S, cy_ * 2 * S, KMIN,
                                    cov);
                for (int s2 = v, t2 = v + 1; t2 < nv; s2++, t2++) P[256 * s2] = P[256 * t2];
                nv--;
                count = 2 * nv;
            }
        }
        if (pass == 0) {
            int bits = 0;
#pragma unroll
            for (int r = 0; r < MW; r++) bits += __builtin_popcount(cov[r]);
            need = nv > 2 || wsum != bits;
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
template <int S, bool SUB, int TX, bool WALK>
__global__ __launch_bounds__(256) void kf_raster(const DevItem* __restrict__ items, const uint2* __restrict__ tile_map,
                                                 int grid3d,
                                                 const uint8_t* __restrict__ in, const uint8_t* __restrict__ gaux,
                                                 uint8_t* __restrict__ gfin, const uint32_t* __restrict__ cell_rows,
                                                 const uint32_t* __restrict__ mask_table,
                                                 const uint16_t* __restrict__ info_table, uint8_t* __restrict__ out)
{
    DPX_TILE_NAMES(TX);
    constexpr int MW = mask_w(S), KMIN = mask_kmin(S), RW = rows_words(S);
    __shared__ uint32_t s_aux32[SUB ? 1 : AY * RWW];
    __shared__ uint32_t s_p32[SUB ? 1 : (GY + 1) * RWW];  // decisions of crossed blocks, by origin
    __shared__ uint16_t s_blk[SUB ? 1 : (GY + 1) * RDB];  // crossed blocks: origin byte index in the window
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
        tile_resolved_load<TX>(it, gfin, x0, y0, s_g32, t);
    } else {
        if (t < (GY + 1) * RWW) s_p32[t] = 0;
        tile_snapshot_load<TX>(it, aux, x0, ay0, s_aux32, t);
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
        __shared__ uint32_t s_walk[WALK ? WalkWindow<TX>::BYTES / 4 : 1];
        tile_resolve<TX, WALK>(it, aux, gfin, x0, y0, ay0, s_aux32, s_p32, s_blk, &s_nblk, s_g32, s_walk, t);
        __syncthreads();
    }
    // candidate cells: key and coverage summary.  Candidate (lx, ly) is the
    // cell at (x0 - 1 + lx, y0 - 1 + ly): resolved row ly, byte lx + 7.
    // (unrolled, so that a thread's table loads are all in flight at once)
#pragma unroll
    for (int q = 0; q < (RCX * RCY + 255) / 256; q++) {
        const int k = t + 256 * q;
        if (k >= RCX * RCY) break;
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

// The raster and smoothing kernels of scale S (1..15), per tile shape (one
// translation unit each: raster_tiles.hip).
// walk: the variants that load the walk window (tile_resolve's WALK)
RasterFn pick_raster_t16(int S, bool sub, bool walk);
RasterFn pick_raster_t32(int S, bool sub, bool walk);
RasterFn pick_raster_t64(int S, bool sub, bool walk);
SmoothFn pick_smooth_t16(int S, bool walk);
SmoothFn pick_smooth_t32(int S, bool walk);
SmoothFn pick_smooth_t64(int S, bool walk);

inline RasterFn pick_raster(int S, bool sub, int tx, bool walk = false)
{
    return tx == 16 ? pick_raster_t16(S, sub, walk) : tx == 64 ? pick_raster_t64(S, sub, walk)
                                                              : pick_raster_t32(S, sub, walk);
}
inline SmoothFn pick_smooth(int S, int tx, bool walk = false)
{
    return tx == 16 ? pick_smooth_t16(S, walk) : tx == 64 ? pick_smooth_t64(S, walk) : pick_smooth_t32(S, walk);
}

}  // namespace dpx
