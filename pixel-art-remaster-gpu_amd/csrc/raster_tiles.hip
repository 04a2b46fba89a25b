// raster_tiles.hip -- the smoothing and raster kernels of one tile shape,
// DPX_TILE_X x (256 / DPX_TILE_X) cells; built once per shape (16, 32, 64).
#include "batch.hpp"

#ifndef DPX_TILE_X
#error "build with -DDPX_TILE_X=16, 32 or 64"
#endif
#define DPX_CAT2(a, b) a##b
#define DPX_CAT(a, b) DPX_CAT2(a, b)

namespace dpx {
namespace {

template <bool SUB, bool WALK>
RasterFn raster_of(int S)
{
    switch (S) {
        case 1: return kf_raster<1, SUB, DPX_TILE_X, WALK>;
        case 2: return kf_raster<2, SUB, DPX_TILE_X, WALK>;
        case 3: return kf_raster<3, SUB, DPX_TILE_X, WALK>;
        case 4: return kf_raster<4, SUB, DPX_TILE_X, WALK>;
        case 5: return kf_raster<5, SUB, DPX_TILE_X, WALK>;
        case 6: return kf_raster<6, SUB, DPX_TILE_X, WALK>;
        case 7: return kf_raster<7, SUB, DPX_TILE_X, WALK>;
        case 8: return kf_raster<8, SUB, DPX_TILE_X, WALK>;
        case 9: return kf_raster<9, SUB, DPX_TILE_X, WALK>;
        case 10: return kf_raster<10, SUB, DPX_TILE_X, WALK>;
        case 11: return kf_raster<11, SUB, DPX_TILE_X, WALK>;
        case 12: return kf_raster<12, SUB, DPX_TILE_X, WALK>;
        case 13: return kf_raster<13, SUB, DPX_TILE_X, WALK>;
        case 14: return kf_raster<14, SUB, DPX_TILE_X, WALK>;
        case 15: return kf_raster<15, SUB, DPX_TILE_X, WALK>;
    }
    return nullptr;
}

template <bool WALK>
SmoothFn smooth_of(int S)
{
    switch (S) {
        case 1: return kf_smooth<1, DPX_TILE_X, WALK>;
        case 2: return kf_smooth<2, DPX_TILE_X, WALK>;
        case 3: return kf_smooth<3, DPX_TILE_X, WALK>;
        case 4: return kf_smooth<4, DPX_TILE_X, WALK>;
        case 5: return kf_smooth<5, DPX_TILE_X, WALK>;
        case 6: return kf_smooth<6, DPX_TILE_X, WALK>;
        case 7: return kf_smooth<7, DPX_TILE_X, WALK>;
        case 8: return kf_smooth<8, DPX_TILE_X, WALK>;
        case 9: return kf_smooth<9, DPX_TILE_X, WALK>;
        case 10: return kf_smooth<10, DPX_TILE_X, WALK>;
        case 11: return kf_smooth<11, DPX_TILE_X, WALK>;
        case 12: return kf_smooth<12, DPX_TILE_X, WALK>;
        case 13: return kf_smooth<13, DPX_TILE_X, WALK>;
        case 14: return kf_smooth<14, DPX_TILE_X, WALK>;
        case 15: return kf_smooth<15, DPX_TILE_X, WALK>;
    }
    return nullptr;
}

}  // namespace

// (the subdivided raster resolves no crossings: kf_smooth did)
RasterFn DPX_CAT(pick_raster_t, DPX_TILE_X)(int S, bool sub, bool walk)
{
    return sub ? raster_of<true, false>(S) : walk ? raster_of<false, true>(S) : raster_of<false, false>(S);
}

SmoothFn DPX_CAT(pick_smooth_t, DPX_TILE_X)(int S, bool walk) { return walk ? smooth_of<true>(S) : smooth_of<false>(S); }

}  // namespace dpx
