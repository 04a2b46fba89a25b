// pixel.hpp -- pixel access and the YUV similarity test of the depixelization path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/depixel.h"

namespace dpx {

// One frame as the kernels see it.  An indexed frame (DPX_INDEXED, ch 1)
// carries its palette of 256 four-byte entries in `pal`; the kernels then see
// the 4-channel frame of those entries.
struct ImgView {
    const uint8_t* p;
    int W, H, stride, ch;
    uint32_t flags;
    const uint8_t* pal;
};

// Output row 0 is the top row: the input's row order, reversed by DPX_FLIP_OUT.
__host__ __device__ inline bool out_topdown(uint32_t flags)
{
    return ((flags & DPX_TOPDOWN) != 0) != ((flags & DPX_FLIP_OUT) != 0);
}

// Bytes per pixel colour: 4 for an indexed frame (its palette entries).
__host__ __device__ inline int colour_channels(const ImgView& im) { return im.pal ? 4 : im.ch; }

// The colour bytes of pixel (i, j): in the frame, or its palette entry.
__host__ __device__ inline const uint8_t* px_ptr(const ImgView& im, int i, int j)
{
    int row = (im.flags & DPX_TOPDOWN) ? (im.H - 1 - j) : j;
    const uint8_t* q = im.p + (size_t)row * im.stride + (size_t)i * im.ch;
    return im.pal ? im.pal + 4u * q[0] : q;
}

// The reference's byte triple (pixel[0], pixel[1], pixel[2] of a BGR buffer,
// graph_functions.cu:192,251) and the alpha byte (255 for 3 channels).
struct Px {
    int c0, c1, c2, a;
};

__host__ __device__ inline Px load_px(const ImgView& im, int i, int j)
{
    const uint8_t* q = px_ptr(im, i, j);
    Px r;
    if (im.flags & DPX_RGB) { r.c0 = q[2]; r.c1 = q[1]; r.c2 = q[0]; }
    else                    { r.c0 = q[0]; r.c1 = q[1]; r.c2 = q[2]; }
    r.a = colour_channels(im) == 4 ? q[3] : 255;
    return r;
}

// RGBtoYUV (graph_functions.cu:80-98) of DATAtoINT(c0,c1,c2) (102-131): r is
// byte 0, b is byte 2.  Luma in IEEE double, (0.299 r + 0.587 g) + 0.114 b
// with no contraction; U and V truncated float products that may be negative
// and borrow into the packed word as the reference's int shifts do.
//
// Luma without floating point where it can be: with n = 299 c0 + 587 c1 +
// 114 c2 the real value is n/1000, and the double expression is within
// 3e-13 of it.  Unless 1000 divides n, n/1000 is at least 0.001 from an
// integer, so the truncated double equals n / 1000.  Only for the exact
// multiples (16,774 of the 2^24 colours, every grey among them) can the
// double land just below the integer; those take the double expression.
// The quotient is a high product: with m = ceil(2^32 / 1000) = 4294968,
// (n * m) >> 32 = floor(n / 1000) while n * (1000 m - 2^32) = 704 n < 2^32,
// i.e. for every n up to 6.1 M (n <= 255,000 here; checked exhaustively).
__host__ __device__ inline uint32_t mul24(uint32_t a, uint32_t b)  // a, b < 2^24
{
#ifdef __HIP_DEVICE_COMPILE__
    return __umul24(a, b);
#else
    return a * b;
#endif
}

__host__ __device__ inline int luma(int c0, int c1, int c2)
{
#pragma clang fp contract(off)
    int n = 299 * c0 + 587 * c1 + 114 * c2;
    int q = (int)(uint32_t)(((uint64_t)(uint32_t)n * 4294968u) >> 32);
    if ((int)mul24((uint32_t)q, 1000u) != n) return q;
    double yd = 0.299 * (double)c0;
    yd = yd + 0.587 * (double)c1;
    yd = yd + 0.114 * (double)c2;
    return (int)yd;
}

__host__ __device__ inline uint32_t yuv_pack(int c0, int c1, int c2)
{
    int y = luma(c0, c1, c2);
    int u = (int)((float)(c2 - y) * 0.492f);
    int v = (int)((float)(c0 - y) * 0.877f);
    return (uint32_t)y * 65536u + (uint32_t)u * 256u + (uint32_t)v;
}

// Lane word of a pixel for a branch-free similarity test: the three masked
// fields of its packed word in 10-bit lanes (Y at bit 20, U at bit 10, V at
// bit 0) and a class in bits 30-31: 1 opaque, 2 transparent (lanes 0), 0 for
// a position outside the frame.
constexpr uint32_t kLaneOpaque = 1u << 30, kLaneTransparent = 2u << 30;

__host__ __device__ inline uint32_t yuv_lanes(uint32_t w)
{
    return ((w >> 16) & 0xFFu) << 20 | ((w >> 8) & 0xFFu) << 10 | (w & 0xFFu);
}

// |dY| <= 5, |dU| <= 7, |dV| <= 6 on the fields, all three lanes at once:
// L = a - b + 256 + t stays inside its lane (7..517, so no lane borrows from
// the next and the class bits subtract exactly, mod 4); a field is within
// threshold iff 256 <= L <= 256 + 2t, i.e. bit 8 of L is set and M = L + 255
// - 2t does not reach bit 9.  Equal classes leave bits 30-31 of L (and M)
// clear.  A pair with one end outside the frame never passes; two ends
// outside the frame do, and callers never ask for such a pair.
struct LaneRef {
    uint32_t not_s;  // ~(a + B1): ~L = b + not_s
    uint32_t m;      // a + B1 + B2: M = m - b
};

__host__ __device__ inline LaneRef lane_ref(uint32_t la)
{
    const uint32_t B1 = (256u + 5) << 20 | (256u + 7) << 10 | (256u + 6);
    const uint32_t B2 = (255u - 10) << 20 | (255u - 14) << 10 | (255u - 12);
    uint32_t s = la + B1;
    return LaneRef{ ~s, s + B2 };
}

__host__ __device__ inline uint32_t lane_similar(const LaneRef& a, uint32_t lb)
{
    uint32_t bad = (0x10040100u & (a.not_s + lb)) | ((a.m - lb) & 0xE0080200u);
    return bad == 0 ? 1u : 0u;
}

__host__ __device__ inline int iabs(int a) { return a < 0 ? -a : a; }

// diff() compare (graph_functions.cu:291-293) on packed words: true = similar.
// Thresholds trY/trU/trV (graph_functions.cu:17-19) on the masked fields.
__host__ __device__ inline bool yuv_similar(uint32_t a, uint32_t b)
{
    int dy = (int)((a & 0x00FF0000u) - (b & 0x00FF0000u));
    int du = (int)((a & 0x0000FF00u) - (b & 0x0000FF00u));
    int dv = (int)((a & 0x000000FFu) - (b & 0x000000FFu));
    return iabs(dy) <= 0x00050000 && iabs(du) <= 0x00000700 && iabs(dv) <= 0x00000006;
}

// Colour class of a pixel for the similarity test: the packed YUV word plus a
// transparency flag (alpha extension: transparent pixels are similar to each
// other, never to opaque ones).
struct Yuv {
    uint32_t w;
    uint32_t transparent;
};

__host__ __device__ inline Yuv yuv_of(const Px& p, uint32_t flags, int ch)
{
    Yuv r;
    r.w = yuv_pack(p.c0, p.c1, p.c2);
    r.transparent = ((flags & DPX_ALPHA) && ch == 4 && p.a == 0) ? 1u : 0u;
    return r;
}

__host__ __device__ inline bool similar(const Yuv& a, const Yuv& b)
{
    if (a.transparent | b.transparent) return a.transparent == b.transparent;
    return yuv_similar(a.w, b.w);
}

// Output colour of a cell: colorPos = (byte2, byte1, byte0, 255)
// (kernel.cu:98-101), alpha from the input for 4-channel frames.
__host__ __device__ inline uint32_t cell_rgba(const Px& p, int ch)
{
    uint32_t a = ch == 4 ? (uint32_t)p.a : 255u;
    return (uint32_t)p.c2 | ((uint32_t)p.c1 << 8) | ((uint32_t)p.c0 << 16) | (a << 24);
}

// checkTJunction (subdivision_functions.cu:170-242) with isColorEqual (157-167).
// The reference indexes its BGR buffer as j*ws + 3i and bounds-checks with the
// width where the stride is meant (line 187).  The same integer expressions
// are evaluated over a virtual buffer: the real bytes for the reference layout
// (3 channels, bottom-up), otherwise a packed 3W-stride view of the logical
// pixels.  Colour reads outside that buffer compare unequal.
struct TjView {
    ImgView im;
    int raw;
    long long ws, size;
};

__host__ __device__ inline TjView tj_view(const ImgView& im)
{
    TjView v;
    v.im = im;
    v.raw = (im.ch == 3) && !im.pal && !(im.flags & DPX_TOPDOWN);
    v.ws = v.raw ? (long long)im.stride : 3LL * im.W;
    v.size = v.ws * im.H;
    return v;
}

__host__ __device__ inline bool tj_equal(const TjView& v, long long a, long long b)
{
    if (v.raw) {
        if (a < 0 || b < 0 || a + 2 >= v.size || b + 2 >= v.size) return false;
        const uint8_t* p = v.im.p;
        return p[a] == p[b] && p[a + 1] == p[b + 1] && p[a + 2] == p[b + 2];
    }
    if (a < 0 || b < 0 || a >= v.size || b >= v.size) return false;
    long long pa = a / 3, pb = b / 3;
    const uint8_t* qa = px_ptr(v.im, (int)(pa % v.im.W), (int)(pa / v.im.W));
    const uint8_t* qb = px_ptr(v.im, (int)(pb % v.im.W), (int)(pb / v.im.W));
    for (int c = 0, n = colour_channels(v.im); c < n; c++)
        if (qa[c] != qb[c]) return false;
    return true;
}

// Is hull vertex (x4,y4) (quarter units) of cell (i,j) kept as a corner?
__host__ __device__ inline bool tjunction(const TjView& v, int i, int j, int x4, int y4)
{
    long long ws = v.ws;
    long long index = (long long)j * ws + 3LL * i;
    if ((index - ws - 1) < 0 || (index + v.im.W + 1) > (v.im.H * ws - 1)) return true;
    long long c0 = index + ws - 3, c1 = index + ws, c2 = index + ws + 3, c3 = index - 3, c4 = index + 3;
    long long c5 = index - ws - 3, c6 = index - ws, c7 = index - ws + 3;
    if (x4 == 0 && y4 == 0 && (!tj_equal(v, c3, c5) || !tj_equal(v, c5, c6))) return true;
    if (x4 == 4 && y4 == 0 && (!tj_equal(v, c4, c7) || !tj_equal(v, c7, c6))) return true;
    if (x4 == 4 && y4 == 4 && (!tj_equal(v, c1, c2) || !tj_equal(v, c2, c4))) return true;
    if (x4 == 0 && y4 == 4 && (!tj_equal(v, c0, c1) || !tj_equal(v, c1, c3))) return true;
    return false;
}

}  // namespace dpx
