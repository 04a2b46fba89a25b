// The product's colour arithmetic (pixel.hpp) against the oracle's
// restatement of the reference's (RGBtoYUV graph_functions.cu:80-98 and the
// compare of diff() 291-293), on the host:
//   1. yuv_pack (integer luma quotient, double only for exact multiples of
//      1000) equals orc_rgb_to_yuv on all 2^24 colours;
//   2. the three-lane subtract of lane_ref/lane_similar equals
//      orc_yuv_similar on every pair of words differing in one field (all
//      256 x 256 value pairs of that field, the other two fields drawn at
//      random, three draws each), on 25 M random colour pairs and 25 M pairs
//      of nearby colours (per-channel deltas within +-8, most of them close
//      to the thresholds), through orc_colors_similar;
//   3. the class bits: transparent words pass only against transparent
//      ones, positions outside the frame never pass against a pixel.
// Built with hipcc (host code) and linked with the oracle by
// test_colour_arithmetic.py.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <random>

#include "depixel_oracle.h"
#include "pixel.hpp"

using namespace dpx;

static uint32_t lanes_of(uint32_t yuv) { return yuv_lanes(yuv) | kLaneOpaque; }
static uint32_t lane_pair(uint32_t a, uint32_t b) { return lane_similar(lane_ref(a), b); }

int main()
{
    long long bad = 0, tested = 0;
    auto report = [&](const char* what, long long n, long long b) {
        printf("%s: %lld tested, %lld mismatches\n", what, n, b);
        tested += n;
        bad += b;
    };
    // 1. every colour
    {
        long long b = 0;
        for (uint32_t c = 0; c < (1u << 24); c++) {
            const int c0 = c & 0xFF, c1 = (c >> 8) & 0xFF, c2 = c >> 16;
            if (yuv_pack(c0, c1, c2) != orc_rgb_to_yuv(c)) {
                if (b < 5) printf("colour %06x: yuv %08x oracle %08x\n", c, yuv_pack(c0, c1, c2), orc_rgb_to_yuv(c));
                b++;
            }
        }
        report("yuv_pack over all 2^24 colours", 1LL << 24, b);
    }
    std::mt19937_64 rng(20261203);
    // 2a. one field at a time, every value pair
    {
        long long n = 0, b = 0;
        for (int f = 0; f < 3; f++) {
            const int sh = 16 - 8 * f;  // Y at 16, U at 8, V at 0
            for (int rep = 0; rep < 3; rep++) {
                const uint32_t other = (uint32_t)rng() & 0xFFFFFFu & ~(0xFFu << sh);
                const uint32_t other_b = rep == 2 ? ((uint32_t)rng() & 0xFFFFFFu & ~(0xFFu << sh)) : other;
                for (uint32_t va = 0; va < 256; va++)
                    for (uint32_t vb = 0; vb < 256; vb++) {
                        const uint32_t a = other | va << sh, bb = other_b | vb << sh;
                        n++;
                        if (lane_pair(lanes_of(a), lanes_of(bb)) != (uint32_t)orc_yuv_similar(a, bb)) {
                            if (b < 5) printf("field %d: %06x vs %06x\n", f, a, bb);
                            b++;
                        }
                    }
            }
        }
        report("lane test, one field (all value pairs)", n, b);
    }
    // 2b. random colour pairs, uniform and nearby
    {
        long long b = 0;
        const long long N = 25000000;
        for (int mode = 0; mode < 2; mode++) {
            for (long long k = 0; k < N; k++) {
                const uint64_t r = rng();
                const uint32_t ca = (uint32_t)r & 0xFFFFFFu;
                uint32_t cb;
                if (mode == 0) {
                    cb = (uint32_t)(r >> 32) & 0xFFFFFFu;
                } else {
                    cb = 0;
                    for (int ch = 0; ch < 3; ch++) {
                        int v = (int)((ca >> (8 * ch)) & 0xFF) + (int)((r >> (32 + 5 * ch)) & 0x1F) % 17 - 8;
                        v = v < 0 ? 0 : (v > 255 ? 255 : v);
                        cb |= (uint32_t)v << (8 * ch);
                    }
                }
                const uint32_t la = lanes_of(yuv_pack(ca & 0xFF, (ca >> 8) & 0xFF, ca >> 16));
                const uint32_t lb = lanes_of(yuv_pack(cb & 0xFF, (cb >> 8) & 0xFF, cb >> 16));
                if (lane_pair(la, lb) != (uint32_t)orc_colors_similar(ca, cb)) {
                    if (b < 5) printf("colours %06x %06x\n", ca, cb);
                    b++;
                }
            }
        }
        report("lane test, 50 M colour pairs", 2 * N, b);
    }
    // 3. classes
    {
        long long n = 0, b = 0;
        for (int k = 0; k < 1000000; k++) {
            const uint32_t w = lanes_of((uint32_t)rng() & 0xFFFFFFu);
            n += 4;
            b += lane_pair(w, kLaneTransparent) != 0;  // opaque vs transparent
            b += lane_pair(kLaneTransparent, w) != 0;
            b += lane_pair(w, 0u) != 0;                // pixel vs outside the frame
            b += lane_pair(kLaneTransparent, 0u) != 0;
        }
        n++;
        b += lane_pair(kLaneTransparent, kLaneTransparent) != 1;
        report("class bits", n, b);
    }
    printf("total %lld tested, mismatches %lld\n", tested, bad);
    return bad != 0;
}
