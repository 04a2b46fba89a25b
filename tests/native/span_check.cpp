// tri_row_bits and tri_rows_or (geom.hpp) against edge_in on every sample of every row, for
// random triangles (vertices on the 1/64 lattice of a cell's local
// [-0.75, 1.75], some on the 1/4 lattice, some exactly on a sample point) at
// every scale 1..15.  Built and run on the host by test_geom_span.py.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdint.h>
#include <random>
#include "geom.hpp"
#ifndef ITERS
#define ITERS 20000
#endif
using namespace dpx;

template <int S>
static void rows_or_s(int ax, int ay, int bx, int by, int cx, int cy, int kmin, uint32_t* cov)
{
    tri_rows_or<2 * S + 2>(ax, ay, bx, by, cx, cy, kmin, cov);
}
static void rows_or(int S, int ax, int ay, int bx, int by, int cx, int cy, int kmin, uint32_t* cov)
{
    switch (S) {
#define C(k) case k: rows_or_s<k>(ax, ay, bx, by, cx, cy, kmin, cov); break;
        C(1) C(2) C(3) C(4) C(5) C(6) C(7) C(8) C(9) C(10) C(11) C(12) C(13) C(14) C(15)
#undef C
    }
}

int main()
{
    std::mt19937_64 rng(12345);
    long long tested = 0, bad = 0;
    for (int S = 1; S <= 15; S++) {
        const int kmin = -(S / 2) - 1, mw = 2 * S + 2;
        for (int it = 0; it < ITERS; it++) {
            int v[6];
            for (int k = 0; k < 6; k++) {
                int x = (int)(rng() % 161) - 48;            // 1/64 units, local -0.75 .. 1.75
                if (rng() % 3 == 0) x = (int)(rng() % 9) * 16 - 48;  // lattice points
                v[k] = x * 2 * S;
            }
            if (rng() % 4 == 0) {  // a vertex exactly on a sample point
                int b = (int)(rng() % mw), r = (int)(rng() % mw);
                v[0] = 128 * (kmin + b) + 64; v[1] = 128 * (kmin + r) + 64;
            }
            int ax = v[0], ay = v[1], bx = v[2], by = v[3], cx = v[4], cy = v[5];
            const int o = (bx - ax) * (cy - ay) - (by - ay) * (cx - ax);
            if (o == 0) continue;
            if (o < 0) { int tx = bx, ty = by; bx = cx; by = cy; cx = tx; cy = ty; }
            uint32_t refs[32] = {};
            for (int r = 0; r < mw; r++) {
                const int sy = (2 * (kmin + r) + 1) * 64;
                uint32_t ref = 0;
                for (int b = 0; b < mw; b++) {
                    const int sx = (2 * (kmin + b) + 1) * 64;
                    if (edge_in(ax, ay, bx, by, sx, sy) && edge_in(bx, by, cx, cy, sx, sy) && edge_in(cx, cy, ax, ay, sx, sy))
                        ref |= 1u << b;
                }
                const uint32_t got = tri_row_bits(ax, ay, bx, by, cx, cy, sy, kmin, mw);
                tested++;
                if (got != ref) {
                    if (bad < 5) printf("S=%d tri (%d,%d) (%d,%d) (%d,%d) row %d: got %x ref %x\n", S, ax, ay, bx, by, cx, cy, r, got, ref);
                    bad++;
                }
                refs[r] = ref;
            }
            // all rows at once, stepping the edges' bounds row by row
            uint32_t cov[32] = {};
            rows_or(S, ax, ay, bx, by, cx, cy, kmin, cov);
            for (int r = 0; r < mw; r++) {
                tested++;
                if (cov[r] != refs[r]) {
                    if (bad < 5) printf("S=%d tri_rows_or (%d,%d) (%d,%d) (%d,%d) row %d: got %x ref %x\n", S, ax, ay, bx, by, cx, cy, r, cov[r], refs[r]);
                    bad++;
                }
            }
        }
    }
    printf("rows tested %lld, mismatches %lld\n", tested, bad);
    return bad != 0;
}
