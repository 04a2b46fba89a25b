// poly_rows_xor (geom.hpp: the coverage of a whole ear-clipped polygon from
// its edges, with its winding-sum check) against the union of the ears'
// tri_rows_or rows, the clipping done as the smoothing kernel does it, at
// every scale 1..15.  Polygons: those in the file given as argv[1] (count,
// then per polygon m and m (x, y) pairs, int16, 1/64 units of a cell) and
// random polygons on the 1/64 lattice (vertices at sorted angles around a
// centre, some snapped to the 1/4 lattice, so that some are not simple).
// For every polygon whose clipping completes: where the check passes the
// rows must equal the union ("unsound" counts those that do not), and on
// the file's polygons the check must pass.  Built and run on the host by
// test_geom_poly.py.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdint.h>
#include <algorithm>
#include <random>
#include <vector>
#include "geom.hpp"
using namespace dpx;

static int px_of(uint32_t p) { return (int)(int16_t)(p & 0xFFFFu); }
static int py_of(uint32_t p) { return (int)p >> 16; }

struct Stats {
    long long polygons = 0, incomplete = 0, refused = 0, refused_equal = 0, unsound = 0;
};

template <int S>
static void run(const std::vector<std::vector<uint32_t>>& polys, Stats& st)
{
    constexpr int MW = 2 * S + 2, KMIN = -(S / 2) - 1;
    for (const auto& poly : polys) {
        std::vector<uint32_t> P = poly;
        const int m = (int)P.size();
        if (m < 3) continue;
        int a2 = 0;
        for (int p = m - 1, q = 0; q < m; p = q++) a2 += px_of(P[p]) * py_of(P[q]) - px_of(P[q]) * py_of(P[p]);
        if (a2 <= 0) std::reverse(P.begin(), P.end());
        const std::vector<uint32_t> ccw = P;
        uint32_t uni[32] = {}, par[32] = {};
        int nv = m, count = 2 * nv;
        for (int v = nv - 1; nv > 2;) {
            if (0 >= (count--)) break;
            int u = v;
            if (nv <= u) u = 0;
            v = u + 1;
            if (nv <= v) v = 0;
            int w = v + 1;
            if (nv <= w) w = 0;
            const int ax = px_of(P[u]), ay = py_of(P[u]), bx = px_of(P[v]), by = py_of(P[v]);
            const int cx = px_of(P[w]), cy = py_of(P[w]);
            bool ear = ((bx - ax) * (cy - ay) - (by - ay) * (cx - ax)) > 0;
            for (int p = 0; ear && p < nv; p++)
                if (p != u && p != v && p != w && inside_tri(ax, ay, bx, by, cx, cy, px_of(P[p]), py_of(P[p])))
                    ear = false;
            if (ear) {
                tri_rows_or<MW>(ax * 2 * S, ay * 2 * S, bx * 2 * S, by * 2 * S, cx * 2 * S, cy * 2 * S, KMIN, uni);
                P.erase(P.begin() + v);
                nv--;
                count = 2 * nv;
            }
        }
        st.polygons++;
        if (nv > 2) {
            st.incomplete++;
            continue;
        }
        const int wsum = poly_rows_xor<MW>(ccw.data(), 1, m, 2 * S, KMIN, par);
        int bits = 0;
        bool equal = true;
        for (int r = 0; r < MW; r++) {
            bits += __builtin_popcount(par[r]);
            equal &= par[r] == uni[r];
        }
        if (wsum != bits) {
            st.refused++;
            st.refused_equal += equal;
        } else if (!equal) {
            st.unsound++;
        }
    }
}

static void run_all(const std::vector<std::vector<uint32_t>>& polys, const char* name)
{
    for (int S = 1; S <= 15; S++) {
        Stats st;
        switch (S) {
#define C(k) \
    case k: run<k>(polys, st); break;
            C(1) C(2) C(3) C(4) C(5) C(6) C(7) C(8) C(9) C(10) C(11) C(12) C(13) C(14) C(15)
#undef C
        }
        printf("%s S %d: polygons %lld incomplete %lld refused %lld (of which equal %lld) unsound %lld\n", name, S,
               st.polygons, st.incomplete, st.refused, st.refused_equal, st.unsound);
    }
}

int main(int argc, char** argv)
{
    std::vector<std::vector<uint32_t>> cells;
    if (argc > 1) {
        FILE* f = fopen(argv[1], "rb");
        if (!f) return 2;
        std::vector<int16_t> buf;
        int16_t x;
        while (fread(&x, 2, 1, f) == 1) buf.push_back(x);
        fclose(f);
        for (size_t i = 0; i < buf.size();) {
            const int m = buf[i++];
            std::vector<uint32_t> p(m);
            for (int k = 0; k < m; k++, i += 2) p[k] = (uint32_t)(uint16_t)buf[i] | ((uint32_t)buf[i + 1] << 16);
            cells.push_back(p);
        }
    }
    std::mt19937_64 rng(7);
    std::vector<std::vector<uint32_t>> rnd;
    for (int it = 0; it < 20000; it++) {
        const int m = 3 + (int)(rng() % 14);
        std::vector<double> ang(m);
        for (auto& a : ang) a = (double)(rng() % 100000) / 100000.0 * 6.283185307179586;
        std::sort(ang.begin(), ang.end());
        std::vector<uint32_t> p;
        for (int k = 0; k < m; k++) {
            const double rr = 10 + (double)(rng() % 60);
            int px = 32 + (int)lround(rr * cos(ang[k])), py = 32 + (int)lround(rr * sin(ang[k]));
            if (rng() % 3 == 0) {
                px = px / 16 * 16;
                py = py / 16 * 16;
            }
            p.push_back((uint32_t)(uint16_t)px | ((uint32_t)py << 16));
        }
        if (rng() % 2) std::reverse(p.begin(), p.end());
        rnd.push_back(p);
    }
    run_all(cells, "cells");
    run_all(rnd, "random");
    return 0;
}
