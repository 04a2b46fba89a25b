// remove_slash (graph.hpp: the crossing heuristic of one crossed 2x2 block)
// compiled on the host, over every crossed block of the graphs in the file
// argv[1] (int32 count, then per graph int32 W, int32 H and W*H graph bytes,
// row y at y*W).  Writes one byte per crossed block to argv[2], in graph
// order and then row-major by lower-left pixel: 1 = "/" removed, 0 = "\"
// removed.  Every graph read must be in the frame; a read outside it ends
// the program with status 2.  Built and run on the host by
// test_decision_native.py, which compares the bytes with the oracle.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <vector>
#include "graph.hpp"
using namespace dpx;

int main(int argc, char** argv)
{
    if (argc != 3) return 1;
    FILE* f = fopen(argv[1], "rb");
    FILE* o = fopen(argv[2], "wb");
    if (!f || !o) return 1;
    int32_t n = 0;
    if (fread(&n, 4, 1, f) != 1) return 1;
    long blocks = 0, reads = 0;
    for (int g = 0; g < n; g++) {
        int32_t W = 0, H = 0;
        if (fread(&W, 4, 1, f) != 1 || fread(&H, 4, 1, f) != 1) return 1;
        std::vector<uint8_t> a((size_t)W * H);
        if (fread(a.data(), 1, a.size(), f) != a.size()) return 1;
        auto get = [&](int x, int y) -> uint32_t {
            if (x < 0 || y < 0 || x >= W || y >= H) {
                fprintf(stderr, "graph %d: read outside the frame at (%d, %d)\n", g, x, y);
                exit(2);
            }
            reads++;
            return a[(size_t)y * W + x];
        };
        for (int j = 0; j + 1 < H; j++)
            for (int i = 0; i + 1 < W; i++) {
                // crossed: "/" from (i,j) (bit 2) and "\" from (i+1,j) (bit 0)
                if (!((a[(size_t)j * W + i] >> 2) & 1u) || !(a[(size_t)j * W + i + 1] & 1u)) continue;
                const uint8_t d = remove_slash(get, i, j) ? 1 : 0;
                fwrite(&d, 1, 1, o);
                blocks++;
            }
    }
    fclose(o);
    printf("graphs %d blocks %ld reads %ld\n", n, blocks, reads);
    return 0;
}
