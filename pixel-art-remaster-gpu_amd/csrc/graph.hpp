// graph.hpp -- similarity-graph crossing resolution (trivial and heuristic).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "geom.hpp"

namespace dpx {

// Number of set bits of `node` other than `edge`, and the lowest of them.
// The reference counts them with an 8-step loop (graph_functions.cu:427-441,
// 462-469, 548-555); popcount gives the same number.
__host__ __device__ inline int others(uint32_t node, int edge)
{
    return __builtin_popcount(node & 0xFFu & ~(1u << edge));
}

// The graph is read through an accessor get(x, y) -> byte, so a kernel can
// serve the bytes of its own tile from LDS and the rest from HBM.  Walks
// follow set bits, which in a graph from the similarity stage never point
// off the frame; an accessor over a caller's graph answers 0 (no links)
// outside it.

// processHeuristics2 (graph_functions.cu:761-849) for the 2x2 block whose
// lower-left pixel is (i,j): i1=(i,j) i2=(i,j+1) i3=(i+1,j) i4=(i+1,j+1).
// true: remove "/" (i1-i4, bits 2 and 5); false: remove "\" (i2-i3, bits 0 and 7).
template <class G>
__host__ __device__ inline bool remove_slash(G get, int i, int j)
{
    uint32_t n1 = get(i, j), n2 = get(i, j + 1), n3 = get(i + 1, j), n4 = get(i + 1, j + 1);
    int o1 = others(n1, 2), o4 = others(n4, 5), o3 = others(n3, 0), o2 = others(n2, 7);
    // curves (checkValence2Edge 421-453)
    if (o1 == 1 && o4 == 1) return false;
    if (o3 == 1 && o2 == 1) return true;
    // islands (checkValence2Vertex 458-479), with the precedence of 805-806
    bool v1 = o1 == 0, v4 = o4 == 0, v3 = o3 == 0, v2 = o2 == 0;
    if ((v1 || v4) && !v3 && !v2) return false;
    if (v3 || (v2 && !v1 && !v4)) return true;
    // Curve lengths; ties remove "/".  calcVal2PathSize (graph_functions.cu:
    // 541-573) follows valence-2 nodes away from an edge, adding one per
    // step to an accumulator while it is <= 30; the reference runs it from
    // each end of "/" into one accumulator ((i,j) away from edge 2, then
    // (i+1,j+1) away from edge 5), then from each end of "\" ((i+1,j) edge
    // 0, (i,j+1) edge 7).  A walk from a node with others != 1 takes no step, and past
    // the curve rule at most one end of each diagonal has others == 1, so
    // each length is one walk: min(steps, 31).  The two walks read only the
    // snapshot, so they advance side by side (two independent reads per
    // step), and stop once the comparison is settled: a finished length a
    // against a running count b >= a (or > a) cannot change the answer.
    int ax = i, ay = j, ae = 2, bx = i + 1, by = j, be = 0;
    uint32_t ar = n1 & ~(1u << 2), br = n3 & ~(1u << 0);
    bool da = false, db = false;
    if (o1 != 1) {
        ax = i + 1; ay = j + 1; ae = 5; ar = n4 & ~(1u << 5);
        da = o4 != 1;
    }
    if (o3 != 1) {
        bx = i; by = j + 1; be = 7; br = n2 & ~(1u << 7);
        db = o2 != 1;
    }
    int seg = 0, seg2 = 0;  // "/" and "\" lengths so far
    for (;;) {
        if (da && seg2 >= seg) return true;
        if (db && seg > seg2) return false;
        if (!da) {
            const int k = __builtin_ctz(ar);
            ax += edx(k); ay += edy(k); ae = 7 - k;
            seg++;
        }
        if (!db) {
            const int k = __builtin_ctz(br);
            bx += edx(k); by += edy(k); be = 7 - k;
            seg2++;
        }
        // (a finished walk stays where it stopped: reading there again is harmless)
        const uint32_t va = get(ax, ay), vb = get(bx, by);
        if (!da) {
            ar = va & ~(1u << ae);
            da = seg > 30 || __builtin_popcount(ar) != 1;
        }
        if (!db) {
            br = vb & ~(1u << be);
            db = seg2 > 30 || __builtin_popcount(br) != 1;
        }
    }
}

// Diagonal bits (0,2,5,7) the pixel at (i,j) keeps after crossCheck_4
// (graph_functions.cu:1211-1274), given its byte and the bytes of its 8
// neighbours (nb[e] = byte of neighbour e, 0 off-frame).  A 2x2 block whose
// four orthogonal edges are all present loses both diagonals.
__host__ __device__ inline uint32_t trivial_clear_mask(uint32_t c, const uint32_t* nb)
{
    uint32_t clear = 0;
    // up-left block: pixels left(3), up-left(0), self, up(1)
    if (bit(nb[3], 1) && bit(nb[3], 4) && bit(nb[0], 6) && bit(nb[0], 4) && bit(c, 3) && bit(c, 1) &&
        bit(nb[1], 6) && bit(nb[1], 3))
        clear |= 1u << 0;
    // up-right block: self, up(1), right(4), up-right(2)
    if (bit(c, 1) && bit(c, 4) && bit(nb[1], 6) && bit(nb[1], 4) && bit(nb[4], 3) && bit(nb[4], 1) &&
        bit(nb[2], 6) && bit(nb[2], 3))
        clear |= 1u << 2;
    // down-left block: down-left(5), left(3), down(6), self
    if (bit(nb[5], 1) && bit(nb[5], 4) && bit(nb[3], 6) && bit(nb[3], 4) && bit(nb[6], 3) && bit(nb[6], 1) &&
        bit(c, 6) && bit(c, 3))
        clear |= 1u << 5;
    // down-right block: down(6), self, down-right(7), right(4)
    if (bit(nb[6], 1) && bit(nb[6], 4) && bit(c, 6) && bit(c, 4) && bit(nb[7], 3) && bit(nb[7], 1) &&
        bit(nb[4], 6) && bit(nb[4], 3))
        clear |= 1u << 7;
    return clear;
}

// Diagonal bits the pixel (i,j) loses in ambiguous_cross_Kernel
// (crossCheck_Heuristics graph_functions.cu:1433-1518), reading the snapshot
// `aux`.  nb[e] = snapshot byte of neighbour e (0 off-frame); c = own byte.
template <class G>
__host__ __device__ inline uint32_t heuristic_clear_mask(G get, int i, int j, uint32_t c, const uint32_t* nb)
{
    uint32_t clear = 0;
    // up-left block, origin (i-1,j): crossed if left has bit 2, up-left bit 7, self bit 0, up bit 5
    if (bit(nb[3], 2) && bit(nb[0], 7) && bit(c, 0) && bit(nb[1], 5))
        if (!remove_slash(get, i - 1, j)) clear |= 1u << 0;
    // down-left block, origin (i-1,j-1)
    if (bit(nb[5], 2) && bit(nb[3], 7) && bit(nb[6], 0) && bit(c, 5))
        if (remove_slash(get, i - 1, j - 1)) clear |= 1u << 5;
    // down-right block, origin (i,j-1)
    if (bit(nb[6], 2) && bit(c, 7) && bit(nb[7], 0) && bit(nb[4], 5))
        if (!remove_slash(get, i, j - 1)) clear |= 1u << 7;
    // up-right block, origin (i,j)
    if (bit(c, 2) && bit(nb[1], 7) && bit(nb[4], 0) && bit(nb[2], 5))
        if (remove_slash(get, i, j)) clear |= 1u << 2;
    return clear;
}

// Border walks (cc_functions.cu): links in clockwise order from up-left are
// bits 0, 1, 2, 4, 7, 6, 5, 3 (getRealLinkIndex 64-91, getClockLinkIndex
// 154-184).  cw_bit(p) is the bit at clockwise position p, cw_pos its inverse.
__host__ __device__ inline int cw_bit(int p)
{
    return (int)((0x35674210u >> (4 * p)) & 0xFu);
}

__host__ __device__ inline int cw_pos(int b)
{
    return (int)((0x45637210u >> (4 * b)) & 0xFu);
}

// The first link of node v clockwise from up-left (getFirstLink 101-113), -1 if none.
__host__ __device__ inline int cw_first_link(uint32_t v)
{
    for (int p = 0; p < 8; p++)
        if (bit(v, cw_bit(p))) return cw_bit(p);
    return -1;
}

// The link of node v next to link a turning clockwise (dir = +1,
// nextNodeClockwise 285-309) or counter-clockwise (dir = -1,
// nextEdgeCounterClockwise 250-275); a itself when it is the only link.
__host__ __device__ inline int cw_turn(uint32_t v, int a, int dir)
{
    int p = cw_pos(a);
    for (int k = 1; k < 8; k++) {
        int b = cw_bit((p + dir * k + 8) & 7);
        if (bit(v, b)) return b;
    }
    return a;
}

}  // namespace dpx
