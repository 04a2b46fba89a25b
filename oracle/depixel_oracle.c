/*
 * depixel_oracle.c -- CPU restatement of marcoc2/pixel-art-remaster-gpu's
 * per-frame depixelization path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product.  Every function names the reference file:line it restates.  The
 * arithmetic deliberately mirrors the reference's float/double expressions
 * (the product computes the same values in exact integer fixed point), so a
 * match between the two is a cross-check of two independent formulations.
 *
 * Canonical choices where the reference's result depends on its compiler or
 * on undefined behaviour (each one is tested):
 *   - luma is IEEE double, (0.299*c0 + 0.587*c1) + 0.114*c2, no FMA
 *     contraction (build flag -ffp-contract=off)                 [T3]
 *   - slope == +-1 is dy == +-dx with dx != 0; sqrt(len2) <= 1 is len2 <= 1
 *     (both exact on the quarter lattice)                        [T12, T13]
 *   - out-of-range graph bytes read by cells_Kernel count as 0   [T9]
 *   - T-junction colour reads outside the image compare unequal [T14]
 *   - a failed ear clipping leaves its unwritten result slots at (0,0)
 *     (local), i.e. degenerate triangles at the pixel origin     [T16]
 *
 * Parity status of each function (how it is pinned) is listed in the header
 * comment of each section and exercised in tests/test_oracle_*.py.
 */
#include "depixel_oracle.h"

#include <stdlib.h>
#include <string.h>
#include <math.h>

#define CS ORC_CELL_SIZE

/* ------------------------------------------------------------------------- */
/* Pixel access with the layout flags                                         */
/* ------------------------------------------------------------------------- */

typedef struct {
    const uint8_t* img;
    int W, H, stride, channels, flags;
} orc_image;

static const uint8_t* px_ptr(const orc_image* im, int i, int j)
{
    int row = (im->flags & ORC_TOPDOWN) ? (im->H - 1 - j) : j;
    return im->img + (size_t)row * im->stride + (size_t)i * im->channels;
}

/* Reference byte triple (c0,c1,c2): the bytes the reference reads as
 * pixel[0], pixel[1], pixel[2] of a BGR buffer (graph_functions.cu:192,251). */
static void px_c012(const orc_image* im, int i, int j, int* c0, int* c1, int* c2, int* a)
{
    const uint8_t* p = px_ptr(im, i, j);
    if (im->flags & ORC_RGB) { *c0 = p[2]; *c1 = p[1]; *c2 = p[0]; }
    else                     { *c0 = p[0]; *c1 = p[1]; *c2 = p[2]; }
    *a = (im->channels == 4) ? p[3] : 255;
}

/* ------------------------------------------------------------------------- */
/* A2  similarity graph                                                       */
/*   graph_Kernel kernel.cu:140-159; diff graph_functions.cu:147-311;          */
/*   DATAtoINT 102-131; RGBtoYUV 80-98; abs_ 134-144; thresholds 11-19         */
/*   Pinned by: pure-colour luma values, Python-float luma over all grey and  */
/*   primary ramps, threshold-boundary pairs, symmetry, ref harness vectors.  */
/* ------------------------------------------------------------------------- */

#define YMASK 0x00FF0000u
#define UMASK 0x0000FF00u
#define VMASK 0x000000FFu
#define TRY   0x00050000
#define TRU   0x00000700
#define TRV   0x00000006

/* DATAtoINT(c0,c1,c2) = c0 | c1<<8 | c2<<16 (graph_functions.cu:102-131) */
static uint32_t data_to_int(int c0, int c1, int c2)
{
    return ((uint32_t)c0 & 0xFFu) | (((uint32_t)c1 << 8) & 0xFF00u) | (((uint32_t)c2 << 16) & 0xFF0000u);
}

/* RGBtoYUV graph_functions.cu:80-98.  r = byte0, g = byte1, b = byte2. */
uint32_t orc_rgb_to_yuv(uint32_t c)
{
    int r = (int)(c & 0xFFu);
    int g = (int)((c & 0xFF00u) >> 8);
    int b = (int)((c & 0xFF0000u) >> 16);
    /* double expression, evaluated left to right, no contraction */
    double yd = 0.299 * (double)(float)r;
    yd = yd + 0.587 * (double)(float)g;
    yd = yd + 0.114 * (double)(float)b;
    int y = (int)yd;
    int u = (int)((float)(b - y) * 0.492f);
    int v = (int)((float)(r - y) * 0.877f);
    /* (y << 16) + (u << 8) + v in two's complement, then unsigned */
    return (uint32_t)y * 65536u + (uint32_t)u * 256u + (uint32_t)v;
}

static int abs_i(int a) { return a < 0 ? -a : a; }

/* The comparison of diff() (graph_functions.cu:291-293) on two packed YUV
 * words: the masked fields' differences against trY/trU/trV; 1 = similar. */
int orc_yuv_similar(uint32_t y1, uint32_t y2)
{
    int dy = (int)((y1 & YMASK) - (y2 & YMASK));
    int du = (int)((y1 & UMASK) - (y2 & UMASK));
    int dv = (int)((y1 & VMASK) - (y2 & VMASK));
    int different = (abs_i(dy) > TRY) || (abs_i(du) > TRU) || (abs_i(dv) > TRV);
    return !different;
}

/* The YUV comparison of diff() (graph_functions.cu:291-293): 1 = similar. */
int orc_colors_similar(uint32_t ca, uint32_t cb)
{
    return orc_yuv_similar(orc_rgb_to_yuv(ca), orc_rgb_to_yuv(cb));
}

static const int EDX[8] = { -1, 0, 1, -1, 1, -1, 0, 1 };
static const int EDY[8] = { 1, 1, 1, 0, 0, -1, -1, -1 };

static int pixel_similar(const orc_image* im, int i, int j, int ni, int nj)
{
    int a0, a1, a2, aa, b0, b1, b2, ba;
    px_c012(im, i, j, &a0, &a1, &a2, &aa);
    px_c012(im, ni, nj, &b0, &b1, &b2, &ba);
    if ((im->flags & ORC_ALPHA) && im->channels == 4) {
        /* alpha extension: transparent pixels are similar to each other and
         * dissimilar to opaque ones; opaque pairs use the reference rule */
        int ta = (aa == 0), tb = (ba == 0);
        if (ta && tb) return 1;
        if (ta != tb) return 0;
    }
    return orc_colors_similar(data_to_int(a0, a1, a2), data_to_int(b0, b1, b2));
}

int orc_similarity_graph(const uint8_t* img, int W, int H, int stride, int channels,
                         int flags, uint8_t* graph)
{
    if (!img || !graph || W <= 0 || H <= 0 || (channels != 3 && channels != 4) ||
        stride < W * channels)
        return ORC_EINVAL;
    orc_image im = { img, W, H, stride, channels, flags };
    for (int j = 0; j < H; j++)
        for (int i = 0; i < W; i++) {
            uint8_t g = 0;
            for (int e = 0; e < 8; e++) {
                int ni = i + EDX[e], nj = j + EDY[e];
                /* off-frame neighbour: diff returns 1 -> bit 0 (graph_functions.cu:307-310) */
                if (ni < 0 || nj < 0 || ni >= W || nj >= H) continue;
                if (pixel_similar(&im, i, j, ni, nj)) g |= (uint8_t)(1u << e);
            }
            graph[(size_t)j * W + i] = g;
        }
    return ORC_OK;
}

/* ------------------------------------------------------------------------- */
/* A3  trivial crossing removal                                               */
/*   trivial_cross_Kernel kernel.cu:162-177; crossCheck_4 graph_functions.cu:1211-1274 */
/*   A block's four orthogonal edges (8 half-edges) all present -> the pixel  */
/*   clears its diagonal bit into that block.  Reads orthogonal bits only,     */
/*   writes diagonal bits only, so the in-place update is order-free.         */
/*   Pinned by: brute force over 3x3 orthogonal neighbourhoods (tests), ref.  */
/* ------------------------------------------------------------------------- */

#define BIT(v, e) (((v) >> (e)) & 1u)

static int in_frame(int i, int j, int W, int H) { return i >= 0 && j >= 0 && i < W && j < H; }

/* block with lower-left pixel (bi,bj): are all 8 orthogonal half-edges set? */
static int block_fully_connected(const uint8_t* g, int W, int H, int bi, int bj)
{
    if (!in_frame(bi, bj, W, H) || !in_frame(bi + 1, bj + 1, W, H)) return 0;
    uint8_t ll = g[(size_t)bj * W + bi], lr = g[(size_t)bj * W + bi + 1];
    uint8_t ul = g[(size_t)(bj + 1) * W + bi], ur = g[(size_t)(bj + 1) * W + bi + 1];
    return BIT(ll, 1) && BIT(ll, 4) && BIT(ul, 6) && BIT(ul, 4) &&
           BIT(lr, 3) && BIT(lr, 1) && BIT(ur, 6) && BIT(ur, 3);
}

int orc_trivial_cross(uint8_t* g, int W, int H)
{
    if (!g || W <= 0 || H <= 0) return ORC_EINVAL;
    for (int j = 0; j < H; j++)
        for (int i = 0; i < W; i++) {
            uint8_t* p = &g[(size_t)j * W + i];
            /* crossCheck_4 order: up-left (bit 0), up-right (2), down-left (5), down-right (7) */
            if (block_fully_connected(g, W, H, i - 1, j))     *p &= (uint8_t)~(1u << 0);
            if (block_fully_connected(g, W, H, i, j))         *p &= (uint8_t)~(1u << 2);
            if (block_fully_connected(g, W, H, i - 1, j - 1)) *p &= (uint8_t)~(1u << 5);
            if (block_fully_connected(g, W, H, i, j - 1))     *p &= (uint8_t)~(1u << 7);
        }
    return ORC_OK;
}

/* ------------------------------------------------------------------------- */
/* A4-A7  heuristic crossing resolution on the post-trivial snapshot          */
/*   ambiguous_cross_Kernel kernel.cu:180-189 (snapshot kernel.cu:415);        */
/*   crossCheck_Heuristics graph_functions.cu:1433-1518;                       */
/*   processHeuristics2 761-849; checkValence2Edge 421-453;                    */
/*   checkValence2Vertex 458-479; calcVal2PathSize 541-573; calc_index 46-76;  */
/*   conected_edge 26-42.                                                      */
/*   Pinned by: closed special cases (checkerboard, islands, curve lengths),  */
/*   symmetry/no-crossing invariants on the reference's resolved graphs,      */
/*   ref harness vectors on seeded frames.                                    */
/* ------------------------------------------------------------------------- */

static int calc_index(int index, int edge, int W)
{
    switch (edge) {
        case 0: return index + W - 1;
        case 1: return index + W;
        case 2: return index + W + 1;
        case 3: return index - 1;
        case 4: return index + 1;
        case 5: return index - W - 1;
        case 6: return index - W;
        case 7: return index - W + 1;
    }
    return 0;
}

/* number of set bits of node other than `edge` (the loops of 421-479, 548-555) */
static int others(uint8_t node, int edge, int* last)
{
    int sum = 0;
    for (int k = edge + 1; k < edge + 8; k++)
        if (BIT(node, k % 8)) { sum++; if (last) *last = k % 8; }
    return sum;
}

/* calcVal2PathSize (541-573): the tail recursion as a loop.  `result` is the
 * accumulator shared by both calls of one diagonal; it stops growing past 31. */
static void val2_path(const uint8_t* g, int N, int index, int edge, int* result, int W)
{
    for (;;) {
        int kn = 0;
        int sum = others(g[index], edge, &kn);
        if (*result > 30) return;
        if (sum != 1) return;
        (*result)++;
        index = calc_index(index, kn, W);
        if (index < 0 || index >= N) return;  /* unreachable on a valid graph */
        edge = 7 - kn;
    }
}

/* processHeuristics2 (761-849) for the block with lower-left pixel (i,j).
 * i1=(i,j) i2=(i,j+1) i3=(i+1,j) i4=(i+1,j+1).
 * Returns 1 -> remove "/" (i1-i4, bits 2/5); 0 -> remove "\" (i2-i3, bits 7/0). */
int orc_heuristic_decision(const uint8_t* g, int W, int H, int i, int j)
{
    int N = W * H;
    int i1 = j * W + i, i2 = (j + 1) * W + i, i3 = j * W + i + 1, i4 = (j + 1) * W + i + 1;
    /* curves: both ends of a diagonal have valence 2 (checkValence2Edge) */
    if (others(g[i1], 2, 0) == 1 && others(g[i4], 5, 0) == 1) return 0;
    if (others(g[i3], 0, 0) == 1 && others(g[i2], 7, 0) == 1) return 1;
    /* islands (checkValence2Vertex = no other edge) */
    int v1 = others(g[i1], 2, 0) == 0, v4 = others(g[i4], 5, 0) == 0;
    int v3 = others(g[i3], 0, 0) == 0, v2 = others(g[i2], 7, 0) == 0;
    if ((v1 || v4) && !v3 && !v2) return 0;
    /* precedence as written at 805-806: A || (B && !C && !D) */
    if (v3 || (v2 && !v1 && !v4)) return 1;
    /* curve lengths; ties remove "/" */
    int seg = 0, seg2 = 0;
    val2_path(g, N, i1, 2, &seg, W);
    val2_path(g, N, i4, 5, &seg, W);
    val2_path(g, N, i3, 0, &seg2, W);
    val2_path(g, N, i2, 7, &seg2, W);
    (void)H;
    return (seg2 < seg) ? 0 : 1;
}

/* does the block with lower-left pixel (bi,bj) hold both diagonals in g? */
static int block_crossed(const uint8_t* g, int W, int H, int bi, int bj)
{
    if (!in_frame(bi, bj, W, H) || !in_frame(bi + 1, bj + 1, W, H)) return 0;
    uint8_t ll = g[(size_t)bj * W + bi], lr = g[(size_t)bj * W + bi + 1];
    uint8_t ul = g[(size_t)(bj + 1) * W + bi], ur = g[(size_t)(bj + 1) * W + bi + 1];
    return BIT(ll, 2) && BIT(ur, 5) && BIT(lr, 0) && BIT(ul, 7);
}

static int graph_valid(const uint8_t* g, int W, int H)
{
    for (int j = 0; j < H; j++)
        for (int i = 0; i < W; i++)
            for (int e = 0; e < 8; e++)
                if (BIT(g[(size_t)j * W + i], e) && !in_frame(i + EDX[e], j + EDY[e], W, H))
                    return 0;
    return 1;
}

int orc_heuristic_cross(uint8_t* g, int W, int H)
{
    if (!g || W <= 0 || H <= 0) return ORC_EINVAL;
    if (!graph_valid(g, W, H)) return ORC_EGRAPH;
    size_t N = (size_t)W * H;
    uint8_t* aux = (uint8_t*)malloc(N);
    if (!aux) return ORC_EINVAL;
    memcpy(aux, g, N);  /* graph_d_aux snapshot, kernel.cu:415 */
    for (int j = 0; j < H; j++)
        for (int i = 0; i < W; i++) {
            uint8_t* p = &g[(size_t)j * W + i];
            /* 1476-1485: block up-left of the pixel, origin (i-1,j); false -> clear bit 0 */
            if (block_crossed(aux, W, H, i - 1, j) && !orc_heuristic_decision(aux, W, H, i - 1, j))
                *p &= (uint8_t)~(1u << 0);
            /* 1487-1496: block down-left, origin (i-1,j-1); true -> clear bit 5 */
            if (block_crossed(aux, W, H, i - 1, j - 1) && orc_heuristic_decision(aux, W, H, i - 1, j - 1))
                *p &= (uint8_t)~(1u << 5);
            /* 1498-1507: block down-right, origin (i,j-1); false -> clear bit 7 */
            if (block_crossed(aux, W, H, i, j - 1) && !orc_heuristic_decision(aux, W, H, i, j - 1))
                *p &= (uint8_t)~(1u << 7);
            /* 1509-1518: block up-right, origin (i,j); true -> clear bit 2 */
            if (block_crossed(aux, W, H, i, j) && orc_heuristic_decision(aux, W, H, i, j))
                *p &= (uint8_t)~(1u << 2);
        }
    free(aux);
    return ORC_OK;
}

/* ------------------------------------------------------------------------- */
/* A8-A9  reshaped cell: point pattern, quirky sort, monotone-chain hull      */
/*   cells_Kernel kernel.cu:192-213; createCellFromPattern                    */
/*   diagram_functions.cu:319-535; convex_hull 238-316; sort 82-129;          */
/*   sort_y 36-80; cross 225-230; multiply 204-212; swap 28-34.               */
/*   Pinned by: the 4096-row table written by the reference's own code (ref   */
/*   harness), plus lattice / extent / orientation properties.               */
/* ------------------------------------------------------------------------- */

static void push(orc_pt* P, int* n, float x, float y) { P[*n].x = x; P[*n].y = y; (*n)++; }

int orc_cell_points(uint8_t node, uint8_t left, uint8_t right, orc_pt* P)
{
    int n = 0;
    /* case 0 (up-left), 328-360 */
    if (BIT(node, 0)) {
        if (BIT(node, 1) && !BIT(node, 3))      push(P, &n, -0.25f, 0.75f);
        else if (BIT(node, 3) && !BIT(node, 1)) push(P, &n, 0.25f, 1.25f);
        else { push(P, &n, -0.25f, 0.75f); push(P, &n, 0.25f, 1.25f); }
    } else {
        if (BIT(left, 2)) push(P, &n, 0.25f, 0.75f); else push(P, &n, 0.0f, 1.0f);
    }
    /* case 1 (up), 363-374 */
    if (BIT(node, 1)) { push(P, &n, 0.0f, 1.0f); push(P, &n, 1.0f, 1.0f); }
    else push(P, &n, 0.5f, 0.75f);
    /* case 2 (up-right), 376-408 */
    if (BIT(node, 2)) {
        if (BIT(node, 1) && !BIT(node, 4))      push(P, &n, 1.25f, 0.75f);
        else if (BIT(node, 4) && !BIT(node, 1)) push(P, &n, 0.75f, 1.25f);
        else { push(P, &n, 0.75f, 1.25f); push(P, &n, 1.25f, 0.75f); }
    } else {
        if (BIT(right, 0)) push(P, &n, 0.75f, 0.75f); else push(P, &n, 1.0f, 1.0f);
    }
    /* case 4 (right), 411-422 */
    if (BIT(node, 4)) { push(P, &n, 1.0f, 1.0f); push(P, &n, 1.0f, 0.0f); }
    else push(P, &n, 0.75f, 0.75f);
    /* case 7 (down-right), 424-456 */
    if (BIT(node, 7)) {
        if (BIT(node, 4) && !BIT(node, 6))      push(P, &n, 0.75f, -0.25f);
        else if (BIT(node, 6) && !BIT(node, 4)) push(P, &n, 1.25f, 0.25f);
        else { push(P, &n, 1.25f, 0.25f); push(P, &n, 0.75f, -0.25f); }
    } else {
        if (BIT(right, 5)) push(P, &n, 0.75f, 0.25f); else push(P, &n, 1.0f, 0.0f);
    }
    /* case 6 (down), 459-470 */
    if (BIT(node, 6)) { push(P, &n, 1.0f, 0.0f); push(P, &n, 0.0f, 0.0f); }
    else push(P, &n, 0.25f, 0.75f);
    /* case 5 (down-left), 472-504 */
    if (BIT(node, 5)) {
        if (BIT(node, 3) && !BIT(node, 6))      push(P, &n, 0.25f, -0.25f);
        else if (BIT(node, 6) && !BIT(node, 3)) push(P, &n, -0.25f, 0.25f);
        else { push(P, &n, 0.25f, -0.25f); push(P, &n, -0.25f, 0.25f); }
    } else {
        if (BIT(left, 7)) push(P, &n, 0.25f, 0.25f); else push(P, &n, 0.0f, 0.0f);
    }
    /* case 3 (left), 507-518 */
    if (BIT(node, 3)) { push(P, &n, 0.0f, 0.0f); push(P, &n, 0.0f, 1.0f); }
    else push(P, &n, 0.25f, 0.5f);
    return n;
}

static void swap_pt(orc_pt* a, orc_pt* b) { orc_pt t = *a; *a = *b; *b = t; }

/* sort_y (36-80): one selection pass moving the minimum-y point(s) of
 * P[start..end] to the front. */
static void sort_y(orc_pt* P, int start, int end)
{
    float minor = 99.0f;
    int tie = 0;
    int tie_index[CS];
    tie_index[0] = start;
    for (int i = start; i <= end; i++) {
        if (P[i].y < minor) { minor = P[i].y; tie = 0; tie_index[tie] = i; }
        else if (P[i].y == minor) { tie++; tie_index[tie] = i; }
    }
    if (tie > 0) {
        for (int i = 0; i <= tie; i++) swap_pt(&P[i + start], &P[tie_index[i]]);
    } else {
        swap_pt(&P[start], &P[tie_index[0]]);
    }
}

/* sort (82-129): repeated selection of the minimum-x group, then sort_y on it */
static void sort_x(orc_pt* P, int size_P)
{
    float minor = 99.0f;
    int tie = 0;
    int tie_index[CS];
    int sorted = 0;
    tie_index[0] = 0;
    while (sorted < size_P) {
        for (int i = sorted; i < size_P; i++) {
            if (P[i].x < minor) { minor = P[i].x; tie = 0; tie_index[tie] = i; }
            else if (P[i].x == minor) { tie++; tie_index[tie] = i; }
        }
        if (tie > 0) {
            for (int i = 0; i <= tie; i++) swap_pt(&P[i + sorted], &P[tie_index[i]]);
            sort_y(P, sorted, sorted + tie);
        } else {
            swap_pt(&P[sorted], &P[tie_index[0]]);
        }
        sorted += tie + 1;
        minor = 99.0f;
    }
}

/* multiply (204-212) and cross (225-230): OA x OB on x100 coordinates, float
 * products of a float and an int-truncated difference, truncated to int. */
static orc_pt mul100(orc_pt p) { orc_pt r; r.x = p.x * 100.0f; r.y = p.y * 100.0f; return r; }

static int cross100(orc_pt O, orc_pt A, orc_pt B)
{
    O = mul100(O); A = mul100(A); B = mul100(B);
    float t1 = (A.x - O.x) * (float)(int)(B.y - O.y);
    float t2 = (A.y - O.y) * (float)(int)(B.x - O.x);
    return (int)(t1 - t2);
}

int orc_convex_hull(const orc_pt* pts, int n, orc_pt* out)
{
    orc_pt P[CS];
    orc_pt H[CS];
    int k = 0;
    if (n <= 0 || n > 16) return ORC_EINVAL;
    memset(P, 0, sizeof(P));
    memset(H, 0, sizeof(H));
    memcpy(P, pts, sizeof(orc_pt) * (size_t)n);
    sort_x(P, n);
    /* lower hull (271-279); the unused `teste` read of H[k-2] is dropped */
    for (int i = 0; i < n; i++) {
        while (k >= 2 && cross100(H[k - 2], H[k - 1], P[i]) <= 0) k--;
        H[k++] = P[i];
    }
    /* upper hull (288-293) */
    for (int i = n - 2, t = k + 1; i >= 0; i--) {
        while (k >= t && cross100(H[k - 2], H[k - 1], P[i]) <= 0) k--;
        H[k++] = P[i];
    }
    /* the reference copies all 45 slots; slots >= k are uninitialised there
     * and zero here (never read downstream) */
    for (int i = 0; i < CS; i++) out[i] = (i < k) ? H[i] : (orc_pt){ 0.0f, 0.0f };
    return k - 1;
}

int orc_cell_from_pattern(uint8_t node, uint8_t left, uint8_t right, orc_pt* out45)
{
    orc_pt P[16];
    int n = orc_cell_points(node, left, right, P);
    return orc_convex_hull(P, n, out45);
}

int orc_cells(const uint8_t* g, int W, int H, orc_pt* diagram, int32_t* edge_count)
{
    if (!g || !diagram || !edge_count || W <= 0 || H <= 0) return ORC_EINVAL;
    size_t N = (size_t)W * H;
    for (size_t idx = 0; idx < N; idx++) {
        /* graph[idx-1], graph[idx+1] without bounds (kernel.cu:204-206);
         * out-of-range bytes count as 0 */
        uint8_t left = idx > 0 ? g[idx - 1] : 0;
        uint8_t right = idx + 1 < N ? g[idx + 1] : 0;
        edge_count[idx] = orc_cell_from_pattern(g[idx], left, right, &diagram[idx * CS]);
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------------- */
/* A10-A14  subdivision (corner cutting) on the hull snapshot                 */
/*   subdivision_Kernel kernel.cu:216-261 (snapshots 434-437);                 */
/*   subdivision subdivision_functions.cu:564-669; checkEdges 541-551;         */
/*   isLinkedEdge 245-424; getQ_i 42-83; getR_i 86-122;                        */
/*   get{R,Q}_i_from_linked_cell 125-154; checkTJunction 170-242;              */
/*   isColorEqual 157-167; getOppositePoint 427-474; getOppositePointCoord     */
/*   477-524; getPointIndex 527-538; midPoint 554-561; mod 3-6;                */
/*   c_neighbor_index 9-39.                                                    */
/*   Pinned by: ref harness vectors on seeded frames; closed forms of the     */
/*   1/4-3/4 and 1/8-7/8 cuts; 1/64-lattice invariant.                        */
/* ------------------------------------------------------------------------- */

static int mod_n(int i, int n) { return (n + (i % n)) % n; }

static int c_neighbor_index(int index, int edge, int W) { return calc_index(index, edge, W); }

static int len_le_1(orc_pt P, orc_pt Pn)
{
    float dx = Pn.x - P.x, dy = Pn.y - P.y;
    float l2 = dx * dx + dy * dy;   /* exact on the quarter lattice */
    return l2 <= 1.0f;              /* sqrt(l2) <= 1.0 (T13) */
}

static orc_pt getQ_i(int i, const orc_pt* d, int n)
{
    int p = i % n;
    if (p < 0) p += n;
    orc_pt P = d[p], Pn = d[(i + 1) % n], Q;
    if (len_le_1(P, Pn)) {
        Q.x = (float)((double)P.x / 2.0 + (double)(P.x + Pn.x) / 4.0);
        Q.y = (float)((double)P.y / 2.0 + (double)(P.y + Pn.y) / 4.0);
    } else {
        Q.x = (float)((7.0 * (double)P.x) / 8.0 + (double)Pn.x / 8.0);
        Q.y = (float)((7.0 * (double)P.y) / 8.0 + (double)Pn.y / 8.0);
    }
    return Q;
}

static orc_pt getR_i(int i, const orc_pt* d, int n)
{
    int p = i % n;
    if (p < 0) p += n;
    orc_pt P = d[p], Pn = d[(i + 1) % n], R;
    if (len_le_1(P, Pn)) {
        R.x = (float)((double)Pn.x / 2.0 + (double)(P.x + Pn.x) / 4.0);
        R.y = (float)((double)Pn.y / 2.0 + (double)(P.y + Pn.y) / 4.0);
    } else {
        R.x = (float)((1.0 * (double)P.x) / 8.0 + (7.0 * (double)Pn.x) / 8.0);
        R.y = (float)((1.0 * (double)P.y) / 8.0 + (7.0 * (double)Pn.y) / 8.0);
    }
    return R;
}

/* own frame -> neighbour-across-`edge` frame (getOppositePoint 427-474) */
static orc_pt opposite_point(orc_pt p, int e) { orc_pt r = { p.x - (float)EDX[e], p.y - (float)EDY[e] }; return r; }
/* neighbour frame -> own frame (getOppositePointCoord 477-524) */
static orc_pt opposite_coord(orc_pt p, int e) { orc_pt r = { p.x + (float)EDX[e], p.y + (float)EDY[e] }; return r; }

static int point_index(orc_pt p, const orc_pt* cell, int n)
{
    for (int i = 0; i < n; i++)
        if (p.x == cell[i].x && p.y == cell[i].y) return i;
    return 0;  /* not found -> 0 (T15) */
}

static orc_pt mid_point(orc_pt a, orc_pt b)
{
    orc_pt r;
    r.x = (float)((double)(a.x + b.x) / 2.0);
    r.y = (float)((double)(a.y + b.y) / 2.0);
    return r;
}

/* isLinkedEdge (245-424).  Returns the link direction or -1 (border). */
static int linked_edge(const orc_pt* c, uint8_t node, int idx, int n)
{
    orc_pt a = c[idx], b = c[mod_n(idx + 1, n)];
    float dx = b.x - a.x, dy = b.y - a.y;
    int slope_p1 = (dx != 0.0f) && (dy == dx);    /* dy/dx == 1  (T12) */
    int slope_m1 = (dx != 0.0f) && (dy == -dx);   /* dy/dx == -1 */
    double mid_y = (double)(a.y + b.y) / 2.0;
    if (a.y == b.y && a.y > 0.5f && BIT(node, 1)) return 1;
    if (a.x == b.x && a.x > 0.5f && BIT(node, 4)) return 4;
    if (a.y == b.y && a.y < 0.5f && BIT(node, 6)) return 6;
    if (a.x == b.x && a.x < 0.5f && BIT(node, 3)) return 3;
    if (slope_p1 && mid_y > 0.5 && BIT(node, 0)) return 0;
    if (slope_m1 && mid_y > 0.5 && BIT(node, 2)) return 2;
    if (slope_p1 && mid_y < 0.5 && BIT(node, 7)) return 7;
    if (slope_m1 && mid_y < 0.5 && BIT(node, 5)) return 5;
    return -1;
}

/* checkTJunction (170-242) with isColorEqual (157-167).  The reference
 * indexes the BGR buffer as j*ws + 3i and tests its bounds with `width`
 * where the stride is meant (187).  The same integer expressions are
 * evaluated here over a virtual buffer: the real one for the reference
 * layout (3 channels, bottom-up), or a packed 3W-stride view of logical
 * pixels otherwise.  Reads outside the buffer compare unequal (T14). */
typedef struct {
    const orc_image* im;
    int raw;      /* 1: real bytes with the real stride */
    long long ws; /* stride of the virtual buffer */
    long long size;
} tj_view;

static int tj_color_equal(const tj_view* v, long long a, long long b)
{
    if (v->raw) {
        if (a < 0 || b < 0 || a + 2 >= v->size || b + 2 >= v->size) return 0;
        const uint8_t* p = v->im->img;
        return p[a] == p[b] && p[a + 1] == p[b + 1] && p[a + 2] == p[b + 2];
    }
    /* packed view: every offset is a multiple of 3 -> logical pixel a/3 */
    if (a < 0 || b < 0 || a >= v->size || b >= v->size) return 0;
    long long pa = a / 3, pb = b / 3;
    int W = v->im->W;
    const uint8_t* qa = px_ptr(v->im, (int)(pa % W), (int)(pa / W));
    const uint8_t* qb = px_ptr(v->im, (int)(pb % W), (int)(pb / W));
    for (int c = 0; c < v->im->channels; c++)
        if (qa[c] != qb[c]) return 0;
    return 1;
}

static int check_tjunction(const tj_view* v, int i, int j, orc_pt p)
{
    long long ws = v->ws;
    long long H = v->im->H, W = v->im->W;
    long long index = (long long)j * ws + (long long)i * 3;
    if ((index - ws - 1) < 0 || (index + W + 1) > (H * ws - 1)) return 1;
    long long c0 = index + ws - 3, c1 = index + ws, c2 = index + ws + 3;
    long long c3 = index - 3, c4 = index + 3;
    long long c5 = index - ws - 3, c6 = index - ws, c7 = index - ws + 3;
    if (p.x == 0.0f && p.y == 0.0f)
        if (!tj_color_equal(v, c3, c5) || !tj_color_equal(v, c5, c6)) return 1;
    if (p.x == 1.0f && p.y == 0.0f)
        if (!tj_color_equal(v, c4, c7) || !tj_color_equal(v, c7, c6)) return 1;
    if (p.x == 1.0f && p.y == 1.0f)
        if (!tj_color_equal(v, c1, c2) || !tj_color_equal(v, c2, c4)) return 1;
    if (p.x == 0.0f && p.y == 1.0f)
        if (!tj_color_equal(v, c0, c1) || !tj_color_equal(v, c1, c3)) return 1;
    return 0;
}

int orc_subdivide(const uint8_t* img, int W, int H, int stride, int channels, int flags,
                  const uint8_t* g, orc_pt* diagram, int32_t* edge_count,
                  uint8_t* edge_status_out, int32_t* link_index_out)
{
    if (!img || !g || !diagram || !edge_count || W <= 0 || H <= 0) return ORC_EINVAL;
    size_t N = (size_t)W * H;
    orc_image im = { img, W, H, stride, channels, flags };
    tj_view v;
    v.im = &im;
    v.raw = (channels == 3) && !(flags & ORC_TOPDOWN);
    v.ws = v.raw ? stride : 3LL * W;
    v.size = v.ws * H;
    /* snapshots diagram_aux / edge_count_aux (kernel.cu:434-437) */
    orc_pt* daux = (orc_pt*)malloc(N * CS * sizeof(orc_pt));
    int32_t* eaux = (int32_t*)malloc(N * sizeof(int32_t));
    if (!daux || !eaux) { free(daux); free(eaux); return ORC_EINVAL; }
    memcpy(daux, diagram, N * CS * sizeof(orc_pt));
    memcpy(eaux, edge_count, N * sizeof(int32_t));

    for (int j = 0; j < H; j++)
        for (int i = 0; i < W; i++) {
            int node_index = j * W + i;
            uint8_t node = g[node_index];
            if (node == 90) continue;  /* internal node (kernel.cu:231) */
            const orc_pt* own = &daux[(size_t)node_index * CS];
            int n = eaux[node_index];
            int status[CS], link[CS];
            for (int e = 0; e < n; e++) {
                link[e] = linked_edge(own, node, e, n);
                status[e] = link[e] >= 0;
            }
            if (edge_status_out)
                for (int e = 0; e < CS; e++)
                    edge_status_out[(size_t)node_index * CS + e] = (uint8_t)(e < n ? status[e] : 0);
            if (link_index_out)
                for (int e = 0; e < CS; e++)
                    link_index_out[(size_t)node_index * CS + e] = e < n ? link[e] : 0;

            orc_pt* cell = &diagram[(size_t)node_index * CS];
            int out = 0;
            for (int k = 0; k < n; k++) {
                int prev = mod_n(k - 1, n);
                if (!status[k] && !status[prev]) {
                    if (check_tjunction(&v, i, j, own[k])) {
                        cell[out++] = own[k];
                    } else {
                        orc_pt Q = getQ_i(k, own, n);
                        orc_pt R = getR_i(k - 1, own, n);
                        cell[out++] = R;
                        cell[out++] = Q;
                    }
                } else if (!status[k] && status[prev]) {
                    orc_pt Q = getQ_i(k, own, n);
                    int le = link[prev];
                    int li = c_neighbor_index(node_index, le, W);
                    const orc_pt* lc = &daux[(size_t)li * CS];
                    int ln = eaux[li];
                    int op = point_index(opposite_point(own[k], le), lc, ln);
                    orc_pt Ra = opposite_coord(getR_i(op - 1, lc, ln), le);
                    Ra = mid_point(Q, Ra);
                    cell[out++] = Ra;
                    cell[out++] = Q;
                } else if (status[k] && !status[prev]) {
                    int le = link[k];
                    int li = c_neighbor_index(node_index, le, W);
                    const orc_pt* lc = &daux[(size_t)li * CS];
                    int ln = eaux[li];
                    int op = point_index(opposite_point(own[k], le), lc, ln);
                    orc_pt Qa = opposite_coord(getQ_i(op, lc, ln), le);
                    orc_pt R = getR_i(k - 1, own, n);
                    Qa = mid_point(R, Qa);
                    cell[out++] = R;
                    cell[out++] = Qa;
                } else {
                    cell[out++] = own[k];
                }
            }
            edge_count[node_index] = out;
        }
    free(daux);
    free(eaux);
    return ORC_OK;
}

/* ------------------------------------------------------------------------- */
/* A15  ear-clipping triangulation                                            */
/*   triangulate_Kernel kernel.cu:264-282; triangulate_polygon                 */
/*   triangulate_functions.cu:220-276; process 108-217; snip 70-105;          */
/*   insideTriangle 41-67; area 6-34; EPSILON 4.                              */
/*   Pinned by: ref harness vectors; n-2 triangles whose areas sum to the    */
/*   polygon area; the ear sequence of a convex polygon.                      */
/* ------------------------------------------------------------------------- */

static const float EPS = 0.0000000001f;

static float poly_area(const orc_pt* c, int n)
{
    float A = 0.0f;
    for (int p = n - 1, q = 0; q < n; p = q++) A += c[p].x * c[q].y - c[q].x * c[p].y;
    return A * 0.5f;
}

static int inside_triangle(float Ax, float Ay, float Bx, float By, float Cx, float Cy,
                           float Px, float Py)
{
    float ax = Cx - Bx, ay = Cy - By, bx = Ax - Cx, by = Ay - Cy, cx = Bx - Ax, cy = By - Ay;
    float apx = Px - Ax, apy = Py - Ay, bpx = Px - Bx, bpy = Py - By, cpx = Px - Cx, cpy = Py - Cy;
    float aCROSSbp = ax * bpy - ay * bpx;
    float cCROSSap = cx * apy - cy * apx;
    float bCROSScp = bx * cpy - by * cpx;
    return (aCROSSbp >= 0.0f) && (bCROSScp >= 0.0f) && (cCROSSap >= 0.0f);
}

static int snip(const orc_pt* c, int u, int v, int w, int n, const int* V)
{
    float Ax = c[V[u]].x, Ay = c[V[u]].y, Bx = c[V[v]].x, By = c[V[v]].y;
    float Cx = c[V[w]].x, Cy = c[V[w]].y;
    if (EPS > (((Bx - Ax) * (Cy - Ay)) - ((By - Ay) * (Cx - Ax)))) return 0;
    for (int p = 0; p < n; p++) {
        if (p == u || p == v || p == w) continue;
        if (inside_triangle(Ax, Ay, Bx, By, Cx, Cy, c[V[p]].x, c[V[p]].y)) return 0;
    }
    return 1;
}

/* process (108-217).  Returns 1 on success; result45 is zero-initialised so a
 * failure leaves degenerate (0,0) slots (T16). */
int orc_triangulate_polygon(const orc_pt* c, int n, orc_pt* result)
{
    int V[CS];
    for (int t = 0; t < CS; t++) { result[t].x = 0.0f; result[t].y = 0.0f; }
    if (n < 3) return 0;
    if (n > CS) return 0;
    if (0.0f < poly_area(c, n)) for (int v = 0; v < n; v++) V[v] = v;
    else for (int v = 0; v < n; v++) V[v] = (n - 1) - v;
    int nv = n, index = 0;
    int count = 2 * nv;
    for (int v = nv - 1; nv > 2;) {
        if (0 >= (count--)) return 0;
        int u = v;
        if (nv <= u) u = 0;
        v = u + 1;
        if (nv <= v) v = 0;
        int w = v + 1;
        if (nv <= w) w = 0;
        if (snip(c, u, v, w, nv, V)) {
            result[index++] = c[V[u]];
            result[index++] = c[V[v]];
            result[index++] = c[V[w]];
            for (int s = v, t = v + 1; t < nv; s++, t++) V[s] = V[t];
            nv--;
            count = 2 * nv;
        }
    }
    return 1;
}

int orc_triangulate(orc_pt* diagram, const int32_t* edge_count, int W, int H, int* n_failed)
{
    if (!diagram || !edge_count || W <= 0 || H <= 0) return ORC_EINVAL;
    int failed = 0;
    for (int j = 0; j < H; j++)
        for (int i = 0; i < W; i++) {
            size_t idx = (size_t)j * W + i;
            orc_pt* cell = &diagram[idx * CS];
            orc_pt result[CS];
            if (!orc_triangulate_polygon(cell, edge_count[idx], result)) failed++;
            /* all 45 slots: result + pixel offset (262-268) */
            for (int t = 0; t < CS; t++) {
                cell[t].x = result[t].x + (float)i;
                cell[t].y = result[t].y + (float)j;
            }
        }
    if (n_failed) *n_failed = failed;
    return ORC_OK;
}

/* ------------------------------------------------------------------------- */
/* A16  VBO packing                                                           */
/*   color_kernel kernel.cu:67-104; position_kernel kernel.cu:109-137.        */
/*   colour = (byte2, byte1, byte0, 255); unused slots at (-100,-100).        */
/* ------------------------------------------------------------------------- */

static uint32_t cell_rgba(const orc_image* im, int i, int j)
{
    int c0, c1, c2, a;
    px_c012(im, i, j, &c0, &c1, &c2, &a);
    if (!(im->channels == 4)) a = 255;
    return (uint32_t)c2 | ((uint32_t)c1 << 8) | ((uint32_t)c0 << 16) | ((uint32_t)a << 24);
}

int orc_export_vbo(const uint8_t* img, int W, int H, int stride, int channels, int flags,
                   const orc_pt* diagram, const int32_t* edge_count, float* pos, uint8_t* color)
{
    if (!img || !diagram || !edge_count || !pos || !color) return ORC_EINVAL;
    orc_image im = { img, W, H, stride, channels, flags };
    for (int j = 0; j < H; j++)
        for (int i = 0; i < W; i++) {
            size_t idx = (size_t)j * W + i;
            uint32_t rgba = cell_rgba(&im, i, j);
            int valid = (edge_count[idx] - 2) * 3;
            for (int t = 0; t < CS; t++) {
                size_t s = idx * CS + (size_t)t;
                if (t < valid) { pos[2 * s] = diagram[s].x; pos[2 * s + 1] = diagram[s].y; }
                else { pos[2 * s] = -100.0f; pos[2 * s + 1] = -100.0f; }
                memcpy(&color[4 * s], &rgba, 4);
            }
        }
    return ORC_OK;
}

/* ------------------------------------------------------------------------- */
/* A17  rasterization in draw order                                           */
/*   glDrawArrays(GL_TRIANGLES) simpleVBO.cpp:236-285 (scale 258, draw 281),   */
/*   clear colour main.cpp:260.  GL's own rasteriser is closed: this is the   */
/*   definition used by the build (parity against GL is unpinned):            */
/*   - output pixel (X,Y) samples world point ((X+0.5)/s, (Y+0.5)/s); Y=0 is  */
/*     the bottom row (the top row with ORC_TOPDOWN);                         */
/*   - a sample on an edge is inside iff it is inside after the symbolic      */
/*     shift (+e, +e^2): for a CCW edge (dx,dy), dy < 0 or (dy == 0, dx > 0);  */
/*   - triangles are drawn in VBO order, later ones overwrite;                */
/*   - uncovered pixels keep the clear colour (0,0,0,255).                    */
/*   Pinned by: nearest-neighbour identity on square-cell images, exact       */
/*   rational point-in-polygon checks in tests.                               */
/* ------------------------------------------------------------------------- */

static int edge_in(long long px, long long py, long long qx, long long qy, long long sx, long long sy)
{
    long long dx = qx - px, dy = qy - py;
    long long e = dx * (sy - py) - dy * (sx - px);
    if (e > 0) return 1;
    if (e < 0) return 0;
    return (dy < 0) || (dy == 0 && dx > 0);
}

static long long to_units(float v, int scale, int* ok)
{
    double q = (double)v * 64.0;
    double r = floor(q);
    if (r != q) *ok = 0;
    return (long long)r * 2LL * scale;
}

int orc_rasterize(const uint8_t* img, int W, int H, int stride, int channels, int flags,
                  const orc_pt* diagram, const int32_t* edge_count, int scale,
                  uint8_t* out, int out_stride)
{
    if (!img || !diagram || !edge_count || !out || scale <= 0) return ORC_EINVAL;
    orc_image im = { img, W, H, stride, channels, flags };
    int OW = W * scale, OH = H * scale;
    for (int y = 0; y < OH; y++)
        for (int x = 0; x < OW; x++) {
            uint8_t* o = out + (size_t)y * out_stride + (size_t)x * 4;
            o[0] = 0; o[1] = 0; o[2] = 0; o[3] = 255;
        }
    int bad = 0;
    for (int j = 0; j < H; j++)
        for (int i = 0; i < W; i++) {
            size_t idx = (size_t)j * W + i;
            uint32_t rgba = cell_rgba(&im, i, j);
            int ntri = edge_count[idx] - 2;
            for (int t = 0; t < ntri && t < CS / 3; t++) {
                const orc_pt* v = &diagram[idx * CS + (size_t)t * 3];
                int ok = 1;
                long long ax = to_units(v[0].x, scale, &ok), ay = to_units(v[0].y, scale, &ok);
                long long bx = to_units(v[1].x, scale, &ok), by = to_units(v[1].y, scale, &ok);
                long long cx = to_units(v[2].x, scale, &ok), cy = to_units(v[2].y, scale, &ok);
                if (!ok) { bad++; continue; }
                long long o = (bx - ax) * (cy - ay) - (by - ay) * (cx - ax);
                if (o == 0) continue;
                if (o < 0) { long long tx = bx, ty = by; bx = cx; by = cy; cx = tx; cy = ty; }
                /* sample bounding box: (2X+1)*64 in [min, max] */
                long long mnx = ax < bx ? (ax < cx ? ax : cx) : (bx < cx ? bx : cx);
                long long mxx = ax > bx ? (ax > cx ? ax : cx) : (bx > cx ? bx : cx);
                long long mny = ay < by ? (ay < cy ? ay : cy) : (by < cy ? by : cy);
                long long mxy = ay > by ? (ay > cy ? ay : cy) : (by > cy ? by : cy);
                long long x0 = (long long)floor(((double)mnx / 64.0 - 1.0) / 2.0) - 1;
                long long x1 = (long long)ceil(((double)mxx / 64.0 - 1.0) / 2.0) + 1;
                long long y0 = (long long)floor(((double)mny / 64.0 - 1.0) / 2.0) - 1;
                long long y1 = (long long)ceil(((double)mxy / 64.0 - 1.0) / 2.0) + 1;
                if (x0 < 0) x0 = 0;
                if (y0 < 0) y0 = 0;
                if (x1 > OW - 1) x1 = OW - 1;
                if (y1 > OH - 1) y1 = OH - 1;
                for (long long Y = y0; Y <= y1; Y++)
                    for (long long X = x0; X <= x1; X++) {
                        long long sx = (2 * X + 1) * 64, sy = (2 * Y + 1) * 64;
                        if (edge_in(ax, ay, bx, by, sx, sy) && edge_in(bx, by, cx, cy, sx, sy) &&
                            edge_in(cx, cy, ax, ay, sx, sy)) {
                            long long row = (flags & ORC_TOPDOWN) ? (OH - 1 - Y) : Y;
                            uint8_t* op = out + (size_t)row * out_stride + (size_t)X * 4;
                            memcpy(op, &rgba, 4);
                        }
                    }
            }
        }
    return bad ? ORC_EINVAL : ORC_OK;
}

/* ------------------------------------------------------------------------- */
/* F4  presentation in the GLUT window                                        */
/*   display() callbacks.cpp:39-72: clear to (0,0,0,1) (main.cpp:260),        */
/*   modelview translate (tx, ty) (callbacks.cpp:32-33, 47), orthographic      */
/*   1:1 window (main.cpp:276), then renderCuda simpleVBO.cpp:185-440:         */
/*   VBO_RENDER draws every cell's triangles with glScalef(s) (258, 281);      */
/*   GL_QUADS draws pixel (j,i) as the triangles (js,is),(js,is+s),(js+s,is+s) */
/*   and (js+s,is+s),(js+s,is),(js,is) in its colour (simpleVBO.cpp:295-321).  */
/*   Definition used by the build (GL's sampling is closed, parity unpinned):  */
/*   - window pixel (X,Y), Y=0 the bottom row, has one sample at offset        */
/*     (1/2,1/2), or with antialiasing four at (3,1),(7,3),(1,5),(5,7) / 8,    */
/*     resolved per channel as (sum + 2) / 4;                                  */
/*   - a sample is covered by the coverage rule of orc_rasterize, primitives   */
/*     in draw order, later ones overwrite; s = scale_x100 / 100.              */
/*   Coordinates: image point p is p * 512 * scale_x100 integer units; a       */
/*   window sample (X + ox/8) maps to (8 (X - tx) + ox) * 6400 units.          */
/*   Pinned by: equality with orc_rasterize for integer scales without         */
/*   translation or antialiasing; nearest-neighbour closed form of GL_QUADS;   */
/*   brute force over exact rationals (tests/test_oracle_present.py).         */
/* ------------------------------------------------------------------------- */

static void present_triangle(long long ax, long long ay, long long bx, long long by, long long cx, long long cy,
                             uint32_t rgba, int tx, int ty, int ww, int wh, int ns, const int* ox, const int* oy,
                             uint32_t* samp)
{
    long long o = (bx - ax) * (cy - ay) - (by - ay) * (cx - ax);
    if (o == 0) return;
    if (o < 0) { long long t1 = bx, t2 = by; bx = cx; by = cy; cx = t1; cy = t2; }
    long long mnx = ax < bx ? (ax < cx ? ax : cx) : (bx < cx ? bx : cx);
    long long mxx = ax > bx ? (ax > cx ? ax : cx) : (bx > cx ? bx : cx);
    long long mny = ay < by ? (ay < cy ? ay : cy) : (by < cy ? by : cy);
    long long mxy = ay > by ? (ay > cy ? ay : cy) : (by > cy ? by : cy);
    long long x0 = (long long)floor(((double)mnx / 6400.0 - 8.0) / 8.0) + tx - 1;
    long long x1 = (long long)ceil((double)mxx / 6400.0 / 8.0) + tx + 1;
    long long y0 = (long long)floor(((double)mny / 6400.0 - 8.0) / 8.0) + ty - 1;
    long long y1 = (long long)ceil((double)mxy / 6400.0 / 8.0) + ty + 1;
    if (x0 < 0) x0 = 0;
    if (y0 < 0) y0 = 0;
    if (x1 > ww - 1) x1 = ww - 1;
    if (y1 > wh - 1) y1 = wh - 1;
    for (long long Y = y0; Y <= y1; Y++)
        for (long long X = x0; X <= x1; X++)
            for (int s = 0; s < ns; s++) {
                long long sx = (8 * (X - tx) + ox[s]) * 6400, sy = (8 * (Y - ty) + oy[s]) * 6400;
                if (edge_in(ax, ay, bx, by, sx, sy) && edge_in(bx, by, cx, cy, sx, sy) &&
                    edge_in(cx, cy, ax, ay, sx, sy))
                    samp[((size_t)Y * ww + (size_t)X) * ns + s] = rgba;
            }
}

int orc_present(const uint8_t* img, int W, int H, int stride, int channels, int flags, const orc_pt* diagram,
                const int32_t* edge_count, int mode, int scale_x100, int tx, int ty, int win_w, int win_h,
                int view_flags, uint8_t* out, int out_stride)
{
    if (!img || !out || W <= 0 || H <= 0 || win_w <= 0 || win_h <= 0 || scale_x100 <= 0) return ORC_EINVAL;
    if (mode != 0 && mode != 1) return ORC_EINVAL;
    if (mode == 0 && (!diagram || !edge_count)) return ORC_EINVAL;
    orc_image im = { img, W, H, stride, channels, flags };
    const int aa = view_flags & 1, ns = aa ? 4 : 1;
    const int sox[4] = { 3, 7, 1, 5 }, soy[4] = { 1, 3, 5, 7 }, cox[1] = { 4 }, coy[1] = { 4 };
    const int* ox = aa ? sox : cox;
    const int* oy = aa ? soy : coy;
    size_t n = (size_t)win_w * win_h * ns;
    uint32_t* samp = (uint32_t*)malloc(n * sizeof(uint32_t));
    if (!samp) return ORC_EINVAL;
    for (size_t k = 0; k < n; k++) samp[k] = 0xFF000000u;
    const long long sc = scale_x100, unit = 512LL * sc;
    int bad = 0;
    for (int j = 0; j < H; j++)
        for (int i = 0; i < W; i++) {
            uint32_t rgba = cell_rgba(&im, i, j);
            if (mode == 1) {
                long long x0 = i * unit, x1 = (i + 1) * unit, y0 = j * unit, y1 = (j + 1) * unit;
                present_triangle(x0, y0, x0, y1, x1, y1, rgba, tx, ty, win_w, win_h, ns, ox, oy, samp);
                present_triangle(x1, y1, x1, y0, x0, y0, rgba, tx, ty, win_w, win_h, ns, ox, oy, samp);
                continue;
            }
            size_t idx = (size_t)j * W + i;
            int ntri = edge_count[idx] - 2;
            for (int t = 0; t < ntri && t < CS / 3; t++) {
                const orc_pt* v = &diagram[idx * CS + (size_t)t * 3];
                long long u[6];
                for (int q = 0; q < 3; q++) {
                    double qx = (double)v[q].x * 64.0, qy = (double)v[q].y * 64.0;
                    if (floor(qx) != qx || floor(qy) != qy) bad++;
                    u[2 * q] = (long long)floor(qx) * 8 * sc;
                    u[2 * q + 1] = (long long)floor(qy) * 8 * sc;
                }
                present_triangle(u[0], u[1], u[2], u[3], u[4], u[5], rgba, tx, ty, win_w, win_h, ns, ox, oy, samp);
            }
        }
    for (int Y = 0; Y < win_h; Y++)
        for (int X = 0; X < win_w; X++) {
            const uint32_t* s = &samp[((size_t)Y * win_w + X) * ns];
            uint32_t c = s[0];
            if (ns == 4) {
                c = 0;
                for (int b = 0; b < 4; b++) {
                    uint32_t sum = 0;
                    for (int k = 0; k < 4; k++) sum += (s[k] >> (8 * b)) & 0xFFu;
                    c |= ((sum + 2) >> 2) << (8 * b);
                }
            }
            int row = (view_flags & 2) ? win_h - 1 - Y : Y;
            memcpy(out + (size_t)row * out_stride + (size_t)X * 4, &c, 4);
        }
    free(samp);
    return bad ? ORC_EINVAL : ORC_OK;
}

/* ------------------------------------------------------------------------- */
/* A1  the per-frame sequence of launch_kernel (kernel.cu:286-526)            */
/* ------------------------------------------------------------------------- */

int orc_depixelize(const uint8_t* img, int W, int H, int stride, int channels, int flags,
                   int scale, int subdivide, uint8_t* out, int out_stride,
                   uint8_t* graph_out, int32_t* edge_count_out, orc_pt* diagram_out,
                   int* n_tri_failed)
{
    if (!img || W <= 0 || H <= 0) return ORC_EINVAL;
    size_t N = (size_t)W * H;
    uint8_t* g = (uint8_t*)malloc(N);
    orc_pt* d = (orc_pt*)calloc(N * CS, sizeof(orc_pt));
    int32_t* ec = (int32_t*)malloc(N * sizeof(int32_t));
    int rc = ORC_EINVAL;
    if (!g || !d || !ec) goto done;
    if ((rc = orc_similarity_graph(img, W, H, stride, channels, flags, g))) goto done;
    if ((rc = orc_trivial_cross(g, W, H))) goto done;
    if ((rc = orc_heuristic_cross(g, W, H))) goto done;
    if ((rc = orc_cells(g, W, H, d, ec))) goto done;
    if (subdivide && (rc = orc_subdivide(img, W, H, stride, channels, flags, g, d, ec, 0, 0))) goto done;
    if ((rc = orc_triangulate(d, ec, W, H, n_tri_failed))) goto done;
    if (out && (rc = orc_rasterize(img, W, H, stride, channels, flags, d, ec, scale, out, out_stride)))
        goto done;
    if (graph_out) memcpy(graph_out, g, N);
    if (edge_count_out) memcpy(edge_count_out, ec, N * sizeof(int32_t));
    if (diagram_out) memcpy(diagram_out, d, N * CS * sizeof(orc_pt));
    rc = ORC_OK;
done:
    free(g);
    free(d);
    free(ec);
    return rc;
}

/* ------------------------------------------------------------------------- */
/* A18  connected components of the similarity graph                          */
/*   The reference walks component borders sequentially (extractBorderPoints  */
/*   cc_functions.cu:348-503, off the live path).  Labels here are the        */
/*   textbook canonical form: every node carries the smallest node index of   */
/*   its component, components joined by the set edge bits.                   */
/*   Pinned by: breadth-first search is the textbook routine; tests compare   */
/*   with scipy.sparse.csgraph.connected_components.                          */
/* ------------------------------------------------------------------------- */

int orc_components(const uint8_t* g, int W, int H, int32_t* labels)
{
    if (!g || !labels || W <= 0 || H <= 0) return ORC_EINVAL;
    if (!graph_valid(g, W, H)) return ORC_EGRAPH;
    int N = W * H;
    int32_t* queue = (int32_t*)malloc((size_t)N * sizeof(int32_t));
    if (!queue) return ORC_EINVAL;
    for (int k = 0; k < N; k++) labels[k] = -1;
    for (int s = 0; s < N; s++) {
        if (labels[s] >= 0) continue;
        int head = 0, tail = 0;
        queue[tail++] = s;
        labels[s] = s;
        while (head < tail) {
            int v = queue[head++];
            int i = v % W, j = v / W;
            for (int e = 0; e < 8; e++) {
                if (!BIT(g[v], e)) continue;
                int u = (j + EDY[e]) * W + (i + EDX[e]);
                if (labels[u] < 0) { labels[u] = s; queue[tail++] = u; }
            }
        }
    }
    free(queue);
    return ORC_OK;
}

/* ------------------------------------------------------------------------- */
/* A18  border chains (extractBorderPoints cc_functions.cu:348-503)           */
/*   A row-major scan over the nodes; a node that is not internal (90) and    */
/*   not yet listed starts a walk along the border of its face: leave by its  */
/*   first link in clockwise order from up-left (getFirstLink 101-113 with    */
/*   getRealLinkIndex 64-91), and at every node arrived at by edge a, leave   */
/*   by the next link clockwise after a (nextNodeClockwise 285-309; a node    */
/*   with one link sends the walk back).  The walk is a chain when it comes   */
/*   back to its start through the link just counter-clockwise of the first   */
/*   one (nextEdgeCounterClockwise 250-275, the loop test at 428-429); it is  */
/*   abandoned when it steps onto an internal, listed or discarded node, and  */
/*   then the node it stepped from is discarded (449-457 store the last node  */
/*   of the walk numCCNodes times).  The nodes of a chain are listed.         */
/*   Where the shipped code and its recorded output disagree the output wins: */
/*   cc_functions.cu:417 lists a walk's start before walking, yet the chains  */
/*   in alex_png.txt:1-21 (the alex graph, kernel.cu:296) are reproduced only */
/*   if an abandoned walk leaves its start unlisted (with the start listed,   */
/*   the walks from 34 and 82 block the chains that start at 35 and 83).      */
/*   Canonical for a node with no links, where the reference indexes a        */
/*   neighbour with edge -1 (undefined): it starts no walk.                   */
/*   Pinned by: alex_png.txt:1-21, all 21 chains and their order.             */
/* ------------------------------------------------------------------------- */

/* clockwise position (from up-left) -> edge bit, and edge bit -> position */
static const int CW_BIT[8] = { 0, 1, 2, 4, 7, 6, 5, 3 };
static const int CW_POS[8] = { 0, 1, 2, 7, 3, 6, 5, 4 };

/* The link of node value v next to edge a, turning clockwise (dir +1) or
   counter-clockwise (dir -1); a itself when it is the only link. */
static int turn_link(uint32_t v, int a, int dir)
{
    int p = CW_POS[a];
    for (int k = 1; k < 8; k++) {
        int b = CW_BIT[(p + dir * k + 8) % 8];
        if (BIT(v, b)) return b;
    }
    return a;
}

int orc_border_chains(const uint8_t* g, int W, int H, int32_t* nodes, int32_t* sizes, int max_nodes,
                      int max_chains, int* n_chains, int* n_nodes)
{
    if (!g || W <= 0 || H <= 0 || !n_chains || !n_nodes) return ORC_EINVAL;
    if (!graph_valid(g, W, H)) return ORC_EGRAPH;
    const long N = (long)W * H;
    uint8_t* mark = (uint8_t*)calloc((size_t)N, 1); /* 1 listed, 2 discarded */
    int32_t* walk = (int32_t*)malloc((size_t)(8 * N + 1) * sizeof(int32_t));
    int rc = ORC_OK;
    int nc = 0, nn = 0;
    if (!mark || !walk) { rc = ORC_EINVAL; goto done; }
    for (long s = 0; s < N; s++) {
        uint32_t v = g[s];
        if (v == 90 || v == 0 || mark[s] == 1) continue;
        int e = -1;
        for (int p = 0; p < 8 && e < 0; p++)
            if (BIT(v, CW_BIT[p])) e = CW_BIT[p];
        const int arrive = turn_link(v, e, -1);
        long len = 0, cur = s;
        int ok = 1;
        walk[len++] = (int32_t)s;
        for (;;) {
            long nxt = cur + (long)EDY[e] * W + EDX[e];
            if (nxt == s && 7 - e == arrive) break;  /* closed */
            cur = nxt;
            e = turn_link(g[cur], 7 - e, +1);
            if (g[cur] == 90 || mark[cur]) {
                mark[walk[len - 1]] = 2;
                ok = 0;
                break;
            }
            if (len > 8 * N) { rc = ORC_EGRAPH; goto done; } /* not a border: the graph is not symmetric */
            walk[len++] = (int32_t)cur;
        }
        if (!ok) continue;
        if (nc >= max_chains || nn + len > max_nodes) { rc = ORC_ESPACE; goto done; }
        for (long k = 0; k < len; k++) {
            nodes[nn + k] = walk[k];
            mark[walk[k]] = 1;
        }
        sizes[nc++] = (int32_t)len;
        nn += (int)len;
    }
done:
    *n_chains = nc;
    *n_nodes = nn;
    free(mark);
    free(walk);
    return rc;
}
