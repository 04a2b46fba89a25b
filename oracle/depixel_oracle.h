/*
 * depixel_oracle.h -- CPU restatement of the per-frame depixelization path of
 * marcoc2/pixel-art-remaster-gpu (kernel.cu:286-526 and the device libraries it
 * includes).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (pixel-art-remaster-gpu_amd/) never links or calls it.
 *
 * Data formats follow the reference:
 *   graph        uint8[W*H], bit e = "similar to neighbour e" (graph_functions.cu:162-172)
 *                e -> (dx,dy): 0:(-1,+1) 1:(0,+1) 2:(+1,+1) 3:(-1,0) 4:(+1,0)
 *                              5:(-1,-1) 6:(0,-1) 7:(+1,-1); +1 = next row up
 *   diagram      orc_pt[W*H*45] (CELL_SIZE, kernel.cu:8), Point = {float x,y} (point.cu:2-11)
 *   edge_count   int32[W*H]
 *   image        8-bit pixels, `channels` = 3 (reference BGR) or 4 (alpha extension),
 *                row stride `stride` bytes, rows bottom-up unless ORC_TOPDOWN.
 */
#ifndef DEPIXEL_ORACLE_H
#define DEPIXEL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float x, y; } orc_pt;

#define ORC_CELL_SIZE 45

/* layout flags (same bit values as the product's DPX_* flags) */
#define ORC_RGB      1  /* byte order R,G,B instead of the reference's B,G,R   */
#define ORC_TOPDOWN  2  /* row 0 of the buffer is the top row                  */
#define ORC_ALPHA    4  /* channels==4: alpha-0 pixels form their own class    */

/* status codes */
#define ORC_OK            0
#define ORC_EINVAL       -1
#define ORC_EGRAPH       -2   /* graph has an edge leaving the frame, or a border walk that never closes */
#define ORC_ESPACE       -3   /* an output array is too small */

/* ---- per-colour helpers (A2) ---- */
uint32_t orc_rgb_to_yuv(uint32_t packed_c0c1c2);
int      orc_colors_similar(uint32_t c_a, uint32_t c_b);
int      orc_yuv_similar(uint32_t yuv_a, uint32_t yuv_b);

/* ---- stages ---- */
int orc_similarity_graph(const uint8_t* img, int W, int H, int stride, int channels,
                         int flags, uint8_t* graph);
int orc_trivial_cross(uint8_t* graph, int W, int H);
int orc_heuristic_cross(uint8_t* graph, int W, int H);
int orc_heuristic_decision(const uint8_t* graph_aux, int W, int H, int bi, int bj);

int orc_cell_points(uint8_t node, uint8_t left, uint8_t right, orc_pt* pts16);
int orc_convex_hull(const orc_pt* pts, int n, orc_pt* out45);
int orc_cell_from_pattern(uint8_t node, uint8_t left, uint8_t right, orc_pt* out45);
int orc_cells(const uint8_t* graph, int W, int H, orc_pt* diagram, int32_t* edge_count);

int orc_subdivide(const uint8_t* img, int W, int H, int stride, int channels, int flags,
                  const uint8_t* graph, orc_pt* diagram, int32_t* edge_count,
                  uint8_t* edge_status, int32_t* link_index);
int orc_triangulate_polygon(const orc_pt* poly, int n, orc_pt* result45);
int orc_triangulate(orc_pt* diagram, const int32_t* edge_count, int W, int H, int* n_failed);

int orc_export_vbo(const uint8_t* img, int W, int H, int stride, int channels, int flags,
                   const orc_pt* diagram, const int32_t* edge_count,
                   float* pos /* W*H*45*2 */, uint8_t* color /* W*H*45*4 */);
int orc_rasterize(const uint8_t* img, int W, int H, int stride, int channels, int flags,
                  const orc_pt* diagram, const int32_t* edge_count, int scale,
                  uint8_t* out, int out_stride);
/* ---- presentation in the viewer window (F4) ----
   mode 0: the cells (VBO_RENDER), 1: one square per pixel (GL_QUADS);
   view_flags bit 0: 4 samples per pixel, bit 1: window row 0 written first
   is the top row.  diagram/edge_count unused (may be NULL) in mode 1. */
int orc_present(const uint8_t* img, int W, int H, int stride, int channels, int flags,
                const orc_pt* diagram, const int32_t* edge_count, int mode, int scale_x100,
                int tx, int ty, int win_w, int win_h, int view_flags, uint8_t* out, int out_stride);
int orc_depixelize(const uint8_t* img, int W, int H, int stride, int channels, int flags,
                   int scale, int subdivide, uint8_t* out, int out_stride,
                   uint8_t* graph_out, int32_t* edge_count_out, orc_pt* diagram_out,
                   int* n_tri_failed);

/* ---- connected components and boundary chains (A18) ---- */
int orc_components(const uint8_t* graph, int W, int H, int32_t* labels);
/* Border chains in scan order: chain k is chain_nodes[sum of earlier sizes ..]
   of chain_sizes[k] node indices (j * W + i in buffer row order).  *n_chains
   and *n_nodes are the totals; ORC_ESPACE if they exceed max_chains/max_nodes. */
int orc_border_chains(const uint8_t* graph, int W, int H, int32_t* chain_nodes,
                      int32_t* chain_sizes, int max_nodes, int max_chains, int* n_chains, int* n_nodes);

#ifdef __cplusplus
}
#endif
#endif
