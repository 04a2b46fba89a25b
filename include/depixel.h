/*
 * depixel.h -- C ABI of the MI355X-native depixelization path.
 *
 * Drop-in for the per-frame call sequence of marcoc2/pixel-art-remaster-gpu
 * (kernel.cu:286-526, `launch_kernel`), plus the stage entry points that
 * sequence is made of, a batched path and a serving loop.  Plain pointers and
 * sizes only; HIP streams travel as `void*` (a hipStream_t, NULL = default).
 *
 * Conventions
 *   - every entry point except launch_kernel returns a dpx_status; the
 *     message of the last failure on the calling thread is dpx_last_error();
 *   - the caller owns every host buffer and every device buffer passed in;
 *     the library owns the device memory behind a dpx_context / dpx_server;
 *   - image buffers: 8-bit pixels, `channels` 3 (the reference's BGR order
 *     unless DPX_RGB) or 4 (BGRA/RGBA, alpha extension), `stride` bytes per
 *     row, rows bottom-up (the reference's flipped OpenCV layout,
 *     Image.cpp:21-38) unless DPX_TOPDOWN; on the batched and serving paths
 *     also `channels` 1 with DPX_INDEXED (one palette index per pixel);
 *   - output images: RGBA8, (scale*W) x (scale*H), same row order as the
 *     input unless DPX_FLIP_OUT;
 *   - graph: 1 byte per pixel, bit e = similar to neighbour e
 *     (e -> (dx,dy): 0:(-1,+1) 1:(0,+1) 2:(+1,+1) 3:(-1,0) 4:(+1,0)
 *      5:(-1,-1) 6:(0,-1) 7:(+1,-1); +1 = next row up);
 *   - diagram: dpx_point[W*H*DPX_CELL_SIZE] cell slots, edge_count int32[W*H].
 */
#ifndef DEPIXEL_H
#define DEPIXEL_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Point (point.cu:2-11): 8-byte {float x, y}, layout-compatible with float2 */
typedef struct { float x, y; } dpx_point;
/* uchar4 of the colour VBO (kernel.cu:67-104) */
typedef struct { uint8_t x, y, z, w; } dpx_uchar4;

#define DPX_CELL_SIZE 45   /* slots per cell, CELL_SIZE kernel.cu:8 */

/* layout / semantics flags */
#define DPX_RGB      1u    /* byte order R,G,B(,A) instead of B,G,R(,A)              */
#define DPX_TOPDOWN  2u    /* buffer row 0 is the top row (output follows the input) */
#define DPX_ALPHA    4u    /* channels==4: alpha-0 pixels are their own colour class */
#define DPX_FLIP_OUT 8u    /* output rows in the opposite order to the input's       */
/* channels==1: each pixel byte indexes a palette of 256 four-byte entries
 * (c0,c1,c2,a in the pixel byte order DPX_RGB selects; a is the alpha an
 * RGBA pixel would carry).  An indexed frame gives exactly the output of the
 * 4-channel frame made of its palette entries.  Batched path (palette at
 * dpx_item.palette_offset) and serving path (dpx_server_set_palette) only. */
#define DPX_INDEXED  16u

typedef enum {
    DPX_OK = 0,
    DPX_EINVAL = -1,     /* bad argument (sizes, pointers, flags)              */
    DPX_ENOMEM = -2,     /* device or host allocation failed                   */
    DPX_EHIP = -3,       /* a HIP runtime call failed                          */
    DPX_ENODEV = -4,     /* no usable GPU                                      */
    DPX_EBUDGET = -5,    /* request larger than the server's HBM budget        */
    DPX_ETIMEOUT = -6,   /* wait timed out (request still pending)             */
    DPX_ECANCELLED = -7, /* request cancelled                                  */
    DPX_ECLOSED = -8,    /* server or stream closed                            */
    DPX_EINJECTED = -9   /* failure injected by dpx_server_inject_failure      */
} dpx_status;

typedef struct {
    int32_t scale;       /* output pixels per input pixel, 1..32 (batched path: 1..15) */
    int32_t subdivide;   /* corner-cutting smoothing on/off (simpleVBO.cpp:41)       */
    uint32_t flags;      /* DPX_RGB | DPX_TOPDOWN | DPX_ALPHA | DPX_FLIP_OUT
                            (| DPX_INDEXED on the serving path)                       */
    int32_t reserved;    /* must be 0                                                 */
} dpx_params;

const char* dpx_last_error(void);
const char* dpx_version(void);

/* ------------------------------------------------------------------------ */
/* Drop-in: replaces launch_kernel, kernel.cu:286-288 (declared for its one  */
/* caller at simpleVBO.cpp:57-60, called from runCuda simpleVBO.cpp:151).    */
/* Same arguments and meaning:                                               */
/*   pos, colorPos  device buffers of W*H*45 entries (or NULL to skip the    */
/*                  VBO packing of kernel.cu:463-475)                         */
/*   time           unused, as in the reference                               */
/*   img_data       host BGR8, bottom-up, stride img_widthstep                */
/*   edge_count_h   host int[W*H], written                                    */
/*   graph_h        host char[W*H], written with the resolved graph           */
/* Returns a malloc'd host dpx_point[W*H*45] that the caller frees (as       */
/* simpleVBO.cpp:437 does).  Where the reference prints and exit(-1)s on a   */
/* CUDA error (kernel.cu:39-62), this returns NULL and sets dpx_last_error.  */
/* ------------------------------------------------------------------------ */
dpx_point* launch_kernel(dpx_point* pos, dpx_uchar4* colorPos, float time, char* img_data,
                         int img_width, int img_height, int img_widthstep, int* edge_count_h,
                         char* graph_h, bool subdivide);

/* ------------------------------------------------------------------------ */
/* Plain per-frame path: the reference's sequence one stage per kernel with  */
/* intermediates in HBM, per-call device allocation and blocking copies,     */
/* plus the raster that replaces glDrawArrays (simpleVBO.cpp:236-285).       */
/* Any of out_rgba / graph_out / edge_count_out / diagram_out may be NULL.   */
/* ------------------------------------------------------------------------ */
int dpx_frame(const uint8_t* img, int width, int height, int stride, int channels,
              const dpx_params* params, uint8_t* out_rgba, int out_stride, uint8_t* graph_out,
              int32_t* edge_count_out, dpx_point* diagram_out);

/* ------------------------------------------------------------------------ */
/* Stage entry points on caller-owned device buffers, in kernel.cu's order.  */
/* All enqueue on `stream` and return without synchronising.                */
/* ------------------------------------------------------------------------ */
/* graph_Kernel kernel.cu:140-159 */
int dpx_stage_similarity_graph(const uint8_t* d_img, int width, int height, int stride,
                               int channels, uint32_t flags, uint8_t* d_graph, void* stream);
/* trivial_cross_Kernel kernel.cu:162-177 (in place) */
int dpx_stage_trivial_cross(uint8_t* d_graph, int width, int height, void* stream);
/* snapshot kernel.cu:415 + ambiguous_cross_Kernel kernel.cu:180-189 (in place;
 * d_graph_aux receives the snapshot).  A link that leaves the frame (none in
 * a graph from the stages above) leads a curve walk to a node without links. */
int dpx_stage_heuristic_cross(uint8_t* d_graph, uint8_t* d_graph_aux, int width, int height,
                              void* stream);
/* cells_Kernel kernel.cu:192-213 */
int dpx_stage_cells(const uint8_t* d_graph, int width, int height, dpx_point* d_diagram,
                    int32_t* d_edge_count, void* stream);
/* snapshots kernel.cu:434-437 + subdivision_Kernel kernel.cu:216-261 */
int dpx_stage_subdivide(const uint8_t* d_img, int width, int height, int stride, int channels,
                        uint32_t flags, const uint8_t* d_graph, dpx_point* d_diagram,
                        dpx_point* d_diagram_aux, int32_t* d_edge_count,
                        int32_t* d_edge_count_aux, void* stream);
/* triangulate_Kernel kernel.cu:264-282 */
int dpx_stage_triangulate(dpx_point* d_diagram, const int32_t* d_edge_count, int width,
                          int height, void* stream);
/* color_kernel kernel.cu:67-104 + position_kernel kernel.cu:109-137 */
int dpx_stage_export_vbo(const uint8_t* d_img, int width, int height, int stride, int channels,
                         uint32_t flags, const dpx_point* d_diagram,
                         const int32_t* d_edge_count, dpx_point* d_pos, dpx_uchar4* d_color,
                         void* stream);
/* glDrawArrays(GL_TRIANGLES) simpleVBO.cpp:281 with glScalef(scale) (258) */
int dpx_stage_rasterize(const uint8_t* d_img, int width, int height, int stride, int channels,
                        uint32_t flags, const dpx_point* d_diagram,
                        const int32_t* d_edge_count, int scale, uint8_t* d_out,
                        int out_stride, void* stream);
/* components of the similarity graph (the reference's border walk,
 * cc_functions.cu:348-503, is off its live path): labels[k] = smallest node
 * index of k's component */
int dpx_stage_components(const uint8_t* d_graph, int width, int height, int32_t* d_labels,
                         void* stream);
/* border chains of a resolved graph: extractBorderPoints cc_functions.cu:348-503
 * (its CClist / CCsizes outputs).  Chain k is d_nodes[sum of d_sizes[0..k-1] ..]
 * of d_sizes[k] node indices j*width+i, chains in the order the row-major scan
 * starts them.  d_scratch: width*height bytes, 4-byte aligned (it holds the
 * scan's marks when they do not fit a block's LDS, frames of more than about
 * 620,000 nodes).  Working memory of about 300 bytes per node is allocated
 * and freed in stream order (hipMallocAsync / hipFreeAsync on `stream`).
 * d_counts[3] receives the number of chains, of nodes and a status (0;
 * DPX_ENOMEM when max_nodes or max_chains is too small; DPX_EINVAL, with no
 * chains, when a link leaves the frame or has no reverse link: the graph
 * is not a resolved, symmetric similarity graph), readable once the stream
 * reaches this point. */
int dpx_stage_border_chains(const uint8_t* d_graph, int width, int height, uint8_t* d_scratch,
                            int32_t* d_nodes, int max_nodes, int32_t* d_sizes, int max_chains,
                            int32_t* d_counts, void* stream);

/* Presentation: what the reference's GLUT window shows (display()
 * callbacks.cpp:39-72 drawing renderCuda simpleVBO.cpp:185-440 into the
 * 1200x800 orthographic window of main.cpp:19-20, 260-276), as an RGBA8
 * window image.  Window pixel (X, Y), Y = 0 the bottom row, samples the
 * image point ((X + ox - tx) / s, (Y + oy - ty) / s) with s = scale_x100/100
 * and (tx, ty) the translation: one sample at (ox, oy) = (1/2, 1/2), or with
 * DPX_VIEW_AA four samples at (3/8, 1/8), (7/8, 3/8), (1/8, 5/8), (5/8, 7/8)
 * averaged per channel, rounding halves up (GL_MULTISAMPLE's sample pattern
 * is driver-defined; this is the standard rotated 4x grid).  A sample takes
 * the colour of the last cell in draw order whose triangles cover it (the
 * rasterize stage's rule), or the clear colour (0,0,0,255) (main.cpp:260).
 * DPX_VIEW_QUADS draws each pixel as its s x s square instead (drawMode
 * GL_QUADS, key 'q', simpleVBO.cpp:288-331); d_diagram and d_edge_count may
 * then be NULL.  With scale_x100 = 100 * scale, no translation, a window of
 * (scale*W) x (scale*H) and no AA, the cells mode equals dpx_stage_rasterize. */
#define DPX_VIEW_CELLS    0   /* drawMode VBO_RENDER, the default (callbacks.cpp:25, key 'p') */
#define DPX_VIEW_QUADS    1   /* drawMode GL_QUADS (key 'q')                                  */
#define DPX_VIEW_AA       1u  /* GL_MULTISAMPLE on (simpleVBO.cpp:238-250, key 'a')           */
#define DPX_VIEW_TOPDOWN  2u  /* window row 0 written first is the top row                    */
typedef struct {
    int32_t mode;                      /* DPX_VIEW_CELLS or DPX_VIEW_QUADS                        */
    int32_t scale_x100;                /* glScalef factor in hundredths, 25..3200 (simpleVBO.cpp:43
                                          default 8.0 = 800; keys step 0.5, mouse drag 0.01)      */
    int32_t translate_x, translate_y;  /* window pixels, -65536..65536 (callbacks.cpp:32-33: 1, 1) */
    int32_t width, height;             /* window, 1..16384 (main.cpp:19-20: 1200 x 800)          */
    uint32_t flags;                    /* DPX_VIEW_AA | DPX_VIEW_TOPDOWN                          */
    int32_t reserved;                  /* must be 0                                               */
} dpx_view;
int dpx_stage_present(const uint8_t* d_img, int width, int height, int stride, int channels,
                      uint32_t flags, const dpx_point* d_diagram, const int32_t* d_edge_count,
                      const dpx_view* view, uint8_t* d_out, int out_stride, void* stream);

/* ------------------------------------------------------------------------ */
/* Batched path: many frames / sprites of uneven sizes through one fused     */
/* kernel chain.  Inputs and outputs live in caller-owned device arenas.     */
/* No counterpart in the reference (it renders one frame per call).          */
/* ------------------------------------------------------------------------ */
typedef struct {
    uint64_t in_offset;   /* byte offset of the frame in d_in                  */
    uint64_t out_offset;  /* byte offset of the RGBA output in d_out           */
    int32_t width, height, stride, channels;
    int32_t scale;        /* 1..15                                             */
    int32_t subdivide;
    uint32_t flags;
    int32_t out_stride;   /* bytes per output row (>= 4*scale*width)           */
    uint64_t palette_offset;  /* DPX_INDEXED: byte offset of the 1 KiB palette
                                 in d_in, a multiple of 4; otherwise ignored  */
} dpx_item;

typedef struct dpx_context dpx_context;

int dpx_context_create(int device, dpx_context** out);
void dpx_context_destroy(dpx_context* ctx);
/* Size the context's device buffers for batches of up to max_items items of
 * up to max_item_pixels pixels each, and build the tables of `scale` (0: none),
 * so that dpx_batch_run never has to grow them (growing waits for the device). */
int dpx_context_reserve(dpx_context* ctx, int max_items, int64_t max_item_pixels, int scale);
/* Enqueue the batch on `stream`; returns once enqueued.  The context's
 * scratch (graph buffers, descriptors) is reused across calls: calls on one
 * context must be serialised by the caller (same stream or explicit sync).
 * A call whose items (and output alignment mod 16) equal the previous call's
 * reuses its launch plan and only launches kernels on `stream`.  Do not
 * capture dpx_batch_run into a HIP graph: a later call with other items
 * rewrites the context's descriptors and may reallocate its scratch, which
 * a captured graph would still point at.  Capture dpx_plan_run instead. */
int dpx_batch_run(dpx_context* ctx, const uint8_t* d_in, uint8_t* d_out, const dpx_item* items,
                  int n_items, void* stream);
/* Raster tile shape per (size class, scale, subdivision): TX x (256 / TX)
 * cells per block, TX = 16, 32 (the default) or 64; the reference's fixed
 * 2x2-thread blocks (kernel.cu:392-394) have no counterpart here.  A context
 * starts with the picks of the tuning file the autotuner writes
 * (pixel-art-remaster-gpu_amd/tuning.json, or the file $DPX_TUNING names);
 * these calls override them.  tile_w 0 restores the default. */
#define DPX_CLASS_SPRITE 0  /* width and height <= 64                 */
#define DPX_CLASS_FRAME  1  /* otherwise width and height <= 512      */
#define DPX_CLASS_LARGE  2  /* larger                                  */
int dpx_context_set_tile(dpx_context* ctx, int size_class, int scale, int subdivide, int tile_w);
int dpx_context_get_tile(dpx_context* ctx, int size_class, int scale, int subdivide, int* tile_w);
/* Walk window per (size class, scale, subdivision), the crossing decisions'
 * halo: with it a tile loads 32 more bytes and rows of the graph snapshot on
 * each side into LDS before deciding, so that no curve walk reads HBM one
 * step at a time; without it walks that leave the tile's small window do.
 * DPX_WALK_AUTO (the default): the tuned choice, i.e. the window where the
 * autotuner measured it faster (tuning.json), else none; DPX_WALK_ALWAYS:
 * every launch of the key; DPX_WALK_NEVER: none.  The results are the same
 * in every mode. */
#define DPX_WALK_AUTO   0
#define DPX_WALK_ALWAYS 1
#define DPX_WALK_NEVER  2
int dpx_context_set_walk(dpx_context* ctx, int size_class, int scale, int subdivide, int walk);
int dpx_context_get_walk(dpx_context* ctx, int size_class, int scale, int subdivide, int* walk);
/* Whether the batches that follow also write the resolved graph of every
 * item (1 byte per pixel, the reference's graph_h): off by default, since
 * the images are the product and the graph writes cost about 3.5% of the
 * raster kernel's time. */
int dpx_context_keep_graph(dpx_context* ctx, int enable);
/* The resolved graphs of the last batch (device pointer + per-item offsets);
 * DPX_EINVAL unless that batch ran with dpx_context_keep_graph on. */
int dpx_batch_graph(dpx_context* ctx, const uint8_t** d_graph, const uint64_t** offsets_host);

/* Per-kernel timing of the batched path: HIP events recorded on the batch's
 * stream around each of its kernels.  Off by default; reading waits for the
 * recorded batches, returns the accumulated times and resets them. */
typedef struct {
    double graph_ms;    /* similarity graph + trivial crossings, summed over the batches */
    double raster_ms;   /* heuristic crossings + cells + raster, summed over the batches */
    uint64_t batches;   /* batches recorded since the last read                           */
} dpx_batch_times;
int dpx_batch_timing(dpx_context* ctx, int enable);
int dpx_batch_timing_read(dpx_context* ctx, dpx_batch_times* out);

/* dpx_batch_run without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  Every buffer beyond d_in and d_out (tables,
 * descriptors, tile maps, the graph snapshot) lives in one static device
 * array of the library, 58 MiB on each GPU it runs on, held until the
 * process ends.  Limits: at most 4096 items, at most 48 MiB of graph bytes
 * (one per pixel, each item rounded up to 256: e.g. 836 frames of 256x224),
 * at most 65,536 graph tiles (64x32 pixels) and 524,288 raster tiles (32x8);
 * larger batches are refused with DPX_EINVAL, and so are items with
 * subdivision (it needs per-cell row buffers).  The resolved graph is not
 * written.  Runs on the current device.  The first call on each device builds
 * the tables and waits for them on `stream`, so it must not be captured.  All
 * calls on one GPU share the array: calls that overlap on one GPU (on
 * different streams, unsynchronised) race. */
int dpx_batch_run_noalloc(const uint8_t* d_in, uint8_t* d_out, const dpx_item* items, int n_items,
                          void* stream);

/* Plans: a batch layout validated once, with descriptors and graph scratch
 * of its own (no counterpart in the reference, which renders one frame per
 * call).  dpx_plan_run launches the batch's kernels for any input and output
 * arena whose address modulo 16 equals out_align modulo 16; it allocates,
 * copies and waits for nothing, so a caller may capture it into a HIP graph,
 * and replays stay valid whatever else runs on the context, until the plan
 * is destroyed.  With DPX_PLAN_GRAPH the library launches the kernels as a
 * HIP graph of its own (built at the first run, its pointers updated when
 * d_in or d_out change).  Runs of one plan must be serialised by the caller;
 * plans must be destroyed before their context. */
#define DPX_PLAN_GRAPH      1u  /* launch through a library-built HIP graph          */
#define DPX_PLAN_KEEP_GRAPH 2u  /* also write the resolved graphs (dpx_plan_graph)   */
typedef struct dpx_plan dpx_plan;
int dpx_plan_create(dpx_context* ctx, const dpx_item* items, int n_items, uint64_t out_align, uint32_t flags,
                    dpx_plan** out);
int dpx_plan_run(dpx_plan* plan, const uint8_t* d_in, uint8_t* d_out, void* stream);
/* The resolved graphs written by the plan's runs (device pointer + per-item
 * offsets); DPX_EINVAL unless the plan was made with DPX_PLAN_KEEP_GRAPH. */
int dpx_plan_graph(dpx_plan* plan, const uint8_t** d_graph, const uint64_t** offsets_host);
void dpx_plan_destroy(dpx_plan* plan);

/* ------------------------------------------------------------------------ */
/* Serving: continuous batching of frames from many streams under an HBM     */
/* budget.  submit is thread-safe; results land in caller host buffers.      */
/* ------------------------------------------------------------------------ */
typedef struct dpx_server dpx_server;

typedef struct {
    uint64_t submitted, completed, failed, rejected, cancelled;
    uint64_t batches;
    double frames_per_s;     /* completed frames / wall time since create or the last reset */
    double p50_ms, p99_ms;   /* submit -> completion latency               */
    double mean_batch;       /* frames per device batch                    */
    uint64_t hbm_in_use, hbm_peak, hbm_budget;
} dpx_server_stats;

int dpx_server_create(int device, uint64_t hbm_budget_bytes, int max_batch, dpx_server** out);
void dpx_server_destroy(dpx_server* srv);
int dpx_server_open_stream(dpx_server* srv, int width, int height, int channels,
                           const dpx_params* params, int* stream_id);
int dpx_server_close_stream(dpx_server* srv, int stream_id);
/* The palette of a DPX_INDEXED stream (channels 1): 256 four-byte entries,
 * uploaded once into the server's HBM arena and kept resident there.  A frame
 * is rendered with the palette set when it was submitted; a replaced palette
 * is released when the last frame that names it completes.  Frames of an
 * indexed stream are refused until a palette is set. */
int dpx_server_set_palette(dpx_server* srv, int stream_id, const uint8_t* palette);
/* img: host frame (stride bytes/row); out: host RGBA buffer kept alive until
 * the ticket completes.  Frames and results inside blocks from dpx_host_alloc
 * move by DMA alone; others are staged through pinned buffers (one host copy
 * each way). */
int dpx_server_submit(dpx_server* srv, int stream_id, const uint8_t* img, int stride,
                      uint8_t* out, int out_stride, uint64_t* ticket);
int dpx_server_wait(dpx_server* srv, uint64_t ticket, int timeout_ms);
/* dpx_server_wait that also returns the request's submit -> completion time
 * (the latency the server's p50/p99 are taken over). */
int dpx_server_wait_timed(dpx_server* srv, uint64_t ticket, int timeout_ms, double* latency_ms);
int dpx_server_cancel(dpx_server* srv, uint64_t ticket);
int dpx_server_stats_get(dpx_server* srv, dpx_server_stats* out);
/* test hook: make the k-th submitted request (0-based) fail on the device side */
int dpx_server_inject_failure(dpx_server* srv, uint64_t request_index);
/* restart the counters and latency window of dpx_server_stats_get */
int dpx_server_stats_reset(dpx_server* srv);
/* pinned host memory for frames and results (no counterpart in the reference) */
int dpx_host_alloc(size_t bytes, void** out);
int dpx_host_free(void* p);

/* ------------------------------------------------------------------------ */
/* Synthetic inputs (deterministic, splitmix64): the workloads the path is   */
/* measured on.  Host-only helpers.                                          */
/*   kind 0: 256x224-style frame: 16-colour palette, 40 rectangles, 20 1-px  */
/*           staircases, 4 checkerboard-dither regions, 30 isolated pixels   */
/*   kind 1: sprite with alpha: 4-12 colours, dark 1-px outline, ~30%        */
/*           transparent background, interior dither (channels must be 4)    */
/*   kind 2: console frame: kind 0's content with 16-64 colours, plus 15% of */
/*           the pixels in 1-px random walks of length 20-200              */
/*   kind 4: large frame: 256 colours, 25% dither, diagonal lattice          */
/* ------------------------------------------------------------------------ */
int dpx_synth_frame(int kind, uint64_t seed, int width, int height, int channels, uint8_t* out,
                    int stride);
/* Frame t (0..1000000) of synthetic game stream `seed`: palette indices
 * (1 byte per pixel, row 0 first) of a background of 16x16 tiles and seven
 * 16-24 px sprites that move by 1-4 px per axis each frame (about a tenth of
 * the 16x16 tiles change between frames at 320x224); `palette` (1 KiB, or
 * NULL) receives the stream's 16-32 colours as (c0, c1, c2, 255) entries,
 * the same for every t.  Entries past the stream's colours are zero. */
int dpx_synth_stream_frame(uint64_t seed, int t, int width, int height, uint8_t* out, int stride,
                           uint8_t* palette);
/* sprite size of sprite `index` of seed `seed` (sizes of the sprite mix) */
int dpx_synth_sprite_size(uint64_t seed, int index, int* width, int* height);

#ifdef __cplusplus
}
#endif
#endif /* DEPIXEL_H */
