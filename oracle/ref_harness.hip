/*
 * ref_harness.hip -- runs the reference's own per-pixel device functions on
 * the GPU so that the oracle can be checked against them.
 *
 * TEST INFRASTRUCTURE ONLY.  Built by `make -C oracle ref` from the
 * reference sources where they lie (-I/root/reference); the output goes to
 * oracle/_ref/ and is never committed.  The reference files are included
 * unchanged, in kernel.cu's order (kernel.cu:6-13, CELL_SIZE from kernel.cu:8).
 *
 * The wrappers below do what kernel.cu's stage kernels do (one thread per
 * pixel, the same calls, the same Jacobi snapshots of kernel.cu:415 and
 * 434-437).  Every device buffer carries guard bytes before and after it, so
 * the reference's unchecked neighbour reads (graph[idx-1], graph rows j-1 and
 * j+1 at the frame border, image reads past the top row) stay inside an
 * allocation.  Guard bytes of the graph are 0; guard bytes of the image hold
 * a fixed pattern.  Results that depend on the guard contents (the T-junction
 * reads past the image) are excluded by the tests that use this harness.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "point.cu"
#define CELL_SIZE 45
#include "graph_functions.cu"
#include "diagram_functions.cu"
#include "triangulate_functions.cu"
#include "subdivision_functions.cu"

namespace {

constexpr int kBlock = 16;

dim3 grid_for(int W, int H) { return dim3((W + kBlock - 1) / kBlock, (H + kBlock - 1) / kBlock); }

struct Guarded {
    char* base = nullptr;
    char* ptr = nullptr;
    size_t guard = 0;
    size_t bytes = 0;
    hipError_t alloc(size_t nbytes, size_t g, int fill_pattern)
    {
        guard = g;
        bytes = nbytes;
        hipError_t e = hipMalloc(&base, nbytes + 2 * g);
        if (e != hipSuccess) return e;
        ptr = base + g;
        if (fill_pattern) {
            char* host = (char*)malloc(nbytes + 2 * g);
            for (size_t k = 0; k < nbytes + 2 * g; k++) host[k] = (char)((k * 37 + 11) & 0xFF);
            e = hipMemcpy(base, host, nbytes + 2 * g, hipMemcpyHostToDevice);
            free(host);
            return e;
        }
        return hipMemset(base, 0, nbytes + 2 * g);
    }
    ~Guarded()
    {
        if (base) (void)hipFree(base);
    }
};

__global__ void h_graph(char* img, int W, int H, int ws, char* graph)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= W || j >= H) return;
    for (int e = 0; e < 8; e++) {
        if (diff(i, j, e, W, H, ws, img)) { SET_BIT(graph[(j * W) + i], e, 0); }
        else                              { SET_BIT(graph[(j * W) + i], e, 1); }
    }
}

__global__ void h_trivial(int W, int H, char* graph)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= W || j >= H) return;
    crossCheck_4(graph, W, H, i, j);
}

__global__ void h_heuristic(int W, int H, char* graph, char* aux)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= W || j >= H) return;
    crossCheck_Heuristics(graph, aux, W, H, i, j, nullptr);
}

__global__ void h_decision(int W, int H, char* aux, const int* blocks, int nblocks, int* out)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nblocks) return;
    out[t] = processHeuristics2(aux, blocks[2 * t], blocks[2 * t + 1], W, H) ? 1 : 0;
}

__global__ void h_cells(int W, int H, char* graph, Point* diagram, int* edge_count)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= W || j >= H) return;
    int idx = j * W + i;
    edge_count[idx] = createCellFromPattern<Point>(graph[idx], graph[idx - 1], graph[idx + 1],
                                                   diagram, idx * CELL_SIZE, i, j);
}

__global__ void h_cell_table(Point* table, int* counts)
{
    int key = blockIdx.x * blockDim.x + threadIdx.x;
    if (key >= 4096) return;
    char node = (char)(key & 0xFF);
    char left = (char)((((key >> 8) & 1) << 2) | (((key >> 9) & 1) << 7));
    char right = (char)((((key >> 10) & 1) << 0) | (((key >> 11) & 1) << 5));
    counts[key] = createCellFromPattern<Point>(node, left, right, table, key * CELL_SIZE, 0, 0);
}

__global__ void h_subdivide(char* img, int W, int ws, int H, char* graph, Point* diagram,
                            Point* diagram_aux, int* edge_count, int* edge_count_aux,
                            bool* edge_status, int* link_index)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= W || j >= H) return;
    int idx = j * W + i;
    if (graph[idx] == 90) return;
    int n = subdivision(img, &diagram[idx * CELL_SIZE], diagram_aux, edge_count_aux, idx, W, ws, H, i, j,
                        graph[idx], &edge_status[idx * CELL_SIZE], &link_index[idx * CELL_SIZE]);
    edge_count[idx] = n;
}

__global__ void h_triangulate(int W, int H, Point* diagram, int* edge_count)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= W || j >= H) return;
    int idx = j * W + i;
    triangulate_polygon(&diagram[idx * CELL_SIZE], edge_count, idx, CELL_SIZE, i, j);
}

int check(hipError_t e) { return e == hipSuccess ? 0 : -(int)e - 1000; }

}  // namespace

#define TRY_HIP(x) do { int _rc = check(x); if (_rc) return _rc; } while (0)

extern "C" {

/* graph_Kernel (kernel.cu:140-159) on a bottom-up BGR buffer of stride ws */
int ref_graph(const uint8_t* img, int W, int H, int ws, uint8_t* graph)
{
    Guarded dimg, dg;
    size_t isz = (size_t)ws * H, n = (size_t)W * H;
    TRY_HIP(dimg.alloc(isz, (size_t)ws + 64, 1));
    TRY_HIP(dg.alloc(n, (size_t)W + 64, 0));
    TRY_HIP(hipMemcpy(dimg.ptr, img, isz, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(h_graph, grid_for(W, H), dim3(kBlock, kBlock), 0, 0, dimg.ptr, W, H, ws, dg.ptr);
    TRY_HIP(hipGetLastError());
    TRY_HIP(hipMemcpy(graph, dg.ptr, n, hipMemcpyDeviceToHost));
    return 0;
}

/* trivial_cross_Kernel (kernel.cu:162-177), in place */
int ref_trivial(uint8_t* graph, int W, int H)
{
    Guarded dg;
    size_t n = (size_t)W * H;
    TRY_HIP(dg.alloc(n, (size_t)W + 64, 0));
    TRY_HIP(hipMemcpy(dg.ptr, graph, n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(h_trivial, grid_for(W, H), dim3(kBlock, kBlock), 0, 0, W, H, dg.ptr);
    TRY_HIP(hipGetLastError());
    TRY_HIP(hipMemcpy(graph, dg.ptr, n, hipMemcpyDeviceToHost));
    return 0;
}

/* snapshot + ambiguous_cross_Kernel (kernel.cu:415-423), in place */
int ref_heuristic(uint8_t* graph, int W, int H)
{
    Guarded dg, da;
    size_t n = (size_t)W * H;
    TRY_HIP(dg.alloc(n, (size_t)W + 64, 0));
    TRY_HIP(da.alloc(n, (size_t)W + 64, 0));
    TRY_HIP(hipMemcpy(dg.ptr, graph, n, hipMemcpyHostToDevice));
    TRY_HIP(hipMemcpy(da.ptr, dg.ptr, n, hipMemcpyDeviceToDevice));
    hipLaunchKernelGGL(h_heuristic, grid_for(W, H), dim3(kBlock, kBlock), 0, 0, W, H, dg.ptr, da.ptr);
    TRY_HIP(hipGetLastError());
    TRY_HIP(hipMemcpy(graph, dg.ptr, n, hipMemcpyDeviceToHost));
    return 0;
}

/* processHeuristics2 (graph_functions.cu:761-849) on chosen block origins.
 * Every block must lie inside the frame (1 <= i+1 < W, 1 <= j+1 < H). */
int ref_decisions(const uint8_t* graph, int W, int H, const int* blocks, int nblocks, int* out)
{
    for (int t = 0; t < nblocks; t++)
        if (blocks[2 * t] < 0 || blocks[2 * t + 1] < 0 || blocks[2 * t] + 1 >= W || blocks[2 * t + 1] + 1 >= H)
            return -1;
    Guarded dg, db, dout;
    size_t n = (size_t)W * H;
    TRY_HIP(dg.alloc(n, (size_t)W + 64, 0));
    TRY_HIP(db.alloc(sizeof(int) * 2 * (size_t)nblocks + 8, 64, 0));
    TRY_HIP(dout.alloc(sizeof(int) * (size_t)nblocks + 8, 64, 0));
    TRY_HIP(hipMemcpy(dg.ptr, graph, n, hipMemcpyHostToDevice));
    TRY_HIP(hipMemcpy(db.ptr, blocks, sizeof(int) * 2 * (size_t)nblocks, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(h_decision, dim3((nblocks + 255) / 256), dim3(256), 0, 0, W, H, dg.ptr,
                       (const int*)db.ptr, nblocks, (int*)dout.ptr);
    TRY_HIP(hipGetLastError());
    TRY_HIP(hipMemcpy(out, dout.ptr, sizeof(int) * (size_t)nblocks, hipMemcpyDeviceToHost));
    return 0;
}

/* cells_Kernel (kernel.cu:192-213): diagram float[W*H*45*2], edge_count int[W*H] */
int ref_cells(const uint8_t* graph, int W, int H, float* diagram, int* edge_count)
{
    Guarded dg, dd, de;
    size_t n = (size_t)W * H;
    TRY_HIP(dg.alloc(n, (size_t)W + 64, 0));
    TRY_HIP(dd.alloc(n * CELL_SIZE * sizeof(Point), 256, 0));
    TRY_HIP(de.alloc(n * sizeof(int), 256, 0));
    TRY_HIP(hipMemcpy(dg.ptr, graph, n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(h_cells, grid_for(W, H), dim3(kBlock, kBlock), 0, 0, W, H, dg.ptr, (Point*)dd.ptr,
                       (int*)de.ptr);
    TRY_HIP(hipGetLastError());
    TRY_HIP(hipMemcpy(diagram, dd.ptr, n * CELL_SIZE * sizeof(Point), hipMemcpyDeviceToHost));
    TRY_HIP(hipMemcpy(edge_count, de.ptr, n * sizeof(int), hipMemcpyDeviceToHost));
    return 0;
}

/* createCellFromPattern for all 4096 (node, left{2,7}, right{0,5}) keys.
 * key bits 0-7 node, 8 left bit 2, 9 left bit 7, 10 right bit 0, 11 right bit 5 */
int ref_cell_table(float* table, int* counts)
{
    Guarded dt, dc;
    TRY_HIP(dt.alloc(4096 * CELL_SIZE * sizeof(Point), 256, 0));
    TRY_HIP(dc.alloc(4096 * sizeof(int), 256, 0));
    hipLaunchKernelGGL(h_cell_table, dim3(16), dim3(256), 0, 0, (Point*)dt.ptr, (int*)dc.ptr);
    TRY_HIP(hipGetLastError());
    TRY_HIP(hipMemcpy(table, dt.ptr, 4096 * CELL_SIZE * sizeof(Point), hipMemcpyDeviceToHost));
    TRY_HIP(hipMemcpy(counts, dc.ptr, 4096 * sizeof(int), hipMemcpyDeviceToHost));
    return 0;
}

/* snapshots + subdivision_Kernel (kernel.cu:434-453) on a cells result */
int ref_subdivide(const uint8_t* img, int W, int H, int ws, const uint8_t* graph, float* diagram,
                  int* edge_count)
{
    Guarded dimg, dg, dd, dda, de, dea, des, dli;
    size_t isz = (size_t)ws * H, n = (size_t)W * H;
    TRY_HIP(dimg.alloc(isz, (size_t)ws + 64, 1));
    TRY_HIP(dg.alloc(n, (size_t)W + 64, 0));
    TRY_HIP(dd.alloc(n * CELL_SIZE * sizeof(Point), 256, 0));
    TRY_HIP(dda.alloc(n * CELL_SIZE * sizeof(Point), 256, 0));
    TRY_HIP(de.alloc(n * sizeof(int), 256, 0));
    TRY_HIP(dea.alloc(n * sizeof(int), 256, 0));
    TRY_HIP(des.alloc(n * CELL_SIZE * sizeof(bool), 256, 0));
    TRY_HIP(dli.alloc(n * CELL_SIZE * sizeof(int), 256, 0));
    TRY_HIP(hipMemcpy(dimg.ptr, img, isz, hipMemcpyHostToDevice));
    TRY_HIP(hipMemcpy(dg.ptr, graph, n, hipMemcpyHostToDevice));
    TRY_HIP(hipMemcpy(dd.ptr, diagram, n * CELL_SIZE * sizeof(Point), hipMemcpyHostToDevice));
    TRY_HIP(hipMemcpy(de.ptr, edge_count, n * sizeof(int), hipMemcpyHostToDevice));
    TRY_HIP(hipMemcpy(dda.ptr, dd.ptr, n * CELL_SIZE * sizeof(Point), hipMemcpyDeviceToDevice));
    TRY_HIP(hipMemcpy(dea.ptr, de.ptr, n * sizeof(int), hipMemcpyDeviceToDevice));
    hipLaunchKernelGGL(h_subdivide, grid_for(W, H), dim3(kBlock, kBlock), 0, 0, dimg.ptr, W, ws, H, dg.ptr,
                       (Point*)dd.ptr, (Point*)dda.ptr, (int*)de.ptr, (int*)dea.ptr, (bool*)des.ptr,
                       (int*)dli.ptr);
    TRY_HIP(hipGetLastError());
    TRY_HIP(hipMemcpy(diagram, dd.ptr, n * CELL_SIZE * sizeof(Point), hipMemcpyDeviceToHost));
    TRY_HIP(hipMemcpy(edge_count, de.ptr, n * sizeof(int), hipMemcpyDeviceToHost));
    return 0;
}

/* triangulate_Kernel (kernel.cu:264-282), in place on the diagram */
int ref_triangulate(float* diagram, const int* edge_count, int W, int H)
{
    Guarded dd, de;
    size_t n = (size_t)W * H;
    for (size_t k = 0; k < n; k++)
        if (edge_count[k] < 0 || edge_count[k] > 16) return -1;
    TRY_HIP(dd.alloc(n * CELL_SIZE * sizeof(Point), 256, 0));
    TRY_HIP(de.alloc(n * sizeof(int), 256, 0));
    TRY_HIP(hipMemcpy(dd.ptr, diagram, n * CELL_SIZE * sizeof(Point), hipMemcpyHostToDevice));
    TRY_HIP(hipMemcpy(de.ptr, edge_count, n * sizeof(int), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(h_triangulate, grid_for(W, H), dim3(kBlock, kBlock), 0, 0, W, H, (Point*)dd.ptr,
                       (int*)de.ptr);
    TRY_HIP(hipGetLastError());
    TRY_HIP(hipMemcpy(diagram, dd.ptr, n * CELL_SIZE * sizeof(Point), hipMemcpyDeviceToHost));
    return 0;
}

int ref_sync(void) { return check(hipDeviceSynchronize()); }

}  // extern "C"
