// server.cpp -- serving loop with continuous batching under an HBM budget.
//
// The server reserves its HBM budget as one device arena when it is created.
// Requests (one host frame each, from any number of open streams) queue in
// arrival order; a request is admitted when its input and output buffers can
// be carved out of the arena (first fit, oldest first).  kLanes lanes, each
// with its own HIP stream, batched-path context and pinned staging buffers,
// take admitted work whenever they are free: a lane forms a batch of up to
// max_batch requests, uploads the frames, runs the fused kernel chain,
// downloads the results and completes the tickets, releasing their arena
// ranges.  While one lane's batch is on the device the others admit the
// requests that arrived meanwhile, so requests overlap on the GPU as memory
// frees up.  Opening a stream sizes every lane's context for full batches of
// its frames, so serving never waits for a context to grow.  Each frame's
// result depends only on that frame (the kernels keep no state across
// frames), so it is the same alone and in any mix.
//
// Palettes of indexed streams stay resident in the arena: setting a palette
// uploads it once into a 1 KiB range, every frame submitted afterwards names
// that range, and the range is released when the stream has moved on to
// another palette (or closed) and no admitted or pending frame still names it.
#include <string.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "common.hpp"

namespace {

using Clock = std::chrono::steady_clock;

constexpr int kLanes = 4;
using dpx::kPaletteBytes;

size_t align256(size_t v) { return (v + 255) / 256 * 256; }

// First-fit allocator of byte ranges inside the device arena.
class Arena {
public:
    void reset(size_t size)
    {
        free_.clear();
        free_[0] = size;
        used_ = 0;
    }
    bool alloc(size_t n, size_t* off)
    {
        n = align256(n);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            size_t o = it->first, rest = it->second - n;
            free_.erase(it);
            if (rest) free_[o + n] = rest;
            *off = o;
            used_ += n;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t n)
    {
        n = align256(n);
        used_ -= n;
        auto it = free_.emplace(off, n).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t used() const { return used_; }

private:
    std::map<size_t, size_t> free_;
    size_t used_ = 0;
};

// A palette resident in the arena, shared by the frames submitted while it
// was the stream's current one.
struct PaletteSlot {
    size_t off = 0;  // arena offset of the 1 KiB
    int users = 0;   // pending or admitted frames that name it
    bool current = true;
};

struct StreamInfo {
    int W, H, ch;
    dpx_params params;
    bool open;
    PaletteSlot* palette = nullptr;  // DPX_INDEXED streams: the current palette
};

struct Request {
    uint64_t ticket = 0;
    uint64_t index = 0;  // submission order
    int stream_id = 0;
    int W = 0, H = 0, ch = 0;
    dpx_params params{};
    const uint8_t* img = nullptr;
    int stride = 0;
    uint8_t* out = nullptr;
    int out_stride = 0;
    size_t in_bytes = 0, out_bytes = 0;
    // DPX_INDEXED: the stream's palette when the frame was submitted
    PaletteSlot* palette = nullptr;
    size_t in_alloc = 0;             // arena bytes of the input range
    size_t in_off = 0, out_off = 0;  // arena ranges while admitted
    Clock::time_point t_submit;
    double latency_ms = 0;  // submit -> completion
    int status = 1;  // 1 pending, 2 admitted, else the final dpx_status
    bool inject = false;
};

struct Lane {
    dpx_context* ctx = nullptr;
    hipStream_t stream = nullptr;
    uint8_t *h_in = nullptr, *h_out = nullptr;
    size_t hin_cap = 0, hout_cap = 0;
    std::thread worker;
};

}  // namespace

struct dpx_server {
    int device = 0;
    uint64_t budget = 0;
    int max_batch = 64;
    uint8_t* d_arena = nullptr;
    Arena arena;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool closing = false;
    std::map<int, StreamInfo> streams;
    int next_stream = 1;
    std::deque<Request*> pending;
    std::map<uint64_t, std::unique_ptr<Request>> requests;
    uint64_t next_ticket = 1, next_index = 0;
    std::vector<uint64_t> inject_at;
    uint64_t peak = 0;
    uint64_t submitted = 0, completed = 0, failed = 0, rejected = 0, cancelled = 0, batches = 0, batched_frames = 0;
    std::vector<double> latencies_ms;
    std::vector<std::unique_ptr<PaletteSlot>> palettes;
    Clock::time_point t_create;
    std::atomic<int64_t> reserved_pixels{ 0 };  // the largest stream frame the lanes are sized for
    Lane lanes[kLanes];
};

namespace {

using namespace dpx;

// Pinned host blocks handed out by dpx_host_alloc: frames and results that
// live in them are copied straight between them and the device.
std::mutex g_pin_mu;
std::map<uintptr_t, size_t> g_pinned;  // base -> bytes

bool in_pinned_block(const void* p, size_t n)
{
    std::lock_guard<std::mutex> lock(g_pin_mu);
    auto it = g_pinned.upper_bound(reinterpret_cast<uintptr_t>(p));
    if (it == g_pinned.begin()) return false;
    --it;
    uintptr_t a = reinterpret_cast<uintptr_t>(p);
    return a >= it->first && a + n <= it->first + it->second;
}

int grow_pinned(uint8_t** p, size_t* cap, size_t need)
{
    if (need <= *cap) return DPX_OK;
    size_t n = std::max(need, *cap * 3 / 2);
    if (*p) (void)hipHostFree(*p);
    *p = nullptr;
    *cap = 0;
    DPX_HIP(hipHostMalloc(p, n, hipHostMallocDefault));
    *cap = n;
    return DPX_OK;
}

// Runs one admitted batch on a lane; returns a status for the whole batch.
// Frames and results in blocks from dpx_host_alloc move by DMA alone; others
// pass through the lane's pinned staging buffers.
int run_batch(dpx_server* srv, Lane& lane, const std::vector<Request*>& batch)
{
    // Once a copy or kernel is queued, every return waits for the lane's
    // stream: the caller then releases the batch's arena ranges and host
    // buffers, which work still in flight would otherwise write into.
    struct DrainOnExit {
        hipStream_t s;
        bool armed = false;
        ~DrainOnExit()
        {
            if (armed) (void)hipStreamSynchronize(s);
        }
    } drain{ lane.stream };
    std::vector<dpx_item> items;
    std::vector<Request*> live;
    std::vector<char> in_direct, out_direct;
    size_t hin = 0, hout = 0;
    for (Request* r : batch) {
        if (r->inject) continue;
        live.push_back(r);
        const size_t row = (size_t)r->W * r->params.scale * 4, rows = (size_t)r->H * r->params.scale;
        in_direct.push_back(in_pinned_block(r->img, r->in_bytes));
        out_direct.push_back(in_pinned_block(r->out, (rows - 1) * r->out_stride + row));
        if (!in_direct.back()) hin += align256(r->in_bytes);
        if (!out_direct.back()) hout += align256(r->out_bytes);
    }
    if (live.empty()) return DPX_OK;
    if (int rc = grow_pinned(&lane.h_in, &lane.hin_cap, hin)) return rc;
    if (int rc = grow_pinned(&lane.h_out, &lane.hout_cap, hout)) return rc;
    size_t ho = 0;
    for (size_t k = 0; k < live.size(); k++) {
        Request* r = live[k];
        const uint8_t* src = r->img;
        if (!in_direct[k]) {
            memcpy(lane.h_in + ho, r->img, r->in_bytes);
            src = lane.h_in + ho;
            ho += align256(r->in_bytes);
        }
        drain.armed = true;
        DPX_HIP(hipMemcpyAsync(srv->d_arena + r->in_off, src, r->in_bytes, hipMemcpyHostToDevice, lane.stream));
        dpx_item it;
        memset(&it, 0, sizeof(it));
        if (r->palette) it.palette_offset = r->palette->off;  // resident in the arena
        it.in_offset = r->in_off;
        it.out_offset = r->out_off;
        it.width = r->W;
        it.height = r->H;
        it.stride = r->stride;
        it.channels = r->ch;
        it.scale = r->params.scale;
        it.subdivide = r->params.subdivide;
        it.flags = r->params.flags;
        it.out_stride = r->W * r->params.scale * 4;
        items.push_back(it);
    }
    if (int rc = dpx_batch_run(lane.ctx, srv->d_arena, srv->d_arena, items.data(), (int)items.size(), lane.stream))
        return rc;
    ho = 0;
    std::vector<size_t> host_off;
    for (size_t k = 0; k < live.size(); k++) {
        Request* r = live[k];
        const size_t row = (size_t)r->W * r->params.scale * 4, rows = (size_t)r->H * r->params.scale;
        if (out_direct[k]) {
            if ((size_t)r->out_stride == row)  // one linear copy
                DPX_HIP(hipMemcpyAsync(r->out, srv->d_arena + r->out_off, r->out_bytes, hipMemcpyDeviceToHost,
                                       lane.stream));
            else
                DPX_HIP(hipMemcpy2DAsync(r->out, (size_t)r->out_stride, srv->d_arena + r->out_off, row, row, rows,
                                         hipMemcpyDeviceToHost, lane.stream));
            host_off.push_back(0);
            continue;
        }
        DPX_HIP(hipMemcpyAsync(lane.h_out + ho, srv->d_arena + r->out_off, r->out_bytes, hipMemcpyDeviceToHost,
                               lane.stream));
        host_off.push_back(ho);
        ho += align256(r->out_bytes);
    }
    drain.armed = false;
    DPX_HIP(hipStreamSynchronize(lane.stream));
    for (size_t k = 0; k < live.size(); k++) {
        if (out_direct[k]) continue;
        Request* r = live[k];
        const size_t row = (size_t)r->W * r->params.scale * 4, rows = (size_t)r->H * r->params.scale;
        const uint8_t* src = lane.h_out + host_off[k];
        for (size_t y = 0; y < rows; y++) memcpy(r->out + y * r->out_stride, src + y * row, row);
    }
    return DPX_OK;
}

// Drops a frame's hold on its palette slot; a slot that is no longer the
// stream's current palette is released with its last frame (caller holds
// the lock).
void drop_palette(dpx_server* srv, PaletteSlot* p)
{
    if (!p) return;
    p->users--;
    if (p->users > 0 || p->current) return;
    srv->arena.release(p->off, kPaletteBytes);
    for (auto it = srv->palettes.begin(); it != srv->palettes.end(); ++it)
        if (it->get() == p) {
            srv->palettes.erase(it);
            break;
        }
}

// Admits the oldest pending requests that fit (caller holds the lock).
void admit(dpx_server* srv, std::vector<Request*>& batch)
{
    while (!srv->pending.empty() && (int)batch.size() < srv->max_batch) {
        Request* r = srv->pending.front();
        size_t a, b;
        if (!srv->arena.alloc(r->in_alloc, &a)) break;
        if (!srv->arena.alloc(r->out_bytes, &b)) {
            srv->arena.release(a, r->in_alloc);
            break;
        }
        r->in_off = a;
        r->out_off = b;
        r->status = 2;
        srv->pending.pop_front();
        batch.push_back(r);
    }
    srv->peak = std::max<uint64_t>(srv->peak, srv->arena.used());
}

void lane_loop(dpx_server* srv, int li)
{
    Lane& lane = srv->lanes[li];
    (void)hipSetDevice(srv->device);
    for (;;) {
        std::vector<Request*> batch;
        {
            std::unique_lock<std::mutex> lock(srv->mu);
            for (;;) {
                if (srv->closing && srv->pending.empty()) return;
                if (!srv->pending.empty()) {
                    admit(srv, batch);
                    if (!batch.empty()) break;
                    // the arena is held by the other lane's batch: wait for it
                    srv->cv_done.wait(lock);
                } else {
                    srv->cv_work.wait(lock);
                }
            }
        }
        int rc = run_batch(srv, lane, batch);
        {
            std::lock_guard<std::mutex> lock(srv->mu);
            auto now = Clock::now();
            srv->batches++;
            srv->batched_frames += batch.size();
            for (Request* r : batch) {
                srv->arena.release(r->in_off, r->in_alloc);
                srv->arena.release(r->out_off, r->out_bytes);
                drop_palette(srv, r->palette);
                r->palette = nullptr;
                r->status = r->inject ? DPX_EINJECTED : rc;
                r->latency_ms = std::chrono::duration<double, std::milli>(now - r->t_submit).count();
                if (r->status == DPX_OK) {
                    srv->completed++;
                    srv->latencies_ms.push_back(r->latency_ms);
                } else {
                    srv->failed++;
                }
            }
        }
        srv->cv_done.notify_all();
        srv->cv_work.notify_all();
    }
}

double percentile(std::vector<double> v, double q)
{
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    double pos = q * (double)(v.size() - 1);
    size_t lo = (size_t)pos;
    size_t hi = std::min(lo + 1, v.size() - 1);
    return v[lo] + (v[hi] - v[lo]) * (pos - (double)lo);
}

}  // namespace

extern "C" int dpx_server_create(int device, uint64_t hbm_budget_bytes, int max_batch, dpx_server** out)
{
    DPX_CHECK(out, DPX_EINVAL, "null output pointer");
    *out = nullptr;
    DPX_CHECK(hbm_budget_bytes >= 4096 && max_batch > 0 && max_batch <= 65536, DPX_EINVAL,
              "bad budget or batch size");
    if (int rc = require_device()) return rc;
    DPX_HIP(hipSetDevice(device));
    auto* srv = new dpx_server();
    srv->device = device;
    srv->budget = align256(hbm_budget_bytes);
    srv->max_batch = max_batch;
    srv->t_create = Clock::now();
    if (hipMalloc(&srv->d_arena, srv->budget) != hipSuccess) {
        dpx_server_destroy(srv);
        return fail(DPX_ENOMEM, "cannot reserve the HBM budget (%llu bytes)", (unsigned long long)srv->budget);
    }
    srv->arena.reset(srv->budget);
    for (Lane& lane : srv->lanes) {
        int rc = dpx_context_create(device, &lane.ctx);
        if (rc == DPX_OK && hipStreamCreateWithFlags(&lane.stream, hipStreamNonBlocking) != hipSuccess)
            rc = fail(DPX_EHIP, "stream creation failed");
        if (rc != DPX_OK) {
            dpx_server_destroy(srv);
            return rc;
        }
    }
    for (int k = 0; k < kLanes; k++) srv->lanes[k].worker = std::thread(lane_loop, srv, k);
    *out = srv;
    return DPX_OK;
}

extern "C" void dpx_server_destroy(dpx_server* srv)
{
    if (!srv) return;
    {
        std::lock_guard<std::mutex> lock(srv->mu);
        srv->closing = true;
    }
    srv->cv_work.notify_all();
    srv->cv_done.notify_all();
    for (Lane& lane : srv->lanes) {
        if (lane.worker.joinable()) lane.worker.join();
        if (lane.stream) (void)hipStreamSynchronize(lane.stream);
        if (lane.h_in) (void)hipHostFree(lane.h_in);
        if (lane.h_out) (void)hipHostFree(lane.h_out);
        if (lane.stream) (void)hipStreamDestroy(lane.stream);
        dpx_context_destroy(lane.ctx);
    }
    (void)hipFree(srv->d_arena);
    delete srv;
}

extern "C" int dpx_server_open_stream(dpx_server* srv, int width, int height, int channels, const dpx_params* params,
                                      int* stream_id)
{
    DPX_CHECK(srv && params && stream_id, DPX_EINVAL, "bad arguments");
    DPX_CHECK(width > 0 && height > 0 && width <= 32768 && height <= 32768, DPX_EINVAL, "bad size %dx%d", width,
              height);
    DPX_CHECK((params->flags & ~kKnownFlags) == 0, DPX_EINVAL, "unknown flags 0x%x", params->flags & ~kKnownFlags);
    const bool indexed = params->flags & DPX_INDEXED;
    DPX_CHECK(indexed ? channels == 1 : (channels == 3 || channels == 4), DPX_EINVAL,
              indexed ? "indexed streams have 1 channel" : "channels must be 3 or 4");
    DPX_CHECK(params->scale >= 1 && params->scale <= 15, DPX_EINVAL, "scale %d outside 1..15", params->scale);
    DPX_CHECK(params->reserved == 0, DPX_EINVAL, "reserved must be 0");
    DPX_CHECK((int64_t)width * channels * height < (int64_t)1 << 31, DPX_EINVAL, "frame of %dx%dx%d exceeds 2 GiB",
              width, height, channels);
    // size every lane for full batches of this stream's frames (outside the
    // server lock: growing a context waits for the device)
    const int64_t pixels = (int64_t)width * height;
    {
        std::lock_guard<std::mutex> lock(srv->mu);
        DPX_CHECK(!srv->closing, DPX_ECLOSED, "server closing");
    }
    // a batch holds at most as many frames as the HBM budget fits
    const uint64_t per_frame = align256((size_t)width * channels * height) +
                               align256((size_t)pixels * params->scale * params->scale * 4);
    const int items = (int)std::max<uint64_t>(1, std::min<uint64_t>(srv->max_batch, srv->budget / per_frame));
    for (Lane& lane : srv->lanes) {
        if (int rc = dpx_context_reserve(lane.ctx, items, std::max(pixels, srv->reserved_pixels.load()),
                                         params->scale))
            return rc;
        // subdivided frames also need cell rows (a word per sample row and one more per cell)
        if (params->subdivide)
            if (int rc = context_reserve_rows(lane.ctx, (size_t)items * align256((size_t)pixels * 4 *
                                                                                 cell_row_words(params->scale))))
                return rc;
    }
    std::lock_guard<std::mutex> lock(srv->mu);
    DPX_CHECK(!srv->closing, DPX_ECLOSED, "server closing");
    srv->reserved_pixels = std::max(pixels, srv->reserved_pixels.load());
    int id = srv->next_stream++;
    StreamInfo& si = srv->streams[id];
    si.W = width;
    si.H = height;
    si.ch = channels;
    si.params = *params;
    si.open = true;
    *stream_id = id;
    return DPX_OK;
}

extern "C" int dpx_server_set_palette(dpx_server* srv, int stream_id, const uint8_t* palette)
{
    DPX_CHECK(srv && palette, DPX_EINVAL, "bad arguments");
    std::unique_lock<std::mutex> lock(srv->mu);
    auto it = srv->streams.find(stream_id);
    DPX_CHECK(it != srv->streams.end() && it->second.open, DPX_ECLOSED, "stream %d is not open", stream_id);
    DPX_CHECK(it->second.params.flags & DPX_INDEXED, DPX_EINVAL, "stream %d is not indexed", stream_id);
    // a 1 KiB range of the arena; while the arena is full, wait for batches
    // to release theirs
    size_t off = 0;
    while (!srv->arena.alloc(kPaletteBytes, &off)) {
        DPX_CHECK(!srv->closing, DPX_ECLOSED, "server closing");
        srv->cv_done.wait(lock);
        it = srv->streams.find(stream_id);
        DPX_CHECK(it != srv->streams.end() && it->second.open, DPX_ECLOSED, "stream %d closed", stream_id);
    }
    srv->peak = std::max<uint64_t>(srv->peak, srv->arena.used());
    if (hipMemcpy(srv->d_arena + off, palette, kPaletteBytes, hipMemcpyHostToDevice) != hipSuccess) {
        srv->arena.release(off, kPaletteBytes);
        return fail(DPX_EHIP, "palette upload failed");
    }
    srv->palettes.push_back(std::make_unique<PaletteSlot>());
    PaletteSlot* slot = srv->palettes.back().get();
    slot->off = off;
    if (PaletteSlot* old = it->second.palette) {
        old->current = false;
        old->users++;  // (dropped at once unless frames still name it)
        drop_palette(srv, old);
    }
    it->second.palette = slot;
    return DPX_OK;
}

extern "C" int dpx_server_close_stream(dpx_server* srv, int stream_id)
{
    DPX_CHECK(srv, DPX_EINVAL, "null server");
    std::lock_guard<std::mutex> lock(srv->mu);
    auto it = srv->streams.find(stream_id);
    DPX_CHECK(it != srv->streams.end() && it->second.open, DPX_EINVAL, "unknown stream %d", stream_id);
    it->second.open = false;  // requests already submitted still complete
    if (PaletteSlot* old = it->second.palette) {
        it->second.palette = nullptr;
        old->current = false;
        old->users++;
        drop_palette(srv, old);
    }
    return DPX_OK;
}

extern "C" int dpx_server_submit(dpx_server* srv, int stream_id, const uint8_t* img, int stride, uint8_t* out,
                                 int out_stride, uint64_t* ticket)
{
    DPX_CHECK(srv && ticket, DPX_EINVAL, "bad arguments");
    std::lock_guard<std::mutex> lock(srv->mu);
    DPX_CHECK(!srv->closing, DPX_ECLOSED, "server closing");
    auto it = srv->streams.find(stream_id);
    DPX_CHECK(it != srv->streams.end() && it->second.open, DPX_ECLOSED, "stream %d is not open", stream_id);
    const StreamInfo& si = it->second;
    DPX_CHECK(img && out, DPX_EINVAL, "null frame or output buffer");
    DPX_CHECK(stride >= si.W * si.ch, DPX_EINVAL, "stride %d < %d", stride, si.W * si.ch);
    // the batched path's limit on an item's input (one bad frame must not
    // fail the other requests of its batch)
    DPX_CHECK((int64_t)stride * si.H < (int64_t)1 << 31, DPX_EINVAL, "stride %d x %d rows exceeds 2 GiB", stride,
              si.H);
    DPX_CHECK(out_stride >= si.W * si.params.scale * 4, DPX_EINVAL, "out_stride %d too small", out_stride);
    auto r = std::make_unique<Request>();
    r->W = si.W;
    r->H = si.H;
    r->ch = si.ch;
    r->params = si.params;
    r->in_bytes = (size_t)stride * si.H;
    r->in_alloc = r->in_bytes;
    if (si.params.flags & DPX_INDEXED)
        DPX_CHECK(si.palette, DPX_EINVAL, "stream %d is indexed and has no palette yet", stream_id);
    r->out_bytes = (size_t)si.W * si.params.scale * 4 * si.H * si.params.scale;
    if (align256(r->in_alloc) + align256(r->out_bytes) > srv->budget) {
        srv->rejected++;
        return fail(DPX_EBUDGET, "frame needs %zu bytes of HBM, budget is %llu", r->in_alloc + r->out_bytes,
                    (unsigned long long)srv->budget);
    }
    r->ticket = srv->next_ticket++;
    r->index = srv->next_index++;
    r->stream_id = stream_id;
    r->img = img;
    r->stride = stride;
    r->out = out;
    r->out_stride = out_stride;
    r->t_submit = Clock::now();
    r->inject = std::find(srv->inject_at.begin(), srv->inject_at.end(), r->index) != srv->inject_at.end();
    if ((r->palette = si.palette)) r->palette->users++;
    *ticket = r->ticket;
    srv->pending.push_back(r.get());
    srv->requests[r->ticket] = std::move(r);
    srv->submitted++;
    srv->cv_work.notify_one();
    return DPX_OK;
}

extern "C" int dpx_server_wait(dpx_server* srv, uint64_t ticket, int timeout_ms)
{
    return dpx_server_wait_timed(srv, ticket, timeout_ms, nullptr);
}

extern "C" int dpx_server_wait_timed(dpx_server* srv, uint64_t ticket, int timeout_ms, double* latency_ms)
{
    DPX_CHECK(srv, DPX_EINVAL, "null server");
    std::unique_lock<std::mutex> lock(srv->mu);
    auto it = srv->requests.find(ticket);
    DPX_CHECK(it != srv->requests.end(), DPX_EINVAL, "unknown ticket %llu", (unsigned long long)ticket);
    Request* r = it->second.get();
    auto done = [&] { return r->status <= 0; };
    if (timeout_ms < 0) srv->cv_done.wait(lock, done);
    else if (!srv->cv_done.wait_for(lock, std::chrono::milliseconds(timeout_ms), done))
        return fail(DPX_ETIMEOUT, "ticket %llu still pending", (unsigned long long)ticket);
    int st = r->status;
    if (latency_ms) *latency_ms = r->latency_ms;
    srv->requests.erase(it);
    if (st != DPX_OK) return fail(st, "request %llu failed (%d)", (unsigned long long)ticket, st);
    return DPX_OK;
}

extern "C" int dpx_server_cancel(dpx_server* srv, uint64_t ticket)
{
    DPX_CHECK(srv, DPX_EINVAL, "null server");
    {
        std::lock_guard<std::mutex> lock(srv->mu);
        auto it = srv->requests.find(ticket);
        DPX_CHECK(it != srv->requests.end(), DPX_EINVAL, "unknown ticket %llu", (unsigned long long)ticket);
        Request* r = it->second.get();
        DPX_CHECK(r->status == 1, DPX_EINVAL, "ticket %llu is already admitted or done", (unsigned long long)ticket);
        srv->pending.erase(std::find(srv->pending.begin(), srv->pending.end(), r));
        drop_palette(srv, r->palette);
        r->palette = nullptr;
        r->status = DPX_ECANCELLED;
        srv->cancelled++;
    }
    srv->cv_done.notify_all();
    srv->cv_work.notify_all();
    return DPX_OK;
}

extern "C" int dpx_server_stats_get(dpx_server* srv, dpx_server_stats* out)
{
    DPX_CHECK(srv && out, DPX_EINVAL, "bad arguments");
    std::lock_guard<std::mutex> lock(srv->mu);
    memset(out, 0, sizeof(*out));
    out->submitted = srv->submitted;
    out->completed = srv->completed;
    out->failed = srv->failed;
    out->rejected = srv->rejected;
    out->cancelled = srv->cancelled;
    out->batches = srv->batches;
    double secs = std::chrono::duration<double>(Clock::now() - srv->t_create).count();
    out->frames_per_s = secs > 0 ? (double)srv->completed / secs : 0.0;
    out->p50_ms = percentile(srv->latencies_ms, 0.50);
    out->p99_ms = percentile(srv->latencies_ms, 0.99);
    out->mean_batch = srv->batches ? (double)srv->batched_frames / (double)srv->batches : 0.0;
    out->hbm_in_use = srv->arena.used();
    out->hbm_peak = srv->peak;
    out->hbm_budget = srv->budget;
    return DPX_OK;
}

extern "C" int dpx_server_stats_reset(dpx_server* srv)
{
    DPX_CHECK(srv, DPX_EINVAL, "null server");
    std::lock_guard<std::mutex> lock(srv->mu);
    srv->submitted = srv->completed = srv->failed = srv->rejected = srv->cancelled = 0;
    srv->batches = srv->batched_frames = 0;
    srv->latencies_ms.clear();
    srv->peak = srv->arena.used();
    srv->t_create = Clock::now();
    return DPX_OK;
}

extern "C" int dpx_host_alloc(size_t bytes, void** out)
{
    DPX_CHECK(out && bytes > 0, DPX_EINVAL, "bad arguments");
    *out = nullptr;
    if (int rc = require_device()) return rc;
    void* p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess || !p)
        return fail(DPX_ENOMEM, "pinned host allocation of %zu bytes failed", bytes);
    {
        std::lock_guard<std::mutex> lock(g_pin_mu);
        g_pinned[reinterpret_cast<uintptr_t>(p)] = bytes;
    }
    *out = p;
    return DPX_OK;
}

extern "C" int dpx_host_free(void* p)
{
    if (!p) return DPX_OK;
    {
        std::lock_guard<std::mutex> lock(g_pin_mu);
        auto it = g_pinned.find(reinterpret_cast<uintptr_t>(p));
        DPX_CHECK(it != g_pinned.end(), DPX_EINVAL, "not a block from dpx_host_alloc");
        g_pinned.erase(it);
    }
    DPX_HIP(hipHostFree(p));
    return DPX_OK;
}

extern "C" int dpx_server_inject_failure(dpx_server* srv, uint64_t request_index)
{
    DPX_CHECK(srv, DPX_EINVAL, "null server");
    std::lock_guard<std::mutex> lock(srv->mu);
    srv->inject_at.push_back(request_index);
    return DPX_OK;
}
