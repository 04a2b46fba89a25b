// synth.cpp -- deterministic synthetic pixel art for tests and measurement.
//
// One generator (splitmix64) feeds the oracle, the per-frame path and the
// batched path, so every workload is reproducible from (kind, seed, size).
// Frames are written in buffer order (row 0 first); the byte order is the
// reference's (c0, c1, c2[, alpha]).
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/depixel.h"

namespace {

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // uniform integer in [lo, hi]
    int uni(int lo, int hi) { return lo + (int)(next() % (uint64_t)(hi - lo + 1)); }
};

struct Canvas {
    int W, H, ch, stride;
    uint8_t* p;
    void set(int x, int y, const uint8_t* c)
    {
        if (x < 0 || y < 0 || x >= W || y >= H) return;
        memcpy(p + (size_t)y * stride + (size_t)x * ch, c, (size_t)ch);
    }
    const uint8_t* get(int x, int y) const { return p + (size_t)y * stride + (size_t)x * ch; }
};

std::vector<uint8_t> make_palette(SplitMix64& r, int n, int ch)
{
    std::vector<uint8_t> pal((size_t)n * ch);
    for (int k = 0; k < n; k++) {
        for (int c = 0; c < 3; c++) pal[(size_t)k * ch + c] = (uint8_t)r.uni(0, 255);
        if (ch == 4) pal[(size_t)k * ch + 3] = 255;
    }
    return pal;
}

void fill_rect(Canvas& cv, int x0, int y0, int w, int h, const uint8_t* c)
{
    for (int y = y0; y < y0 + h; y++)
        for (int x = x0; x < x0 + w; x++) cv.set(x, y, c);
}

void rects(Canvas& cv, SplitMix64& r, const std::vector<uint8_t>& pal, int npal, int count, int lo, int hi)
{
    for (int k = 0; k < count; k++) {
        int w = std::min(r.uni(lo, hi), cv.W), h = std::min(r.uni(lo, hi), cv.H);
        int x = r.uni(0, cv.W - w), y = r.uni(0, cv.H - h);
        fill_rect(cv, x, y, w, h, &pal[(size_t)r.uni(0, npal - 1) * cv.ch]);
    }
}

// 1-px diagonal lines and staircases
void lines(Canvas& cv, SplitMix64& r, const std::vector<uint8_t>& pal, int npal, int count)
{
    for (int k = 0; k < count; k++) {
        int x = r.uni(0, cv.W - 1), y = r.uni(0, cv.H - 1);
        int len = r.uni(8, 64);
        int sx = r.uni(0, 1) ? 1 : -1, sy = r.uni(0, 1) ? 1 : -1;
        int run = r.uni(1, 3);  // 1: pure diagonal, 2-3: staircase with runs
        const uint8_t* c = &pal[(size_t)r.uni(0, npal - 1) * cv.ch];
        for (int s = 0; s < len; s++) {
            cv.set(x, y, c);
            if (run == 1) { x += sx; y += sy; }
            else { x += sx; if ((s + 1) % run == 0) y += sy; }
        }
    }
}

// 2x2-period checkerboard of two palette colours
void dither(Canvas& cv, SplitMix64& r, const std::vector<uint8_t>& pal, int npal, int x0, int y0, int w, int h)
{
    const uint8_t* a = &pal[(size_t)r.uni(0, npal - 1) * cv.ch];
    const uint8_t* b = &pal[(size_t)r.uni(0, npal - 1) * cv.ch];
    for (int y = y0; y < y0 + h; y++)
        for (int x = x0; x < x0 + w; x++) cv.set(x, y, ((x + y) & 1) ? a : b);
}

void dithers(Canvas& cv, SplitMix64& r, const std::vector<uint8_t>& pal, int npal, int count, int lo, int hi)
{
    for (int k = 0; k < count; k++) {
        int w = std::min(r.uni(lo, hi), cv.W), h = std::min(r.uni(lo, hi), cv.H);
        dither(cv, r, pal, npal, r.uni(0, cv.W - w), r.uni(0, cv.H - h), w, h);
    }
}

void isolated(Canvas& cv, SplitMix64& r, const std::vector<uint8_t>& pal, int npal, int count)
{
    for (int k = 0; k < count; k++)
        cv.set(r.uni(0, cv.W - 1), r.uni(0, cv.H - 1), &pal[(size_t)r.uni(0, npal - 1) * cv.ch]);
}

// 8-connected 1-px random walks covering about `frac` of the frame
void walks(Canvas& cv, SplitMix64& r, const std::vector<uint8_t>& pal, int npal, double frac)
{
    long target = (long)(frac * cv.W * cv.H), done = 0;
    while (done < target) {
        int x = r.uni(0, cv.W - 1), y = r.uni(0, cv.H - 1);
        int len = r.uni(20, 200);
        const uint8_t* c = &pal[(size_t)r.uni(0, npal - 1) * cv.ch];
        int d = r.uni(0, 7);
        for (int s = 0; s < len && done < target; s++) {
            cv.set(x, y, c);
            done++;
            if (r.uni(0, 3) == 0) d = (d + (r.uni(0, 1) ? 1 : 7)) & 7;
            static const int DX[8] = { 1, 1, 0, -1, -1, -1, 0, 1 };
            static const int DY[8] = { 0, 1, 1, 1, 0, -1, -1, -1 };
            x = std::clamp(x + DX[d], 0, cv.W - 1);
            y = std::clamp(y + DY[d], 0, cv.H - 1);
        }
    }
}

// diagonal crossings on a 3-px lattice inside a window
void lattice(Canvas& cv, SplitMix64& r, const std::vector<uint8_t>& pal, int npal, int x0, int y0, int w, int h)
{
    const uint8_t* a = &pal[(size_t)r.uni(0, npal - 1) * cv.ch];
    const uint8_t* b = &pal[(size_t)r.uni(0, npal - 1) * cv.ch];
    for (int y = y0; y < y0 + h; y++)
        for (int x = x0; x < x0 + w; x++) {
            int u = (x + y) % 3, v = ((x - y) % 3 + 3) % 3;
            if (u == 0) cv.set(x, y, a);
            else if (v == 0) cv.set(x, y, b);
        }
}

void sprite(Canvas& cv, SplitMix64& r)
{
    int npal = r.uni(4, 12);
    auto pal = make_palette(r, npal, 4);
    uint8_t clear[4] = { 0, 0, 0, 0 };
    fill_rect(cv, 0, 0, cv.W, cv.H, clear);
    // body: an ellipse covering about 70% of the box
    double cx = (cv.W - 1) / 2.0, cy = (cv.H - 1) / 2.0;
    double ax = cv.W / 2.0 * 0.944, ay = cv.H / 2.0 * 0.944;
    std::vector<uint8_t> body((size_t)cv.W * cv.H, 0);
    for (int y = 0; y < cv.H; y++)
        for (int x = 0; x < cv.W; x++) {
            double u = (x - cx) / ax, v = (y - cy) / ay;
            if (u * u + v * v <= 1.0) {
                body[(size_t)y * cv.W + x] = 1;
                cv.set(x, y, &pal[0]);
            }
        }
    // interior detail: rectangles and dither, clipped to the body
    int nr = r.uni(2, 6);
    for (int k = 0; k < nr; k++) {
        // (at most the sprite's size: a sprite one pixel wide or tall)
        int w = std::min(cv.W, r.uni(2, std::max(2, cv.W / 2))), h = std::min(cv.H, r.uni(2, std::max(2, cv.H / 2)));
        int x0 = r.uni(0, cv.W - w), y0 = r.uni(0, cv.H - h);
        const uint8_t* c = &pal[(size_t)r.uni(0, npal - 1) * 4];
        const uint8_t* c2 = &pal[(size_t)r.uni(0, npal - 1) * 4];
        bool dith = r.uni(0, 2) == 0;
        for (int y = y0; y < y0 + h; y++)
            for (int x = x0; x < x0 + w; x++)
                if (body[(size_t)y * cv.W + x]) cv.set(x, y, (dith && ((x + y) & 1)) ? c2 : c);
    }
    // dark 1-px outline on the body's border
    uint8_t dark[4] = { (uint8_t)r.uni(0, 40), (uint8_t)r.uni(0, 40), (uint8_t)r.uni(0, 40), 255 };
    for (int y = 0; y < cv.H; y++)
        for (int x = 0; x < cv.W; x++) {
            if (!body[(size_t)y * cv.W + x]) continue;
            bool edge = x == 0 || y == 0 || x == cv.W - 1 || y == cv.H - 1 ||
                        !body[(size_t)y * cv.W + x - 1] || !body[(size_t)y * cv.W + x + 1] ||
                        !body[(size_t)(y - 1) * cv.W + x] || !body[(size_t)(y + 1) * cv.W + x];
            if (edge) cv.set(x, y, dark);
        }
}

const int kSpriteW[6] = { 8, 16, 16, 24, 32, 48 };
const int kSpriteH[6] = { 8, 16, 32, 24, 32, 64 };

}  // namespace

extern "C" int dpx_synth_sprite_size(uint64_t seed, int index, int* width, int* height)
{
    if (!width || !height || index < 0) return DPX_EINVAL;
    SplitMix64 r(seed * 0x100000001B3ull + (uint64_t)index * 0x9E3779B1ull + 17);
    int k = r.uni(0, 5);
    *width = kSpriteW[k];
    *height = kSpriteH[k];
    return DPX_OK;
}

extern "C" int dpx_synth_frame(int kind, uint64_t seed, int width, int height, int channels, uint8_t* out,
                               int stride)
{
    if (!out || width <= 0 || height <= 0 || (channels != 3 && channels != 4) || stride < width * channels)
        return DPX_EINVAL;
    if (kind == 1 && channels != 4) return DPX_EINVAL;
    memset(out, 0, (size_t)stride * height);
    Canvas cv{ width, height, channels, stride, out };
    SplitMix64 r(seed ^ ((uint64_t)kind << 56) ^ 0xD1B54A32D192ED03ull);
    double area = (double)width * height / (256.0 * 224.0);
    switch (kind) {
        case 0: {
            auto pal = make_palette(r, 16, channels);
            fill_rect(cv, 0, 0, width, height, &pal[0]);
            rects(cv, r, pal, 16, std::max(1, (int)(40 * area)), 4, 64);
            lines(cv, r, pal, 16, std::max(1, (int)(20 * area)));
            dithers(cv, r, pal, 16, std::max(1, (int)(4 * area)), 8, 32);
            isolated(cv, r, pal, 16, std::max(1, (int)(30 * area)));
            return DPX_OK;
        }
        case 1:
            sprite(cv, r);
            return DPX_OK;
        case 2: {
            int npal = r.uni(16, 64);
            auto pal = make_palette(r, npal, channels);
            fill_rect(cv, 0, 0, width, height, &pal[0]);
            rects(cv, r, pal, npal, std::max(1, (int)(40 * area)), 4, 64);
            walks(cv, r, pal, npal, 0.15);
            return DPX_OK;
        }
        case 4: {
            auto pal = make_palette(r, 256, channels);
            fill_rect(cv, 0, 0, width, height, &pal[0]);
            rects(cv, r, pal, 256, std::max(1, (int)(40 * area)), 4, 96);
            // about 25% of the area dithered
            long want = (long)(0.25 * width * height), got = 0;
            while (got < want) {
                int w = std::min(r.uni(16, 96), width), h = std::min(r.uni(16, 96), height);
                dither(cv, r, pal, 256, r.uni(0, width - w), r.uni(0, height - h), w, h);
                got += (long)w * h;
            }
            // about 10% of the area as a 3-px diagonal lattice
            want = (long)(0.10 * width * height);
            got = 0;
            while (got < want) {
                int w = std::min(r.uni(24, 96), width), h = std::min(r.uni(24, 96), height);
                lattice(cv, r, pal, 256, r.uni(0, width - w), r.uni(0, height - h), w, h);
                got += (long)w * h;
            }
            isolated(cv, r, pal, 256, std::max(1, (int)(30 * area)));
            return DPX_OK;
        }
        default:
            return DPX_EINVAL;
    }
}
