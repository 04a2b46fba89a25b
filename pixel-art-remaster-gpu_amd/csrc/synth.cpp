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

// ---------------------------------------------------------------------------
// Game streams: palette-indexed frames of a scene whose background is a map
// of 16x16 tiles (drawn from a tileset of the same content as kind 0 frames)
// and whose sprites move by 1-4 pixels per axis from one frame to the next.
// Frame t is a pure function of (seed, t): the sprites' paths are replayed
// from frame 0.
// ---------------------------------------------------------------------------
constexpr int kStreamSprites = 7;

struct StreamSprite {
    int w, h, x, y, vx, vy;
    std::vector<uint8_t> idx;  // w*h palette indices, 0 = transparent
};

int stream_palette(SplitMix64& r, uint8_t* pal1k)
{
    const int npal = r.uni(16, 32);
    if (pal1k) {
        memset(pal1k, 0, 1024);
        for (int k = 0; k < npal; k++) {
            for (int c = 0; c < 3; c++) pal1k[4 * k + c] = (uint8_t)r.uni(0, 255);
            pal1k[4 * k + 3] = 255;
        }
    } else {
        for (int k = 0; k < 3 * npal; k++) r.next();
    }
    return npal;
}

// Velocity component: +-1..4 pixels per frame.
int stream_speed(SplitMix64& r) { return (r.uni(0, 1) ? 1 : -1) * r.uni(1, 4); }

void stream_scene(uint64_t seed, int W, int H, std::vector<uint8_t>& bg, std::vector<StreamSprite>& sprites,
                  uint8_t* pal1k)
{
    SplitMix64 r(seed * 0x2545F4914F6CDD1Dull + 0x3C6EF372FE94F82Bull);
    const int npal = stream_palette(r, pal1k);
    // identity palette in index form, so the frame drawers below write indices
    std::vector<uint8_t> ids((size_t)npal);
    for (int k = 0; k < npal; k++) ids[k] = (uint8_t)k;
    // tileset: 8 tiles of 16x16 with rectangles, a staircase, dither, dots
    constexpr int T = 16, NT = 8;
    std::vector<uint8_t> tiles((size_t)NT * T * T);
    for (int k = 0; k < NT; k++) {
        Canvas tc{ T, T, 1, T, &tiles[(size_t)k * T * T] };
        fill_rect(tc, 0, 0, T, T, &ids[r.uni(0, npal - 1)]);
        rects(tc, r, ids, npal, r.uni(0, 3), 2, 12);
        if (r.uni(0, 1)) lines(tc, r, ids, npal, 1);
        if (r.uni(0, 3) == 0) dithers(tc, r, ids, npal, 1, 4, 12);
        isolated(tc, r, ids, npal, r.uni(0, 3));
    }
    bg.assign((size_t)W * H, 0);
    Canvas cv{ W, H, 1, W, bg.data() };
    const int tx = (W + T - 1) / T, ty = (H + T - 1) / T;
    for (int j = 0; j < ty; j++)
        for (int i = 0; i < tx; i++) {
            const uint8_t* tile = &tiles[(size_t)r.uni(0, NT - 1) * T * T];
            for (int y = 0; y < T; y++)
                for (int x = 0; x < T; x++) cv.set(i * T + x, j * T + y, &tile[y * T + x]);
        }
    // platforms across tiles
    rects(cv, r, ids, npal, std::max(1, (int)(6.0 * W * H / (320.0 * 224.0))), 16, 96);
    // sprites: an outlined body with detail, transparent around it
    sprites.assign(kStreamSprites, StreamSprite{});
    for (StreamSprite& s : sprites) {
        s.w = r.uni(0, 1) ? 16 : 24;
        s.h = r.uni(0, 2) ? 16 : 24;
        s.w = std::min(s.w, W);
        s.h = std::min(s.h, H);
        s.x = r.uni(0, W - s.w);
        s.y = r.uni(0, H - s.h);
        s.vx = stream_speed(r);
        s.vy = stream_speed(r);
        s.idx.assign((size_t)s.w * s.h, 0);
        const int body = r.uni(1, npal - 1), detail = r.uni(1, npal - 1), outline = r.uni(1, npal - 1);
        const double cx = (s.w - 1) / 2.0, cy = (s.h - 1) / 2.0, ax = s.w / 2.0 * 0.95, ay = s.h / 2.0 * 0.95;
        auto inside = [&](int x, int y) {
            if (x < 0 || y < 0 || x >= s.w || y >= s.h) return false;
            const double u = (x - cx) / ax, v = (y - cy) / ay;
            return u * u + v * v <= 1.0;
        };
        const int dx0 = r.uni(0, s.w / 2), dy0 = r.uni(0, s.h / 2);
        for (int y = 0; y < s.h; y++)
            for (int x = 0; x < s.w; x++) {
                if (!inside(x, y)) continue;
                const bool edge = !inside(x - 1, y) || !inside(x + 1, y) || !inside(x, y - 1) || !inside(x, y + 1);
                const bool det = x >= dx0 && x < dx0 + s.w / 2 && y >= dy0 && y < dy0 + s.h / 2 && ((x + y) & 1);
                // index 0 stands for transparent: shift the body's indices by one
                s.idx[(size_t)y * s.w + x] = (uint8_t)(1 + (edge ? outline : (det ? detail : body)));
            }
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

extern "C" int dpx_synth_stream_frame(uint64_t seed, int t, int width, int height, uint8_t* out, int stride,
                                      uint8_t* palette)
{
    if (!out || width <= 0 || height <= 0 || stride < width || t < 0 || t > 1000000) return DPX_EINVAL;
    std::vector<uint8_t> bg;
    std::vector<StreamSprite> sprites;
    stream_scene(seed, width, height, bg, sprites, palette);
    // replay the sprites' paths: every frame each moves by its velocity
    // (1-4 px per axis), bouncing off the borders, and picks a new velocity
    // now and then
    SplitMix64 r(seed * 0x9E3779B97F4A7C15ull + 0x7F4A7C159E3779B9ull);
    for (int f = 1; f <= t; f++)
        for (StreamSprite& s : sprites) {
            if (r.uni(0, 23) == 0) {
                s.vx = stream_speed(r);
                s.vy = stream_speed(r);
            }
            int nx = s.x + s.vx, ny = s.y + s.vy;
            if (nx < 0 || nx > width - s.w) { s.vx = -s.vx; nx = s.x + s.vx; }
            if (ny < 0 || ny > height - s.h) { s.vy = -s.vy; ny = s.y + s.vy; }
            s.x = std::clamp(nx, 0, width - s.w);
            s.y = std::clamp(ny, 0, height - s.h);
        }
    for (int y = 0; y < height; y++) memcpy(out + (size_t)y * stride, &bg[(size_t)y * width], (size_t)width);
    for (const StreamSprite& s : sprites)
        for (int y = 0; y < s.h; y++)
            for (int x = 0; x < s.w; x++) {
                const uint8_t v = s.idx[(size_t)y * s.w + x];
                if (v) out[(size_t)(s.y + y) * stride + s.x + x] = (uint8_t)(v - 1);
            }
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
            // kind 0's content with 16-64 colours, and 15% of the pixels in
            // long 1-px random walks
            int npal = r.uni(16, 64);
            auto pal = make_palette(r, npal, channels);
            fill_rect(cv, 0, 0, width, height, &pal[0]);
            rects(cv, r, pal, npal, std::max(1, (int)(40 * area)), 4, 64);
            lines(cv, r, pal, npal, std::max(1, (int)(20 * area)));
            dithers(cv, r, pal, npal, std::max(1, (int)(4 * area)), 8, 32);
            isolated(cv, r, pal, npal, std::max(1, (int)(30 * area)));
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
