// common.hpp -- status codes, error reporting and launch helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>

#include "../../include/depixel.h"

namespace dpx {

// Records a message as the calling thread's last error and returns `code`.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Flags the library knows; the stage and per-frame paths take no indexed frames.
constexpr uint32_t kKnownFlags = DPX_RGB | DPX_TOPDOWN | DPX_ALPHA | DPX_FLIP_OUT | DPX_INDEXED;

// Bytes of a DPX_INDEXED palette: 256 four-byte entries.
constexpr size_t kPaletteBytes = 1024;

// Returns DPX_ENODEV when the process sees no GPU.
int require_device();

// Sizes a context's cell-row buffer (subdivided items) to at least `bytes`.
int context_reserve_rows(dpx_context* ctx, size_t bytes);

// Cell-row words per cell at scale S: 2S + 2 sample rows and a summary.
constexpr size_t cell_row_words(int S) { return 2 * (size_t)S + 3; }

}  // namespace dpx

#define DPX_HIP(call)                                                                                   \
    do {                                                                                                \
        hipError_t dpx_e_ = (call);                                                                     \
        if (dpx_e_ != hipSuccess)                                                                       \
            return ::dpx::fail(dpx_e_ == hipErrorOutOfMemory ? DPX_ENOMEM : DPX_EHIP, "%s: %s (%s:%d)", \
                               #call, hipGetErrorString(dpx_e_), __FILE__, __LINE__);                   \
    } while (0)

#define DPX_CHECK(cond, ...)                          \
    do {                                              \
        if (!(cond)) return ::dpx::fail(__VA_ARGS__); \
    } while (0)
