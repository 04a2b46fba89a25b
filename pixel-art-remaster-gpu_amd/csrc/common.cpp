// common.cpp -- last-error storage and device probing.
#include "common.hpp"

#include <stdio.h>

namespace {
thread_local char g_last_error[512] = "";
}

namespace dpx {

int fail(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return code;
}

int require_device()
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) return fail(DPX_ENODEV, "no GPU visible to this process (%s)", hipGetErrorString(e));
    return DPX_OK;
}

}  // namespace dpx

extern "C" const char* dpx_last_error(void) { return g_last_error; }

extern "C" const char* dpx_version(void) { return "depixel-mi355x 0.1 (gfx950)"; }
