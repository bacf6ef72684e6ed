// rt_internal.h - shared by the HIP translation units of the library (not part of the C ABI)
#pragma once
#include <hip/hip_runtime.h>
extern hipStream_t g_bk_stream;
int bk_check(hipError_t e, const char *what);
int bk_fail(const char *what);                   // an invalid argument: message, -1 (the device stays usable)
#define BK_LAUNCH_CHECK(name) bk_check(hipGetLastError(), name)
static inline unsigned bk_grid1(long long n, int bs, unsigned cap = 65535u * 4) {
    long long g = (n + bs - 1) / bs;
    if (g < 1) g = 1;
    if (g > (long long)cap) g = cap;
    return (unsigned)g;
}
