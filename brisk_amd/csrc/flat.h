// flat.h - flattened grids for batched launches. The workgroups of every item of a launch form
// one 1-D grid: item q owns workgroups [pre[q], pre[q+1]), and a workgroup finds its item by a
// binary search over the staged prefix array. A launch over items of very different sizes (a
// serving batch) then costs what its items need, not (largest item) x (number of items).
#pragma once
#include <stdint.h>
#include <string.h>
#include <vector>
#include "brisk_dev.h"

// per-item workgroup counts -> the prefix array staged on the device; returns the total
// (0: nothing to launch) or -1 (staging failed or more than 2^31 - 1 workgroups)
static inline int64_t flat_stage(const std::vector<int64_t> &cnt, const int **dpre) {
    const size_t n = cnt.size();
    std::vector<int> pre(n + 1);
    int64_t tot = 0;
    for (size_t q = 0; q < n; q++) {
        pre[q] = (int)tot;
        tot += cnt[q] > 0 ? cnt[q] : 0;
        if (tot > 0x7fffffff) return -1;
    }
    pre[n] = (int)tot;
    *dpre = (const int *)bk_stage(pre.data(), sizeof(int) * (n + 1));
    return *dpre ? tot : -1;
}
static inline int64_t flat_cap(int64_t need, int64_t cap) { return need < cap ? need : cap; }
// the item array and its prefix array staged together (one copy); returns as flat_stage
static inline int64_t flat_stage_items(const void *items, size_t bytes, const std::vector<int64_t> &cnt, const void **ditems,
                                       const int **dpre) {
    const size_t n = cnt.size(), po = (bytes + 15) & ~(size_t)15;
    std::vector<char> buf(po + sizeof(int) * (n + 1));
    memcpy(buf.data(), items, bytes);
    int *pre = (int *)(buf.data() + po);
    int64_t tot = 0;
    for (size_t q = 0; q < n; q++) {
        pre[q] = (int)tot;
        tot += cnt[q] > 0 ? cnt[q] : 0;
        if (tot > 0x7fffffff) return -1;
    }
    pre[n] = (int)tot;
    const char *d = (const char *)bk_stage(buf.data(), buf.size());
    if (!d) return -1;
    *ditems = d;
    *dpre = (const int *)(d + po);
    return tot;
}

// the item of workgroup b: pre[q] <= b < pre[q + 1] (pre[0] = 0, pre[n] = total > b)
__device__ __forceinline__ int flat_item(const int *__restrict__ pre, int n, int b) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pre[mid] <= b) lo = mid; else hi = mid;
    }
    return lo;
}
// row-major enumeration of the lower triangle: t = a (a + 1) / 2 + b with b <= a
__device__ __forceinline__ void tri_index(int64_t t, int &a, int &b) {
    int64_t x = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while (x * (x + 1) / 2 > t) x--;
    while ((x + 1) * (x + 2) / 2 <= t) x++;
    a = (int)x;
    b = (int)(t - x * (x + 1) / 2);
}
