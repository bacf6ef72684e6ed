// rt.hip - device runtime of the solver library: one stream, device allocations, a pinned
// staging ring for the small item arrays the batched launches read, events and errors.
#include <hip/hip_runtime.h>
#include <thread>
#include <stdio.h>
#include <string.h>
#include <time.h>
#include "brisk_dev.h"
#include "rt_internal.h"

static int g_init = 0;
static int g_failed = 0;
static char g_err[256] = "";
hipStream_t g_bk_stream = nullptr;   // the stream launches go to (the library stream unless forked)
static hipStream_t g_main_stream = nullptr, g_side_stream = nullptr;
static hipEvent_t g_fork_ev = nullptr, g_join_ev = nullptr, g_side_ev = nullptr;
static int g_forked = 0;
static hipEvent_t g_sync_ev;

// staging ring: host pinned + device, recycled at every synchronisation
static char *g_ring_h = nullptr, *g_ring_d = nullptr;
// transfers up to KCOPY_MAX bytes are kernels on the library stream that read or write pinned
// host memory directly (k_pull, k_push): a DMA transfer between kernels costs the compute queue
// far more idle time than the few microseconds of such a kernel
#define KCOPY_MAX ((size_t)1 << 20)
static char *g_bounce = nullptr;      // pinned: destination of small device-to-host reads
#define BOUNCE_BYTES ((size_t)32 << 20)   // its size: a batch of small reads lands in it before one wait
// larger transfers go through two pinned buffers of XFER_CHUNK bytes in turn, so that the host's
// copy of one chunk (to or from pageable memory) overlaps the DMA of the next: a DMA from
// pageable memory runs at a few GB/s, from pinned memory at the link's rate
#define XFER_CHUNK ((size_t)16 << 20)
static char *g_xfer[2] = { nullptr, nullptr };
static hipEvent_t g_xfer_ev[2];
// a host copy of one staging chunk, split over up to 8 threads (one thread copies a few GB/s)
static void pcopy(void *dst, const void *src, size_t len) {
    const size_t part = (size_t)2 << 20;
    int np = (int)(len / part);
    if (np > 8) np = 8;
    const unsigned hw = std::thread::hardware_concurrency();
    if (hw && np > (int)hw) np = (int)hw;
    if (np < 2) { memcpy(dst, src, len); return; }
    std::thread th[8];
    for (int p = 1; p < np; p++) {
        const size_t lo = len * p / np, hi = len * (p + 1) / np;
        th[p] = std::thread([=] { memcpy((char *)dst + lo, (const char *)src + lo, hi - lo); });
    }
    memcpy(dst, src, len / np);
    for (int p = 1; p < np; p++) th[p].join();
}
static size_t g_ring_cap = 0, g_ring_pos = 0;
#define CACHE_CAP ((size_t)32 << 20)
#define CACHE_CHUNK ((size_t)64 << 10)
#define CACHE_TAB 16384
typedef struct { uint64_t hash; uint32_t bytes, off; uint64_t wrap; } CacheEnt;
static char *g_cache_h = nullptr, *g_cache_d = nullptr;
static size_t g_cache_pos = 0;
static uint64_t g_cache_wrap = 1, g_epoch = 1;
static CacheEnt g_cache_tab[CACHE_TAB];
static uint64_t g_chunk_use[CACHE_CAP / CACHE_CHUNK];
static double g_cache_hits = 0, g_cache_miss = 0;

int bk_fail(const char *what) {
    snprintf(g_err, sizeof(g_err), "%s", what);
    fprintf(stderr, "brisk gpu: %s\n", g_err);
    return -1;
}
int bk_check(hipError_t e, const char *what) {
    if (e == hipSuccess) return 0;
    g_failed = 1;
    snprintf(g_err, sizeof(g_err), "%s: %s", what, hipGetErrorString(e));
    fprintf(stderr, "brisk gpu: %s\n", g_err);
    return -1;
}

extern "C" int bk_init(void) {
    if (g_init) return g_failed ? -1 : 0;
    int n = 0;
    if (bk_check(hipGetDeviceCount(&n), "hipGetDeviceCount")) return -1;
    if (n < 1) { snprintf(g_err, sizeof(g_err), "no GPU"); g_failed = 1; return -1; }
    // the host waits on the stream many times per iteration: spin rather than sleep (a
    // request that fails because a context already exists is ignored)
    (void)hipSetDeviceFlags(hipDeviceScheduleSpin);
    (void)hipGetLastError();
    if (bk_check(hipSetDevice(0), "hipSetDevice")) return -1;
    if (bk_check(hipStreamCreateWithFlags(&g_main_stream, hipStreamNonBlocking), "hipStreamCreate")) return -1;
    g_bk_stream = g_main_stream;
    if (bk_check(hipEventCreateWithFlags(&g_sync_ev, hipEventDisableTiming), "hipEventCreate")) return -1;
    g_ring_cap = (size_t)64 << 20;
    if (bk_check(hipHostMalloc((void **)&g_ring_h, g_ring_cap, hipHostMallocDefault), "hipHostMalloc ring")) return -1;
    if (bk_check(hipMalloc((void **)&g_ring_d, g_ring_cap), "hipMalloc ring")) return -1;
    g_ring_pos = 0;
    if (bk_check(hipHostMalloc((void **)&g_cache_h, CACHE_CAP, hipHostMallocDefault), "hipHostMalloc cache")) return -1;
    if (bk_check(hipMalloc((void **)&g_cache_d, CACHE_CAP), "hipMalloc cache")) return -1;
    if (bk_check(hipHostMalloc((void **)&g_bounce, BOUNCE_BYTES, hipHostMallocDefault), "hipHostMalloc bounce")) return -1;
    for (int b = 0; b < 2; b++) {
        if (bk_check(hipHostMalloc((void **)&g_xfer[b], XFER_CHUNK, hipHostMallocDefault), "hipHostMalloc transfer")) return -1;
        if (bk_check(hipEventCreateWithFlags(&g_xfer_ev[b], hipEventDisableTiming), "hipEventCreate transfer")) return -1;
    }
    // the device address of the no-allocation symmetric product's scratch array (k_dense.hip)
    if (bk_symv_scratch_init()) return -1;
    g_init = 1;
    return 0;
}

extern "C" const char *bk_error(void) { return g_err; }
extern "C" int bk_failed(void) { return g_failed; }

extern "C" void *bk_malloc(size_t bytes) {
    void *p = nullptr;
    if (bytes == 0) bytes = 256;
    if (bk_check(hipMalloc(&p, bytes), "hipMalloc")) return nullptr;
    return p;
}
extern "C" void bk_free(void *p) { if (p) (void)hipFree(p); }
extern "C" void *bk_host_alloc(size_t bytes) {
    void *p = nullptr;
    if (bytes == 0) bytes = 256;
    if (bk_check(hipHostMalloc(&p, bytes, hipHostMallocDefault), "hipHostMalloc")) return nullptr;
    return p;
}
extern "C" void bk_host_free(void *p) { if (p) (void)hipHostFree(p); }

static int g_prof_on_host = 0;
static inline int g_prof_host_on(void) { return g_prof_on_host; }
// host-side counters (read as profile classes 62: waits on the stream, 63: staged copies)
static double g_sync_ms = 0, g_sync_n = 0, g_stage_n = 0, g_stage_bytes = 0;
// A second stream for independent work of a batch (the f32 and the f64 factorizations):
// bk_fork makes it wait for the work queued so far, bk_side directs launches to it, bk_join
// makes the library stream wait for it. Both streams have the default priority.
extern "C" int bk_fork(void) {
    if (g_forked) return 0;
    if (!g_side_stream) {
        if (bk_check(hipStreamCreateWithFlags(&g_side_stream, hipStreamNonBlocking), "hipStreamCreate side")) return -1;
        if (bk_check(hipEventCreateWithFlags(&g_fork_ev, hipEventDisableTiming), "hipEventCreate") ||
            bk_check(hipEventCreateWithFlags(&g_join_ev, hipEventDisableTiming), "hipEventCreate") ||
            bk_check(hipEventCreateWithFlags(&g_side_ev, hipEventDisableTiming), "hipEventCreate")) return -1;
    }
    if (bk_check(hipEventRecord(g_fork_ev, g_main_stream), "hipEventRecord fork") ||
        bk_check(hipStreamWaitEvent(g_side_stream, g_fork_ev, 0), "hipStreamWaitEvent fork")) return -1;
    g_forked = 1;
    return 0;
}
extern "C" void bk_side(int on) { g_bk_stream = (on && g_forked) ? g_side_stream : g_main_stream; }
extern "C" int bk_join(void) {
    g_bk_stream = g_main_stream;
    if (!g_forked) return 0;
    g_forked = 0;
    if (bk_check(hipEventRecord(g_join_ev, g_side_stream), "hipEventRecord join")) return -1;
    return bk_check(hipStreamWaitEvent(g_main_stream, g_join_ev, 0), "hipStreamWaitEvent join");
}
extern "C" int bk_sync(void) {
    const double t0 = g_prof_host_on() ? bk_now() : 0;
    // an event after the queued work, polled without sleeping: the host resumes within
    // microseconds of the GPU instead of after a wake-up (both streams while forked)
    int r = bk_check(hipEventRecord(g_sync_ev, g_main_stream), "hipEventRecord");
    if (!r && g_forked) r = bk_check(hipEventRecord(g_side_ev, g_side_stream), "hipEventRecord side");
    if (!r) {
        hipError_t e;
        while ((e = hipEventQuery(g_sync_ev)) == hipErrorNotReady) { }
        r = bk_check(e, "hipEventQuery");
        if (!r && g_forked) {
            while ((e = hipEventQuery(g_side_ev)) == hipErrorNotReady) { }
            r = bk_check(e, "hipEventQuery side");
        }
    }
    if (t0 > 0) { g_sync_ms += 1e3 * (bk_now() - t0); g_sync_n += 1; }
    g_ring_pos = 0;
    g_epoch++;
    return r;
}
// byte copy by a kernel: 16- or 8-byte words when both ends are aligned to them, then the tail
__global__ void k_kcopy(unsigned char *__restrict__ dst, const unsigned char *__restrict__ src, size_t bytes) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x, nt = (size_t)gridDim.x * blockDim.x;
    const uintptr_t al = (uintptr_t)dst | (uintptr_t)src;
    size_t head = 0;
    if ((al & 15) == 0) {
        const size_t n16 = bytes / 16;
        for (size_t q = t; q < n16; q += nt) ((uint4 *)dst)[q] = ((const uint4 *)src)[q];
        head = n16 * 16;
    } else if ((al & 7) == 0) {
        const size_t n8 = bytes / 8;
        for (size_t q = t; q < n8; q += nt) ((unsigned long long *)dst)[q] = ((const unsigned long long *)src)[q];
        head = n8 * 8;
    }
    for (size_t q = head + t; q < bytes; q += nt) dst[q] = src[q];
}
static int kcopy(void *dst, const void *src, size_t bytes, const char *what) {
    const size_t g = (bytes + 4095) / 4096;
    hipLaunchKernelGGL(k_kcopy, dim3((unsigned)(g < 64 ? (g ? g : 1) : 64)), dim3(256), 0, g_bk_stream, (unsigned char *)dst,
                       (const unsigned char *)src, bytes);
    return bk_check(hipGetLastError(), what);
}
// host (pinned) -> device on the stream
static int pull(void *dst, const void *pinned, size_t bytes, const char *what) {
    if (bytes <= KCOPY_MAX) return kcopy(dst, pinned, bytes, what);
    return bk_check(hipMemcpyAsync(dst, pinned, bytes, hipMemcpyHostToDevice, g_bk_stream), what);
}
extern "C" int bk_h2d(void *dst, const void *src, size_t bytes) {
    if (!bytes) return 0;
    if (bytes <= KCOPY_MAX) return bk_h2d_async(dst, src, bytes) || bk_sync() ? -1 : 0;
    // chunk k: the host copies it into buffer k % 2 (once that buffer's previous DMA is done),
    // then its DMA is queued on the stream
    for (size_t off = 0, k = 0; off < bytes; off += XFER_CHUNK, k++) {
        const int b = (int)(k & 1);
        const size_t len = bytes - off < XFER_CHUNK ? bytes - off : XFER_CHUNK;
        if (k >= 2 && bk_check(hipEventSynchronize(g_xfer_ev[b]), "hipEventSynchronize H2D")) return -1;
        pcopy(g_xfer[b], (const char *)src + off, len);
        if (bk_check(hipMemcpyAsync((char *)dst + off, g_xfer[b], len, hipMemcpyHostToDevice, g_bk_stream), "hipMemcpy H2D") ||
            bk_check(hipEventRecord(g_xfer_ev[b], g_bk_stream), "hipEventRecord H2D")) return -1;
    }
    return bk_sync();
}
extern "C" int bk_d2h(void *dst, const void *src, size_t bytes) {
    if (!bytes) return bk_sync();
    if (bytes <= KCOPY_MAX) {
        if (kcopy(g_bounce, src, bytes, "d2h") || bk_sync()) return -1;
        memcpy(dst, g_bounce, bytes);
        return 0;
    }
    // chunk k is DMAed into buffer k % 2 while the host copies chunk k - 1 out of the other
    size_t prev_off = 0, prev_len = 0;
    for (size_t off = 0, k = 0; off < bytes; off += XFER_CHUNK, k++) {
        const int b = (int)(k & 1);
        const size_t len = bytes - off < XFER_CHUNK ? bytes - off : XFER_CHUNK;
        if (bk_check(hipMemcpyAsync(g_xfer[b], (const char *)src + off, len, hipMemcpyDeviceToHost, g_bk_stream), "hipMemcpy D2H") ||
            bk_check(hipEventRecord(g_xfer_ev[b], g_bk_stream), "hipEventRecord D2H")) return -1;
        if (k > 0) {
            if (bk_check(hipEventSynchronize(g_xfer_ev[b ^ 1]), "hipEventSynchronize D2H")) return -1;
            pcopy((char *)dst + prev_off, g_xfer[b ^ 1], prev_len);
        }
        prev_off = off; prev_len = len;
    }
    const int lb = (int)(((bytes + XFER_CHUNK - 1) / XFER_CHUNK - 1) & 1);
    if (bk_check(hipEventSynchronize(g_xfer_ev[lb]), "hipEventSynchronize D2H")) return -1;
    pcopy((char *)dst + prev_off, g_xfer[lb], prev_len);
    return bk_sync();
}
// n device-to-host reads with one wait per 32 MB: the small ones (up to KCOPY_MAX each) are copied
// by kernels into the pinned bounce buffer one after another, then the stream is waited for once
// and each is copied out; a larger one goes through bk_d2h on its own
extern "C" int bk_d2h_many(int n, void *const *dst, const void *const *src, const size_t *bytes) {
    int i = 0;
    while (i < n) {
        size_t used = 0;
        int j = i;
        for (; j < n; j++) {
            const size_t al = (bytes[j] + 255) & ~(size_t)255;
            if (bytes[j] > KCOPY_MAX || used + al > BOUNCE_BYTES) break;
            if (bytes[j] && kcopy(g_bounce + used, src[j], bytes[j], "d2h")) return -1;
            used += al;
        }
        if (j == i) {
            if (bk_d2h(dst[i], src[i], bytes[i])) return -1;
            i++;
            continue;
        }
        if (bk_sync()) return -1;
        size_t off = 0;
        for (int t = i; t < j; t++) {
            if (bytes[t]) memcpy(dst[t], g_bounce + off, bytes[t]);
            off += (bytes[t] + 255) & ~(size_t)255;
        }
        i = j;
    }
    return 0;
}
extern "C" int bk_d2d(void *dst, const void *src, size_t bytes) {
    if (!bytes) return 0;
    return bk_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, g_bk_stream), "hipMemcpy D2D");
}

__global__ void k_zero(unsigned long long *p, size_t nw) {
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < nw; q += (size_t)gridDim.x * blockDim.x) p[q] = 0ull;
}
__global__ void k_zero_bytes(unsigned char *p, size_t nb) {
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < nb; q += (size_t)gridDim.x * blockDim.x) p[q] = 0;
}
extern "C" int bk_zero(void *dst, size_t bytes) {
    if (!bytes) return 0;
    if (((uintptr_t)dst & 7) == 0 && (bytes & 7) == 0) {
        size_t nw = bytes / 8;
        unsigned grid = (unsigned)((nw + 255) / 256 < 4096 ? (nw + 255) / 256 : 4096);
        hipLaunchKernelGGL(k_zero, dim3(grid), dim3(256), 0, g_bk_stream, (unsigned long long *)dst, nw);
    } else {
        unsigned grid = (unsigned)((bytes + 255) / 256 < 4096 ? (bytes + 255) / 256 : 4096);
        hipLaunchKernelGGL(k_zero_bytes, dim3(grid), dim3(256), 0, g_bk_stream, (unsigned char *)dst, bytes);
    }
    return bk_check(hipGetLastError(), "k_zero");
}

// ---- staged item arrays: a content-addressed cache in device memory -------------------------
// Most item arrays repeat exactly from one launch to the next (the same solves, the same
// operands); a repeated array is found by its hash and reused without a copy. The pool is a ring
// of 64 KB chunks; a chunk used since the last synchronisation is never overwritten (the ring
// synchronises first), so a queued kernel never sees its items change.
static uint64_t stage_hash(const void *p, size_t n) {
    const unsigned char *b = (const unsigned char *)p;
    uint64_t h = 0x9e3779b97f4a7c15ull ^ (uint64_t)n;
    size_t i = 0;
    for (; i + 8 <= n; i += 8) {
        uint64_t w;
        memcpy(&w, b + i, 8);
        h = (h ^ w) * 0xff51afd7ed558ccdull;
        h ^= h >> 32;
    }
    for (; i < n; i++) h = (h ^ b[i]) * 0x100000001b3ull;
    return h ^ (h >> 29);
}
// an entry is live while the ring has not passed over its bytes since it was written
static int cache_live(const CacheEnt &e) {
    if (!e.bytes) return 0;
    if (e.wrap == g_cache_wrap) return 1;
    return e.wrap + 1 == g_cache_wrap && e.off >= g_cache_pos;
}
static void *stage_cached(const void *host, size_t bytes) {
    const uint64_t h = stage_hash(host, bytes);
    const size_t need = (bytes + 255) & ~(size_t)255;
    size_t slot = h % CACHE_TAB;
    for (int probe = 0; probe < 8; probe++, slot = (slot + 1) % CACHE_TAB) {
        CacheEnt &e = g_cache_tab[slot];
        if (!cache_live(e)) continue;
        if (e.hash == h && e.bytes == bytes && !memcmp(g_cache_h + e.off, host, bytes)) {
            // data of the previous lap, ahead of the ring: the ring must not pass over it before a synchronisation
            if (e.wrap != g_cache_wrap)
                for (size_t c = e.off / CACHE_CHUNK; c <= (e.off + need - 1) / CACHE_CHUNK; c++) g_chunk_use[c] = g_epoch;
            g_cache_hits += 1;
            return g_cache_d + e.off;
        }
    }
    // miss: the next bytes of the ring (synchronise first if they were used in this epoch)
    // a new lap starts after a synchronisation (the last lap's arrays may be in use); within a
    // lap, data of the previous lap read in this epoch is not overwritten before one
    if (g_cache_pos + need > CACHE_CAP) { bk_sync(); g_cache_pos = 0; g_cache_wrap++; }
    for (size_t c = g_cache_pos / CACHE_CHUNK; c <= (g_cache_pos + need - 1) / CACHE_CHUNK; c++)
        if (g_chunk_use[c] == g_epoch) { bk_sync(); break; }
    const size_t off = g_cache_pos;
    char *hp = g_cache_h + off, *dp = g_cache_d + off;
    memcpy(hp, host, bytes);
    if (pull(dp, hp, bytes, "stage")) return nullptr;
    g_cache_pos += need;
    // insert (replacing a dead entry in the probe window, else the first)
    slot = h % CACHE_TAB;
    size_t put = slot;
    for (int probe = 0; probe < 8; probe++) {
        const size_t s2 = (slot + probe) % CACHE_TAB;
        if (!cache_live(g_cache_tab[s2])) { put = s2; break; }
    }
    CacheEnt ne = { h, (uint32_t)bytes, (uint32_t)off, g_cache_wrap };
    g_cache_tab[put] = ne;
    g_cache_miss += 1;
    g_stage_n += 1; g_stage_bytes += (double)bytes;
    return dp;
}

extern "C" void *bk_stage(const void *host, size_t bytes) {
    if (g_cache_d && bytes > 0 && bytes <= CACHE_CHUNK) return stage_cached(host, bytes);
    size_t need = (bytes + 255) & ~(size_t)255;
    if (need == 0) need = 256;
    if (need > g_ring_cap) { snprintf(g_err, sizeof(g_err), "staging request of %zu bytes too large", bytes); g_failed = 1; return nullptr; }
    if (g_ring_pos + need > g_ring_cap) bk_sync();            // every staged copy has completed: recycle
    char *h = g_ring_h + g_ring_pos, *d = g_ring_d + g_ring_pos;
    memcpy(h, host, bytes);
    if (pull(d, h, bytes, "stage")) return nullptr;
    g_stage_n += 1; g_stage_bytes += (double)bytes;
    g_ring_pos += need;
    return d;
}

extern "C" int bk_h2d_async(void *dst, const void *src, size_t bytes) {
    if (!bytes) return 0;
    // small uploads go through the pinned ring (no wait); large ones are copied synchronously
    if (bytes > g_ring_cap / 4) return bk_h2d(dst, src, bytes);
    const size_t need = (bytes + 255) & ~(size_t)255;
    if (g_ring_pos + need > g_ring_cap) bk_sync();
    char *h = g_ring_h + g_ring_pos;
    memcpy(h, src, bytes);
    if (pull(dst, h, bytes, "hipMemcpy H2D async")) return -1;
    g_ring_pos += need;
    return 0;
}

extern "C" double bk_now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}
extern "C" size_t bk_mem_free(void) {
    size_t f = 0, t = 0;
    if (hipMemGetInfo(&f, &t) != hipSuccess) return 0;
    return f;
}
extern "C" void *bk_event_new(void) {
    hipEvent_t e;
    if (bk_check(hipEventCreate(&e), "hipEventCreate")) return nullptr;
    return (void *)e;
}
extern "C" void bk_event_record(void *e) { bk_check(hipEventRecord((hipEvent_t)e, g_bk_stream), "hipEventRecord"); }
extern "C" float bk_event_ms(void *a, void *b) {
    float ms = 0;
    bk_check(hipEventSynchronize((hipEvent_t)b), "hipEventSynchronize");
    bk_check(hipEventElapsedTime(&ms, (hipEvent_t)a, (hipEvent_t)b), "hipEventElapsedTime");
    return ms;
}
extern "C" void bk_event_free(void *e) { if (e) (void)hipEventDestroy((hipEvent_t)e); }
// ---- optional per-class kernel timing: an event pair around each instrumented launch on the
// library stream; read back (and summed per class) by bk_prof_read
#define PROF_MAX 32768
#define PROF_CLS 64
static int g_prof_on = 0, g_prof_n = 0;
static hipEvent_t g_prof_ev[2 * PROF_MAX];
static int g_prof_cls[PROF_MAX];
static double g_prof_flops[PROF_MAX], g_prof_bytes[PROF_MAX];
static double g_acc_ms[PROF_CLS], g_acc_n[PROF_CLS], g_acc_flops[PROF_CLS], g_acc_bytes[PROF_CLS];
static int g_prof_evinit = 0;
static void prof_flush(void) {
    if (!g_prof_n) return;
    (void)hipStreamSynchronize(g_main_stream);
    if (g_side_stream) (void)hipStreamSynchronize(g_side_stream);
    for (int i = 0; i < g_prof_n; i++) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, g_prof_ev[2 * i], g_prof_ev[2 * i + 1]) == hipSuccess) {
            const int c = g_prof_cls[i];
            g_acc_ms[c] += ms; g_acc_n[c] += 1; g_acc_flops[c] += g_prof_flops[i]; g_acc_bytes[c] += g_prof_bytes[i];
        }
    }
    g_prof_n = 0;
}
extern "C" void bk_prof_enable(int on) {
    if (on && !g_prof_evinit) {
        for (int i = 0; i < 2 * PROF_MAX; i++) (void)hipEventCreate(&g_prof_ev[i]);
        g_prof_evinit = 1;
    }
    g_prof_on = on;
    g_prof_on_host = on;
}
extern "C" void bk_prof_reset(void) {
    prof_flush();
    g_sync_ms = g_sync_n = g_stage_n = g_stage_bytes = 0; g_cache_hits = g_cache_miss = 0;
    memset(g_acc_ms, 0, sizeof(g_acc_ms)); memset(g_acc_n, 0, sizeof(g_acc_n));
    memset(g_acc_flops, 0, sizeof(g_acc_flops)); memset(g_acc_bytes, 0, sizeof(g_acc_bytes));
}
extern "C" int bk_prof_begin(void) {
    if (!g_prof_on) return -1;
    if (g_prof_n == PROF_MAX) prof_flush();
    const int i = g_prof_n++;
    (void)hipEventRecord(g_prof_ev[2 * i], g_bk_stream);
    return i;
}
extern "C" void bk_prof_end(int tok, int cls, double flops, double bytes) {
    if (tok < 0 || cls < 0 || cls >= PROF_CLS) return;
    (void)hipEventRecord(g_prof_ev[2 * tok + 1], g_bk_stream);
    g_prof_cls[tok] = cls; g_prof_flops[tok] = flops; g_prof_bytes[tok] = bytes;
}
extern "C" int bk_prof_on(void) { return g_prof_on; }
extern "C" void bk_prof_host(int cls, double seconds) {
    if (cls < 0 || cls >= PROF_CLS) return;
    g_acc_ms[cls] += 1e3 * seconds; g_acc_n[cls] += 1;
}
extern "C" void bk_prof_read(int cls, double *out4) {
    prof_flush();
    if (cls == 62) { out4[0] = g_sync_ms; out4[1] = g_sync_n; out4[2] = 0; out4[3] = 0; return; }
    if (cls == 63) { out4[0] = 0; out4[1] = g_stage_n; out4[2] = g_cache_hits; out4[3] = g_stage_bytes; return; }
    if (cls < 0 || cls >= PROF_CLS) { out4[0] = out4[1] = out4[2] = out4[3] = 0; return; }
    out4[0] = g_acc_ms[cls]; out4[1] = g_acc_n[cls]; out4[2] = g_acc_flops[cls]; out4[3] = g_acc_bytes[cls];
}

static char g_devname[256] = "";
extern "C" const char *bk_device_name(void) {
    if (!g_devname[0]) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, 0) == hipSuccess) snprintf(g_devname, sizeof(g_devname), "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    }
    return g_devname;
}
