// k_ipm.hip - the interior-point method's own device kernels (CDNA4, wave64):
//   * constraint operators A(D) (gather, one wave per constraint) and A*(y) (scatter by position,
//     one thread per position, the constraints of a position summed in a fixed order);
//   * Schur complement assembly, route 0 (reference solver.c:2626-2660): one thread per pair of
//     constraints over their common blocks, and the Hadamard form M = D_v (X o Zi) D_v for
//     constraints that are one diagonal entry each (max-cut; reference solver.c:2631-2634);
//   * segmented dot products with a fixed summation order (the result does not depend on how
//     many other solves share the launch) and fused linear combinations;
//   * equilibration of M into the f32 or f64 factor input, symmetrisation, transposes;
//   * Lanczos steps with full reorthogonalisation for the step-length bound (reference
//     solver.c:6083-6225).
// All operands are offsets from the one device heap (doubles unless stated).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <vector>
#include "brisk_dev.h"
#include "rt_internal.h"
#include "flat.h"

#define BK_LZK 30                     // Lanczos steps at most (the reference's lanczos_k)

static char *g_heap = nullptr;
static size_t g_heap_size = 0;
static double *g_partial = nullptr;        // dot-product chunk partials
static size_t g_partial_cap = 0;

extern "C" int bk_heap_init(size_t bytes) {
    if (g_heap && g_heap_size >= bytes) return 0;
    if (g_heap) { bk_sync(); (void)hipFree(g_heap); g_heap = nullptr; g_heap_size = 0; }
    if (bk_check(hipMalloc((void **)&g_heap, bytes), "hipMalloc heap")) return -1;
    g_heap_size = bytes;
    return 0;
}
extern "C" void *bk_heap(void) { return g_heap; }
extern "C" size_t bk_heap_size(void) { return g_heap_size; }

#define H64(off) (((double *)g_heapd) + (off))

// ---- dot products ---------------------------------------------------------------------------
#define DCH 4096
__global__ __launch_bounds__(256) void k_dot_chunks(const DotItem *__restrict__ it, const int64_t *__restrict__ cbase,
                                                    const int *__restrict__ pre, int nit, const double *__restrict__ H,
                                                    double *__restrict__ part) {
    const int q = flat_item(pre, nit, blockIdx.x);
    const DotItem d = it[q];
    const int64_t c = blockIdx.x - pre[q];
    __shared__ double sh[256];
    const int t = threadIdx.x;
    double s = 0;
    const double *x = H + d.xo, *y = H + d.yo;
    if ((c + 1) * DCH <= d.len) {               // a full chunk: no guards, the loads overlap
#pragma unroll
        for (int u = 0; u < DCH / 256; u++) {
            const int64_t e = c * DCH + t + 256 * u;
            s += x[e] * y[e];
        }
    } else {
#pragma unroll 4
        for (int u = 0; u < DCH / 256; u++) {
            const int64_t e = c * DCH + t + 256 * u;
            if (e < d.len) s += x[e] * y[e];
        }
    }
    sh[t] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) sh[t] += sh[t + w];
        __syncthreads();
    }
    if (t == 0) part[cbase[q] + c] = sh[0];
}
// results in [moff, moff + mlen) go straight to the host mirror (pinned, device-visible)
// one wave per item: lane l adds the chunks l, l + 64, ... in order, then a butterfly over the
// lanes (a fixed order that depends on the chunk count only)
__global__ __launch_bounds__(256) void k_dot_final(const DotItem *__restrict__ it, const int64_t *__restrict__ cbase, int nit,
                                                   double *__restrict__ H, const double *__restrict__ part, int64_t moff, int64_t mlen,
                                                   double *__restrict__ mirror) {
    const int q = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (q >= nit) return;
    const DotItem d = it[q];
    const int64_t nch = (d.len + DCH - 1) / DCH;
    double s = 0;
    for (int64_t c = lane; c < nch; c += 64) s += part[cbase[q] + c];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane) return;
    if (mirror && d.out >= moff && d.out < moff + mlen) mirror[d.out - moff] = s;
    else H[d.out] = s;
}
static int64_t g_mirror_off = 0, g_mirror_len = 0;
static double *g_mirror = nullptr;
extern "C" void bk_dot_mirror(int64_t off, int64_t len, double *host) { g_mirror_off = off; g_mirror_len = len; g_mirror = host; }
extern "C" int bk_dots(const DotItem *it, int n_it) {
    if (n_it <= 0) return 0;
    std::vector<int64_t> cb(n_it), cnt(n_it);
    int64_t tot = 0;
    for (int q = 0; q < n_it; q++) {
        cb[q] = tot;
        const int64_t nch = (it[q].len + DCH - 1) / DCH;
        tot += nch;
        cnt[q] = nch;
    }
    if ((size_t)tot > g_partial_cap) {
        bk_sync();
        if (g_partial) (void)hipFree(g_partial);
        g_partial_cap = (size_t)tot * 2 + 4096;
        if (bk_check(hipMalloc((void **)&g_partial, g_partial_cap * sizeof(double)), "hipMalloc partials")) return -1;
    }
    // items, chunk bases and the prefix array in one staged copy
    std::vector<char> buf(sizeof(DotItem) * n_it + sizeof(int64_t) * n_it);
    memcpy(buf.data(), it, sizeof(DotItem) * n_it);
    memcpy(buf.data() + sizeof(DotItem) * n_it, cb.data(), sizeof(int64_t) * n_it);
    const void *dv = nullptr;
    const int *pre = nullptr;
    const int64_t nb = flat_stage_items(buf.data(), buf.size(), cnt, &dv, &pre);
    if (nb < 0) return -1;
    const DotItem *d = (const DotItem *)dv;
    const int64_t *dcb = (const int64_t *)((const char *)dv + sizeof(DotItem) * n_it);
    const int tok = bk_prof_begin();
    if (nb > 0) hipLaunchKernelGGL(k_dot_chunks, dim3((unsigned)nb), dim3(256), 0, g_bk_stream, d, dcb, pre, n_it, (const double *)g_heap, g_partial);
    if (tok >= 0) { double by = 0; for (int q = 0; q < n_it; q++) by += 16.0 * it[q].len; bk_prof_end(tok, 13, 0, by); }
    if (BK_LAUNCH_CHECK("k_dot_chunks")) return -1;
    hipLaunchKernelGGL(k_dot_final, dim3((n_it + 3) / 4), dim3(256), 0, g_bk_stream, d, dcb, n_it, (double *)g_heap, g_partial,
                       g_mirror_off, g_mirror_len, g_mirror);
    return BK_LAUNCH_CHECK("k_dot_final");
}

// ---- linear combinations ----------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_lincomb(const LinItem *__restrict__ it, const int *__restrict__ pre, int nit,
                                                 double *__restrict__ H) {
    const int q = flat_item(pre, nit, blockIdx.x);
    const LinItem L = it[q];
    const int64_t lb = blockIdx.x - pre[q], nb = pre[q + 1] - pre[q];
    for (int64_t e = lb * 256 + threadIdx.x; e < L.len; e += nb * 256) {
        double v = 0;
#pragma unroll
        for (int t = 0; t < 4; t++)
            if (L.in[t] >= 0) v += L.c[t] * H[L.in[t] + e];
        H[L.out + e] = v;
    }
}
// out[pos[i]] = val[i]: a sparse array into a zeroed dense one (C's nonzeros at setup)
__global__ __launch_bounds__(256) void k_scatter(const int64_t *__restrict__ pos, const double *__restrict__ val, int64_t n,
                                                 double *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[pos[i]] = val[i];
}
extern "C" int bk_scatter_f64(const int64_t *pos, const double *val, int64_t n, double *out) {
    if (n <= 0) return 0;
    const int64_t nb = (n + 255) / 256;
    if (nb > 0x7fffffff) return bk_fail("scatter: too many entries");
    hipLaunchKernelGGL(k_scatter, dim3((unsigned)nb), dim3(256), 0, g_bk_stream, pos, val, n, out);
    return BK_LAUNCH_CHECK("k_scatter");
}
static int64_t lin_stage(const LinItem *it, int n_it, const LinItem **d, const int **pre) {
    std::vector<int64_t> cnt(n_it);
    for (int q = 0; q < n_it; q++) cnt[q] = flat_cap((it[q].len + 255) / 256, 2048);
    return flat_stage_items(it, sizeof(LinItem) * n_it, cnt, (const void **)d, pre);
}
extern "C" int bk_lincomb(const LinItem *it, int n_it) {
    if (n_it <= 0) return 0;
    const LinItem *d = nullptr;
    const int *pre = nullptr;
    const int64_t nb = lin_stage(it, n_it, &d, &pre);
    if (nb < 0) return -1;
    if (nb == 0) return 0;
    const int tok = bk_prof_begin();
    hipLaunchKernelGGL(k_lincomb, dim3((unsigned)nb), dim3(256), 0, g_bk_stream, d, pre, n_it, (double *)g_heap);
    if (tok >= 0) { double by = 0; for (int q = 0; q < n_it; q++) { int t = 1; for (int u = 0; u < 4; u++) t += it[q].in[u] >= 0; by += 8.0 * t * it[q].len; } bk_prof_end(tok, 14, 0, by); }
    return BK_LAUNCH_CHECK("k_lincomb");
}

// out[q] = c0 in0[q] in1[q] (+ c1 in2[q] when in[2] >= 0)
__global__ __launch_bounds__(256) void k_vmul(const LinItem *__restrict__ it, const int *__restrict__ pre, int nit,
                                              double *__restrict__ H) {
    const int q = flat_item(pre, nit, blockIdx.x);
    const LinItem L = it[q];
    const int64_t lb = blockIdx.x - pre[q], nb = pre[q + 1] - pre[q];
    for (int64_t e = lb * 256 + threadIdx.x; e < L.len; e += nb * 256) {
        double v = L.c[0] * H[L.in[0] + e] * H[L.in[1] + e];
        if (L.in[2] >= 0) v += L.c[1] * H[L.in[2] + e];
        H[L.out + e] = v;
    }
}
extern "C" int bk_vmul(const LinItem *it, int n_it) {
    if (n_it <= 0) return 0;
    const LinItem *d = nullptr;
    const int *pre = nullptr;
    const int64_t nb = lin_stage(it, n_it, &d, &pre);
    if (nb < 0) return -1;
    if (nb == 0) return 0;
    hipLaunchKernelGGL(k_vmul, dim3((unsigned)nb), dim3(256), 0, g_bk_stream, d, pre, n_it, (double *)g_heap);
    return BK_LAUNCH_CHECK("k_vmul");
}

// ---- dense route helpers ------------------------------------------------------------------------
// M (f64, full symmetric) = sum over the split-K partials of their lower triangles, in chunk order
template <typename T>
__global__ void k_splitk_reduce(const T *__restrict__ part, int nchunk, int64_t stride, int m, double *__restrict__ M) {
    const int64_t tot = (int64_t)m * m;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % m), j = (int)(e / m);
        if (i < j) continue;
        double s = 0;
        for (int c = 0; c < nchunk; c++) s += (double)part[c * stride + e];
        M[e] = s;
        M[j + (int64_t)i * m] = s;
    }
}
extern "C" int bk_splitk_reduce(int prec, int64_t part_off, int nchunk, int64_t stride, int m, int64_t M_off) {
    const unsigned g = bk_grid1((int64_t)m * m, 256, 8192);
    if (prec) hipLaunchKernelGGL((k_splitk_reduce<double>), dim3(g), dim3(256), 0, g_bk_stream, ((const double *)g_heap) + part_off, nchunk, stride, m, ((double *)g_heap) + M_off);
    else hipLaunchKernelGGL((k_splitk_reduce<float>), dim3(g), dim3(256), 0, g_bk_stream, ((const float *)g_heap) + part_off, nchunk, stride, m, ((double *)g_heap) + M_off);
    return BK_LAUNCH_CHECK("k_splitk_reduce");
}
// f32 copy of an f64 array (offsets: doubles for the source, floats for the destination)
__global__ void k_d2f(const double *__restrict__ a, float *__restrict__ b, int64_t n) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) b[e] = (float)a[e];
}
extern "C" int bk_d2f(int64_t n, int64_t src_off, int64_t dst_off_f32) {
    hipLaunchKernelGGL(k_d2f, dim3(bk_grid1(n, 256, 8192)), dim3(256), 0, g_bk_stream, ((const double *)g_heap) + src_off, ((float *)g_heap) + dst_off_f32, n);
    return BK_LAUNCH_CHECK("k_d2f");
}
// a(r, c) *= d[c] in place and b(r, c) = (float)a(r, c), columns of `rows` elements: the grid's
// y index walks the columns, x the rows of one
__global__ void k_colscale_d2f(double *__restrict__ a, float *__restrict__ b, const double *__restrict__ d, int64_t rows, int cols) {
    for (int c = blockIdx.y; c < cols; c += gridDim.y) {
        const double s = d[c];
        double *ac = a + (int64_t)c * rows;
        float *bc = b + (int64_t)c * rows;
        for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += (int64_t)gridDim.x * blockDim.x) {
            const double v = ac[r] * s;
            ac[r] = v;
            bc[r] = (float)v;
        }
    }
}
extern "C" int bk_colscale_d2f(int64_t rows, int cols, int64_t off, int64_t dst_off_f32, int64_t d_off) {
    if (rows <= 0 || cols <= 0) return 0;
    const dim3 grid(bk_grid1(rows, 256, 64), (unsigned)(cols < 8192 ? cols : 8192));
    hipLaunchKernelGGL(k_colscale_d2f, grid, dim3(256), 0, g_bk_stream, ((double *)g_heap) + off,
                       ((float *)g_heap) + dst_off_f32, ((const double *)g_heap) + d_off, rows, cols);
    return BK_LAUNCH_CHECK("k_colscale_d2f");
}

// ---- per-matrix helpers ---------------------------------------------------------------------
#define FLAT_MITEM                                                          \
    const int q = flat_item(pre, nit, blockIdx.x);                          \
    const MItem m = it[q];                                                  \
    const int64_t lb = blockIdx.x - pre[q], nb = pre[q + 1] - pre[q];
// G = (G + G')/2 over pairs of 32 x 32 tiles (I, J), I >= J, both staged in LDS so that every
// read and write is coalesced; each entry pair gets 0.5 (g_rc + g_cr). With opt = 1 the strict
// upper triangle takes the lower one instead (G(c, r) = G(r, c), r > c; the lower tile only read)
__global__ void k_symmetrize(const MItem *__restrict__ it, const int *__restrict__ pre, int nit, double *__restrict__ H) {
    FLAT_MITEM
    const int n = m.n, tiles = (n + 31) / 32;
    double *G = H + m.off;
    __shared__ double ta[32][33], tb[32][33];                 // ta[c][r] = G(32 I + r, 32 J + c), tb the mirror tile
    const int lx = threadIdx.x & 31, ly = threadIdx.x >> 5;   // 256 threads: 32 x 8
    for (int64_t tt = lb; tt < (int64_t)tiles * (tiles + 1) / 2; tt += nb) {
        int I = (int)((sqrt(8.0 * (double)tt + 1.0) - 1.0) * 0.5);
        while ((int64_t)I * (I + 1) / 2 > tt) I--;
        while ((int64_t)(I + 1) * (I + 2) / 2 <= tt) I++;
        const int J = (int)(tt - (int64_t)I * (I + 1) / 2);
        const int r0 = 32 * I, c0 = 32 * J;
        for (int y = ly; y < 32; y += 8) {
            if (r0 + lx < n && c0 + y < n) ta[y][lx] = G[(r0 + lx) + (int64_t)(c0 + y) * n];    // G(r0 + lx, c0 + y)
            if (!m.opt && c0 + lx < n && r0 + y < n) tb[y][lx] = G[(c0 + lx) + (int64_t)(r0 + y) * n];    // G(c0 + lx, r0 + y)
        }
        __syncthreads();
        if (m.opt) {
            for (int y = ly; y < 32; y += 8)
                if (c0 + lx < n && r0 + y < n && (I > J || y > lx)) G[(c0 + lx) + (int64_t)(r0 + y) * n] = ta[lx][y];
            __syncthreads();
            continue;
        }
        for (int y = ly; y < 32; y += 8) {
            // G(r0 + lx, c0 + y) and its mirror G(c0 + y, r0 + lx) = tb[lx][y]
            if (r0 + lx < n && c0 + y < n && (I > J || lx > y))
                G[(r0 + lx) + (int64_t)(c0 + y) * n] = 0.5 * (ta[y][lx] + tb[lx][y]);
            // G(c0 + lx, r0 + y) and its mirror G(r0 + y, c0 + lx) = ta[lx][y]
            if (c0 + lx < n && r0 + y < n && (I > J || y > lx))
                G[(c0 + lx) + (int64_t)(r0 + y) * n] = 0.5 * (tb[y][lx] + ta[lx][y]);
        }
        __syncthreads();
    }
}
template <typename T>
__global__ void k_transpose(const MItem *__restrict__ it, const int *__restrict__ pre, int nit, T *__restrict__ H) {
    FLAT_MITEM
    const int n = m.n;
    // 32 x 32 tiles through LDS: coalesced reads and writes
    __shared__ T tile[32][33];
    const int tiles = (n + 31) / 32;
    for (int64_t tt = lb; tt < (int64_t)tiles * tiles; tt += nb) {
        const int tr = (int)(tt % tiles) * 32, tc = (int)(tt / tiles) * 32;
        const int lx = threadIdx.x & 31, ly = threadIdx.x >> 5;     // 256 threads: 32 x 8
        for (int y = ly; y < 32; y += 8) {
            const int r = tr + lx, c = tc + y;
            if (r < n && c < n) tile[y][lx] = H[m.off + r + (int64_t)c * n];
        }
        __syncthreads();
        for (int y = ly; y < 32; y += 8) {
            const int r = tc + lx, c = tr + y;                       // dst(r, c) = src(c, r)
            if (r < n && c < n) H[m.off4 + r + (int64_t)c * n] = tile[lx][y];
        }
        __syncthreads();
    }
}
__global__ void k_add_diag(const MItem *__restrict__ it, const int *__restrict__ pre, int nit, double *__restrict__ H, int64_t s_off) {
    FLAT_MITEM
    const double s = H[s_off + m.slot];
    for (int64_t i = lb * blockDim.x + threadIdx.x; i < m.n; i += nb * blockDim.x) H[m.off + i + i * m.n] += s;
}
// Out = Diag(V) X per block: out(r, c) = V(r, r) X(r, c) (V at off, X at off2, out at off3): the
// product Diag(dZ) X of a block whose dual direction is diagonal
__global__ void k_rowscale(const MItem *__restrict__ it, const int *__restrict__ pre, int nit, double *__restrict__ H) {
    FLAT_MITEM
    const int n = m.n;
    for (int64_t e = lb * blockDim.x + threadIdx.x; e < (int64_t)n * n; e += nb * blockDim.x) {
        const int64_t r = e % n;
        H[m.off3 + e] = H[m.off + r + r * n] * H[m.off2 + e];
    }
}
__global__ void k_equil_d(const MItem *__restrict__ it, const int *__restrict__ pre, int nit, double *__restrict__ H) {
    FLAT_MITEM
    for (int64_t i = lb * blockDim.x + threadIdx.x; i < m.n; i += nb * blockDim.x) {
        const double d = H[m.off + i + i * m.n];
        H[m.off2 + i] = d > 0 ? 1.0 / sqrt(d) : 1.0;
    }
}
template <typename T>
__global__ void k_equil_f(const MItem *__restrict__ it, const int *__restrict__ pre, int nit, const double *__restrict__ H,
                          T *__restrict__ F, int64_t reg_off) {
    FLAT_MITEM
    const int n = m.n;
    const double reg = H[reg_off + m.slot];
    const double *D = H + m.off2;
    for (int64_t e = lb * blockDim.x + threadIdx.x; e < (int64_t)n * n; e += nb * blockDim.x) {
        const int r = (int)(e % n), c = (int)(e / n);
        T v = 0;
        if (r >= c) {
            double x = H[m.off + e] * D[r] * D[c];
            if (r == c) x += reg;
            v = fabs(x) < 1e-30 ? (T)0 : (T)x;
        }
        F[m.off3 + e] = v;
    }
}
// workgroups per matrix item: kind 1 one per 256 of the order, 2 one per 256 entries (at most
// 1024), 3 one per 32 x 32 tile (at most 512)
static int64_t mitem_stage(const MItem *it, int n_it, int kind, const MItem **d, const int **pre) {
    std::vector<int64_t> cnt(n_it);
    for (int q = 0; q < n_it; q++) {
        const int64_t n = it[q].n, t = (n + 31) / 32;
        cnt[q] = kind == 1 ? flat_cap((n + 255) / 256, 1024) : kind == 2 ? flat_cap(t * (t + 1) / 2, 512) : flat_cap(t * t, 512);
    }
    return flat_stage_items(it, sizeof(MItem) * n_it, cnt, (const void **)d, pre);
}
#define MITEM_LAUNCH(kind_)                                                    \
    if (n_it <= 0) return 0;                                                   \
    const MItem *d = nullptr;                                                  \
    const int *pre = nullptr;                                                  \
    const int64_t nb = mitem_stage(it, n_it, kind_, &d, &pre);                \
    if (nb < 0) return -1;                                                     \
    if (nb == 0) return 0;
extern "C" int bk_symmetrize(const MItem *it, int n_it) {
    MITEM_LAUNCH(2)
    hipLaunchKernelGGL(k_symmetrize, dim3((unsigned)nb), dim3(256), 0, g_bk_stream, d, pre, n_it, (double *)g_heap);
    return BK_LAUNCH_CHECK("k_symmetrize");
}
extern "C" int bk_transpose_f64(const MItem *it, int n_it) {
    MITEM_LAUNCH(3)
    hipLaunchKernelGGL((k_transpose<double>), dim3((unsigned)nb), dim3(256), 0, g_bk_stream, d, pre, n_it, (double *)g_heap);
    return BK_LAUNCH_CHECK("k_transpose");
}
extern "C" int bk_transpose_f32(const MItem *it, int n_it) {
    MITEM_LAUNCH(3)
    hipLaunchKernelGGL((k_transpose<float>), dim3((unsigned)nb), dim3(256), 0, g_bk_stream, d, pre, n_it, (float *)g_heap);
    return BK_LAUNCH_CHECK("k_transpose");
}
extern "C" int bk_rowscale(const MItem *it, int n_it) {
    MITEM_LAUNCH(3)
    hipLaunchKernelGGL(k_rowscale, dim3((unsigned)nb), dim3(256), 0, g_bk_stream, d, pre, n_it, (double *)g_heap);
    return BK_LAUNCH_CHECK("k_rowscale");
}
extern "C" int bk_add_diag(const MItem *it, int n_it, int64_t s_off) {
    MITEM_LAUNCH(1)
    hipLaunchKernelGGL(k_add_diag, dim3((unsigned)nb), dim3(256), 0, g_bk_stream, d, pre, n_it, (double *)g_heap, s_off);
    return BK_LAUNCH_CHECK("k_add_diag");
}
extern "C" int bk_equil(const MItem *it, int n_it, int prec, int64_t reg_off) {
    MITEM_LAUNCH(1)
    hipLaunchKernelGGL(k_equil_d, dim3((unsigned)nb), dim3(256), 0, g_bk_stream, d, pre, n_it, (double *)g_heap);
    if (BK_LAUNCH_CHECK("k_equil_d")) return -1;
    const MItem *d2 = nullptr;
    const int *pre2 = nullptr;
    const int64_t nb2 = mitem_stage(it, n_it, 2, &d2, &pre2);
    if (nb2 < 0) return -1;
    if (prec == 0) hipLaunchKernelGGL((k_equil_f<float>), dim3((unsigned)nb2), dim3(256), 0, g_bk_stream, d2, pre2, n_it, (const double *)g_heap, (float *)g_heap, reg_off);
    else hipLaunchKernelGGL((k_equil_f<double>), dim3((unsigned)nb2), dim3(256), 0, g_bk_stream, d2, pre2, n_it, (const double *)g_heap, (double *)g_heap, reg_off);
    return BK_LAUNCH_CHECK("k_equil_f");
}

// ---- constraint operators -------------------------------------------------------------------
__device__ __forceinline__ const int *HI(const double *H, int64_t off) { return (const int *)(((const char *)H) + off); }

// A(D): one wave per constraint, lanes stride its entries, butterfly in a fixed pattern
__global__ __launch_bounds__(256) void k_aop(const ConItem *__restrict__ it, const int *__restrict__ pre, int nit, double *__restrict__ H) {
    const int q = flat_item(pre, nit, blockIdx.x);
    const ConItem C = it[q];
    const int i = (blockIdx.x - pre[q]) * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= C.m) return;
    const int *cptr = HI(H, C.cptr);
    const SEnt *ent = (const SEnt *)(((const char *)H) + C.ent);
    const BlkRef *blk = (const BlkRef *)(((const char *)H) + C.blk);
    const double *D = H + C.D;
    double s = 0;
    for (int e = cptr[i] + lane; e < cptr[i + 1]; e += 64) {
        const SEnt E = ent[e];
        const BlkRef B = blk[E.k];
        double x = D[B.off + E.r + (int64_t)E.c * B.n];
        if (E.r != E.c) x += D[B.off + E.c + (int64_t)E.r * B.n];
        s += E.v * x;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) {
        double *o = H + C.out + i;
        *o = C.beta != 0 ? C.beta * *o + C.alpha * s : C.alpha * s;
    }
}
extern "C" int bk_aop(const ConItem *it, int n_it) {
    if (n_it <= 0) return 0;
    std::vector<int64_t> cnt(n_it);
    for (int q = 0; q < n_it; q++) cnt[q] = (it[q].m + 3) / 4;
    const ConItem *d = nullptr;
    const int *pre = nullptr;
    const int64_t nb = flat_stage_items(it, sizeof(ConItem) * n_it, cnt, (const void **)&d, &pre);
    if (nb < 0) return -1;
    if (nb == 0) return 0;
    const int tok = bk_prof_begin();
    hipLaunchKernelGGL(k_aop, dim3((unsigned)nb), dim3(256), 0, g_bk_stream, d, pre, n_it, (double *)g_heap);
    if (tok >= 0) bk_prof_end(tok, 11, 0, 0);
    return BK_LAUNCH_CHECK("k_aop");
}

// A*(y) added at the positions the constraints touch: D[pos] += alpha sum_i y_i v_i
__global__ __launch_bounds__(256) void k_aty(const ConItem *__restrict__ it, const int *__restrict__ pre, int nit, double *__restrict__ H) {
    const int q = flat_item(pre, nit, blockIdx.x);
    const ConItem C = it[q];
    const int p = (blockIdx.x - pre[q]) * 256 + threadIdx.x;
    if (p >= C.npos) return;
    const int *pptr = HI(H, C.pptr), *ppos = HI(H, C.ppos), *pcon = HI(H, C.pcon);
    const double *pval = H + C.pval, *y = H + C.y;
    double s = 0;
    for (int t = pptr[p]; t < pptr[p + 1]; t++) s += y[pcon[t]] * pval[t];
    double *D = H + C.D;
    const int o1 = ppos[2 * p], o2 = ppos[2 * p + 1];
    D[o1] += C.alpha * s;
    if (o2 >= 0) D[o2] += C.alpha * s;
}
extern "C" int bk_aty(const ConItem *it, int n_it) {
    if (n_it <= 0) return 0;
    std::vector<int64_t> cnt(n_it);
    for (int q = 0; q < n_it; q++) cnt[q] = (it[q].npos + 255) / 256;
    const ConItem *d = nullptr;
    const int *pre = nullptr;
    const int64_t nb = flat_stage_items(it, sizeof(ConItem) * n_it, cnt, (const void **)&d, &pre);
    if (nb < 0) return -1;
    if (nb == 0) return 0;
    hipLaunchKernelGGL(k_aty, dim3((unsigned)nb), dim3(256), 0, g_bk_stream, d, pre, n_it, (double *)g_heap);
    return BK_LAUNCH_CHECK("k_aty");
}

// Schur complement, Hadamard form: every constraint a single diagonal entry v_a at (r_a, r_a) of
// block k_a: M_ab = v_a v_b X[r_a, r_b] Zi[r_a, r_b] when k_a == k_b (else 0). A workgroup covers
// 256 rows a of one column b: X and Zi columns streamed, M written coalesced.
__global__ __launch_bounds__(256) void k_schur_diag(const ConItem *__restrict__ it, const int *__restrict__ pre, int nit,
                                                    double *__restrict__ H) {
    const int q = flat_item(pre, nit, blockIdx.x);
    const ConItem C = it[q];
    const int cb = (C.m + 255) / 256, lb = blockIdx.x - pre[q];
    const int b = lb / cb, a = (lb % cb) * 256 + threadIdx.x;
    if (b >= C.m || a >= C.m) return;
    const SEnt *ent = (const SEnt *)(((const char *)H) + C.ent);
    const BlkRef *blk = (const BlkRef *)(((const char *)H) + C.blk);
    const SEnt Ea = ent[a], Eb = ent[b];
    double v = 0;
    if (Ea.k == Eb.k) {
        const BlkRef B = blk[Ea.k];
        const int64_t o = B.off + Ea.r + (int64_t)Eb.r * B.n;
        v = Ea.v * Eb.v * H[C.X + o] * H[C.Zi + o];
    }
    H[C.M + a + (int64_t)b * C.m] = v;
}
// Schur complement, route 0: M_ab = sum over the entries p of A_a and q of A_b in a common block
// of v_p v_q X[c_p, r_q] Zi[c_q, r_p] over the symmetric expansions of both (reference
// solver.c:2626-2660). A workgroup computes a 16 x 16 tile of pairs of the lower triangle of tiles
// (16 x 16 threads, rows a along the fast thread index). The tile's 32 constraints are read once
// into LDS (entry range, first entry, its block); a pair of single-entry constraints (a Lovasz
// theta edge, for instance) is then four products from X and Zi with no further index loads,
// other pairs merge their entries (grouped by block in ascending order) over the common blocks.
// The tile is stored at (a, b) and, through LDS, transposed at (b, a): both stores are 16-row
// coalesced runs.
__device__ __forceinline__ double pair_term(const SEnt &P, const SEnt &Q, const double *Xk, const double *Zk, int n) {
    // expansions (r, c) and, off the diagonal, (c, r) of both entries
    double acc = Xk[P.c + (int64_t)Q.r * n] * Zk[Q.c + (int64_t)P.r * n];
    if (Q.r != Q.c) acc += Xk[P.c + (int64_t)Q.c * n] * Zk[Q.r + (int64_t)P.r * n];
    if (P.r != P.c) {
        acc += Xk[P.r + (int64_t)Q.r * n] * Zk[Q.c + (int64_t)P.c * n];
        if (Q.r != Q.c) acc += Xk[P.r + (int64_t)Q.c * n] * Zk[Q.r + (int64_t)P.c * n];
    }
    return P.v * Q.v * acc;
}
// <A_l, P>: the entries of a (light) group against a heavy group's product P = X A_h Zi
__device__ __forceinline__ double group_dot(const SEnt *__restrict__ e, int cnt, const double *__restrict__ Pm, int n) {
    double s = 0;
    for (int w = 0; w < cnt; w++) {
        const SEnt E = e[w];
        double x = Pm[E.r + (int64_t)E.c * n];
        if (E.r != E.c) x += Pm[E.c + (int64_t)E.r * n];
        s += E.v * x;
    }
    return s;
}
// M_ab over the common blocks of constraints a and b: their group lists (ascending block) merged;
// a heavy group's stored product against the other's entries, else the entry pairs
__device__ __forceinline__ double pair_sum(const int *__restrict__ gptr, const GrpRef *__restrict__ grp, const SEnt *__restrict__ ent,
                                           const BlkRef *__restrict__ blk, const double *__restrict__ H, const double *X,
                                           const double *Zi, int a, int b) {
    double s = 0;
    int pa = gptr[a], ea = gptr[a + 1], pb = gptr[b], eb = gptr[b + 1];
    while (pa < ea && pb < eb) {
        const GrpRef Ga = grp[pa], Gb = grp[pb];
        if (Ga.k < Gb.k) { pa++; continue; }
        if (Gb.k < Ga.k) { pb++; continue; }
        const BlkRef B = blk[Ga.k];
        if (Ga.poff >= 0 && (Gb.poff < 0 || Gb.cnt <= Ga.cnt)) s += group_dot(ent + Gb.beg, Gb.cnt, H + Ga.poff, B.n);
        else if (Gb.poff >= 0) s += group_dot(ent + Ga.beg, Ga.cnt, H + Gb.poff, B.n);
        else {
            const double *Xk = X + B.off, *Zk = Zi + B.off;
            for (int u = 0; u < Ga.cnt; u++) {
                const SEnt P = ent[Ga.beg + u];
                for (int w = 0; w < Gb.cnt; w++) s += pair_term(P, ent[Gb.beg + w], Xk, Zk, B.n);
            }
        }
        pa++; pb++;
    }
    return s;
}
// the same pairs, listed: one thread per pair (a, b), a >= b, of constraints that share a block
// (by column b, rows a ascending), the others left at the zeros written at setup. For problems
// of many blocks, where few pairs share one (64 blocks of 60: 6%); the sums are pair_sum's.
__global__ __launch_bounds__(256) void k_schur_pairs(const ConItem *__restrict__ it, const int *__restrict__ pre, int nit,
                                                     double *__restrict__ H) {
    const int q = flat_item(pre, nit, blockIdx.x);
    const ConItem C = it[q];
    const int64_t idx = (int64_t)(blockIdx.x - pre[q]) * 256 + threadIdx.x;
    if (idx >= C.npair) return;
    const int *pr = HI(H, C.pairs);
    const int a = pr[2 * idx], b = pr[2 * idx + 1];
    const double s = pair_sum(HI(H, C.gptr), (const GrpRef *)(((const char *)H) + C.grp), (const SEnt *)(((const char *)H) + C.ent),
                              (const BlkRef *)(((const char *)H) + C.blk), H, H + C.X, H + C.Zi, a, b);
    H[C.M + a + (int64_t)b * C.m] = s;
    if (a != b) H[C.M + b + (int64_t)a * C.m] = s;
}
__global__ __launch_bounds__(256) void k_schur_r0(const ConItem *__restrict__ it, const int *__restrict__ pre, int nit,
                                                  int64_t ntiles, double *__restrict__ H) {
    __shared__ double tile[16][17];
    __shared__ int gbeg[32], gcnt[32];
    __shared__ GrpRef g0[32];
    __shared__ SEnt first[32];
    __shared__ BlkRef fblk[32];
    const int i = threadIdx.x & 15, j = threadIdx.x >> 4;
    // a persistent grid: each workgroup walks the tiles of every item with a stride
    for (int64_t vb = blockIdx.x; vb < ntiles; vb += gridDim.x) {
        const int q = flat_item(pre, nit, (int)vb);
        const ConItem C = it[q];
        int ta, tb;
        tri_index(vb - pre[q], ta, tb);
        const int a = ta * 16 + i, b = tb * 16 + j;
        const int *gptr = HI(H, C.gptr);
        const GrpRef *grp = (const GrpRef *)(((const char *)H) + C.grp);
        const SEnt *ent = (const SEnt *)(((const char *)H) + C.ent);
        const BlkRef *blk = (const BlkRef *)(((const char *)H) + C.blk);
        __syncthreads();                 // the previous tile's LDS reads are done
        if (threadIdx.x < 32) {
            const int c = threadIdx.x < 16 ? ta * 16 + threadIdx.x : tb * 16 + threadIdx.x - 16;
            int b0 = 0, nc = 0;
            if (c < C.m) { b0 = gptr[c]; nc = gptr[c + 1] - b0; }
            gbeg[threadIdx.x] = b0; gcnt[threadIdx.x] = nc;
            if (nc > 0) {
                const GrpRef G = grp[b0];
                g0[threadIdx.x] = G;
                const SEnt e = ent[G.beg];
                first[threadIdx.x] = e; fblk[threadIdx.x] = blk[e.k];
            }
        }
        __syncthreads();
        const double *X = H + C.X, *Zi = H + C.Zi;
        double s = 0;
        if (a < C.m && b < C.m) {
            const int na = gcnt[i], nb = gcnt[16 + j];
            if (na == 1 && nb == 1 && g0[i].cnt == 1 && g0[16 + j].cnt == 1 && g0[i].poff < 0 && g0[16 + j].poff < 0) {
                // two single-entry constraints: four products
                const SEnt P = first[i], Q = first[16 + j];
                if (P.k == Q.k) {
                    const BlkRef B = fblk[i];
                    s = pair_term(P, Q, X + B.off, Zi + B.off, B.n);
                }
            } else {
                // merge the two group lists (ascending block) over the common blocks
                s = pair_sum(gptr, grp, ent, blk, H, X, Zi, a, b);
            }
            H[C.M + a + (int64_t)b * C.m] = s;
        }
        if (ta != tb) {                  // a diagonal tile holds both triangles already (uniform per workgroup)
            tile[i][j] = s;
            __syncthreads();
            const int r = tb * 16 + i, c = ta * 16 + j;      // M[r, c] = M[c, r] = s(a = c, b = r) = tile[j][i]
            if (r < C.m && c < C.m) H[C.M + r + (int64_t)c * C.m] = tile[j][i];
        }
    }
}

// T = A Zi_k for the heavy groups: T[r, j] = sum over the expanded entries (r, c, v) of A of
// v Zi[c, j]; workgroups stride the n^2 outputs (rows fastest), the entries staged in LDS
__global__ __launch_bounds__(256) void k_grp_t(const GrpItem *__restrict__ it, const int *__restrict__ pre, int nit,
                                               double *__restrict__ H) {
    const int q = flat_item(pre, nit, blockIdx.x);
    const GrpItem G = it[q];
    const int64_t lb = blockIdx.x - pre[q], nbk = pre[q + 1] - pre[q];
    const int n = G.n;
    const SEnt *ent = (const SEnt *)(((const char *)H) + G.ent);
    const double *Zi = H + G.zi;
    double *T = H + G.t;
    __shared__ SEnt se[256];
    for (int64_t o0 = lb * 256; o0 < (int64_t)n * n; o0 += nbk * 256) {
        const int64_t o = o0 + threadIdx.x;
        const int r = (int)(o % n), jc = (int)(o / n);
        double acc = 0;
        for (int c0 = 0; c0 < G.cnt; c0 += 256) {
            const int nc = min(256, G.cnt - c0);
            __syncthreads();
            if (threadIdx.x < nc) se[threadIdx.x] = ent[c0 + threadIdx.x];
            __syncthreads();
            if (o < (int64_t)n * n) {
                for (int w = 0; w < nc; w++) {
                    const SEnt E = se[w];
                    if (E.r == r) acc += E.v * Zi[E.c + (int64_t)jc * n];
                    if (E.c == r && E.r != E.c) acc += E.v * Zi[E.r + (int64_t)jc * n];
                }
            }
        }
        if (o < (int64_t)n * n) T[o] = acc;
    }
}
extern "C" int bk_grp_t(const GrpItem *it, int n_it) {
    if (n_it <= 0) return 0;
    std::vector<int64_t> cnt(n_it);
    for (int q = 0; q < n_it; q++) cnt[q] = flat_cap(((int64_t)it[q].n * it[q].n + 255) / 256, 64);
    const GrpItem *d = nullptr;
    const int *pre = nullptr;
    const int64_t nb = flat_stage_items(it, sizeof(GrpItem) * n_it, cnt, (const void **)&d, &pre);
    if (nb < 0) return -1;
    if (nb == 0) return 0;
    const int tok = bk_prof_begin();
    hipLaunchKernelGGL(k_grp_t, dim3((unsigned)nb), dim3(256), 0, g_bk_stream, d, pre, n_it, (double *)g_heap);
    if (tok >= 0) bk_prof_end(tok, 10, 0, 0);
    return BK_LAUNCH_CHECK("k_grp_t");
}
extern "C" int bk_schur(const ConItem *it, int n_it) {
    if (n_it <= 0) return 0;
    std::vector<int64_t> cd(n_it), cr(n_it), cp(n_it);
    int anyd = 0, anyr = 0, anyp = 0;
    for (int q = 0; q < n_it; q++) {
        const int64_t m = it[q].m, t = (m + 15) / 16;
        const int pl_ = !it[q].diag && it[q].npair > 0;
        cd[q] = it[q].diag ? ((m + 255) / 256) * m : 0;
        cr[q] = it[q].diag || pl_ ? 0 : t * (t + 1) / 2;
        cp[q] = pl_ ? (it[q].npair + 255) / 256 : 0;
        anyd |= it[q].diag != 0; anyr |= cr[q] > 0; anyp |= pl_;
    }
    // items and the three prefix arrays in one staged copy
    const size_t po = (sizeof(ConItem) * n_it + 15) & ~(size_t)15, pl = sizeof(int) * (n_it + 1);
    std::vector<char> buf(po + 3 * pl);
    memcpy(buf.data(), it, sizeof(ConItem) * n_it);
    int64_t tot[3] = { 0, 0, 0 };
    for (int w = 0; w < 3; w++) {
        int *pre = (int *)(buf.data() + po + w * pl);
        for (int q = 0; q < n_it; q++) { pre[q] = (int)tot[w]; tot[w] += w == 2 ? cp[q] : w ? cr[q] : cd[q]; }
        pre[n_it] = (int)tot[w];
        if (tot[w] > 0x7fffffff) return -1;
    }
    const char *dv = (const char *)bk_stage(buf.data(), buf.size());
    if (!dv) return -1;
    const ConItem *d = (const ConItem *)dv;
    const int tok = bk_prof_begin();
    if (anyd && tot[0] > 0) {
        hipLaunchKernelGGL(k_schur_diag, dim3((unsigned)tot[0]), dim3(256), 0, g_bk_stream, d, (const int *)(dv + po), n_it, (double *)g_heap);
        if (BK_LAUNCH_CHECK("k_schur_diag")) return -1;
    }
    if (anyr && tot[1] > 0) {
        const unsigned g = (unsigned)(tot[1] < 8192 ? tot[1] : 8192);
        hipLaunchKernelGGL(k_schur_r0, dim3(g), dim3(256), 0, g_bk_stream, d, (const int *)(dv + po + pl), n_it, tot[1], (double *)g_heap);
        if (BK_LAUNCH_CHECK("k_schur_r0")) return -1;
    }
    if (anyp && tot[2] > 0) {
        hipLaunchKernelGGL(k_schur_pairs, dim3((unsigned)tot[2]), dim3(256), 0, g_bk_stream, d, (const int *)(dv + po + 2 * pl), n_it,
                           (double *)g_heap);
        if (BK_LAUNCH_CHECK("k_schur_pairs")) return -1;
    }
    if (tok >= 0) {
        double by = 0, fl = 0;
        for (int q = 0; q < n_it; q++) {
            const double m = it[q].m;
            if (it[q].diag) { by += 8.0 * m * m + 2.0 * 8.0 * m * m; fl += 3.0 * m * m; }   // X, Zi entries read once, M written
            else { by += 8.0 * m * m; fl += 0; }
        }
        bk_prof_end(tok, 10, fl, by);
    }
    return 0;
}

// ---- Lanczos --------------------------------------------------------------------------------
__device__ double block_sum(double v, double *sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) sh[t] += sh[t + w];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}
__global__ __launch_bounds__(256) void k_lz_start(const LzItem *__restrict__ it, int nit, double *__restrict__ H) {
    const int q = blockIdx.x;
    if (q >= nit) return;
    const LzItem L = it[q];
    __shared__ double sh[256];
    double *Q = H + L.Q;
    double s = 0;
    for (int i = threadIdx.x; i < L.n; i += 256) { const double v = 1.0 + 0.5 * sin(1.7 * i + 0.3); Q[i] = v; s += v * v; }
    const double nrm = 1.0 / sqrt(block_sum(s, sh));
    for (int i = threadIdx.x; i < L.n; i += 256) Q[i] *= nrm;
}
// sum over the 4 waves of a workgroup of per-lane values, for np values at once (fixed order)
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// G(r, r) = (Zi V X)(r, r) = sum_t Zi(t, r) V(t, t) X(t, r) per block, V diagonal (Zi at off, V at
// off2, X at off3, G at off4): one wave per row, the columns r of Zi and X read contiguously
__global__ void k_zvx_diag(const MItem *__restrict__ it, const int *__restrict__ pre, int nit, double *__restrict__ H) {
    FLAT_MITEM
    const int n = m.n, lane = threadIdx.x & 63;
    for (int64_t r = lb * 4 + (threadIdx.x >> 6); r < n; r += nb * 4) {
        const double *zi = H + m.off + r * n, *x = H + m.off3 + r * n, *v = H + m.off2;
        double s0 = 0, s1 = 0;
        int t = lane;
        for (; t + 64 < n; t += 128) {
            s0 += zi[t] * v[t + (int64_t)t * n] * x[t];
            s1 += zi[t + 64] * v[(t + 64) + (int64_t)(t + 64) * n] * x[t + 64];
        }
        if (t < n) s0 += zi[t] * v[t + (int64_t)t * n] * x[t];
        const double sum = wave_sum(s0 + s1);
        if (lane == 0) H[m.off4 + r + r * n] = sum;
    }
}
extern "C" int bk_zvx_diag(const MItem *it, int n_it) {
    if (n_it <= 0) return 0;
    std::vector<int64_t> cnt(n_it);
    for (int q = 0; q < n_it; q++) cnt[q] = flat_cap((it[q].n + 3) / 4, 1024);
    const MItem *d = nullptr;
    const int *pre = nullptr;
    const int64_t nb = flat_stage_items(it, sizeof(MItem) * n_it, cnt, (const void **)&d, &pre);
    if (nb <= 0) return (int)nb;
    hipLaunchKernelGGL(k_zvx_diag, dim3((unsigned)nb), dim3(256), 0, g_bk_stream, d, pre, n_it, (double *)g_heap);
    return BK_LAUNCH_CHECK("k_zvx_diag");
}
// ---- fused Lanczos for blocks of order <= 256: one workgroup per block, one row per thread -----
#define LZF 256
// number of eigenvalues below x of the m x m tridiagonal (a, b) (Sturm sequence: the pivots of
// the LDL' factor of T - x I). 1 / q is the hardware reciprocal refined by one Newton step
// (relative error below 1e-13, against the 2e-16 of a division, at a third of its latency): a
// pivot's sign is that of the exact recurrence of a matrix within about 1e-13 ||T|| of T
__device__ __forceinline__ int sturm_count(const double *a, const double *b, int m, double x) {
    int cnt = 0;
    double q = a[0] - x;
    if (q < 0) cnt++;
    for (int k = 1; k < m; k++) {
        const double bb = b[k - 1] * b[k - 1], d = a[k] - x;
        if (fabs(q) < 1e-300) q = q < 0 ? -1e-300 : 1e-300;
        double r = __builtin_amdgcn_rcp(q);
        r = __builtin_fma(r, __builtin_fma(-q, r, 1.0), r);
        q = __builtin_fma(-bb, r, d);
        if (q < 0) cnt++;
    }
    return cnt;
}
// the smallest Ritz value of the leading m x m tridiagonal (al, be) after kk of K steps of a block
// of order n: bisection with 64 Sturm counts per round (one wave, lane = lane index), then two
// inverse iterations below it for the last eigenvector entry (lane 0, in the LDS workspace ws of
// 2 BK_LZK doubles). out = theta - |beta z_last| (or -1e300 for coefficients that are not
// finite); stop when converged to tol, at a breakdown or at K; lane 0's out and stop are final.
__device__ void lz_ritz(const double *al, const double *be, int kk, int K, int n, double tol, int lane, double *ws, double &out_, int &stop_) {
    int m = kk, brk = 0;
    double scale = 0;
    for (int u = 0; u < kk; u++) {
        if (fabs(al[u]) > scale) scale = fabs(al[u]);
        if (!(be[u] > 1e-12 * (scale + 1e-300))) { m = u + 1; brk = 1; break; }
    }
    int bad = 0;
    for (int u = 0; u < m; u++) if (!isfinite(al[u]) || !isfinite(be[u])) bad = 1;
    double out = -1e300;
    int stop = 1;
    if (!bad) {
        double lo = INFINITY, hi = -INFINITY;
        for (int u = 0; u < m; u++) {
            const double r = (u > 0 ? fabs(be[u - 1]) : 0.0) + (u + 1 < m ? fabs(be[u]) : 0.0);
            lo = fmin(lo, al[u] - r); hi = fmax(hi, al[u] + r);
        }
        for (int round = 0; round < 16; round++) {
            const double w_ = hi - lo;
            if (!(w_ > 2e-16 * fmax(fabs(lo), fabs(hi))) || w_ <= 1e-300) break;
            const double x = lo + w_ * (double)(lane + 1) / 65.0;
            const int c = sturm_count(al, be, m, x);
            const unsigned long long bal = __ballot(c >= 1);
            if (!bal) {
                // lambda_min lies above the 64th point: in [lo + 64 w / 65, hi]
                lo = lo + w_ * 64.0 / 65.0;
                continue;
            }
            const int f = __ffsll((long long)bal) - 1;
            const double nlo = f > 0 ? lo + w_ * (double)f / 65.0 : lo;
            hi = lo + w_ * (double)(f + 1) / 65.0;
            lo = nlo;
        }
        const double th = 0.5 * (lo + hi);
        double res = 0;
        if (!brk && m < n && lane == 0) {
            // (T - sigma I) x = r, sigma just below lambda_min (an SPD tridiagonal solve), in the
            // workspace ws (2 BK_LZK doubles) by lane 0
            const double sig = lo - (hi - lo) - 1e-15 * (scale + 1e-300);
            double *x = ws, *dd = ws + BK_LZK;
            for (int u = 0; u < m; u++) x[u] = 1.0;
            for (int iter = 0; iter < 2; iter++) {
                dd[0] = al[0] - sig;
                for (int u = 1; u < m; u++) {
                    if (fabs(dd[u - 1]) < 1e-300) dd[u - 1] = 1e-300;
                    const double l = be[u - 1] / dd[u - 1];
                    dd[u] = al[u] - sig - l * be[u - 1];
                    x[u] -= l * x[u - 1];
                }
                if (fabs(dd[m - 1]) < 1e-300) dd[m - 1] = 1e-300;
                x[m - 1] /= dd[m - 1];
                for (int u = m - 2; u >= 0; u--) x[u] = (x[u] - be[u] * x[u + 1]) / dd[u];
                double nx = 0;
                for (int u = 0; u < m; u++) nx += x[u] * x[u];
                nx = 1.0 / sqrt(nx);
                for (int u = 0; u < m; u++) x[u] *= nx;
            }
            res = fabs(be[m - 1] * x[m - 1]);
        }
        out = th - res;
        stop = brk || kk >= K || res <= tol * fabs(th) + 1e-14 * scale;
    }
    out_ = out; stop_ = stop;
}
// a product row . xs over the row's columns c in [row, n) (kind 1), [0, n) (0) or [0, row] (2):
// thread (row, part) forms the part's share of the columns, the row's thread adds the LZP parts
// in order (every thread returns; only the first LZF hold rows)
#define LZP 4
__device__ __forceinline__ double lzf_mv(const double *__restrict__ M, int n, int kind, const double *xs, double (*pr)[LZF]) {
    const int t = threadIdx.x, row = t % LZF, part = t / LZF;
    const int lo = kind == 1 ? row : 0, hi = kind == 2 ? row + 1 : n, q = (n + LZP - 1) / LZP;
    const int c0 = max(lo, part * q), c1 = min(hi, part * q + q);
    double s0 = 0, s1 = 0;
    if (row < n) {
        int c = c0;
        for (; c + 1 < c1; c += 2) { s0 += M[row + (int64_t)c * n] * xs[c]; s1 += M[row + (int64_t)(c + 1) * n] * xs[c + 1]; }
        if (c < c1) s0 += M[row + (int64_t)c * n] * xs[c];
    }
    pr[part][row] = s0 + s1;
    __syncthreads();
    double v = 0;
    if (t < LZF) {
#pragma unroll
        for (int u = 0; u < LZP; u++) v += pr[u][t];
    }
    __syncthreads();
    return v;
}
// the squared norm of a vector held one entry per row thread (t < LZF), fixed order: across
// each of the 4 row waves, then their four sums
__device__ __forceinline__ double lzf_norm2(double v, double *red4) {
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    if (wv < 4) {
        const double s = wave_sum(t < LZF ? v * v : 0.0);
        if (lane == 0) red4[wv] = s;
    }
    __syncthreads();
    const double r = (red4[0] + red4[1]) + (red4[2] + red4[3]);
    __syncthreads();
    return r;
}
__global__ __launch_bounds__(LZF * LZP) void k_lz_fused(const LzfItem *__restrict__ it, int nit, double *__restrict__ H) {
    const int qi = blockIdx.x;
    if (qi >= nit) return;
    const LzfItem L = it[qi];
    const int n = L.n, K = L.K, t = threadIdx.x, lane = t & 63;
    __shared__ double xs[LZF];
    __shared__ double pr[LZP][LZF];
    __shared__ double Qs[BK_LZK][LZF];    // the basis, q_p(i) at Qs[p][i]
    __shared__ double ws[LZF];
    __shared__ double cs[32];
    __shared__ double red4[4];
    __shared__ double al[BK_LZK], be[BK_LZK], rw[2 * BK_LZK];
    __shared__ double lam_sh;
    __shared__ int stop_sh;
    const double *Li = H + L.Li, *LiT = H + L.LiT, *V = H + L.V;
    // q_0: the reference's start vector, normalised
    double q0 = t < n ? 1.0 + 0.5 * sin(1.7 * t + 0.3) : 0.0;
    q0 *= 1.0 / sqrt(lzf_norm2(q0, red4));
    if (t < LZF) Qs[0][t] = q0;
    // the coefficient groups: 32 groups of 32 lanes, group p summing q_p(i) w(i) over the rows
    // i = l, l + 32, ... (fixed order: eight rows at most, then across the group's lanes)
    const int gp = t >> 5, gl = t & 31;
    double lam = 0;
    for (int j = 0; j < K; j++) {
        // w = T q_j: the current basis vector through LDS, one output row per thread
        __syncthreads();
        if (t < LZF) xs[t] = Qs[j][t];
        __syncthreads();
        double w;
        if (L.mode == 0) w = lzf_mv(V, n, 0, xs, pr);
        else {
            const double t1 = lzf_mv(LiT, n, 1, xs, pr);    // Li' q (upper: columns c >= row)
            if (t < LZF) xs[t] = t1;
            __syncthreads();
            const double t2 = lzf_mv(V, n, 0, xs, pr);      // V t1
            if (t < LZF) xs[t] = t2;
            __syncthreads();
            w = lzf_mv(Li, n, 2, xs, pr);                   // Li t2 (lower)
        }
        // classical Gram-Schmidt twice; the first pass's coefficient of q_j is alpha_j
        double alpha = 0;
        for (int pass = 0; pass < 2; pass++) {
            if (t < LZF) ws[t] = t < n ? w : 0.0;
            __syncthreads();
            if (gp <= j) {
                double c = 0;
                for (int i = gl; i < n; i += 32) c += Qs[gp][i] * ws[i];
#pragma unroll
                for (int o = 16; o > 0; o >>= 1) c += __shfl_xor(c, o, 32);
                if (gl == 0) cs[gp] = c;
            }
            __syncthreads();
            if (pass == 0) alpha = cs[j];
            if (t < n)
                for (int p = 0; p <= j; p++) w -= cs[p] * Qs[p][t];
        }
        const double beta = sqrt(lzf_norm2(t < n ? w : 0.0, red4));
        if (t == 0) { al[j] = alpha; be[j] = beta; }
        if (j + 1 < K && t < LZF) {
            const double inv = beta > 1e-300 ? 1.0 / beta : 0.0;
            Qs[j + 1][t] = t < n ? w * inv : 0.0;
        }
        const int kk = j + 1;
        if (!(kk == K || (kk >= 6 && kk % 3 == 0))) continue;
        __syncthreads();
        // the smallest Ritz value of the leading m x m tridiagonal (wave 0)
        if (t < 64) {
            double out;
            int stop;
            lz_ritz(al, be, kk, K, n, L.tol, lane, rw, out, stop);
            if (t == 0) { lam_sh = out; stop_sh = stop; }
        }
        __syncthreads();
        lam = lam_sh;
        if (stop_sh) break;
    }
    if (t == 0) H[L.out] = lam;
}
extern "C" int bk_lanczos_fused(const LzfItem *it, int n_it) {
    if (n_it <= 0) return 0;
    for (int q = 0; q < n_it; q++) if (it[q].n > LZF || it[q].K > BK_LZK || it[q].K < 1) return -1;
    const LzfItem *d = (const LzfItem *)bk_stage(it, sizeof(LzfItem) * n_it);
    if (!d) return -1;
    const int tok = bk_prof_begin();
    hipLaunchKernelGGL(k_lz_fused, dim3(n_it), dim3(LZF * LZP), 0, g_bk_stream, d, n_it, (double *)g_heap);
    if (tok >= 0) bk_prof_end(tok, 12, 0, 0);
    return BK_LAUNCH_CHECK("k_lz_fused");
}
extern "C" int bk_lanczos_start(const LzItem *it, int n_it) {
    if (n_it <= 0) return 0;
    const LzItem *d = (const LzItem *)bk_stage(it, sizeof(LzItem) * n_it);
    hipLaunchKernelGGL(k_lz_start, dim3(n_it), dim3(256), 0, g_bk_stream, d, n_it, (double *)g_heap);
    return BK_LAUNCH_CHECK("k_lz_start");
}
// ---- Lanczos step for blocks of order > 256 over many workgroups ----------------------------
// Workgroup b of item q owns rows [256 c, 256 c + 256), c = b - pre[q], one row per thread.
// Partial sums go to slots of LZS doubles per workgroup (coefficients of q_0..q_j at 0..31, the
// squared norm at 32); every workgroup of the item adds the item's slots in workgroup order, so
// the coefficients are the same in all of them and independent of the batch.
#define LZS 33
// per-workgroup sums of np values (p <= j) over the 4 waves, fixed order, into red[p] (LDS)
__device__ __forceinline__ void lzm_wg_sums(double (&c)[32], int j, double (*red)[32], double *out) {
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
#pragma unroll
    for (int p = 0; p < 32; p++) {
        if (p <= j) {
            const double v = wave_sum(c[p]);
            if (lane == 0) red[wv][p] = v;
        }
    }
    __syncthreads();
    if (t <= j) out[t] = (red[0][t] + red[1][t]) + (red[2][t] + red[3][t]);
    __syncthreads();
}
// sum of a column of workgroup slots in workgroup order, 8 loads at a time (index clamped,
// value masked): the loads of a batch are in flight together, the order of the sum unchanged
__device__ __forceinline__ double lzm_slot_sum(const double *__restrict__ in, int b0, int nb, int col) {
    double v = 0;
    for (int u0 = 0; u0 < nb; u0 += 8) {
        double x[8];
#pragma unroll
        for (int u = 0; u < 8; u++) x[u] = in[(int64_t)(b0 + min(u0 + u, nb - 1)) * LZS + col];
#pragma unroll
        for (int u = 0; u < 8; u++) if (u0 + u < nb) v += x[u];
    }
    return v;
}
// slots of pass 0: c_p = sum over the rows of q_p(i) w(i)
__global__ __launch_bounds__(256) void k_lzm_part(const LzItem *__restrict__ items, const int *__restrict__ pre, int nit, int j,
                                                  double *__restrict__ H, double *__restrict__ part) {
    const int q = flat_item(pre, nit, blockIdx.x);
    const LzItem L = items[q];
    if (j >= L.K) return;
    __shared__ double red[4][32];
    __shared__ double cf[32];
    const int t = threadIdx.x, i = 256 * (blockIdx.x - pre[q]) + t, ic = min(i, L.n - 1);
    const double *Q = H + L.Q, *w = H + L.w;
    double c[32];
    const double wi = i < L.n ? w[i] : 0.0;
    // every basis entry loaded (index clamped into the basis, value masked): 32 loads in flight
#pragma unroll
    for (int p = 0; p < 32; p++) c[p] = Q[(int64_t)min(p, j) * L.n + ic];
#pragma unroll
    for (int p = 0; p < 32; p++) c[p] = (p <= j && i < L.n) ? c[p] * wi : 0.0;
    lzm_wg_sums(c, j, red, cf);
    if (t <= j) part[(int64_t)blockIdx.x * LZS + t] = cf[t];
}
// one Gram-Schmidt pass: w -= sum_p cf_p q_p with cf the item's summed slots of `in`; pass 0
// records alpha_j = cf_j and writes the next pass's slots, pass 1 the squared-norm slot
__global__ __launch_bounds__(256) void k_lzm_apply(const LzItem *__restrict__ items, const int *__restrict__ pre, int nit, int j, int pass,
                                                   double *__restrict__ H, const double *__restrict__ in, double *__restrict__ outp) {
    const int q = flat_item(pre, nit, blockIdx.x);
    const LzItem L = items[q];
    if (j >= L.K) return;
    __shared__ double red[4][32];
    __shared__ double cf[32], cs[32];
    const int t = threadIdx.x, b0 = pre[q], nb = pre[q + 1] - b0, i = 256 * (blockIdx.x - b0) + t;
    if (t <= j) cf[t] = lzm_slot_sum(in, b0, nb, t);
    __syncthreads();
    double *Q = H + L.Q, *w = H + L.w;
    if (pass == 0 && blockIdx.x == b0 && t == 0) H[L.ab + j] = cf[j];
    // this row's basis entries q_0..q_j, all loads in flight (index clamped, value masked), used by
    // the update and by the next pass's partial sums
    const int ic = min(i, L.n - 1);
    double qv[32];
#pragma unroll
    for (int p = 0; p < 32; p++) qv[p] = Q[(int64_t)min(p, j) * L.n + ic];
#pragma unroll
    for (int p = 0; p < 32; p++) qv[p] = p <= j ? qv[p] : 0.0;
    double v = 0;
    if (i < L.n) {
        v = w[i];
#pragma unroll
        for (int p = 0; p < 32; p++) if (p <= j) v -= cf[p] * qv[p];
        w[i] = v;
    }
    if (pass == 0) {
        double c[32];
#pragma unroll
        for (int p = 0; p < 32; p++) c[p] = (p <= j && i < L.n) ? qv[p] * v : 0.0;
        lzm_wg_sums(c, j, red, cs);
        if (t <= j) outp[(int64_t)blockIdx.x * LZS + t] = cs[t];
    } else {
        double c[32];
        c[0] = v * v;
        lzm_wg_sums(c, 0, red, cs);
        if (t == 0) outp[(int64_t)blockIdx.x * LZS + 32] = cs[0];
    }
}
// beta_j = |w| from the item's norm slots; q_{j+1} = w / beta_j
__global__ __launch_bounds__(256) void k_lzm_fin(const LzItem *__restrict__ items, const int *__restrict__ pre, int nit, int j,
                                                 double *__restrict__ H, const double *__restrict__ in) {
    const int q = flat_item(pre, nit, blockIdx.x);
    const LzItem L = items[q];
    if (j >= L.K) return;
    __shared__ double bsh;
    const int t = threadIdx.x, b0 = pre[q], nb = pre[q + 1] - b0, i = 256 * (blockIdx.x - b0) + t;
    if (t == 0) {
        const double v = lzm_slot_sum(in, b0, nb, 32);
        bsh = sqrt(v);
        if (blockIdx.x == b0) H[L.ab + 32 + j] = bsh;
    }
    __syncthreads();
    const double beta = bsh;
    if (j + 1 < L.K && i < L.n) {
        const double inv = beta > 1e-300 ? 1.0 / beta : 0.0;
        H[L.Q + (int64_t)(j + 1) * L.n + i] = H[L.w + i] * inv;
    }
}
// ---- the same step in one workgroup per block, for blocks of order <= LZC_N ----------------------
// w is staged in LDS; the coefficients of a pass come from 32 groups of 32 lanes (group p sums
// q_p(i) w(i) over the rows i = lane, lane + 32, ... with LZC_U accumulators, then across its 32
// lanes; fixed order), the updates from one thread per row, in the order q = 0, 1, ... One launch replaces the four of the
// workgroup-slot form, and the basis is read from L2 by one CU (n (j + 1) doubles per pass).
#define LZC_N 4096
#define LZC_T 1024
#define LZC_U 16          // loads in flight per lane in the coefficients' sums
__global__ __launch_bounds__(LZC_T) void k_lzm_cgs(const LzItem *__restrict__ items, const int *__restrict__ sel, int j,
                                                  double *__restrict__ H) {
    const LzItem L = items[sel[blockIdx.x]];
    if (j >= L.K) return;
    __shared__ double ws[LZC_N];
    __shared__ double cs[32];
    __shared__ double red[LZC_T / 64];
    const int n = L.n, t = threadIdx.x, lane = t & 63, wv = t >> 6, p = t >> 5, l = t & 31;
    double *Q = H + L.Q, *w = H + L.w;
    for (int i = t; i < n; i += LZC_T) ws[i] = w[i];
    __syncthreads();
    for (int pass = 0; pass < 2; pass++) {
        if (p <= j) {
            // LZC_U loads of the basis vector in flight per lane (LZC_U accumulators, summed pairwise)
            const double *qp = Q + (int64_t)p * n;
            double a[LZC_U];
#pragma unroll
            for (int u = 0; u < LZC_U; u++) a[u] = 0;
            int i = l;
            for (; i + 32 * (LZC_U - 1) < n; i += 32 * LZC_U) {
#pragma unroll
                for (int u = 0; u < LZC_U; u++) a[u] += qp[i + 32 * u] * ws[i + 32 * u];
            }
            if (i < n) {        // the tail (< 32 LZC_U rows) as one batch, indices clamped and masked
                double b[LZC_U];
#pragma unroll
                for (int u = 0; u < LZC_U; u++) b[u] = qp[min(i + 32 * u, n - 1)];
#pragma unroll
                for (int u = 0; u < LZC_U; u++)
                    if (i + 32 * u < n) a[u] += b[u] * ws[i + 32 * u];
            }
#pragma unroll
            for (int h = LZC_U / 2; h > 0; h /= 2)
#pragma unroll
                for (int u = 0; u < h; u++) a[u] += a[u + h];
            double c = a[0];
#pragma unroll
            for (int o = 16; o > 0; o >>= 1) c += __shfl_xor(c, o, 32);
            if (l == 0) cs[p] = c;
        }
        __syncthreads();
        if (pass == 0 && t == 0) H[L.ab + j] = cs[j];
        // the update, in the order q = 0, 1, ..., j
        for (int i = t; i < n; i += LZC_T) {
            double v = ws[i];
            for (int q = 0; q <= j; q++) v -= cs[q] * Q[(int64_t)q * n + i];
            ws[i] = v;
        }
        __syncthreads();
    }
    double s2 = 0;
    for (int i = t; i < n; i += LZC_T) s2 += ws[i] * ws[i];
    s2 = wave_sum(s2);
    if (lane == 0) red[wv] = s2;
    __syncthreads();
    double tot = 0;
#pragma unroll
    for (int u = 0; u < LZC_T / 64; u++) tot += red[u];
    const double beta = sqrt(tot);
    if (t == 0) H[L.ab + 32 + j] = beta;
    const double inv = beta > 1e-300 ? 1.0 / beta : 0.0;
    for (int i = t; i < n; i += LZC_T) {
        w[i] = ws[i];
        if (j + 1 < L.K) Q[(int64_t)(j + 1) * n + i] = ws[i] * inv;
    }
}
static double *g_lzm_part = nullptr;
static int64_t g_lzm_cap = 0;
extern "C" int bk_lanczos_step(const LzItem *it_all, int n_all, int j) {
    if (n_all <= 0) return 0;
    // blocks of order <= LZC_N: one workgroup each; the rest in the workgroup-slot form below
    std::vector<LzItem> big;
    std::vector<int> small;
    for (int q = 0; q < n_all; q++) {
        if (it_all[q].K > 32) return -1;
        if (it_all[q].n <= LZC_N) small.push_back(q);
        else big.push_back(it_all[q]);
    }
    if (!small.empty()) {
        std::vector<char> buf(sizeof(LzItem) * n_all + sizeof(int) * small.size() + 16);
        memcpy(buf.data(), it_all, sizeof(LzItem) * n_all);
        const size_t so = (sizeof(LzItem) * n_all + 15) & ~(size_t)15;
        buf.resize(so + sizeof(int) * small.size());
        memcpy(buf.data() + so, small.data(), sizeof(int) * small.size());
        const char *d = (const char *)bk_stage(buf.data(), buf.size());
        if (!d) return -1;
        const int tok = bk_prof_begin();
        hipLaunchKernelGGL(k_lzm_cgs, dim3((unsigned)small.size()), dim3(LZC_T), 0, g_bk_stream, (const LzItem *)d, (const int *)(d + so), j,
                           (double *)g_heap);
        if (tok >= 0) bk_prof_end(tok, 12, 0, 0);
        if (BK_LAUNCH_CHECK("k_lzm_cgs")) return -1;
    }
    if (big.empty()) return 0;
    const LzItem *it = big.data();
    const int n_it = (int)big.size();
    std::vector<int64_t> cnt(n_it);
    for (int q = 0; q < n_it; q++) cnt[q] = it[q].n > 0 ? (it[q].n + 255) / 256 : 0;
    const LzItem *d;
    const int *pre;
    const int64_t nb = flat_stage_items(it, sizeof(LzItem) * n_it, cnt, (const void **)&d, &pre);
    if (nb < 0) return -1;
    if (nb == 0) return 0;
    if (2 * nb * LZS > g_lzm_cap) {          // two slot sets: the passes alternate between them
        if (bk_sync()) return -1;
        if (g_lzm_part) (void)hipFree(g_lzm_part);
        g_lzm_cap = 2 * nb * LZS + 4096;
        if (bk_check(hipMalloc((void **)&g_lzm_part, sizeof(double) * g_lzm_cap), "hipMalloc lanczos slots")) { g_lzm_cap = 0; g_lzm_part = nullptr; return -1; }
    }
    double *pa = g_lzm_part, *pb = g_lzm_part + nb * LZS, *Hd = (double *)g_heap;
    const int tok = bk_prof_begin();
    hipLaunchKernelGGL(k_lzm_part, dim3((unsigned)nb), dim3(256), 0, g_bk_stream, d, pre, n_it, j, Hd, pa);
    hipLaunchKernelGGL(k_lzm_apply, dim3((unsigned)nb), dim3(256), 0, g_bk_stream, d, pre, n_it, j, 0, Hd, (const double *)pa, pb);
    hipLaunchKernelGGL(k_lzm_apply, dim3((unsigned)nb), dim3(256), 0, g_bk_stream, d, pre, n_it, j, 1, Hd, (const double *)pb, pa);
    hipLaunchKernelGGL(k_lzm_fin, dim3((unsigned)nb), dim3(256), 0, g_bk_stream, d, pre, n_it, j, Hd, (const double *)pa);
    if (tok >= 0) bk_prof_end(tok, 12, 0, 0);
    return BK_LAUNCH_CHECK("k_lz_step");
}
