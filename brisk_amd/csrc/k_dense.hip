// k_dense.hip - dense batched linear algebra on CDNA4 matrix cores (gfx950):
//   * batched GEMM on MFMA (f64: v_mfma_f64_16x16x4f64, f32: v_mfma_f32_16x16x4f32), LDS-staged
//     64x64x16 tiles, 4 waves per workgroup each owning a 32x32 quadrant, triangular K-range
//     trimming and a lower-only epilogue for SYRK-shaped updates;
//   * blocked right-looking Cholesky of a set of matrices (64-wide panels: an LDS kernel factors
//     and inverts the diagonal block, MFMA GEMMs do the panel and the trailing update);
//   * triangular inverse by recursive doubling (log2(n/64) levels of batched GEMMs);
//   * batched matrix-vector products (one wave per output element for the transposed form).
// The dense Cholesky replaces the reference's task-DAG potrf (sparsechol.c:3759-3859) and its
// triangular solves (sparsechol.c:4058-4093); M's solves become products with the explicit
// inverse factor, which are HBM-bound streams instead of latency-bound substitutions.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <math.h>
#include <vector>
#include <algorithm>
#include <string.h>
#include "brisk_dev.h"
#include "rt_internal.h"
#include "chol_diag.h"
#include "flat.h"

#define GBM 64
#define GBN 64
#define GBK 16

template <typename T> struct MF;
// out(g, r): the row of the 16x16 result held by register r of a lane in lane group g = lane/16
// (the column is lane % 16). The f64 instruction interleaves the rows of the four lane groups.
template <> struct MF<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mma(double a, double b, v4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    static __device__ __forceinline__ int out(int g, int r) { return g + 4 * r; }
};
template <> struct MF<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mma(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    static __device__ __forceinline__ int out(int g, int r) { return 4 * g + r; }
};

// C = alpha op(A) op(B) + beta C for one BM x BN tile of one item. Eight waves in a 2 x 4 grid
// (two waves per SIMD, so that one's LDS reads and barrier waits overlap the other's products),
// each owning a (BM/2) x (BN/4) part of 16 x 16 MFMA tiles. The MFMA is fed with the B tile
// as its A operand and the A tile as its B operand, so each output register row is a run of 16
// consecutive rows of column-major C (coalesced stores). LDS holds two K slabs: slab s + 1 is
// fetched into registers before slab s is multiplied and written to the other buffer after it,
// so each slab costs one barrier. A slab is 16 deep (32 for the f32 128 x 128 tile). Each
// thread's load addresses are fixed for the launch, a row (column) outside the item clamped to
// the item's last one and the value zeroed when it is written to LDS; a slab adds one offset
// common to the workgroup, so the loads cost one address add each, carry no branches and stay
// in flight across the products (per-slab index and mask arithmetic had cost 5% of f64 and 13%
// of f32 throughput at n = 2000). Rows of a slab are padded to BM + 1 (BN + 1) elements: 16
// consecutive k rows written by consecutive threads then fall in distinct banks.
#define GEMM_NT(BM_) 512                           // threads of a k_gemm workgroup (either tile size)
template <typename T, int TA, int TB, int BM, int BN>
__global__ __launch_bounds__(GEMM_NT(BM)) void k_gemm(const GemmItem *__restrict__ items, const int *__restrict__ pre, int nitems,
                                              const T *__restrict__ A, const T *__restrict__ B, T *__restrict__ C, T alpha,
                                              T beta, const int *__restrict__ info) {
    typedef typename MF<T>::v4 v4;
    constexpr int BK = BM == 128 && sizeof(T) == 4 ? 2 * GBK : GBK, NT = GEMM_NT(BM);
    constexpr int WM = BM / 2, WN = BN / (NT / 128), FM = WM / 16, FN = WN / 16;
    constexpr int LA = BM * BK / NT, LB = BN * BK / NT;
    static_assert(NT % BM == 0 && NT % BN == 0 && NT % BK == 0 && LA <= 32 && LB <= 32 && FN >= 1, "tile shape");
    const int item = flat_item(pre, nitems, blockIdx.x);
    const GemmItem it = items[item];
    if (it.guard >= 0 && info && info[it.guard] != 0) return;
    const int tilesM = (it.M + BM - 1) / BM;
    const int t = blockIdx.x - pre[item];
    int tm, tn;
    if ((it.tri & 1) && it.M == it.N) tri_index(t, tm, tn);   // lower-only square C: its lower tiles only
    else if (BM == 128) {
        // XCD-aware order: workgroups are dealt to the eight XCDs round-robin, so tile t runs on
        // XCD t % 8 (for an item whose first tile is dealt to XCD 0). The eighth of the item's tiles
        // on one XCD is made one run of the order below, which walks groups of 4 tile rows column
        // by column: a 16 x 16 item gives each XCD a 4 x 8 block of tiles, whose A and B strips it
        // reads through its own L2 (2 x 16 strips before). Both steps are bijections.
        const int nt = tilesM * ((it.N + BN - 1) / BN);
        const int u = nt % 8 == 0 ? (t % 8) * (nt / 8) + t / 8 : t;
        const int gw = 4 * ((it.N + BN - 1) / BN), fm = (u / gw) * 4, gs = min(tilesM - fm, 4);
        tm = fm + (u % gw) % gs; tn = (u % gw) / gs;
    } else { tm = t % tilesM; tn = t / tilesM; }
    const int r0 = tm * BM, c0 = tn * BN;
    if ((it.tri & 1) && c0 > r0 + BM - 1) return;
    int kb = 0, ke = it.K;
    if (it.tri & 2) kb = max(kb, r0);            // op(A) upper triangular: A(i,k) = 0 for k < i
    if (it.tri & 4) ke = min(ke, r0 + BM);       // op(A) lower triangular: A(i,k) = 0 for k > i
    if (it.tri & 8) kb = max(kb, c0);            // op(B) lower triangular: B(k,j) = 0 for k < j
    if (it.tri & 16) ke = min(ke, c0 + BN);      // op(B) upper triangular: B(k,j) = 0 for k > j
    kb = (kb / BK) * BK;
    __shared__ T As[2][BK][BM + 1];
    __shared__ T Bs[2][BK][BN + 1];
    const T *a = A + it.a, *b = B + it.b;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, wm = wave & 1, wn = wave >> 1;
    const int M = it.M, N = it.N, lda = it.lda, ldb = it.ldb;
    v4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < FN; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) acc[i][j][r] = 0;
    // load geometry of element e = tid + NT r: op(A)(i, k) and op(B)(k, j), with the coordinate
    // that advances with r stepping SA (SB) per r
    constexpr int SA = TA == 0 ? NT / BM : NT / BK, SB = TB == 0 ? NT / BK : NT / BN;
    const int ia = TA == 0 ? tid % BM : tid / BK, ka = TA == 0 ? tid / BM : tid % BK;
    const int jb = TB == 0 ? tid / BK : tid % BN, kq = TB == 0 ? tid % BK : tid / BN;
    uint32_t va = 0, vb = 0;                     // elements whose row (column) lies inside the item
#pragma unroll
    for (int r = 0; r < LA; r++) { const int i = TA == 0 ? ia : ia + SA * r; va |= (r0 + i < M ? 1u : 0u) << r; }
#pragma unroll
    for (int r = 0; r < LB; r++) { const int j = TB == 0 ? jb + SB * r : jb; vb |= (c0 + j < N ? 1u : 0u) << r; }
    // each load's address at k = 0, its row (column) clamped into the item (the value is zeroed
    // when it is written to LDS); a slab adds one offset common to the workgroup
    const T *pa[LA], *pb[LB];
#pragma unroll
    for (int r = 0; r < LA; r++)
        pa[r] = TA == 0 ? a + min(r0 + ia, M - 1) + (int64_t)(ka + SA * r) * lda
                        : a + ka + (int64_t)min(r0 + ia + SA * r, M - 1) * lda;
#pragma unroll
    for (int r = 0; r < LB; r++)
        pb[r] = TB == 0 ? b + kq + (int64_t)min(c0 + jb + SB * r, N - 1) * ldb
                        : b + min(c0 + jb, N - 1) + (int64_t)(kq + SB * r) * ldb;
    T ra[LA], rb[LB];
    uint32_t ma = 0, mb = 0;                     // loads of the slab in flight that are inside the item
    auto fetch = [&](int k0) {
        const int64_t oa = TA == 0 ? (int64_t)k0 * lda : k0, ob = TB == 0 ? k0 : (int64_t)k0 * ldb;
        if (k0 + BK <= ke) {
#pragma unroll
            for (int r = 0; r < LA; r++) ra[r] = pa[r][oa];
#pragma unroll
            for (int r = 0; r < LB; r++) rb[r] = pb[r][ob];
            ma = va; mb = vb;
        } else {
            // the last, partial slab: a k beyond the range reads the item's first element instead
            ma = 0; mb = 0;
#pragma unroll
            for (int r = 0; r < LA; r++) {
                const bool ok = (TA == 0 ? k0 + ka + SA * r : k0 + ka) < ke;
                ma |= (ok ? 1u : 0u) << r;
                ra[r] = *(ok ? pa[r] + oa : a);
            }
#pragma unroll
            for (int r = 0; r < LB; r++) {
                const bool ok = (TB == 0 ? k0 + kq : k0 + kq + SB * r) < ke;
                mb |= (ok ? 1u : 0u) << r;
                rb[r] = *(ok ? pb[r] + ob : b);
            }
            ma &= va; mb &= vb;
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int r = 0; r < LA; r++) {
            const int e = tid + NT * r;
            const T x = ((ma >> r) & 1) ? ra[r] : (T)0;
            if (TA == 0) As[buf][e / BM][e % BM] = x; else As[buf][e % BK][e / BK] = x;
        }
#pragma unroll
        for (int r = 0; r < LB; r++) {
            const int e = tid + NT * r;
            const T x = ((mb >> r) & 1) ? rb[r] : (T)0;
            if (TB == 0) Bs[buf][e % BK][e / BK] = x; else Bs[buf][e / BN][e % BN] = x;
        }
    };
    const int nk = kb < ke ? (ke - kb + BK - 1) / BK : 0;
    if (nk > 0) { fetch(kb); store(0); }
    __syncthreads();
    for (int s = 0; s < nk; s++) {
        const int buf = s & 1;
        if (s + 1 < nk) fetch(kb + (s + 1) * BK);
        {
            // the operands of step kk + 1 are read from LDS while step kk's products issue
            T av[2][FM], bv[2][FN];
            const int l16 = lane & 15, g4 = lane >> 4;
#pragma unroll
            for (int q = 0; q < FM; q++) av[0][q] = As[buf][g4][wm * WM + q * 16 + l16];
#pragma unroll
            for (int q = 0; q < FN; q++) bv[0][q] = Bs[buf][g4][wn * WN + q * 16 + l16];
#pragma unroll
            for (int kk = 0; kk < BK / 4; kk++) {
                const int cu = kk & 1, nx = cu ^ 1;
                if (kk + 1 < BK / 4) {
                    const int kr = 4 * (kk + 1) + g4;
#pragma unroll
                    for (int q = 0; q < FM; q++) av[nx][q] = As[buf][kr][wm * WM + q * 16 + l16];
#pragma unroll
                    for (int q = 0; q < FN; q++) bv[nx][q] = Bs[buf][kr][wn * WN + q * 16 + l16];
                }
#pragma unroll
                for (int si = 0; si < FM; si++)
#pragma unroll
                    for (int sj = 0; sj < FN; sj++) acc[si][sj] = MF<T>::mma(bv[cu][sj], av[cu][si], acc[si][sj]);
            }
        }
        if (s + 1 < nk) store(buf ^ 1);
        __syncthreads();
    }
    T *c = C + it.c;
    const int ldc = it.ldc;
    // with beta != 0 every C entry of the thread is read first, all loads in flight (an entry
    // outside the item reads its nearest entry inside, and is not stored)
    T cv[FM][FN][4];
    if (beta != (T)0) {
#pragma unroll
        for (int si = 0; si < FM; si++)
#pragma unroll
            for (int sj = 0; sj < FN; sj++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int gi = min(r0 + wm * WM + si * 16 + (lane & 15), M - 1);
                    const int gj = min(c0 + wn * WN + sj * 16 + MF<T>::out(lane >> 4, r), N - 1);
                    cv[si][sj][r] = c[gi + (int64_t)gj * ldc];
                }
    }
#pragma unroll
    for (int si = 0; si < FM; si++)
#pragma unroll
        for (int sj = 0; sj < FN; sj++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int gi = r0 + wm * WM + si * 16 + (lane & 15);
                const int gj = c0 + wn * WN + sj * 16 + MF<T>::out(lane >> 4, r);
                if (gi < M && gj < N && (!(it.tri & 1) || gi >= gj)) {
                    T v = alpha * acc[si][sj][r];
                    if (beta != (T)0) v += beta * cv[si][sj][r];
                    c[gi + (int64_t)gj * ldc] = v;
                }
            }
}

// algorithmic flops of one item: 2 per multiply-add where both operands can be nonzero
static double gemm_flops(const GemmItem &g) {
    const double M = g.M, N = g.N, K = g.K;
    if ((g.tri & 1) && (g.tri & 2) && (g.tri & 8)) return K * (K + 1) * (K + 2) / 3.0;   // C lower, k >= i (sum of 2 (i + 1)(n - i))
    if (g.tri & 1) return M * (M + 1) * K;                                   // C lower only (M == N)
    if ((g.tri & 2) && (g.tri & 8)) return 2.0 * K * K * K / 3.0 + K * K;   // k >= max(i, j), M == N == K
    if (g.tri & 4) return N * M * (M + 1);                                   // op(A) lower, K == M
    if (g.tri & 16) return M * N * (N + 1);                                  // op(B) upper, K == N
    if (g.tri & 8) return 2.0 * M * N * K - M * N * (N - 1);                 // op(B) lower
    if (g.tri & 2) return 2.0 * M * N * K - M * N * (M - 1);                 // op(A) upper
    return 2.0 * M * N * K;
}
// A launch plan: the items and, for each tile size, the prefix array of their workgroups, laid
// out in one host buffer so that a whole sequence of GEMMs (the steps of a factorization) is
// staged to the device in a single copy.
struct GemmPlan {
    size_t items, pre[2];          // byte offsets in the buffer: items, prefix (64 x 64), prefix (128 x 128)
    int n;
    int64_t nb[2];                 // workgroups of each launch
    double flops, bytes;
};
static size_t buf_align(std::vector<char> &buf) {
    const size_t o = (buf.size() + 15) & ~(size_t)15;
    buf.resize(o);
    return o;
}
template <typename T>
static GemmPlan gemm_plan(std::vector<char> &buf, const GemmItem *items, int nitems) {
    GemmPlan P;
    P.n = nitems; P.flops = 0; P.bytes = 0;
    P.items = buf_align(buf);
    buf.insert(buf.end(), (const char *)items, (const char *)(items + nitems));
    // 128 x 128 tiles for the items large enough to fill them, provided the launch then still has
    // a tile for every CU (fewer 128-wide tiles run K-serial on a fraction of the chip: 64 x 64
    // tiles are faster there); 64 x 64 for the rest
    int64_t nbig = 0;
    for (int q = 0; q < nitems; q++) {
        const GemmItem &g = items[q];
        if (g.M < 192 || g.N < 192) continue;
        const int64_t tm = (g.M + 127) / 128, tn = (g.N + 127) / 128;
        nbig += ((g.tri & 1) && g.M == g.N) ? tm * (tm + 1) / 2 : tm * tn;
    }
    const bool use_big = nbig >= 256;
    for (int big = 0; big < 2; big++) {
        const int TB = big ? 128 : 64;
        P.pre[big] = buf_align(buf);
        buf.resize(P.pre[big] + sizeof(int) * (nitems + 1));
        int64_t tot = 0;
        for (int q = 0; q < nitems; q++) {
            const GemmItem &g = items[q];
            ((int *)(buf.data() + P.pre[big]))[q] = (int)tot;
            if ((use_big && g.M >= 192 && g.N >= 192) != (big == 1) || g.M <= 0 || g.N <= 0) continue;
            const int64_t tm = (g.M + TB - 1) / TB, tn = (g.N + TB - 1) / TB;
            tot += ((g.tri & 1) && g.M == g.N) ? tm * (tm + 1) / 2 : tm * tn;
        }
        ((int *)(buf.data() + P.pre[big]))[nitems] = (int)tot;
        P.nb[big] = tot;
    }
    for (int q = 0; q < nitems; q++) {
        P.flops += gemm_flops(items[q]);
        P.bytes += (double)sizeof(T) * ((double)items[q].M * items[q].K + (double)items[q].K * items[q].N + 2.0 * items[q].M * items[q].N);
    }
    return P;
}
template <typename T>
static int gemm_exec(const char *dbuf, const GemmPlan &P, int ta, int tb, T alpha, T beta, const T *A, const T *B, T *C, const int *info) {
    if (P.nb[0] + P.nb[1] == 0) return 0;
    if (P.nb[0] > 0x7fffffff || P.nb[1] > 0x7fffffff) return -1;
    const GemmItem *d = (const GemmItem *)(dbuf + P.items);
    const int tok = bk_prof_begin();
    for (int big = 1; big >= 0; big--) {
        if (!P.nb[big]) continue;
        const int *pre = (const int *)(dbuf + P.pre[big]);
        const dim3 grid((unsigned)P.nb[big]);
        const int nitems = P.n;
#define GL(TA_, TB_, BM_) hipLaunchKernelGGL((k_gemm<T, TA_, TB_, BM_, BM_>), grid, dim3(GEMM_NT(BM_)), 0, g_bk_stream, d, pre, nitems, A, B, C, alpha, beta, info)
        if (big) {
            if (!ta && !tb) GL(0, 0, 128); else if (!ta && tb) GL(0, 1, 128); else if (ta && !tb) GL(1, 0, 128); else GL(1, 1, 128);
        } else {
            if (!ta && !tb) GL(0, 0, 64); else if (!ta && tb) GL(0, 1, 64); else if (ta && !tb) GL(1, 0, 64); else GL(1, 1, 64);
        }
#undef GL
        if (BK_LAUNCH_CHECK("k_gemm")) return -1;
    }
    if (tok >= 0) bk_prof_end(tok, (sizeof(T) == 8 ? 0 : 4) + 2 * ta + tb, P.flops, P.bytes);
    return 0;
}
template <typename T>
static int gemm_launch(int ta, int tb, T alpha, T beta, const T *A, const T *B, T *C, const GemmItem *items, int nitems, const int *info) {
    if (nitems <= 0) return 0;
    std::vector<char> buf;
    const GemmPlan P = gemm_plan<T>(buf, items, nitems);
    if (P.nb[0] + P.nb[1] == 0) return 0;
    const char *d = (const char *)bk_stage(buf.data(), buf.size());
    if (!d) return -1;
    return gemm_exec<T>(d, P, ta, tb, alpha, beta, A, B, C, info);
}
extern "C" int bk_gemm_f64(int ta, int tb, double alpha, double beta, const double *A, const double *B, double *C,
                           const GemmItem *items, int nitems, const int *info) {
    return gemm_launch<double>(ta, tb, alpha, beta, A, B, C, items, nitems, info);
}
extern "C" int bk_gemm_f32(int ta, int tb, float alpha, float beta, const float *A, const float *B, float *C,
                           const GemmItem *items, int nitems, const int *info) {
    return gemm_launch<float>(ta, tb, alpha, beta, A, B, C, items, nitems, info);
}

// ---- Cholesky: the diagonal block ------------------------------------------------------------
#define POTRF_PW256_N 3685                 // by default 256-column panels from this order (n^3 >= 5e10)
// the panel width of a matrix: its item's opt (128 or 256), else by its order, so that a matrix
// is factored the same way in any batch
__host__ __device__ inline int potrf_pw(const MItem &m) {
    const int o = m.opt & 1023;
    return o == 128 || o == 256 ? o : m.n < POTRF_PW256_N ? 128 : 256;
}
// One workgroup per matrix whose last 64-column step this is (no rows below the block), or whose
// steps are split (BK_POTRF_SPLIT: the panel then by a GEMM with Dinv'): factors the nb x nb
// diagonal block at (k0, k0) in LDS, records the first failing column and the smallest squared
// pivot, writes L_kk back (strict upper part of the block zeroed) and its inverse to Dinv.
template <typename T>
__global__ __launch_bounds__(512) void k_potrf_diag(const MItem *__restrict__ items, int nit, int k0, T *__restrict__ A,
                                                    T *__restrict__ Dinv, int *__restrict__ info, T *__restrict__ pmin) {
    const int q = blockIdx.x;
    if (q >= nit) return;
    const MItem it = items[q];
    const int n = it.n;
    if (k0 >= n || (k0 + 64 < n && !(it.opt & BK_POTRF_SPLIT)) || info[it.slot] != 0) return;
    const int nb = min(64, n - k0);
    T *Ab = A + it.off + k0 + (int64_t)k0 * n;
    T *Di = Dinv + it.off2 + (int64_t)(k0 / 64) * 64 * 64;
    T pm;
    const int f = diag_factor_inverse<T>(Ab, n, nb, Ab, n, Di, pm);
    if (threadIdx.x == 0) {
        if (f) info[it.slot] = k0 + f;
        else pmin[it.slot] = (k0 == 0) ? pm : min(pmin[it.slot], pm);
    }
}
// The diagonal block and the panel below it in one launch, for the matrices with rows below the
// block at step k0: workgroup t of a matrix brings the 64 x 64 block and the panel tile of rows
// k0 + 64 (t + 1) .. up to date with the panel's earlier steps (columns K0 .. k0 of the factor, K0
// the panel's first column), factors and inverts the block itself (every workgroup the same
// arithmetic on the same input, so all agree bit for bit) and multiplies the tile by L_kk^{-T} in
// place. Workgroup 0 writes L_kk to the
// staging area W (64 x 64 per step at the matrix's scratch, copied into the matrix by
// k_potrf_lcopy after the last step: the block is read by every workgroup of this launch and must
// not change under them), Dinv, the failing column and the smallest pivot. The panel's rows are
// read whole before they are written, by the one workgroup that owns them.
template <typename T>
__global__ __launch_bounds__(512) void k_potrf_panel(const MItem *__restrict__ items, const int *__restrict__ pre, int nit, int k0,
                                                     T *__restrict__ A, T *__restrict__ Dinv, T *__restrict__ W,
                                                     int *__restrict__ info, T *__restrict__ pmin) {
    const int q = flat_item(pre, nit, blockIdx.x);
    const MItem it = items[q];
    const int n = it.n, t = blockIdx.x - pre[q];
    if (k0 + 64 >= n || (it.opt & BK_POTRF_SPLIT) || info[it.slot] != 0) return;
    const T *Ab = A + it.off + k0 + (int64_t)k0 * n;
    const int r0 = k0 + 64 * (t + 1), prow = min(64, n - r0), K = k0 % potrf_pw(it), K0 = k0 - K;
    T *Pt = A + it.off + r0 + (int64_t)k0 * n;
    T *Lo = t == 0 ? W + it.off3 + (int64_t)(k0 / 64) * 64 * 64 : nullptr;
    T *Di = t == 0 ? Dinv + it.off2 + (int64_t)(k0 / 64) * 64 * 64 : nullptr;
    T pm;
    const int f = diag_factor_inverse<T, true>(Ab, n, 64, Lo, 64, Di, pm, Pt, n, prow, A + it.off + k0 + (int64_t)K0 * n,
                                               A + it.off + r0 + (int64_t)K0 * n, K);
    if (t == 0 && threadIdx.x == 0) {
        if (f) info[it.slot] = k0 + f;
        else pmin[it.slot] = (k0 == 0) ? pm : min(pmin[it.slot], pm);
    }
}
// the staged diagonal factors L_kk of the steps with a panel into the matrices' diagonal blocks
// (after the factorization; 64 x 64 each, one workgroup per block; a matrix whose factorization
// failed keeps whatever its steps left)
template <typename T>
__global__ void k_potrf_lcopy(const MItem *__restrict__ items, const int *__restrict__ pre, int nit, T *__restrict__ A,
                              const T *__restrict__ W) {
    const int q = flat_item(pre, nit, blockIdx.x);
    const MItem it = items[q];
    const int n = it.n, k0 = 64 * (blockIdx.x - pre[q]);
    if (k0 + 64 >= n || (it.opt & BK_POTRF_SPLIT)) return;
    const T *src = W + it.off3 + (int64_t)(k0 / 64) * 64 * 64;
    T *dst = A + it.off + k0 + (int64_t)k0 * n;
    for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) dst[(e & 63) + (int64_t)(e >> 6) * n] = src[e];
}

// the strict upper triangle of every matrix set to zero
template <typename T>
__global__ void k_zero_upper(const MItem *__restrict__ items, const int *__restrict__ pre, int nit, T *__restrict__ A) {
    const int q = flat_item(pre, nit, blockIdx.x);
    const MItem it = items[q];
    const int n = it.n;
    const int64_t tot = (int64_t)n * n, lb = blockIdx.x - pre[q], nb = pre[q + 1] - pre[q];
    for (int64_t e = lb * blockDim.x + threadIdx.x; e < tot; e += nb * blockDim.x) {
        const int r = (int)(e % n), c = (int)(e / n);
        if (r < c) A[it.off + e] = (T)0;
    }
}

// ---- captured launch sequences -----------------------------------------------------------
// A plan (the staged bytes of a whole factorization) and its operand pointers identify a launch
// sequence. The first time a sequence is seen it runs from the staging ring; the second time its
// plan is copied to a buffer of its own and the sequence is captured into a graph (kernels only:
// no copies or memsets), which later calls replay. Eight entries, least recently used replaced.
struct GraphEnt {
    uint64_t key;
    std::vector<char> plan;           // the plan's bytes (compared on a hash match)
    const void *ops[5];
    int elt;
    hipGraph_t graph;
    hipGraphExec_t exec;
    void *dplan;
    uint64_t last;
};
static GraphEnt g_graphs[8];
static uint64_t g_graph_clock = 0;
static uint64_t plan_key(const std::vector<char> &buf, const void *const *ops, int elt) {
    uint64_t h = 0xcbf29ce484222325ull ^ (uint64_t)buf.size() ^ ((uint64_t)elt << 56);
    for (size_t i = 0; i < buf.size(); i++) h = (h ^ (unsigned char)buf[i]) * 0x100000001b3ull;
    for (int k = 0; k < 5; k++) h = (h ^ (uint64_t)(uintptr_t)ops[k]) * 0x100000001b3ull;
    return h;
}
static GraphEnt *graph_find(const std::vector<char> &buf, const void *const *ops, int elt) {
    const uint64_t h = plan_key(buf, ops, elt);
    for (auto &g : g_graphs)
        if (g.key == h && g.elt == elt && g.plan.size() == buf.size() && !memcmp(g.plan.data(), buf.data(), buf.size()) &&
            !memcmp(g.ops, ops, sizeof(g.ops))) {
            g.last = ++g_graph_clock;
            return &g;
        }
    return nullptr;
}
static void graph_drop(GraphEnt &g) {
    if (g.exec || g.dplan) bk_sync();            // nothing of it may still be queued
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    if (g.dplan) (void)hipFree(g.dplan);
    g.exec = nullptr; g.graph = nullptr; g.dplan = nullptr; g.key = 0; g.plan.clear();
}
static void graph_seen(const std::vector<char> &buf, const void *const *ops, int elt) {
    GraphEnt *v = &g_graphs[0];
    for (auto &g : g_graphs) if (g.last < v->last) v = &g;
    graph_drop(*v);
    v->key = plan_key(buf, ops, elt); v->plan = buf; memcpy(v->ops, ops, sizeof(v->ops)); v->elt = elt;
    v->last = ++g_graph_clock;
}
template <typename F>
static int graph_capture(GraphEnt *g, const std::vector<char> &buf, F &&launch) {
    if (bk_check(hipMalloc(&g->dplan, buf.size()), "hipMalloc graph plan") ||
        bk_check(hipMemcpy(g->dplan, buf.data(), buf.size(), hipMemcpyHostToDevice), "hipMemcpy graph plan")) { graph_drop(*g); return -1; }
    int rc = bk_check(hipStreamBeginCapture(g_bk_stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    if (!rc) {
        const int lr = launch((const char *)g->dplan);
        const int er = bk_check(hipStreamEndCapture(g_bk_stream, &g->graph), "hipStreamEndCapture");
        rc = lr || er;
    }
    if (!rc) rc = bk_check(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    if (rc) { graph_drop(*g); return -1; }
    return bk_check(hipGraphLaunch(g->exec, g_bk_stream), "hipGraphLaunch");
}

// Blocked right-looking Cholesky in panels of 128 or 256 columns, each factored left-looking in
// 64-column steps: the 64 columns are first brought up to date with the panel's earlier steps
// (K <= 192: inside the fused diagonal-block and panel launch, else a GEMM before it), then the
// diagonal block is factored and inverted in LDS and the rows below are multiplied by its
// inverse; the trailing matrix takes the whole panel's update as one lower-triangle GEMM with
// K = the panel width (a half or a quarter of the launches and of the trailing traffic of
// 64-wide steps).
template <typename T>
static int potrf_run(const MItem *it, int n_it, T *A, T *Dinv, T *W, int *info, T *pmin) {
    if (n_it <= 0) return 0;
    int nmax = 0;
    for (int q = 0; q < n_it; q++) nmax = it[q].n > nmax ? it[q].n : nmax;
    // every step's GEMMs planned at once: one staging copy for the whole factorization
    std::vector<char> buf;
    const size_t it_off = buf_align(buf);
    buf.insert(buf.end(), (const char *)it, (const char *)(it + n_it));
    std::vector<int64_t> zc(n_it);
    for (int q = 0; q < n_it; q++) zc[q] = flat_cap(((int64_t)it[q].n * it[q].n + 255) / 256, 1024);
    const size_t zpre_off = buf_align(buf);
    buf.resize(zpre_off + sizeof(int) * (n_it + 1));
    int64_t ztot = 0;
    for (int q = 0; q < n_it; q++) { ((int *)(buf.data() + zpre_off))[q] = (int)ztot; ztot += zc[q]; }
    ((int *)(buf.data() + zpre_off))[n_it] = (int)ztot;
    // the staged diagonal factors' copy: one workgroup per 64-column step with a panel
    const size_t lc_off = buf_align(buf);
    buf.resize(lc_off + sizeof(int) * (n_it + 1));
    int64_t lc_tot = 0;
    for (int q = 0; q < n_it; q++) {
        ((int *)(buf.data() + lc_off))[q] = (int)lc_tot;
        lc_tot += it[q].n > 64 && !(it[q].opt & BK_POTRF_SPLIT) ? (it[q].n - 1) / 64 : 0;
    }
    ((int *)(buf.data() + lc_off))[n_it] = (int)lc_tot;
    struct Step { int k0; GemmPlan upd, panel, trail; int has_trail; size_t ppre; int64_t npanel; int diag; };
    std::vector<Step> st;
    std::vector<GemmItem> gu, gp, gt;
    // panel width (potrf_pw): 128 columns for latency (two n = 2000 blocks in FP64: 1.37 -> 1.30
    // ms against 256), 256 where the trailing updates run at higher throughput with K = 256 (four
    // n = 5000 blocks in FP32: 4.87 -> 4.60 ms; the serving stream with 256 from order 1024: 365
    // -> 374 solves/s)
    for (int k0 = 0; k0 < nmax; k0 += 64) {
        gu.clear(); gp.clear(); gt.clear();
        Step S;
        S.k0 = k0;
        // the fused diagonal-block and panel launch: one workgroup per 64-row panel tile of the
        // matrices with rows below the block (flat prefix per item); the one-workgroup diagonal
        // launch for the rest (their last step, or split steps)
        S.ppre = buf_align(buf);
        buf.resize(S.ppre + sizeof(int) * (n_it + 1));
        S.npanel = 0; S.diag = 0;
        for (int q = 0; q < n_it; q++) {
            const int n = it[q].n, split = (it[q].opt & BK_POTRF_SPLIT) != 0;
            ((int *)(buf.data() + S.ppre))[q] = (int)S.npanel;
            if (k0 + 64 < n && !split) S.npanel += (n - k0 - 64 + 63) / 64;
            else if (k0 < n) S.diag = 1;
        }
        ((int *)(buf.data() + S.ppre))[n_it] = (int)S.npanel;
        for (int q = 0; q < n_it; q++) {
            const int n = it[q].n;
            if (k0 >= n) continue;
            const int PW = potrf_pw(it[q]), K0 = k0 - k0 % PW, sub = (k0 - K0) / 64;
            const int fused = k0 + 64 < n && !(it[q].opt & BK_POTRF_SPLIT);
            GemmItem g;
            if (sub > 0 && !fused) {
                // columns k0 .. k0+64 (rows k0 ..) -= L[k0.., K0..k0] L[k0..k0+64, K0..k0]'
                g.M = n - k0; g.N = min(64, n - k0); g.K = k0 - K0;
                g.lda = n; g.ldb = n; g.ldc = n;
                g.a = it[q].off + k0 + (int64_t)K0 * n; g.b = g.a;
                g.c = it[q].off + k0 + (int64_t)k0 * n;
                g.guard = it[q].slot; g.tri = 0;
                gu.push_back(g);
            }
            if (k0 + 64 < n && (it[q].opt & BK_POTRF_SPLIT)) {
                // split step: the panel, in place, A[k0+64:, k0:k0+64] := A[k0+64:, k0:k0+64]
                // Dinv_k' (Dinv_k' upper triangular); the 64 x 64 tiles span the panel's full
                // width, so every workgroup reads its rows completely before it writes them
                g.M = n - k0 - 64; g.N = 64; g.K = 64;
                g.lda = n; g.ldb = 64; g.ldc = n;
                g.a = it[q].off + (k0 + 64) + (int64_t)k0 * n;
                g.b = it[q].off2 + (int64_t)(k0 / 64) * 64 * 64;
                g.c = g.a;
                g.guard = it[q].slot; g.tri = 16;
                gp.push_back(g);
            }
            const int e = K0 + PW;
            if (sub == PW / 64 - 1 && e < n) {
                // trailing: A[e:, e:] -= L[e:, K0:e] L[e:, K0:e]' (lower triangle)
                g.M = n - e; g.N = n - e; g.K = PW;
                g.lda = n; g.ldb = n; g.ldc = n;
                g.a = it[q].off + e + (int64_t)K0 * n; g.b = g.a;
                g.c = it[q].off + e + (int64_t)e * n;
                g.guard = it[q].slot; g.tri = 1;
                gt.push_back(g);
            }
        }
        S.upd = gemm_plan<T>(buf, gu.data(), (int)gu.size());
        S.panel = gemm_plan<T>(buf, gp.data(), (int)gp.size());
        S.trail = gemm_plan<T>(buf, gt.data(), (int)gt.size());
        S.has_trail = !gt.empty();
        st.push_back(S);
    }
    // the launch sequence, reading every plan from the device copy d
    auto launch = [&](const char *d) -> int {
        const MItem *d_it = (const MItem *)(d + it_off);
        if (ztot > 0) {
            hipLaunchKernelGGL((k_zero_upper<T>), dim3((unsigned)ztot), dim3(256), 0, g_bk_stream, d_it, (const int *)(d + zpre_off), n_it, A);
            if (BK_LAUNCH_CHECK("k_zero_upper")) return -1;
        }
        for (size_t u = 0; u < st.size(); u++) {
            const int k0 = st[u].k0;
            if (gemm_exec<T>(d, st[u].upd, 0, 1, (T)-1, (T)1, A, A, A, info)) return -1;
            const int tok = bk_prof_begin();
            if (st[u].npanel > 0) {
                hipLaunchKernelGGL((k_potrf_panel<T>), dim3((unsigned)st[u].npanel), dim3(512), 0, g_bk_stream, d_it,
                                   (const int *)(d + st[u].ppre), n_it, k0, A, Dinv, W, info, pmin);
                if (BK_LAUNCH_CHECK("k_potrf_panel")) return -1;
            }
            if (st[u].diag) {
                hipLaunchKernelGGL((k_potrf_diag<T>), dim3(n_it), dim3(512), 0, g_bk_stream, d_it, n_it, k0, A, Dinv, info, pmin);
                if (BK_LAUNCH_CHECK("k_potrf_diag")) return -1;
            }
            if (tok >= 0) {
                double fl = 0;
                for (int q = 0; q < n_it; q++) {
                    const double b = it[q].n > k0 ? fmin(64.0, it[q].n - k0) : 0, rows = it[q].n > k0 + 64 ? it[q].n - k0 - 64 : 0;
                    fl += 2.0 * b * b * b / 3.0 + 2.0 * rows * 64.0 * 64.0;
                }
                bk_prof_end(tok, 8, fl, 0);
            }
            if (gemm_exec<T>(d, st[u].panel, 0, 1, (T)1, (T)0, A, Dinv, A, info)) return -1;
            if (st[u].has_trail && gemm_exec<T>(d, st[u].trail, 0, 1, (T)-1, (T)1, A, A, A, info)) return -1;
        }
        if (lc_tot > 0) {
            hipLaunchKernelGGL((k_potrf_lcopy<T>), dim3((unsigned)lc_tot), dim3(256), 0, g_bk_stream, d_it, (const int *)(d + lc_off), n_it, A,
                               (const T *)W);
            if (BK_LAUNCH_CHECK("k_potrf_lcopy")) return -1;
        }
        return 0;
    };
    // a factorization whose plan and operands repeat (the same solves, one per iteration) is
    // replayed from a captured graph; with the kernel timers on every launch runs eagerly, so
    // that each one is timed
    if (!bk_prof_on()) {
        const void *ops[5] = { A, Dinv, W, info, pmin };
        GraphEnt *ge = graph_find(buf, ops, sizeof(T));
        if (ge && ge->exec) return bk_check(hipGraphLaunch(ge->exec, g_bk_stream), "hipGraphLaunch potrf");
        if (ge) return graph_capture(ge, buf, launch);
        graph_seen(buf, ops, sizeof(T));
    }
    const char *d = (const char *)bk_stage(buf.data(), buf.size());
    if (!d) return -1;
    return launch(d);
}
extern "C" int bk_potrf_f64(const MItem *it, int n_it, double *A, double *Dinv, double *W, int *info, double *pmin) {
    return potrf_run<double>(it, n_it, A, Dinv, W, info, pmin);
}
extern "C" int bk_potrf_f32(const MItem *it, int n_it, float *A, float *Dinv, float *W, int *info, float *pmin) {
    return potrf_run<float>(it, n_it, A, Dinv, W, info, pmin);
}

// ---- triangular inverse by recursive doubling ----------------------------------------------
// Linv starts as the block diagonal of the 64 x 64 inverses; at level s the lower-left block of
// every 2s x 2s diagonal block is -C^{-1} B A^{-1} (A, C the inverted s x s diagonal blocks, B the
// block of L between them): two batched GEMMs per level, T = B A^{-1} into scratch, then the
// product with C^{-1}.
template <typename T>
__global__ void k_trtri_init(const MItem *__restrict__ items, const int *__restrict__ pre, int nit, const T *__restrict__ Dinv,
                             T *__restrict__ Linv, const int *__restrict__ info) {
    const int q = flat_item(pre, nit, blockIdx.x);
    const MItem it = items[q];
    if (info && info[it.slot] != 0) return;
    const int n = it.n;
    const int64_t tot = (int64_t)n * n, lb = blockIdx.x - pre[q], nb = pre[q + 1] - pre[q];
    for (int64_t e = lb * blockDim.x + threadIdx.x; e < tot; e += nb * blockDim.x) {
        const int r = (int)(e % n), c = (int)(e / n);
        T v = 0;
        if ((r >> 6) == (c >> 6) && r >= c) v = Dinv[it.off2 + (int64_t)(c >> 6) * 4096 + (r & 63) + 64 * (c & 63)];
        Linv[it.off4 + e] = v;
    }
}
template <typename T>
static int trtri_run(const MItem *it, int n_it, const T *L, const T *Dinv, T *Linv, T *W, const int *info, int smax = 1 << 30) {
    if (n_it <= 0) return 0;
    int nmax = 0;
    for (int q = 0; q < n_it; q++) nmax = it[q].n > nmax ? it[q].n : nmax;
    if (nmax > smax) nmax = smax;                   // levels s < smax only: diagonal blocks of order smax
    // all levels planned at once: one staging copy
    std::vector<char> buf;
    const size_t it_off = buf_align(buf);
    buf.insert(buf.end(), (const char *)it, (const char *)(it + n_it));
    const size_t zpre_off = buf_align(buf);
    buf.resize(zpre_off + sizeof(int) * (n_it + 1));
    int64_t ztot = 0;
    for (int q = 0; q < n_it; q++) {
        ((int *)(buf.data() + zpre_off))[q] = (int)ztot;
        ztot += flat_cap(((int64_t)it[q].n * it[q].n + 255) / 256, 1024);
    }
    ((int *)(buf.data() + zpre_off))[n_it] = (int)ztot;
    // the staged diagonal factors' copy: one workgroup per 64-column step with a panel
    const size_t lc_off = buf_align(buf);
    buf.resize(lc_off + sizeof(int) * (n_it + 1));
    int64_t lc_tot = 0;
    for (int q = 0; q < n_it; q++) {
        ((int *)(buf.data() + lc_off))[q] = (int)lc_tot;
        lc_tot += it[q].n > 64 && !(it[q].opt & BK_POTRF_SPLIT) ? (it[q].n - 1) / 64 : 0;
    }
    ((int *)(buf.data() + lc_off))[n_it] = (int)lc_tot;
    std::vector<GemmPlan> p1, p2;
    std::vector<GemmItem> g1, g2;
    for (int s = 64; s < nmax; s *= 2) {
        g1.clear(); g2.clear();
        for (int q = 0; q < n_it; q++) {
            const int n = it[q].n;
            for (int p0 = 0; p0 + s < n; p0 += 2 * s) {
                const int s2 = min(s, n - p0 - s);          // order of the second diagonal block
                GemmItem g;
                // T = B * A^{-1}: B = L[p0+s : p0+s+s2, p0 : p0+s], A^{-1} = Linv[p0:p0+s, p0:p0+s] (lower)
                g.M = s2; g.N = s; g.K = s;
                g.lda = n; g.ldb = n; g.ldc = n;
                g.a = it[q].off + (p0 + s) + (int64_t)p0 * n;
                g.b = it[q].off4 + p0 + (int64_t)p0 * n;
                g.c = it[q].off3 + (p0 + s) + (int64_t)p0 * n;
                g.guard = it[q].slot; g.tri = 8;
                g1.push_back(g);
                // Linv[p0+s.., p0..] = -C^{-1} T, C^{-1} = Linv[p0+s : p0+s+s2, same] (lower)
                g.M = s2; g.N = s; g.K = s2;
                g.lda = n; g.ldb = n; g.ldc = n;
                g.a = it[q].off4 + (p0 + s) + (int64_t)(p0 + s) * n;
                g.b = it[q].off3 + (p0 + s) + (int64_t)p0 * n;
                g.c = it[q].off4 + (p0 + s) + (int64_t)p0 * n;
                g.guard = it[q].slot; g.tri = 4;
                g2.push_back(g);
            }
        }
        p1.push_back(gemm_plan<T>(buf, g1.data(), (int)g1.size()));
        p2.push_back(gemm_plan<T>(buf, g2.data(), (int)g2.size()));
    }
    const char *d = (const char *)bk_stage(buf.data(), buf.size());
    if (!d) return -1;
    if (ztot > 0) hipLaunchKernelGGL((k_trtri_init<T>), dim3((unsigned)ztot), dim3(256), 0, g_bk_stream, (const MItem *)(d + it_off),
                                     (const int *)(d + zpre_off), n_it, Dinv, Linv, info);
    if (BK_LAUNCH_CHECK("k_trtri_init")) return -1;
    for (size_t l = 0; l < p1.size(); l++) {
        // first GEMM: A-operand from L, B-operand from Linv; second: both from Linv / W
        if (gemm_exec<T>(d, p1[l], 0, 0, (T)1, (T)0, L, Linv, W, info)) return -1;
        if (gemm_exec<T>(d, p2[l], 0, 0, (T)-1, (T)0, Linv, W, Linv, info)) return -1;
    }
    return 0;
}
extern "C" int bk_trtri_f64(const MItem *it, int n_it, const double *L, const double *Dinv, double *Linv, double *W, const int *info) {
    return trtri_run<double>(it, n_it, L, Dinv, Linv, W, info);
}
extern "C" int bk_trtri_f32(const MItem *it, int n_it, const float *L, const float *Dinv, float *Linv, float *W, const int *info) {
    return trtri_run<float>(it, n_it, L, Dinv, Linv, W, info);
}
extern "C" int bk_trtri_diag_f64(const MItem *it, int n_it, const double *L, const double *Dinv, double *Linv, double *W, const int *info) {
    return trtri_run<double>(it, n_it, L, Dinv, Linv, W, info, 256);
}
extern "C" int bk_trtri_diag_f32(const MItem *it, int n_it, const float *L, const float *Dinv, float *Linv, float *W, const int *info) {
    return trtri_run<float>(it, n_it, L, Dinv, Linv, W, info, 256);
}

// ---- blocked triangular solves with the factor ---------------------------------------------
// L v1 = v0 forward, then L' v0 = v1 backward, in 256-wide block columns, one launch per block
// column and direction over every item of the batch. Every workgroup of a launch forms the
// block's solution itself from the inverted diagonal block (a 256 x 256 product read from L2),
// so the block and its effect on the rest need no second launch:
//   forward step k:   y_k = Di_k b_k,    b_j -= L_jk y_k   (j > k)
//   backward step k:  z_k = Di_k' y_k,   y_j -= L_kj' z_k  (j < k)
// A workgroup owns 64 rows (forward) or 64 columns of L (backward) of one item; the chunks
// inside block k write its solution. Vectors f64, products and sums in f64, fixed order.
// Items are sorted by order, largest first, so the items with a block k are a prefix.
// a row's dot product with the entries r0 .. r0 + NB - 1 of a strided run of the matrix,
// restricted to [lo, hi) (lo < hi): every load is issued (its index clamped into [lo, hi), an
// address in bounds) and the entries outside are masked afterwards, so the NB loads are in
// flight together instead of each behind a branch (the matrix comes from L2)
template <typename T, int NB>
__device__ __forceinline__ double trsv_dot(const T *__restrict__ p, int64_t stride, int r0, int lo, int hi, const double *__restrict__ x) {
    T v[NB];
#pragma unroll
    for (int u = 0; u < NB; u++) v[u] = p[(int64_t)min(max(r0 + u, lo), hi - 1) * stride];
    double s0 = 0, s1 = 0;
#pragma unroll
    for (int u = 0; u < NB; u += 2) {
        const double a = (r0 + u >= lo && r0 + u < hi) ? (double)v[u] : 0.0;
        const double b = (r0 + u + 1 >= lo && r0 + u + 1 < hi) ? (double)v[u + 1] : 0.0;
        s0 += a * x[r0 + u];
        s1 += b * x[r0 + u + 1];
    }
    return s0 + s1;
}
// loads in flight per batch: 32 of f32, 16 of f64 (within the 128 registers of a 1024-thread group)
#define TRSV_NB (sizeof(T) == 8 ? 16 : 32)
#ifndef TRSV_WG1
#define TRSV_WG1 512
#endif
#ifndef TRSV_G
#define TRSV_G 4
#endif
// 1024 threads: the block's 256 x 256 product as thread (row, quarter) over 64 columns each, the
// update as 16 groups of 16 columns (forward) or 16 waves of 4 columns (backward); partial sums
// added in a fixed order
template <typename T>
__global__ __launch_bounds__(1024) void k_trsv_fwd(const TrsvItem *__restrict__ items, const int *__restrict__ wpre, int nact, int k,
                                                   int G, const T *__restrict__ A, double *__restrict__ V) {
    // the item: the last i with wpre[i] <= blockIdx; its workgroups take G chunks each, from 4 k
    const int g = blockIdx.x;
    int lo = 0, hi = nact - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (wpre[mid] <= g) lo = mid; else hi = mid - 1;
    }
    const TrsvItem it = items[lo];
    const int first = 4 * k + G * (g - wpre[lo]);
    const int m = it.m, kb = 256 * k, nbk = min(256, m - kb), t = threadIdx.x;
    __shared__ double bs[256 + 32], ys[256 + 32], part[16][256];
    double *b = V + it.v0;
    if (t < 256) bs[t] = t < nbk ? b[kb + t] : 0.0;
    else if (t < 288) bs[t] = 0.0;
    __syncthreads();
    {
        // y_r = sum_{c <= r} Di[kb + r, kb + c] b_c: thread (r, q) the columns 64 q .. 64 q + 63
        const int r = t & 255, q = t >> 8, c0 = 64 * q, c1 = min(c0 + 64, r + 1);
        double s = 0;
        if (r < nbk && c0 < c1) {
            const T *d = A + it.Di + (kb + r) + (int64_t)kb * m;
#pragma unroll
            for (int u = 0; u < 64; u += TRSV_NB) s += trsv_dot<T, TRSV_NB>(d, m, c0 + u, c0, c1, bs);
        }
        part[q][r] = s;
    }
    __syncthreads();
    if (t < 256) ys[t] = t < nbk ? (part[0][t] + part[1][t]) + (part[2][t] + part[3][t]) : 0.0;
    else if (t < 288) ys[t] = 0.0;
    __syncthreads();
    for (int chunk = first; chunk < first + G && chunk < it.c; chunk++) {
        const int r0 = 64 * chunk;
        if (chunk < 4 * k + 4) {
            if (t < 64 && r0 + t < m) V[it.v1 + r0 + t] = ys[r0 + t - kb];
            continue;
        }
        // b[rows] -= L[rows, kb .. kb + nbk) y: thread (row, group h) over the columns 16 h .. 16 h + 15
        const int lane = t & 63, h = t >> 6, row = r0 + lane, c0 = 16 * h, c1 = min(c0 + 16, nbk);
        double s = 0;
        if (row < m && c0 < c1) {
            const T *l = A + it.L + row + (int64_t)kb * m;
            s = trsv_dot<T, 16>(l, m, c0, c0, c1, ys);
        }
        part[h][lane] = s;
        __syncthreads();
        if (t < 64 && row < m) {
            double a = 0;
#pragma unroll
            for (int u = 0; u < 16; u++) a += part[u][t];
            b[row] -= a;
        }
        __syncthreads();
    }
}
template <typename T>
__global__ __launch_bounds__(1024) void k_trsv_bwd(const TrsvItem *__restrict__ items, int nact, int k, int G,
                                                   const T *__restrict__ A, double *__restrict__ V) {
    // per item ceil((4 k + 4) / G) workgroups of G column chunks each
    const int per = (4 * k + 4 + G - 1) / G, q = blockIdx.x / per, first = G * (blockIdx.x % per);
    if (q >= nact) return;
    const TrsvItem it = items[q];
    if (first >= it.c) return;
    const int m = it.m, kb = 256 * k, nbk = min(256, m - kb), t = threadIdx.x, lane = t & 63, w = t >> 6;
    __shared__ double ys[256 + 32], zs[256 + 32], part[4][256];
    double *y = V + it.v1;
    if (t < 256) ys[t] = t < nbk ? y[kb + t] : 0.0;
    else if (t < 288) ys[t] = 0.0;
    __syncthreads();
    {
        // z_r = sum_{c >= r} DiT[kb + r, kb + c] y_c, DiT = Di': thread (r, qq) the columns 64 qq .. 64 qq + 63
        const int r = t & 255, qq = t >> 8, c0 = max(64 * qq, r), c1 = min(64 * qq + 64, nbk);
        double s = 0;
        if (r < nbk && c0 < c1) {
            const T *d = A + it.DiT + (kb + r) + (int64_t)kb * m;
#pragma unroll
            for (int u = 0; u < 64; u += TRSV_NB) s += trsv_dot<T, TRSV_NB>(d, m, 64 * qq + u, c0, c1, ys);
        }
        part[qq][r] = s;
    }
    __syncthreads();
    if (t < 256) zs[t] = t < nbk ? (part[0][t] + part[1][t]) + (part[2][t] + part[3][t]) : 0.0;
    else if (t < 288) zs[t] = 0.0;
    __syncthreads();
    for (int chunk = first; chunk < first + G && chunk < it.c && chunk < 4 * k + 4; chunk++) {
        const int r0 = 64 * chunk;
        if (chunk >= 4 * k) {
            if (t < 64 && r0 + t < m) V[it.v0 + r0 + t] = zs[r0 + t - kb];
            continue;
        }
        // y[j] -= sum_r L[kb + r, j] z_r for the columns j = r0 + 4 w .. r0 + 4 w + 3 (j < kb): the
        // wave reads its 4 columns (4 rows per lane each, all 16 loads issued together), then a
        // butterfly per column
        T v[4][4];
#pragma unroll
        for (int jj = 0; jj < 4; jj++) {
            const T *col = A + it.L + kb + (int64_t)(r0 + 4 * w + jj) * m;
#pragma unroll
            for (int u = 0; u < 4; u++) v[jj][u] = col[min(lane + 64 * u, nbk - 1)];      // masked below
        }
#pragma unroll
        for (int jj = 0; jj < 4; jj++) {
            double s = 0;
#pragma unroll
            for (int u = 0; u < 4; u++) s += (lane + 64 * u < nbk ? (double)v[jj][u] : 0.0) * zs[lane + 64 * u];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
            if (lane == 0) y[r0 + 4 * w + jj] -= s;
        }
    }
}
// DiT = Di' on the diagonal blocks: one 64 x 64 tile of the lower part of a block per workgroup
template <typename T>
__global__ __launch_bounds__(256) void k_trsv_prep(const TrsvItem *__restrict__ items, const int *__restrict__ pre, int nit,
                                                   T *__restrict__ A) {
    const int q = flat_item(pre, nit, blockIdx.x);
    const TrsvItem it = items[q];
    const int e = blockIdx.x - pre[q], blk = e / 16, ta = (e % 16) / 4, tb = e % 4;   // tile (ta, tb) of block blk
    const int m = it.m, kb = 256 * blk, nbk = min(256, m - kb);
    const int i0 = kb + 64 * ta, j0 = kb + 64 * tb;
    if (64 * ta >= nbk || 64 * tb >= nbk) return;
    __shared__ T tile[64][65];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int ni = min(64, kb + nbk - i0), nj = min(64, kb + nbk - j0);
    for (int c = w; c < nj; c += 4)                     // tile[c][r] = Di[i0 + r, j0 + c]
        tile[c][lane] = lane < ni ? A[it.Di + (i0 + lane) + (int64_t)(j0 + c) * m] : (T)0;
    __syncthreads();
    for (int r = w; r < ni; r += 4)                     // DiT[j0 + c, i0 + r] = Di[i0 + r, j0 + c]
        if (lane < nj) A[it.DiT + (j0 + lane) + (int64_t)(i0 + r) * m] = tile[lane][r];
}
template <typename T>
static int trsv_sorted(const TrsvItem *it, int n_it, std::vector<TrsvItem> &v) {
    v.assign(it, it + n_it);
    std::stable_sort(v.begin(), v.end(), [](const TrsvItem &a, const TrsvItem &b) { return a.m > b.m; });
    int64_t pre = 0;
    for (auto &x : v) {
        if (x.m <= 0) return -1;
        x.c = (x.m + 63) / 64;
        x.pre = pre;
        pre += x.c;
    }
    return 0;
}
template <typename T>
static int trsv_prep_run(const TrsvItem *it, int n_it, T *A) {
    if (n_it <= 0) return 0;
    std::vector<TrsvItem> v;
    if (trsv_sorted<T>(it, n_it, v)) return bk_fail("trsv: an item of order <= 0");
    std::vector<int64_t> cnt(n_it);
    for (int q = 0; q < n_it; q++) cnt[q] = 16 * (int64_t)((v[q].m + 255) / 256);
    const TrsvItem *d = nullptr;
    const int *pre = nullptr;
    const int64_t nb = flat_stage_items(v.data(), sizeof(TrsvItem) * n_it, cnt, (const void **)&d, &pre);
    if (nb < 0) return -1;
    hipLaunchKernelGGL((k_trsv_prep<T>), dim3((unsigned)nb), dim3(256), 0, g_bk_stream, d, pre, n_it, A);
    return BK_LAUNCH_CHECK("k_trsv_prep");
}
template <typename T>
static int trsv_run(const TrsvItem *it, int n_it, const T *A, double *V) {
    if (n_it <= 0) return 0;
    std::vector<TrsvItem> v;
    if (trsv_sorted<T>(it, n_it, v)) return bk_fail("trsv: an item of order <= 0");
    const int nbk = (v[0].m + 255) / 256;
    std::vector<int> act(nbk);                          // items with a block k: a prefix
    for (int k = 0, a = n_it; k < nbk; k++) {
        while (a > 0 && v[a - 1].m <= 256 * k) a--;
        act[k] = a;
    }
    // Chunks per workgroup: every workgroup of a launch first forms the block's solution from the
    // 256 x 256 inverted diagonal block, so a launch of many workgroups re-reads that block many
    // times. Above TRSV_WG1 one-chunk workgroups, each workgroup takes 4 chunks. The forward
    // launches read their workgroup prefix per item from the staged buffer after the items.
    std::vector<char> buf((const char *)v.data(), (const char *)(v.data() + n_it));
    std::vector<int> gf(nbk), gb(nbk);
    std::vector<size_t> wo(nbk);
    std::vector<int64_t> gridf(nbk);
    for (int k = 0; k < nbk; k++) {
        const int a = act[k];
        int64_t tot = 0;
        for (int i = 0; i < a; i++) tot += v[i].c - 4LL * k;
        gf[k] = tot > TRSV_WG1 ? TRSV_G : 1;
        gb[k] = (int64_t)a * (4 * k + 4) > TRSV_WG1 ? TRSV_G : 1;
        wo[k] = buf_align(buf);
        buf.resize(wo[k] + sizeof(int) * (a + 1));
        int64_t acc = 0;
        for (int i = 0; i <= a; i++) {
            ((int *)(buf.data() + wo[k]))[i] = (int)acc;
            if (i < a) acc += (v[i].c - 4LL * k + gf[k] - 1) / gf[k];
        }
        gridf[k] = acc;
        if (acc <= 0 || acc > 0x7fffffff) return bk_fail("trsv: bad forward grid");
    }
    const char *dd = (const char *)bk_stage(buf.data(), buf.size());
    if (!dd) return -1;
    const TrsvItem *d = (const TrsvItem *)dd;
    const int tok = bk_prof_begin();
    for (int k = 0; k < nbk; k++) {
        hipLaunchKernelGGL((k_trsv_fwd<T>), dim3((unsigned)gridf[k]), dim3(1024), 0, g_bk_stream, d, (const int *)(dd + wo[k]),
                           act[k], k, gf[k], A, V);
        if (BK_LAUNCH_CHECK("k_trsv_fwd")) return -1;
    }
    for (int k = nbk - 1; k >= 0; k--) {
        const int64_t grid = (int64_t)act[k] * ((4 * k + 4 + gb[k] - 1) / gb[k]);
        if (grid <= 0 || grid > 0x7fffffff) return bk_fail("trsv: bad backward grid");
        hipLaunchKernelGGL((k_trsv_bwd<T>), dim3((unsigned)grid), dim3(1024), 0, g_bk_stream, d, act[k], k, gb[k], A, V);
        if (BK_LAUNCH_CHECK("k_trsv_bwd")) return -1;
    }
    if (tok >= 0) {
        double fl = 0, by = 0;
        for (auto &x : v) { fl += 2.0 * x.m * (double)x.m; by += (double)sizeof(T) * x.m * (double)x.m + 4.0 * 8 * x.m; }
        bk_prof_end(tok, 16, fl, by);
    }
    return 0;
}
extern "C" int bk_trsv_prep_f64(const TrsvItem *it, int n_it, double *A) { return trsv_prep_run<double>(it, n_it, A); }
extern "C" int bk_trsv_prep_f32(const TrsvItem *it, int n_it, float *A) { return trsv_prep_run<float>(it, n_it, A); }
extern "C" int bk_trsv_f64(const TrsvItem *it, int n_it, const double *A, double *V) { return trsv_run<double>(it, n_it, A, V); }
extern "C" int bk_trsv_f32(const TrsvItem *it, int n_it, const float *A, double *V) { return trsv_run<float>(it, n_it, A, V); }

// ---- batched matrix-vector products -------------------------------------------------------
// trans = 1: y_j = alpha sum_i A[i, j] x_i + beta y_j, one to four waves per output (column of A
// read contiguously, split in segments), fixed-order lane partials + butterfly, segments summed
// in order: deterministic.
// trans = 0: y_i = alpha sum_j A[i, j] x_j + beta y_i, one thread per row (workgroups of one
// wave), GVN_U partial sums over the columns (rows read coalesced across the wave).
// loads in flight per lane (gemv_t: columns of a few thousand rows are common, and 8 serve them
// best) or per thread (gemv_n), and their partial sums' fixed order (pairwise)
#define GVT_U 8
#define GVN_U 16
template <typename T, int U>
__device__ __forceinline__ T gv_sum(const T (&p)[U]) {
    T s[U];
#pragma unroll
    for (int u = 0; u < U; u++) s[u] = p[u];
#pragma unroll
    for (int h = U / 2; h > 0; h /= 2)
#pragma unroll
        for (int u = 0; u < h; u++) s[u] = s[u] + s[u + h];
    return s[0];
}
// waves of a workgroup per output of the transposed form: a column of a thousand rows or more is
// split over 4 waves (each of its segments then takes one or two rounds of GVT_U loads per lane,
// not several in sequence), of 512 or more over 2; by the item's own rows, so that a product is
// summed the same way in any batch
__host__ __device__ inline int gvt_wpc(int M) { return M >= 1024 ? 4 : M >= 512 ? 2 : 1; }
template <typename TA, typename TX, typename TY, typename TACC>
__global__ __launch_bounds__(256) void k_gemv_t(const GemvItem *__restrict__ items, const int *__restrict__ pre, int nit,
                                                const TA *__restrict__ A, const TX *__restrict__ x, TY *__restrict__ y, TACC alpha,
                                                TACC beta) {
    const int q = flat_item(pre, nit, blockIdx.x);
    const GemvItem it = items[q];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wpc = gvt_wpc(it.M), seg = wave % wpc;                      // uniform in the workgroup
    const int j = (blockIdx.x - pre[q]) * (4 / wpc) + wave / wpc;
    __shared__ TACC part[4];
    const bool live = j < it.N;
    TACC s = 0;
    if (live) {
        const TA *col = A + it.a + (int64_t)j * it.lda;
        const TX *xv = x + it.x;
        const int r0 = it.tri == 1 ? j : 0, r1 = it.tri == 2 ? j + 1 : it.M;     // triangular A: only its nonzero rows
        // this wave's segment of the rows (segments of a multiple of 64 rows)
        const int per = (((r1 - r0) + wpc - 1) / wpc + 63) & ~63;
        const int i0 = min(r1, r0 + seg * per), i1 = min(r1, i0 + per);
        // GVT_U loads of the segment in flight per lane (independent partial sums, added in a
        // fixed order)
        TACC p[GVT_U];
#pragma unroll
        for (int u = 0; u < GVT_U; u++) p[u] = 0;
        int i = i0 + lane;
        for (; i + 64 * (GVT_U - 1) < i1; i += 64 * GVT_U) {
#pragma unroll
            for (int u = 0; u < GVT_U; u++) p[u] += (TACC)col[i + 64 * u] * (TACC)xv[i + 64 * u];
        }
        if (i < i1) {       // the tail (< 64 GVT_U rows) as one batch, addresses clamped to i1 - 1 and masked
            TA a[GVT_U];
            TX b[GVT_U];
#pragma unroll
            for (int u = 0; u < GVT_U; u++) { const int ii = min(i + 64 * u, i1 - 1); a[u] = col[ii]; b[u] = xv[ii]; }
#pragma unroll
            for (int u = 0; u < GVT_U; u++) p[u] += i + 64 * u < i1 ? (TACC)a[u] * (TACC)b[u] : (TACC)0;
        }
        s = gv_sum<TACC, GVT_U>(p);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    }
    if (wpc > 1) {          // the segments' sums in order, by the column's first wave
        if (lane == 0) part[wave] = s;
        __syncthreads();
        if (seg == 0)
            for (int u = 1; u < wpc; u++) s += part[wave + u];
    }
    if (live && seg == 0 && lane == 0) {
        TY *yp = y + it.y + j;
        TACC v = alpha * s;
        if (beta != (TACC)0) v += beta * (TACC)*yp;
        *yp = (TY)v;
    }
}
// the workgroups of the transposed form for an item, and its algorithmic bytes and flops (a
// triangular matrix: its nonzero half)
static inline int64_t gvt_blocks(const GemvItem &g) { const int cpb = 4 / gvt_wpc(g.M); return (g.N + cpb - 1) / cpb; }
static inline double gv_elems(const GemvItem &g) { return g.tri ? 0.5 * (double)g.N * (g.N + 1) : (double)g.M * g.N; }
template <typename TA, typename TX, typename TY, typename TACC>
__global__ __launch_bounds__(64) void k_gemv_n(const GemvItem *__restrict__ items, const int *__restrict__ pre, int nit,
                                                const TA *__restrict__ A, const TX *__restrict__ x, TY *__restrict__ y, TACC alpha,
                                                TACC beta) {
    const int q = flat_item(pre, nit, blockIdx.x);
    const GemvItem it = items[q];
    const int i = (blockIdx.x - pre[q]) * 64 + threadIdx.x;
    if (i >= it.M) return;
    const TA *row = A + it.a + i;
    const TX *xv = x + it.x;
    // GVN_U columns' loads in flight per thread (independent partial sums over j = GVN_U q + u,
    // added in a fixed order)
    TACC p[GVN_U];
#pragma unroll
    for (int u = 0; u < GVN_U; u++) p[u] = 0;
    int j = 0;
    for (; j + GVN_U - 1 < it.N; j += GVN_U) {
        TA a[GVN_U];
#pragma unroll
        for (int u = 0; u < GVN_U; u++) a[u] = row[(int64_t)(j + u) * it.lda];
#pragma unroll
        for (int u = 0; u < GVN_U; u++) p[u] += (TACC)a[u] * (TACC)xv[j + u];
    }
    for (; j < it.N; j++) p[0] += (TACC)row[(int64_t)j * it.lda] * (TACC)xv[j];
    TACC s = gv_sum<TACC, GVN_U>(p);
    TY *yp = y + it.y + i;
    TACC v = alpha * s;
    if (beta != (TACC)0) v += beta * (TACC)*yp;
    *yp = (TY)v;
}
template <typename TA, typename TX, typename TY, typename TACC>
static int gemv_run(int trans, TACC alpha, TACC beta, const TA *A, const TX *x, TY *y, const GemvItem *it, int n_it) {
    if (n_it <= 0) return 0;
    std::vector<int64_t> cnt(n_it);
    for (int q = 0; q < n_it; q++) cnt[q] = trans ? gvt_blocks(it[q]) : (it[q].M + 63) / 64;
    const GemvItem *d = nullptr;
    const int *pre = nullptr;
    const int64_t nb = flat_stage_items(it, sizeof(GemvItem) * n_it, cnt, (const void **)&d, &pre);
    if (nb < 0) return -1;
    if (nb == 0) return 0;
    const dim3 grid((unsigned)nb);
    const int tok = bk_prof_begin();
    if (trans) hipLaunchKernelGGL((k_gemv_t<TA, TX, TY, TACC>), grid, dim3(256), 0, g_bk_stream, d, pre, n_it, A, x, y, alpha, beta);
    else hipLaunchKernelGGL((k_gemv_n<TA, TX, TY, TACC>), grid, dim3(64), 0, g_bk_stream, d, pre, n_it, A, x, y, alpha, beta);
    if (tok >= 0) {
        double fl = 0, by = 0;
        for (int q = 0; q < n_it; q++) { fl += 2.0 * gv_elems(it[q]); by += (double)sizeof(TA) * gv_elems(it[q]) + sizeof(TX) * (double)(trans ? it[q].M : it[q].N) + sizeof(TY) * (double)(trans ? it[q].N : it[q].M) * (beta != (TACC)0 ? 2 : 1); }
        bk_prof_end(tok, 9, fl, by);
    }
    return BK_LAUNCH_CHECK("k_gemv");
}
extern "C" int bk_gemv_f64(int trans, double alpha, double beta, const double *A, const double *x, double *y, const GemvItem *it, int n_it) {
    return gemv_run<double, double, double, double>(trans, alpha, beta, A, x, y, it, n_it);
}
// nstage products y = A' x in sequence from one staging (stage s: the items it[s n_it ..
// s n_it + n_it), offsets into the heap H): the three factors of the Lanczos operator per step
extern "C" int bk_gemv_chain_f64(double *H, const GemvItem *it, int n_it, int nstage) {
    if (n_it <= 0 || nstage <= 0) return 0;
    std::vector<char> buf;
    std::vector<size_t> io(nstage), po(nstage);
    std::vector<int64_t> nb(nstage);
    for (int st = 0; st < nstage; st++) {
        io[st] = buf_align(buf);
        buf.insert(buf.end(), (const char *)(it + (size_t)st * n_it), (const char *)(it + (size_t)(st + 1) * n_it));
        po[st] = buf_align(buf);
        buf.resize(po[st] + sizeof(int) * (n_it + 1));
        int64_t tot = 0;
        for (int q = 0; q < n_it; q++) {
            ((int *)(buf.data() + po[st]))[q] = (int)tot;
            tot += gvt_blocks(it[(size_t)st * n_it + q]);
        }
        ((int *)(buf.data() + po[st]))[n_it] = (int)tot;
        if (tot > 0x7fffffff) return -1;
        nb[st] = tot;
    }
    const char *d = (const char *)bk_stage(buf.data(), buf.size());
    if (!d) return -1;
    for (int st = 0; st < nstage; st++) {
        if (nb[st] == 0) continue;
        const int tok = bk_prof_begin();
        hipLaunchKernelGGL((k_gemv_t<double, double, double, double>), dim3((unsigned)nb[st]), dim3(256), 0, g_bk_stream,
                           (const GemvItem *)(d + io[st]), (const int *)(d + po[st]), n_it, H, H, H, 1.0, 0.0);
        if (tok >= 0) {
            double fl = 0, by = 0;
            for (int q = 0; q < n_it; q++) { const GemvItem &g = it[(size_t)st * n_it + q]; fl += 2.0 * gv_elems(g); by += 8.0 * gv_elems(g) + 8.0 * (g.M + g.N); }
            bk_prof_end(tok, 9, fl, by);
        }
        if (BK_LAUNCH_CHECK("k_gemv")) return -1;
    }
    return 0;
}
extern "C" int bk_gemv_f32(int trans, float alpha, float beta, const float *A, const float *x, float *y, const GemvItem *it, int n_it) {
    return gemv_run<float, float, float, float>(trans, alpha, beta, A, x, y, it, n_it);
}
extern "C" int bk_gemv_mixed(int trans, double alpha, double beta, const float *A, const double *x, double *y, const GemvItem *it, int n_it) {
    return gemv_run<float, double, double, double>(trans, alpha, beta, A, x, y, it, n_it);
}

// ---- symmetric matrix-vector product from the lower triangle ----------------------------------
// The lower 64 x 64 tiles (I, J), I >= J, of an item are enumerated row-major over the lower
// triangle and dealt to workgroups in chunks of SYMV_CH consecutive tiles. A workgroup stages a
// tile in LDS (coalesced column reads; the next tile's loads are in flight meanwhile) and all 256
// threads form quarter sums: thread (i, h) the row part over columns 16h..16h+15 of row i and the
// column part over rows 16h..16h+15 of column i. The quarters are added in order and the tile
// writes r(i) = sum_c A(64I + i, 64J + c) x(64J + c) to slot (I, J) and, for I > J,
// c(j) = sum_r A(64I + r, 64J + j) x(64I + r) to slot (J, I). A diagonal tile reads its lower
// triangle only and mirrors it.
#define SYMV_CH 4
__device__ __forceinline__ void symv_tiles(const SymvItem it, int64_t wg, const double *__restrict__ A, const double *__restrict__ X,
                                           double *__restrict__ P) {
    const int n = it.n, nT = (n + 63) / 64, tid = threadIdx.x, li = tid & 63, h = tid >> 6;
    const int64_t ntile = (int64_t)nT * (nT + 1) / 2, t0 = wg * SYMV_CH;
    const int64_t t1 = t0 + SYMV_CH < ntile ? t0 + SYMV_CH : ntile;
    __shared__ double Ts[64][65];    // Ts[c][r] = A(64 I + r, 64 J + c)
    __shared__ double xs[2][64];     // x of the column block J, of the row block I
    __shared__ double red[2][4][64]; // quarter sums of the row and the column parts
    const double *a = A + it.a;
    double *p = P + it.p;
    double ra[16];
    auto load = [&](int64_t t) {
        int I, J;
        tri_index(t, I, J);
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const int e = tid + 256 * u, r = e & 63, c = e >> 6, gi = 64 * I + r, gj = 64 * J + c;
            ra[u] = (gi < n && gj < n) ? a[gi + (int64_t)gj * it.ld] : 0.0;
        }
    };
    if (t0 < t1) load(t0);
    for (int64_t t = t0; t < t1; t++) {
        int I, J;
        tri_index(t, I, J);
#pragma unroll
        for (int u = 0; u < 16; u++) { const int e = tid + 256 * u; Ts[e >> 6][e & 63] = ra[u]; }
        if (tid < 64) xs[0][tid] = 64 * J + tid < n ? X[it.x + 64 * J + tid] : 0.0;
        else if (tid < 128) xs[1][tid - 64] = 64 * I + tid - 64 < n ? X[it.x + 64 * I + tid - 64] : 0.0;
        __syncthreads();
        if (t + 1 < t1) load(t + 1);           // the next tile's loads stay in flight during the sums
        double sr = 0, sc = 0;
        if (I > J) {
#pragma unroll
            for (int k = 0; k < 16; k++) {
                const int c = 16 * h + k;
                sr += Ts[c][li] * xs[0][c];
                sc += Ts[li][c] * xs[1][c];
            }
        } else {
#pragma unroll
            for (int k = 0; k < 16; k++) {
                const int c = 16 * h + k;
                sr += (c <= li ? Ts[c][li] : Ts[li][c]) * xs[0][c];
            }
        }
        red[0][h][li] = sr;
        red[1][h][li] = sc;
        __syncthreads();
        if (tid < 64) p[((int64_t)I * nT + J) * 64 + tid] = (red[0][0][tid] + red[0][1][tid]) + (red[0][2][tid] + red[0][3][tid]);
        else if (tid < 128 && I > J) {
            const int j = tid - 64;
            p[((int64_t)J * nT + I) * 64 + j] = (red[1][0][j] + red[1][1][j]) + (red[1][2][j] + red[1][3][j]);
        }
        __syncthreads();
    }
}
__global__ __launch_bounds__(256) void k_symv_tiles(const SymvItem *__restrict__ items, const int *__restrict__ pre, int nit,
                                                    const double *__restrict__ A, const double *__restrict__ X, double *__restrict__ P) {
    const int q = flat_item(pre, nit, blockIdx.x);
    symv_tiles(items[q], blockIdx.x - pre[q], A, X, P);
}
// y(64 a + i) = alpha sum_b P(a, b)(i) + beta y(64 a + i): one 64-thread workgroup per row block,
// the slots summed in the order b = 0 .. T - 1
__device__ __forceinline__ void symv_reduce(const SymvItem it, int a, const double *__restrict__ P, double *__restrict__ Y, double alpha,
                                            double beta) {
    const int n = it.n, nT = (n + 63) / 64, i = threadIdx.x, row = 64 * a + i;
    if (row >= n) return;
    const double *p = P + it.p + (int64_t)a * nT * 64 + i;
    // the slots in order, loaded 8 at a time (index clamped, value masked) so that the loads of a
    // batch are in flight together; the sum's order is unchanged
    double s = 0;
    for (int b0 = 0; b0 < nT; b0 += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = p[(int64_t)min(b0 + u, nT - 1) * 64];
#pragma unroll
        for (int u = 0; u < 8; u++) if (b0 + u < nT) s += v[u];
    }
    double *y = Y + it.y + row;
    *y = beta != 0.0 ? alpha * s + beta * *y : alpha * s;
}
__global__ __launch_bounds__(64) void k_symv_reduce(const SymvItem *__restrict__ items, const int *__restrict__ pre, int nit,
                                                    const double *__restrict__ P, double *__restrict__ Y, double alpha, double beta) {
    const int q = flat_item(pre, nit, blockIdx.x);
    symv_reduce(items[q], blockIdx.x - pre[q], P, Y, alpha, beta);
}
extern "C" int bk_symv_f64(double alpha, double beta, const double *A, const double *x, double *y, double *p, const SymvItem *it, int n_it) {
    if (n_it <= 0) return 0;
    std::vector<int64_t> ct(n_it), cr(n_it);
    double by = 0;
    for (int q = 0; q < n_it; q++) {
        const int64_t nT = (it[q].n + 63) / 64;
        ct[q] = it[q].n > 0 ? (nT * (nT + 1) / 2 + SYMV_CH - 1) / SYMV_CH : 0;
        cr[q] = it[q].n > 0 ? nT : 0;
        by += 8.0 * ((double)it[q].n * (it[q].n + 1) / 2 + 3.0 * it[q].n);
    }
    const SymvItem *d;
    const int *pt;
    const int64_t nt = flat_stage_items(it, sizeof(SymvItem) * n_it, ct, (const void **)&d, &pt);
    if (nt < 0) return -1;
    if (nt == 0) return 0;
    const int *pr;
    const int64_t nr = flat_stage(cr, &pr);
    if (nr <= 0) return -1;
    const int tok = bk_prof_begin();
    hipLaunchKernelGGL(k_symv_tiles, dim3((unsigned)nt), dim3(256), 0, g_bk_stream, d, pt, n_it, A, x, p);
    hipLaunchKernelGGL(k_symv_reduce, dim3((unsigned)nr), dim3(64), 0, g_bk_stream, d, pr, n_it, (const double *)p, y, alpha, beta);
    if (tok >= 0) {
        double fl = 0;
        for (int q = 0; q < n_it; q++) fl += 2.0 * it[q].n * it[q].n;
        bk_prof_end(tok, 15, fl, by);
    }
    return BK_LAUNCH_CHECK("symv");
}

// ---- the symmetric product without device allocation (bk_symv_f64_noalloc) --------------------
// For callers that cannot allocate device memory (code under graph capture, tight serving loops):
// the same two kernels as bk_symv_f64, with the partial-sum slots in the file-scope array
// `scratch` instead of a caller workspace, and the item descriptors and workgroup prefixes
// passed by value as kernel arguments instead of staged device copies, so that a call makes no
// allocation and no copy. Every GPU the library runs on holds its own copy of the array until
// the process ends; the slots a call uses are all written by its first kernel before its second
// reads them, so nothing needs clearing, but calls that overlap on one GPU race on the array.
#define SYMV_NA_SCRATCH (4 << 20)                  // doubles: 32 MiB of partial-sum slots
#define SYMV_NA_ITEMS 16                           // items per call (their descriptors are kernel arguments)
static __device__ __align__(256) double scratch[SYMV_NA_SCRATCH];
struct SymvBatch {
    SymvItem it[SYMV_NA_ITEMS];
    int pre[SYMV_NA_ITEMS + 1];                    // workgroups of item q: [pre[q], pre[q + 1])
    int nit;
};
__global__ __launch_bounds__(256) void k_symv_tiles_na(const SymvBatch b, const double *__restrict__ A, const double *__restrict__ X,
                                                       double *__restrict__ P) {
    int q = 0;
    while (q + 1 < b.nit && b.pre[q + 1] <= (int)blockIdx.x) q++;
    symv_tiles(b.it[q], blockIdx.x - b.pre[q], A, X, P);
}
__global__ __launch_bounds__(64) void k_symv_reduce_na(const SymvBatch b, const double *__restrict__ P, double *__restrict__ Y, double alpha,
                                                       double beta) {
    int q = 0;
    while (q + 1 < b.nit && b.pre[q + 1] <= (int)blockIdx.x) q++;
    symv_reduce(b.it[q], blockIdx.x - b.pre[q], P, Y, alpha, beta);
}
// the device address of `scratch` for each device the library has run on (the array's name in
// host code is a host-side stand-in, never a device address)
static double *g_symv_scratch[64];
extern "C" int bk_symv_scratch_init(void) {
    int dev = 0;
    if (bk_check(hipGetDevice(&dev), "hipGetDevice") || dev < 0 || dev >= 64) return -1;
    if (g_symv_scratch[dev]) return 0;
    void *ptr = nullptr;
    if (bk_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress scratch") || !ptr) return -1;
    g_symv_scratch[dev] = (double *)ptr;
    return 0;
}
extern "C" int bk_symv_f64_noalloc(double alpha, double beta, const double *A, const double *x, double *y, const SymvItem *it, int n_it) {
    if (n_it <= 0) return 0;
    if (!it || n_it > SYMV_NA_ITEMS) return bk_fail("bk_symv_f64_noalloc: at most 16 items per call");
    int dev = 0;
    if (bk_check(hipGetDevice(&dev), "hipGetDevice") || dev < 0 || dev >= 64) return -1;
    if (!g_symv_scratch[dev] && bk_symv_scratch_init()) return -1;
    SymvBatch bt, br;
    memset(&bt, 0, sizeof(bt));
    int64_t slots = 0, nt = 0, nr = 0;
    for (int q = 0; q < n_it; q++) {
        const int n = it[q].n;
        if (n < 0) return bk_fail("bk_symv_f64_noalloc: an item of negative order");
        const int64_t T = (n + 63) / 64;
        bt.it[q] = it[q];
        bt.it[q].p = slots;                          // the item's slots within scratch
        slots += T * T * 64;
        if (slots > SYMV_NA_SCRATCH)
            return bk_fail("bk_symv_f64_noalloc: the items need more than the 32 MiB of partial-sum slots (one item of order <= 16384)");
        bt.pre[q] = (int)nt;
        nt += n > 0 ? (T * (T + 1) / 2 + SYMV_CH - 1) / SYMV_CH : 0;
    }
    bt.pre[n_it] = (int)nt;
    bt.nit = n_it;
    br = bt;
    for (int q = 0; q < n_it; q++) { br.pre[q] = (int)nr; nr += it[q].n > 0 ? (it[q].n + 63) / 64 : 0; }
    br.pre[n_it] = (int)nr;
    if (nt == 0) return 0;
    double *P = g_symv_scratch[dev];
    hipLaunchKernelGGL(k_symv_tiles_na, dim3((unsigned)nt), dim3(256), 0, g_bk_stream, bt, A, x, P);
    hipLaunchKernelGGL(k_symv_reduce_na, dim3((unsigned)nr), dim3(64), 0, g_bk_stream, br, (const double *)P, y, alpha, beta);
    return BK_LAUNCH_CHECK("symv noalloc");
}
