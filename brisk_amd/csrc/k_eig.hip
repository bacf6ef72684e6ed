// k_eig.hip - batched symmetric eigendecomposition A = V diag(w) V' by parallel cyclic Jacobi
// (CDNA4, wave64), one workgroup per matrix.
//
// A sweep visits every pair of indices once, in n - 1 rounds (n rounded up to even) of n / 2
// disjoint pairs (the circle ordering: index n - 1 fixed, the others rotating by one position a
// round). In a round every pair (p, q) first takes its rotation J = [c s; -s c] that zeroes
// A(p, q) (tau = (A(q,q) - A(p,p)) / (2 A(p,q)), t = sign(tau) / (|tau| + sqrt(1 + tau^2)),
// c = 1 / sqrt(1 + t^2), s = t c), then every 2 x 2 block of A between two pairs (a, b) is
// replaced by Ja' B Jb by the one thread that owns it (the blocks partition A, so the round needs
// no other ordering), and every row of V by V J. A rotation is skipped when |A(p, q)| is at most
// 2^-52 sqrt(|A(p,p) A(q,q)|); a sweep with no rotation ends the iteration. The rotations keep
// V orthogonal to rounding and give every eigenvalue with an absolute error of a few units of
// eps |A|, clusters included (the eigenvectors of a cluster come out orthogonal by construction).
//
// Matrices of order <= 64 are held in LDS (A and V, 66 KB, 256 threads); larger ones are
// transformed in place in the heap (A destroyed, 1024 threads). Eigenvalues are not sorted.
#include <hip/hip_runtime.h>
#include <math.h>
#include <vector>
#include "brisk_dev.h"
#include "rt_internal.h"
#include "flat.h"

#define EIG_LDS_N 64          // orders held in LDS
#define EIG_MAX_N 4096        // largest order (the pair arrays of the in-heap form: 2048 pairs)
#define EIG_SWEEPS 40

// the index at position k of round r of the circle ordering of N (even) indices
__device__ __forceinline__ int eig_pos(int k, int r, int N) { return k == N - 1 ? N - 1 : (k + r) % (N - 1); }

template <int NT, bool LDS>
__device__ void eig_jacobi(const EigItem it, double *__restrict__ H, double *As, double *Vs, int ld, double *cs, double *sn, int *pi,
                           int *qi) {
    const int n = it.n, N = n + (n & 1), np = N / 2, tid = threadIdx.x;
    __shared__ int rot;
    double *A = LDS ? As : H + it.a;
    double *V = LDS ? Vs : H + it.v;
    if (LDS) {
        for (int e = tid; e < n * n; e += NT) { const int r = e % n, c = e / n; As[r + c * ld] = H[it.a + e]; }
    }
    for (int e = tid; e < n * n; e += NT) { const int r = e % n, c = e / n; V[r + c * ld] = r == c ? 1.0 : 0.0; }
    const int64_t nblk = (int64_t)np * (np + 1) / 2;
    for (int sweep = 0; sweep < EIG_SWEEPS; sweep++) {
        if (tid == 0) rot = 0;
        __syncthreads();
        for (int r = 0; r < N - 1; r++) {
            for (int k = tid; k < np; k += NT) {
                int p = eig_pos(k, r, N), q = eig_pos(N - 1 - k, r, N);
                if (p > q) { const int t = p; p = q; q = t; }
                double c = 1.0, s = 0.0;
                if (q < n) {
                    const double app = A[p + p * ld], aqq = A[q + q * ld], apq = A[p + q * ld];
                    if (fabs(apq) > 1e-300 && fabs(apq) > 2.220446049250313e-16 * sqrt(fabs(app) * fabs(aqq))) {
                        const double tau = (aqq - app) / (2.0 * apq);
                        const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                        c = 1.0 / sqrt(1.0 + t * t);
                        s = t * c;
                        rot = 1;
                    }
                }
                pi[k] = p; qi[k] = q; cs[k] = c; sn[k] = s;
            }
            __syncthreads();
            // the 2 x 2 blocks (a, b), a >= b: B' = Ja' B Jb (entries at the odd order's missing index absent)
            for (int64_t t = tid; t < nblk; t += NT) {
                int a, b;
                tri_index(t, a, b);
                const double ca = cs[a], sa = sn[a], cb = cs[b], sb = sn[b];
                if (sa == 0.0 && sb == 0.0) continue;
                const int pa = pi[a], qa = qi[a], pb = pi[b], qb = qi[b];
                if (a == b) {
                    const double app = A[pa + pa * ld], aqq = A[qa + qa * ld], apq = A[pa + qa * ld];
                    A[pa + pa * ld] = ca * ca * app - 2.0 * ca * sa * apq + sa * sa * aqq;
                    A[qa + qa * ld] = sa * sa * app + 2.0 * ca * sa * apq + ca * ca * aqq;
                    A[pa + qa * ld] = 0.0;
                    A[qa + pa * ld] = 0.0;
                    continue;
                }
                const bool qav = qa < n, qbv = qb < n;
                const double b00 = A[pa + pb * ld], b01 = qbv ? A[pa + qb * ld] : 0.0;
                const double b10 = qav ? A[qa + pb * ld] : 0.0, b11 = qav && qbv ? A[qa + qb * ld] : 0.0;
                // B Jb
                const double e00 = b00 * cb - b01 * sb, e01 = b00 * sb + b01 * cb;
                const double e10 = b10 * cb - b11 * sb, e11 = b10 * sb + b11 * cb;
                // Ja' (B Jb)
                const double f00 = ca * e00 - sa * e10, f01 = ca * e01 - sa * e11;
                const double f10 = sa * e00 + ca * e10, f11 = sa * e01 + ca * e11;
                A[pa + pb * ld] = f00; A[pb + pa * ld] = f00;
                if (qbv) { A[pa + qb * ld] = f01; A[qb + pa * ld] = f01; }
                if (qav) { A[qa + pb * ld] = f10; A[pb + qa * ld] = f10; }
                if (qav && qbv) { A[qa + qb * ld] = f11; A[qb + qa * ld] = f11; }
            }
            // V J
            for (int64_t t = tid; t < (int64_t)n * np; t += NT) {
                const int i = (int)(t % n), k = (int)(t / n);
                const double s = sn[k];
                if (s == 0.0) continue;
                const double c = cs[k];
                const int p = pi[k], q = qi[k];
                const double vp = V[i + p * ld], vq = V[i + q * ld];
                V[i + p * ld] = c * vp - s * vq;
                V[i + q * ld] = s * vp + c * vq;
            }
            __syncthreads();
        }
        const int any = rot;
        __syncthreads();
        if (!any) break;
    }
    for (int i = tid; i < n; i += NT) H[it.w + i] = A[i + i * ld];
    if (LDS)
        for (int e = tid; e < n * n; e += NT) { const int r = e % n, c = e / n; H[it.v + e] = Vs[r + c * ld]; }
}

__global__ __launch_bounds__(256) void k_syev_lds(const EigItem *__restrict__ items, double *__restrict__ H) {
    __shared__ double As[EIG_LDS_N * (EIG_LDS_N + 1)], Vs[EIG_LDS_N * (EIG_LDS_N + 1)];
    __shared__ double cs[EIG_LDS_N / 2], sn[EIG_LDS_N / 2];
    __shared__ int pi[EIG_LDS_N / 2], qi[EIG_LDS_N / 2];
    eig_jacobi<256, true>(items[blockIdx.x], H, As, Vs, EIG_LDS_N + 1, cs, sn, pi, qi);
}
__global__ __launch_bounds__(1024) void k_syev_glb(const EigItem *__restrict__ items, double *__restrict__ H) {
    __shared__ double cs[EIG_MAX_N / 2], sn[EIG_MAX_N / 2];
    __shared__ int pi[EIG_MAX_N / 2], qi[EIG_MAX_N / 2];
    const EigItem it = items[blockIdx.x];
    eig_jacobi<1024, false>(it, H, nullptr, nullptr, it.n, cs, sn, pi, qi);
}

extern "C" int bk_syev_f64(const EigItem *it, int n_it, double *H) {
    if (n_it <= 0) return 0;
    std::vector<EigItem> small, large;
    for (int q = 0; q < n_it; q++) {
        if (it[q].n <= 0) continue;
        if (it[q].n > EIG_MAX_N) return bk_fail("syev: order above 4096");
        (it[q].n <= EIG_LDS_N ? small : large).push_back(it[q]);
    }
    const int tok = bk_prof_begin();
    if (!small.empty()) {
        const EigItem *d = (const EigItem *)bk_stage(small.data(), sizeof(EigItem) * small.size());
        if (!d) return -1;
        hipLaunchKernelGGL(k_syev_lds, dim3((unsigned)small.size()), dim3(256), 0, g_bk_stream, d, H);
        if (BK_LAUNCH_CHECK("k_syev_lds")) return -1;
    }
    if (!large.empty()) {
        const EigItem *d = (const EigItem *)bk_stage(large.data(), sizeof(EigItem) * large.size());
        if (!d) return -1;
        hipLaunchKernelGGL(k_syev_glb, dim3((unsigned)large.size()), dim3(1024), 0, g_bk_stream, d, H);
        if (BK_LAUNCH_CHECK("k_syev_glb")) return -1;
    }
    if (tok >= 0) {
        double fl = 0;
        for (int q = 0; q < n_it; q++) fl += 8.0 * (double)it[q].n * it[q].n * it[q].n;
        bk_prof_end(tok, 17, fl, 0);
    }
    return 0;
}

// ---- scalings of the Nesterov-Todd frame --------------------------------------------------------
// mode 0: A(i, j) *= u(j); 1: A(i, j) *= u(i); 2: A(i, j) *= 2 / (u(i) + u(j)); 3 (vector, n
// entries of w at a): u(i) = sqrt(w(i)), u(n + i) = w(i)^(-1/4), u(2n + i) = w(i)^(1/4), NaN where
// w(i) <= 0 or is not finite; 4 (vector): u(i) = w(i)^(1/4), 0 where w(i) <= 0. One thread per entry
// over a flattened grid.
__global__ __launch_bounds__(256) void k_nt_scale(const ScaleItem *__restrict__ items, const int *__restrict__ pre, int nit,
                                                  double *__restrict__ H) {
    const int q = flat_item(pre, nit, blockIdx.x);
    const ScaleItem it = items[q];
    const int64_t e = (int64_t)(blockIdx.x - pre[q]) * 256 + threadIdx.x, n = it.n;
    if (it.mode == 3) {
        if (e >= n) return;
        const double w = H[it.a + e];
        const bool ok = w > 0 && w < INFINITY;
        const double r = ok ? sqrt(w) : NAN, r4 = ok ? sqrt(r) : NAN;
        H[it.u + e] = r; H[it.u + n + e] = ok ? 1.0 / r4 : NAN; H[it.u + 2 * n + e] = r4;
        return;
    }
    if (it.mode == 4) {
        if (e >= n) return;
        const double w = H[it.a + e];
        H[it.u + e] = w > 0 && w < INFINITY ? sqrt(sqrt(w)) : 0.0;
        return;
    }
    if (e >= n * n) return;
    const int i = (int)(e % n), j = (int)(e / n);
    double *a = H + it.a + e;
    const double *u = H + it.u;
    if (it.mode == 0) *a *= u[j];
    else if (it.mode == 1) *a *= u[i];
    else *a *= 2.0 / (u[i] + u[j]);
}
extern "C" int bk_nt_scale(const ScaleItem *it, int n_it, double *H) {
    if (n_it <= 0) return 0;
    std::vector<int64_t> cnt(n_it);
    for (int q = 0; q < n_it; q++) cnt[q] = ((it[q].mode >= 3 ? (int64_t)it[q].n : (int64_t)it[q].n * it[q].n) + 255) / 256;
    const ScaleItem *d;
    const int *pre;
    const int64_t nb = flat_stage_items(it, sizeof(ScaleItem) * n_it, cnt, (const void **)&d, &pre);
    if (nb < 0) return -1;
    if (nb == 0) return 0;
    hipLaunchKernelGGL(k_nt_scale, dim3((unsigned)nb), dim3(256), 0, g_bk_stream, d, pre, n_it, H);
    return BK_LAUNCH_CHECK("k_nt_scale");
}
