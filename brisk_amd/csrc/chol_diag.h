// chol_diag.h - factor and invert one 64 x 64 diagonal block in a workgroup of 512 threads, and
// optionally multiply a 64-row panel tile by the inverse's transpose (included by k_dense.hip).
//
// Layout: thread (lane r, wave w), eight waves (two per SIMD, whose dependency chains
// interleave), holds row r of the block at the 8 columns c = 8u + w in registers. The factor runs
// in 16 groups of four columns with one barrier each: the four waves holding the group's columns
// publish them to LDS, every thread factors the 4 x 4 diagonal piece itself
// (reciprocal square roots by v_rsq plus Newton steps, no divisions), forms the four values
// l(r, j0..j0+3) of its own row and writes them to its wave's private copy in LDS, from which
// the wave reads the rows it needs for the rank-4 update of its own columns (an in-order
// exchange within the wave: no second barrier).
//
// The inverse X = L^{-1} is built by recursive doubling from the 4 x 4 diagonal blocks: for
// block size S every pair of S x S diagonal blocks A, C with the block B = L[C, A] between them
// gets X[C, A] = -X[C, C] (B X[A, A]). Sizes 4 and 8 run on the vector units with fully
// unrolled inner products, sizes 16 and 32 on MFMA with one 16 x 16 tile per wave.
#pragma once

template <typename T> struct DiagMF;
// out(g, r): the row of the 16 x 16 MFMA result held by register r of a lane in lane group
// g = lane / 16 (the column is lane % 16)
template <> struct DiagMF<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mma(double a, double b, v4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    static __device__ __forceinline__ int out(int g, int r) { return g + 4 * r; }
    // 1 / sqrt(x): the hardware estimate and two Newton steps y <- y (3/2 - x y^2 / 2)
    static __device__ __forceinline__ double rsq(double x) {
        double y = __builtin_amdgcn_rsq(x);
        const double h = 0.5 * x;
#pragma unroll
        for (int s = 0; s < 2; s++) { const double hy = h * y; y = y * __builtin_fma(-hy, y, 1.5); }
        return y;
    }
};
template <> struct DiagMF<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mma(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    static __device__ __forceinline__ int out(int g, int r) { return 4 * g + r; }
    static __device__ __forceinline__ float rsq(float x) {
        const float y = __builtin_amdgcn_rsqf(x), hy = 0.5f * x * y;
        return y * __builtin_fmaf(-hy, y, 1.5f);
    }
};

// LDS written by this wave is visible to the wave's later reads (LDS executes one wave's
// instructions in order; the fences keep the compiler from reordering across this point)
__device__ __forceinline__ void diag_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// one doubling level for S <= 8 (32 S elements per phase, one per thread); the upper parts of L
// and X are zero in LDS, so the inner products run over the whole block without bounds
template <typename T, int S>
__device__ __forceinline__ void diag_inv_simt(T (*Ls)[65], T (*Xs)[65], T (*Ts)[65], int tid) {
    constexpr int TOT = 32 * S;
    const int pr = tid / (S * S), e = tid % (S * S), r = e % S, c = e / S;
    const int p = pr * 2 * S, row = p + S + r, cl = p + c;
    if (tid < TOT) {
        T a = 0;
#pragma unroll
        for (int k = 0; k < S; k++) a += Ls[p + k][row] * Xs[cl][p + k];
        Ts[cl][row] = a;
    }
    __syncthreads();
    if (tid < TOT) {
        T a = 0;
#pragma unroll
        for (int k = 0; k < S; k++) a += Xs[p + S + k][row] * Ts[cl][p + S + k];
        Xs[cl][row] = -a;
    }
    __syncthreads();
}
// one doubling level for S = 16, 32 on MFMA: wave q computes one 16 x 16 tile (pairs x tiles =
// 2 x 1 for S = 16, 1 x 4 for S = 32). Operands: a = left(row, k), b = right(k, col), so
// mma(b, a) accumulates the tile with its row index in lane % 16.
template <typename T, int S>
__device__ __forceinline__ void diag_inv_mfma(T (*Ls)[65], T (*Xs)[65], T (*Ts)[65], int tid) {
    typedef typename DiagMF<T>::v4 v4;
    constexpr int NT = S / 16, JOBS = (32 / S) * NT * NT;
    const int lane = tid & 63, wave = tid >> 6, l16 = lane & 15, g4 = lane >> 4;
    const int pr = wave / (NT * NT), tt = wave % (NT * NT), ti = tt % NT, tj = tt / NT;
    const int p = pr * 2 * S, row0 = p + S + 16 * ti, col0 = p + 16 * tj;
    if (wave < JOBS) {
        v4 acc = {0, 0, 0, 0};
#pragma unroll
        for (int kk = 0; kk < S; kk += 4) {
            const int k = p + kk + g4;
            acc = DiagMF<T>::mma(Xs[col0 + l16][k], Ls[k][row0 + l16], acc);      // (L[C, A] X[A, A])
        }
#pragma unroll
        for (int r = 0; r < 4; r++) Ts[col0 + DiagMF<T>::out(g4, r)][row0 + l16] = acc[r];
    }
    __syncthreads();
    if (wave < JOBS) {
        v4 acc = {0, 0, 0, 0};
#pragma unroll
        for (int kk = 0; kk < S; kk += 4) {
            const int k = p + S + kk + g4;
            acc = DiagMF<T>::mma(Ts[col0 + l16][k], Xs[k][row0 + l16], acc);      // X[C, C] T
        }
#pragma unroll
        for (int r = 0; r < 4; r++) Xs[col0 + DiagMF<T>::out(g4, r)][row0 + l16] = -acc[r];
    }
    __syncthreads();
}

// Factor the nb x nb block at Ab (leading dimension n, lower triangle read) as L L', write L
// (strict upper part zeroed) to Lo (leading dimension ldo; Ab itself for the factor in place, or
// nullptr) and L^{-1} (64 x 64, zero outside the nb x nb lower part) to Di (or nullptr). Rows and
// columns beyond nb are factored as the identity. With PANEL, the 64-column tile at P (leading
// dimension ldp, prow <= 64 rows) is then multiplied in place by L^{-T} on MFMA: the panel step
// of a blocked factorization, L_I = A_I L^{-T}. With K > 0 (nb = 64) the block and the tile are
// first brought up to date with K earlier columns of the factor, on MFMA: the block takes
// -Lk Lk' and the tile -Lp Lk', Lk (64 x K) and Lp (prow x K) at leading dimension n (the
// left-looking update of a step inside a panel). Returns 0, or the 1-based column whose pivot is
// not positive (then nothing is written); pm_out is the smallest squared pivot.
template <typename T, bool PANEL = false>
__device__ int diag_factor_inverse(const T *Ab, int n, int nb, T *Lo, int ldo, T *Di, T &pm_out, T *P = nullptr, int ldp = 0,
                                   int prow = 0, const T *Lk = nullptr, const T *Lp = nullptr, int K = 0) {
    __shared__ T Ls[64][65];       // Ls[c][r] = L(r, c)
    __shared__ T Xs[64][65];       // Xs[c][r] = X(r, c), X = L^{-1}
    __shared__ T Ts[64][65];
    __shared__ T colb[2][4][64];   // the group's four columns (double-buffered by group parity)
    __shared__ T lb[8][64][4];     // per wave: l(r, j0..j0+3) of every row r
    __shared__ T rdg[64];          // 1 / L(j, j)
    typedef typename DiagMF<T>::v4 v4;
    const int tid = threadIdx.x, r = tid & 63, w = tid >> 6;
    // the MFMA tiles of wave w: rows 16 (w & 3) .., columns 16 (2 (w >> 2) + h) .., h = 0, 1
    const int l16 = r & 15, g4 = r >> 4, row0 = 16 * (w & 3);
    // thread (r, w) holds row r at the 8 columns c = 8 u + w; group g's columns 4 g + t are held
    // by the waves 4 (g & 1) + t, at u = g / 2
    T v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
        const int c = 8 * u + w;
        v[u] = (r < nb && c < nb) ? (r >= c ? Ab[r + (int64_t)c * n] : (T)0) : (r == c ? (T)1 : (T)0);
    }
    // the panel tile's rows (PANEL), loaded now so that their latency overlaps the factorization
    T pt[8];
#pragma unroll
    for (int u = 0; u < 8; u++) pt[u] = PANEL && r < prow ? P[r + (int64_t)(8 * u + w) * ldp] : (T)0;
    v4 pu[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};       // the tile's update Lp Lk' (PANEL, K > 0)
    if (K > 0) {
        // 64 columns of Lk at a time into Ls (Ls[k][r] = Lk(r, k)) and of Lp into Xs, every load
        // of a chunk in flight at once; then Ts[c][r] = (Lk Lk')(r, c) on MFMA with operands
        // a = Lk(column, k), b = Lk(row, k) (Lp(row, k) for the tile): the result's row index is
        // lane % 16
        v4 acc[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
        for (int kc = 0; kc < K; kc += 64) {
            T la[8], lp[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int64_t e = r + (int64_t)(kc + 8 * u + w) * n;
                la[u] = Lk[e];
                if (PANEL) lp[u] = r < prow ? Lp[e] : (T)0;
            }
            if (kc > 0) __syncthreads();
#pragma unroll
            for (int u = 0; u < 8; u++) {
                Ls[8 * u + w][r] = la[u];
                if (PANEL) Xs[8 * u + w][r] = lp[u];
            }
            __syncthreads();
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int col0 = 16 * (2 * (w >> 2) + h);
#pragma unroll
                for (int kk = 0; kk < 64; kk += 4) {
                    const T a = Ls[kk + g4][col0 + l16];
                    acc[h] = DiagMF<T>::mma(a, Ls[kk + g4][row0 + l16], acc[h]);
                    if (PANEL) pu[h] = DiagMF<T>::mma(a, Xs[kk + g4][row0 + l16], pu[h]);
                }
            }
        }
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int col0 = 16 * (2 * (w >> 2) + h);
#pragma unroll
            for (int q = 0; q < 4; q++) Ts[col0 + DiagMF<T>::out(g4, q)][row0 + l16] = acc[h][q];
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int c = 8 * u + w;
            if (r >= c) v[u] -= Ts[c][r];
        }
    }
    T pm = (T)1e30;
    int fail = 0;
    const int half = w >> 2, wq = w & 3;
#pragma unroll
    for (int g = 0; g < 16; g++) {
        const int j0 = 4 * g, bf = g & 1, hu = g >> 1;
        if (half == (g & 1)) colb[bf][wq][r] = v[hu];
        __syncthreads();
        T D[4][4], cr[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            cr[t] = colb[bf][t][r];
#pragma unroll
            for (int s = t; s < 4; s++) D[s][t] = colb[bf][t][j0 + s];
        }
        // the 4 x 4 diagonal piece (the same arithmetic in every thread, so every thread takes
        // the same branch on a failed pivot)
        T L4[4][4], y[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            T p = D[t][t];
#pragma unroll
            for (int q = 0; q < t; q++) p -= L4[t][q] * L4[t][q];
            if (!(p > (T)0) && !fail) fail = j0 + t + 1;
            if (j0 + t < nb) pm = p < pm ? p : pm;
            y[t] = DiagMF<T>::rsq(p);
            L4[t][t] = p * y[t];
#pragma unroll
            for (int s = t + 1; s < 4; s++) {
                T a = D[s][t];
#pragma unroll
                for (int q = 0; q < t; q++) a -= L4[s][q] * L4[t][q];
                L4[s][t] = a * y[t];
            }
        }
        if (fail) break;
        if (w == 0 && r < 4) rdg[j0 + r] = r == 0 ? y[0] : r == 1 ? y[1] : r == 2 ? y[2] : y[3];
        // this row's l(r, j0..j0+3); the piece's own rows take L4, rows above the columns 0
        T l[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            T a = cr[t];
#pragma unroll
            for (int q = 0; q < t; q++) a -= l[q] * L4[t][q];
            l[t] = a * y[t];
        }
#pragma unroll
        for (int s = 0; s < 4; s++)
            if (r == j0 + s) {
#pragma unroll
                for (int t = 0; t < 4; t++) l[t] = t <= s ? L4[s][t] : (T)0;
            }
#pragma unroll
        for (int t = 0; t < 4; t++) l[t] = r >= j0 + t ? l[t] : (T)0;
#pragma unroll
        for (int t = 0; t < 4; t++) lb[w][r][t] = l[t];
        diag_wave_sync();
        // the group's columns are final; the other half's column at u = g / 2 is a later one when
        // g is even (column 4 g + 4 + wq), an earlier one when g is odd
        if (half == (g & 1)) v[hu] = wq == 0 ? l[0] : wq == 1 ? l[1] : wq == 2 ? l[2] : l[3];
        else if (!(g & 1)) {
            const int k = 8 * hu + w;
            T a = v[hu];
#pragma unroll
            for (int t = 0; t < 4; t++) a -= l[t] * lb[w][k][t];
            v[hu] = a;
        }
        // rank-4 update of this thread's later columns k: a(r, k) -= sum_t l(r, t) l(k, t)
#pragma unroll
        for (int h = hu + 1; h < 8; h++) {
            const int k = 8 * h + w;
            T a = v[h];
#pragma unroll
            for (int t = 0; t < 4; t++) a -= l[t] * lb[w][k][t];
            v[h] = a;
        }
    }
    __syncthreads();
    if (fail) return fail;
#pragma unroll
    for (int u = 0; u < 8; u++) {
        const int c = 8 * u + w;
        const T x = r >= c ? v[u] : (T)0;
        v[u] = x;
        Ls[c][r] = x;
        Xs[c][r] = (T)0;
    }
    __syncthreads();
    // the 4 x 4 diagonal blocks of X: thread (b, t) forms column t of block b by substitution
    if (tid < 64) {
        const int p = 4 * (tid >> 2), t = tid & 3;
        T x[4];
#pragma unroll
        for (int s = 0; s < 4; s++) {
            T a = 0;
#pragma unroll
            for (int q = 0; q < s; q++) a += Ls[p + q][p + s] * x[q];
            x[s] = s < t ? (T)0 : s == t ? rdg[p + s] : -rdg[p + s] * a;
        }
#pragma unroll
        for (int s = 0; s < 4; s++) Xs[p + t][p + s] = x[s];
    }
    __syncthreads();
    diag_inv_simt<T, 4>(Ls, Xs, Ts, tid);
    diag_inv_simt<T, 8>(Ls, Xs, Ts, tid);
    diag_inv_mfma<T, 16>(Ls, Xs, Ts, tid);
    diag_inv_mfma<T, 32>(Ls, Xs, Ts, tid);
#pragma unroll
    for (int u = 0; u < 8; u++) {
        const int c = 8 * u + w;
        if (Lo && r < nb && c < nb) Lo[r + (int64_t)c * ldo] = v[u];
        if (Di) Di[r + 64 * c] = (r < nb && c < nb && r >= c) ? Xs[c][r] : (T)0;
    }
    pm_out = pm;
    if (PANEL) {
        // the tile into Ts (Ts[c][r] = P(r, c), rows past prow zero), less its update, then
        // P := P X' with wave w forming its two 16 x 16 output tiles
#pragma unroll
        for (int u = 0; u < 8; u++) Ts[8 * u + w][r] = pt[u];
        __syncthreads();
        if (K > 0) {
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int col0 = 16 * (2 * (w >> 2) + h);
#pragma unroll
                for (int q = 0; q < 4; q++) Ts[col0 + DiagMF<T>::out(g4, q)][row0 + l16] -= pu[h][q];
            }
            __syncthreads();
        }
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int col0 = 16 * (2 * (w >> 2) + h);
            v4 acc = {0, 0, 0, 0};
#pragma unroll
            for (int kk = 0; kk < 64; kk += 4) acc = DiagMF<T>::mma(Xs[kk + g4][col0 + l16], Ts[kk + g4][row0 + l16], acc);
            if (row0 + l16 < prow) {
#pragma unroll
                for (int q = 0; q < 4; q++) P[(row0 + l16) + (int64_t)(col0 + DiagMF<T>::out(g4, q)) * ldp] = acc[q];
            }
        }
    }
    return 0;
}
