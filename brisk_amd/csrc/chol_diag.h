// chol_diag.h - factor and invert one 64 x 64 diagonal block in a workgroup of 256 threads
// (included by k_dense.hip). Four pivots per barrier: every thread publishes one value of the
// block's four columns, then eliminates the 4 x 4 diagonal piece redundantly and applies the
// rank-4 update to the 16 entries it owns (row t % 64, columns t / 64 + 4u), in registers. The
// inverse of the factor is built by recursive doubling: for block size s every pair of s x s
// diagonal blocks A, C with the block B between them gets X[C, A] = -X[C, C] B X[A, A].
#pragma once
// one doubling level: T = L[C, A] X[A, A], then X[C, A] = -X[C, C] T, for the 64 / (2S) pairs
template <typename T, int S>
__device__ __forceinline__ void diag_inv_level(T (*Ls)[65], T (*Xs)[65], T (*Ts)[65], int tid) {
    constexpr int TOT = 32 * S;
    for (int e = tid; e < TOT; e += 256) {
        const int pr = e / (S * S), w = e % (S * S), r = w % S, c = w / S;
        const int p = pr * 2 * S, row = p + S + r, cl = p + c;
        T a0 = 0, a1 = 0;
        int k = c;
        for (; k + 1 < S; k += 2) { a0 += Ls[p + k][row] * Xs[cl][p + k]; a1 += Ls[p + k + 1][row] * Xs[cl][p + k + 1]; }
        if (k < S) a0 += Ls[p + k][row] * Xs[cl][p + k];
        Ts[cl][row] = a0 + a1;
    }
    __syncthreads();
    for (int e = tid; e < TOT; e += 256) {
        const int pr = e / (S * S), w = e % (S * S), r = w % S, c = w / S;
        const int p = pr * 2 * S, row = p + S + r, cl = p + c;
        T a0 = 0, a1 = 0;
        int k = 0;
        for (; k + 1 <= r; k += 2) { a0 += Xs[p + S + k][row] * Ts[cl][p + S + k]; a1 += Xs[p + S + k + 1][row] * Ts[cl][p + S + k + 1]; }
        if (k <= r) a0 += Xs[p + S + k][row] * Ts[cl][p + S + k];
        Xs[cl][row] = -(a0 + a1);
    }
    __syncthreads();
}
template <typename T>
__device__ int diag_factor_inverse(T *Ab, int n, int nb, T *Di, T &pm_out) {
    __shared__ T cbuf[2][4][64];
    __shared__ T Ls[64][65];     // Ls[c][r] = L(r, c)
    __shared__ T Xs[64][65];     // Xs[c][r] = X(r, c), X = L^{-1}
    __shared__ T Ts[64][65];
    __shared__ int sfail;
    const int tid = threadIdx.x, ri = tid & 63, cb = tid >> 6;
    T v[16];
#pragma unroll
    for (int u = 0; u < 16; u++) {
        const int c = cb + 4 * u;
        v[u] = (ri < nb && c < nb && ri >= c) ? Ab[ri + (int64_t)c * n] : (T)0;
    }
    if (tid == 0) sfail = 0;
    T pm = (T)1e30;
    T dsave[16];              // pivots d_j of this thread's columns (for the final scaling)
#pragma unroll
    for (int u = 0; u < 16; u++) dsave[u] = (T)1;
    const int nblk4 = (nb + 3) / 4;
#pragma unroll 1
    for (int J = 0; J < 16; J++) {
        if (J >= nblk4) break;
        const int j0 = 4 * J, buf = J & 1;
        // publish columns j0 .. j0+3 (column j0 + cb is this thread's v[J]); rows >= nb read 0
        T vj = (T)0;
#pragma unroll
        for (int u = 0; u < 16; u++) if (u == J) vj = v[u];
        cbuf[buf][cb][ri] = vj;
        __syncthreads();
        // the 4 x 4 diagonal piece, eliminated locally (redundantly in every thread)
        T D[4][4];
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int s = 0; s < 4; s++) D[t][s] = (j0 + s < nb) ? cbuf[buf][t][j0 + s] : (T)0;   // D[t][s] = c_t[j0+s]
        // values of the 4 columns at this thread's row and at its 16 columns
        T cr[4], cl[4][16];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            cr[t] = cbuf[buf][t][ri];
#pragma unroll
            for (int u = 0; u < 16; u++) cl[t][u] = cbuf[buf][t][cb + 4 * u];
        }
        T d[4], id[4];
        int fl = 0;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const bool live = j0 + t < nb;
            d[t] = live ? D[t][t] : (T)1;
            if (live && !(d[t] > (T)0) && !fl) fl = j0 + t + 1;
            id[t] = (T)1 / d[t];
            // eliminate column t from the later columns of the block (at rows j0+s, ri and the 16 columns)
#pragma unroll
            for (int w = t + 1; w < 4; w++) {
                const T g = D[t][w] * id[t];          // c_t[j0+w] / d_t
#pragma unroll
                for (int s = 0; s < 4; s++) D[w][s] -= g * D[t][s];
                cr[w] -= g * cr[t];
#pragma unroll
                for (int u = 0; u < 16; u++) cl[w][u] -= g * cl[t][u];
            }
        }
        if (fl) { if (tid == 0) sfail = fl; break; }
#pragma unroll
        for (int t = 0; t < 4; t++) if (j0 + t < nb) pm = d[t] < pm ? d[t] : pm;
        // this thread's block column (j0 + cb) as it stands at its pivot, and its pivot
#pragma unroll
        for (int u = 0; u < 16; u++) if (u == J) {
            T cv = cr[0];
#pragma unroll
            for (int t = 1; t < 4; t++) if (t == cb) cv = cr[t];
            v[u] = (ri >= j0 + cb) ? cv : (T)0;
            T dv = d[0];
#pragma unroll
            for (int t = 1; t < 4; t++) if (t == cb) dv = d[t];
            dsave[u] = dv;
        }
        // trailing update of the columns beyond the block: A(ri, l) -= sum_t c_t[ri] c_t[l] / d_t
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const int l = cb + 4 * u;
            if (u > J && l <= ri) {
                T acc = v[u];
#pragma unroll
                for (int t = 0; t < 4; t++) acc -= cr[t] * id[t] * cl[t][u];
                v[u] = acc;
            }
        }
    }
    __syncthreads();
    if (sfail) return sfail;
    // L(r, c) = a(r, c) / sqrt(d_c)
#pragma unroll
    for (int u = 0; u < 16; u++) {
        const int c = cb + 4 * u;
        const T x = (ri >= c && ri < nb && c < nb) ? v[u] / sqrt(dsave[u]) : (T)0;
        v[u] = x;
        Ls[c][ri] = x;
        Xs[c][ri] = (T)0;
    }
    __syncthreads();
    if (tid < 64) Xs[tid][tid] = tid < nb ? (T)1 / Ls[tid][tid] : (T)1;
    __syncthreads();
    // inverse by recursive doubling (block sizes 1, 2, .., 32 as compile-time levels)
    diag_inv_level<T, 1>(Ls, Xs, Ts, tid);
    diag_inv_level<T, 2>(Ls, Xs, Ts, tid);
    diag_inv_level<T, 4>(Ls, Xs, Ts, tid);
    diag_inv_level<T, 8>(Ls, Xs, Ts, tid);
    diag_inv_level<T, 16>(Ls, Xs, Ts, tid);
    diag_inv_level<T, 32>(Ls, Xs, Ts, tid);
#pragma unroll
    for (int u = 0; u < 16; u++) {
        const int c = cb + 4 * u;
        if (ri < nb && c < nb) Ab[ri + (int64_t)c * n] = v[u];
        Di[ri + 64 * c] = (ri < nb && c < nb && ri >= c) ? Xs[c][ri] : (T)0;
    }
    pm_out = pm;
    return 0;
}
