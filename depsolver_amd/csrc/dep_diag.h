// dep_diag.h -- LU of one 64 x 64 complex diagonal block and the inverses of
// its triangular factors, in LDS, one barrier per elimination step.
//
// Elimination without pivoting (the reference's gaussBksb exchanges rows only
// for |pivot| <= 1e-9, never reached here): at step p every thread updating
// (r, c > p) forms its own multiplier D[r][p] / D[p][p], and the thread of
// column p stores the multiplier in the staging array X, so no value is read
// and written in the same step.  The inverses use Gauss-Jordan sweeps on
// [L | I] (unit lower, forward) and [U | I] (backward); each row is final one
// step before it is needed and is written straight to the output.
#pragma once

#include "dep_common.h"

#define DIAG_NB 64

template <class T>
struct dcx {
    T r, i;
};
template <class T>
__device__ __forceinline__ dcx<T> dmul(dcx<T> a, dcx<T> b)
{
    return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}
template <class T>
__device__ __forceinline__ dcx<T> drecip(dcx<T> a)
{
    const T d = a.r * a.r + a.i * a.i;
    return {a.r / d, -a.i / d};
}

// D holds the block on entry and L\U on exit (L unit lower below the
// diagonal); X is scratch.  Linv, Uinv: NBxNB complex interleaved outputs
// (row-major, stride DIAG_NB).  Returns min |U_pp| through *piv (thread 0).
// Must be called by all 256 threads of the block.
template <class T>
__device__ void diag_factor_invert(int nb, T (*Dr)[DIAG_NB + 1], T (*Di)[DIAG_NB + 1],
                                   T (*Xr)[DIAG_NB + 1], T (*Xi)[DIAG_NB + 1], T* __restrict__ Linv,
                                   T* __restrict__ Uinv, T* __restrict__ piv)
{
    const int t = threadIdx.x;
    // ---- LU: one barrier per step; multipliers staged in X
    for (int p = 0; p < nb - 1; p++) {
        const dcx<T> ip = drecip(dcx<T>{Dr[p][p], Di[p][p]});
        const int m = nb - 1 - p, w = nb - p;
        for (int idx = t; idx < m * w; idx += 256) {
            const int r = p + 1 + idx / w, c = p + idx % w;
            const dcx<T> l = dmul(dcx<T>{Dr[r][p], Di[r][p]}, ip);
            if (c == p) {
                Xr[r][p] = l.r;
                Xi[r][p] = l.i;
            } else {
                const dcx<T> u = dmul(l, dcx<T>{Dr[p][c], Di[p][c]});
                Dr[r][c] -= u.r;
                Di[r][c] -= u.i;
            }
        }
        __syncthreads();
    }
    for (int idx = t; idx < nb * nb; idx += 256) {
        const int r = idx / nb, c = idx % nb;
        if (r > c) {
            Dr[r][c] = Xr[r][c];
            Di[r][c] = Xi[r][c];
        }
    }
    if (t == 0) {
        T mn = Dr[0][0] * Dr[0][0] + Di[0][0] * Di[0][0];
        for (int p = 1; p < nb; p++) mn = fmin(mn, Dr[p][p] * Dr[p][p] + Di[p][p] * Di[p][p]);
        *piv = sqrt(mn);
    }
    __syncthreads();
    // ---- Linv: X = I; step p: X[r][j] -= L[r][p] X[p][j] (r > p, j <= p); row p final at step p
    for (int idx = t; idx < DIAG_NB * DIAG_NB; idx += 256) {
        const int r = idx / DIAG_NB, c = idx % DIAG_NB;
        Xr[r][c] = (r == c) ? T(1) : T(0);
        Xi[r][c] = T(0);
        Linv[2 * idx] = (r == c) ? T(1) : T(0);
        Linv[2 * idx + 1] = T(0);
    }
    __syncthreads();
    for (int p = 0; p < nb; p++) {
        const int w = p + 1, m = nb - 1 - p;
        for (int idx = t; idx < (m + 1) * w; idx += 256) {
            const int r = p + idx / w, j = idx % w;
            if (r == p) {
                Linv[2 * (p * DIAG_NB + j)] = Xr[p][j];
                Linv[2 * (p * DIAG_NB + j) + 1] = Xi[p][j];
            } else {
                const dcx<T> lx = dmul(dcx<T>{Dr[r][p], Di[r][p]}, dcx<T>{Xr[p][j], Xi[p][j]});
                Xr[r][j] -= lx.r;
                Xi[r][j] -= lx.i;
            }
        }
        __syncthreads();
    }
    // ---- Uinv: X = I; step p (descending): x = X[p][j] / U[p][p] is row p of the
    // inverse; X[r][j] -= U[r][p] x for r < p
    for (int idx = t; idx < DIAG_NB * DIAG_NB; idx += 256) {
        const int r = idx / DIAG_NB, c = idx % DIAG_NB;
        Xr[r][c] = (r == c) ? T(1) : T(0);
        Xi[r][c] = T(0);
        Uinv[2 * idx] = T(0);
        Uinv[2 * idx + 1] = T(0);
    }
    __syncthreads();
    for (int p = nb - 1; p >= 0; p--) {
        const dcx<T> ip = drecip(dcx<T>{Dr[p][p], Di[p][p]});
        const int w = nb - p;
        for (int idx = t; idx < (p + 1) * w; idx += 256) {
            const int r = idx / w, j = p + idx % w;
            const dcx<T> x = dmul(dcx<T>{Xr[p][j], Xi[p][j]}, ip);
            if (r == p) {
                Uinv[2 * (p * DIAG_NB + j)] = x.r;
                Uinv[2 * (p * DIAG_NB + j) + 1] = x.i;
            } else {
                const dcx<T> ux = dmul(dcx<T>{Dr[r][p], Di[r][p]}, x);
                Xr[r][j] -= ux.r;
                Xi[r][j] -= ux.i;
            }
        }
        __syncthreads();
    }
}
