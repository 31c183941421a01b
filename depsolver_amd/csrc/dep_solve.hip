// dep_solve.hip -- FP64 dense solve of the complex collocation system.
//
// Replaces gaussBksb (src/gaussBksb.c:39-108): Gaussian elimination without
// pivoting (the reference exchanges rows only when |pivot| <= 1e-9, which
// these matrices never reach) followed by back-substitution.  Here the
// complex N x N system Z = F - iG is factored directly (the reference works
// on the real 2N x 2N embedding [[F, G], [-G, F]], twice the flops) with a
// blocked right-looking LU on split real/imaginary planes:
//   1. diagonal block: unblocked LU in LDS, then the inverses of L11, U11
//   2. L21 = A21 U11^-1 and U12 = L11^-1 A12 as small in-place products
//   3. trailing update A22 -= L21 U12 (complex GEMM)
// Triangular solves use the stored inverse diagonal blocks.
//
// Z itself is never needed in FP64 for matrix-vector products: Z x is
// evaluated as w .* (P x) + shift .* x from the real geometric blocks P, which
// halves the bytes streamed per product (dep_residual below).
#include <stdlib.h>
#include <string.h>

#include "dep_common.h"
#include "dep_internal.h"

#define LU_NB 64

namespace {

struct cplx {
    double r, i;
};
__device__ __forceinline__ cplx cmul(cplx a, cplx b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ __forceinline__ cplx cdiv(cplx a, cplx b)
{
    const double d = b.r * b.r + b.i * b.i;
    return {(a.r * b.r + a.i * b.i) / d, (a.i * b.r - a.r * b.i) / d};
}

// Z = w * P + shift on the diagonal (FP64 split planes)
__global__ void k_weight(int N, const double* __restrict__ P, long ldP, const double* __restrict__ wr,
                         const double* __restrict__ wi, const double* __restrict__ shift,
                         double* __restrict__ Zr, double* __restrict__ Zi, long ldZ)
{
    const int i = blockIdx.y;
    const double a = wr[i], b = wi[i];
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < N; c += gridDim.x * blockDim.x) {
        const double p = P[(long)i * ldP + c];
        Zr[(long)i * ldZ + c] = a * p + (c == i ? shift[i] : 0.0);
        Zi[(long)i * ldZ + c] = b * p;
    }
}

// Unblocked LU (Doolittle, unit L) of the nb x nb diagonal block at (k0,k0),
// then Linv (unit lower) and Uinv (upper) into inv[2][nb][nb] complex.
__global__ __launch_bounds__(256) void k_lu_diag(int nb, double* __restrict__ Zr,
                                                 double* __restrict__ Zi, long ldZ, int k0,
                                                 double* __restrict__ inv, double* __restrict__ piv)
{
    __shared__ double Dr[LU_NB][LU_NB + 1], Di[LU_NB][LU_NB + 1];
    __shared__ double Ir[LU_NB][LU_NB + 1], Ii[LU_NB][LU_NB + 1];
    const int t = threadIdx.x;
    for (int idx = t; idx < nb * nb; idx += 256) {
        const int r = idx / nb, c = idx % nb;
        Dr[r][c] = Zr[(long)(k0 + r) * ldZ + k0 + c];
        Di[r][c] = Zi[(long)(k0 + r) * ldZ + k0 + c];
    }
    __syncthreads();
    for (int p = 0; p < nb; p++) {
        const cplx pv = {Dr[p][p], Di[p][p]};
        // column below the pivot
        for (int r = p + 1 + t; r < nb; r += 256) {
            const cplx l = cdiv({Dr[r][p], Di[r][p]}, pv);
            Dr[r][p] = l.r;
            Di[r][p] = l.i;
        }
        __syncthreads();
        const int m = nb - p - 1;
        for (int idx = t; idx < m * m; idx += 256) {
            const int r = p + 1 + idx / m, c = p + 1 + idx % m;
            const cplx l = {Dr[r][p], Di[r][p]}, u = {Dr[p][c], Di[p][c]};
            const cplx lu = cmul(l, u);
            Dr[r][c] -= lu.r;
            Di[r][c] -= lu.i;
        }
        __syncthreads();
    }
    if (t == 0) {
        double mn = 1e300;
        for (int p = 0; p < nb; p++) mn = fmin(mn, sqrt(Dr[p][p] * Dr[p][p] + Di[p][p] * Di[p][p]));
        piv[k0 / LU_NB] = mn;
    }
    // write the factors back
    for (int idx = t; idx < nb * nb; idx += 256) {
        const int r = idx / nb, c = idx % nb;
        Zr[(long)(k0 + r) * ldZ + k0 + c] = Dr[r][c];
        Zi[(long)(k0 + r) * ldZ + k0 + c] = Di[r][c];
    }
    // Linv by Gauss-Jordan on [L | I] (unit lower): X = I; X[r][:] -= L[r][p] X[p][:] for r > p
    for (int idx = t; idx < LU_NB * LU_NB; idx += 256) {
        const int r = idx / LU_NB, c = idx % LU_NB;
        Ir[r][c] = (r == c) ? 1.0 : 0.0;
        Ii[r][c] = 0.0;
    }
    __syncthreads();
    for (int p = 0; p < nb - 1; p++) {
        const int w = p + 1, m = nb - 1 - p;
        for (int idx = t; idx < m * w; idx += 256) {
            const int r = p + 1 + idx / w, j = idx % w;
            const cplx lx = cmul(cplx{Dr[r][p], Di[r][p]}, cplx{Ir[p][j], Ii[p][j]});
            Ir[r][j] -= lx.r;
            Ii[r][j] -= lx.i;
        }
        __syncthreads();
    }
    for (int idx = t; idx < nb * nb; idx += 256) {
        const int r = idx / nb, c = idx % nb;
        inv[2 * (r * LU_NB + c)] = Ir[r][c];
        inv[2 * (r * LU_NB + c) + 1] = Ii[r][c];
    }
    __syncthreads();
    // Uinv by backward Gauss-Jordan on [U | I]
    for (int idx = t; idx < LU_NB * LU_NB; idx += 256) {
        const int r = idx / LU_NB, c = idx % LU_NB;
        Ir[r][c] = (r == c) ? 1.0 : 0.0;
        Ii[r][c] = 0.0;
    }
    __syncthreads();
    for (int p = nb - 1; p >= 0; p--) {
        const cplx d = {Dr[p][p], Di[p][p]};
        for (int j = p + t; j < nb; j += 256) {
            const cplx v = cdiv(cplx{Ir[p][j], Ii[p][j]}, d);
            Ir[p][j] = v.r;
            Ii[p][j] = v.i;
        }
        __syncthreads();
        const int w = nb - p;
        for (int idx = t; idx < p * w; idx += 256) {
            const int r = idx / w, j = p + idx % w;
            const cplx ux = cmul(cplx{Dr[r][p], Di[r][p]}, cplx{Ir[p][j], Ii[p][j]});
            Ir[r][j] -= ux.r;
            Ii[r][j] -= ux.i;
        }
        __syncthreads();
    }
    double* Ui = inv + 2 * LU_NB * LU_NB;
    for (int idx = t; idx < nb * nb; idx += 256) {
        const int r = idx / nb, c = idx % nb;
        Ui[2 * (r * LU_NB + c)] = Ir[r][c];
        Ui[2 * (r * LU_NB + c) + 1] = Ii[r][c];
    }
}

// rows [r0, N): A[i, k0:k0+nb] <- A[i, k0:k0+nb] * Uinv   (16 rows per block)
__global__ __launch_bounds__(256) void k_apply_uinv(int N, int nb, int r0, double* __restrict__ Zr,
                                                    double* __restrict__ Zi, long ldZ, int k0,
                                                    const double* __restrict__ Uinv)
{
    __shared__ double Ar[16][LU_NB], Ai[16][LU_NB];
    __shared__ double Ur[LU_NB][LU_NB + 1], Ui[LU_NB][LU_NB + 1];
    const int t = threadIdx.x;
    const int rb = r0 + blockIdx.x * 16;
    for (int idx = t; idx < nb * nb; idx += 256) {
        const int r = idx / nb, c = idx % nb;
        Ur[r][c] = Uinv[2 * (r * LU_NB + c)];
        Ui[r][c] = Uinv[2 * (r * LU_NB + c) + 1];
    }
    for (int idx = t; idx < 16 * nb; idx += 256) {
        const int r = idx / nb, c = idx % nb;
        const int gi = rb + r;
        Ar[r][c] = gi < N ? Zr[(long)gi * ldZ + k0 + c] : 0.0;
        Ai[r][c] = gi < N ? Zi[(long)gi * ldZ + k0 + c] : 0.0;
    }
    __syncthreads();
    for (int idx = t; idx < 16 * nb; idx += 256) {
        const int r = idx / nb, c = idx % nb;
        double sr = 0.0, si = 0.0;
        for (int q = 0; q <= c; q++) {
            sr += Ar[r][q] * Ur[q][c] - Ai[r][q] * Ui[q][c];
            si += Ar[r][q] * Ui[q][c] + Ai[r][q] * Ur[q][c];
        }
        const int gi = rb + r;
        if (gi < N) {
            Zr[(long)gi * ldZ + k0 + c] = sr;
            Zi[(long)gi * ldZ + k0 + c] = si;
        }
    }
}

// columns [c0, N): A[k0:k0+nb, j] <- Linv * A[k0:k0+nb, j]   (16 columns per block)
__global__ __launch_bounds__(256) void k_apply_linv(int N, int nb, int c0, double* __restrict__ Zr,
                                                    double* __restrict__ Zi, long ldZ, int k0,
                                                    const double* __restrict__ Linv)
{
    __shared__ double Br[LU_NB][16], Bi[LU_NB][16];
    __shared__ double Lr[LU_NB][LU_NB + 1], Li[LU_NB][LU_NB + 1];
    const int t = threadIdx.x;
    const int cb = c0 + blockIdx.x * 16;
    for (int idx = t; idx < nb * nb; idx += 256) {
        const int r = idx / nb, c = idx % nb;
        Lr[r][c] = Linv[2 * (r * LU_NB + c)];
        Li[r][c] = Linv[2 * (r * LU_NB + c) + 1];
    }
    for (int idx = t; idx < 16 * nb; idx += 256) {
        const int r = idx / 16, c = idx % 16;
        const int gc = cb + c;
        Br[r][c] = gc < N ? Zr[(long)(k0 + r) * ldZ + gc] : 0.0;
        Bi[r][c] = gc < N ? Zi[(long)(k0 + r) * ldZ + gc] : 0.0;
    }
    __syncthreads();
    for (int idx = t; idx < 16 * nb; idx += 256) {
        const int r = idx / 16, c = idx % 16;
        double sr = 0.0, si = 0.0;
        for (int q = 0; q <= r; q++) {
            sr += Lr[r][q] * Br[q][c] - Li[r][q] * Bi[q][c];
            si += Lr[r][q] * Bi[q][c] + Li[r][q] * Br[q][c];
        }
        const int gc = cb + c;
        if (gc < N) {
            Zr[(long)(k0 + r) * ldZ + gc] = sr;
            Zi[(long)(k0 + r) * ldZ + gc] = si;
        }
    }
}

// trailing update C[r0:, r0:] -= A[r0:, k0:k0+nb] * B[k0:k0+nb, r0:]; 64x64 tile / 256 threads
__global__ __launch_bounds__(256) void k_gemm_update(int N, int nb, int k0, int r0,
                                                     double* __restrict__ Zr,
                                                     double* __restrict__ Zi, long ldZ)
{
    constexpr int BM = 64, BN = 64, BK = 16;
    __shared__ double Asr[BK][BM + 1], Asi[BK][BM + 1];
    __shared__ double Bsr[BK][BN + 1], Bsi[BK][BN + 1];
    const int t = threadIdx.x;
    const int tm = t / 16, tn = t % 16;           // 4x4 outputs per thread
    const int m0 = r0 + blockIdx.y * BM, n0 = r0 + blockIdx.x * BN;
    double cr[4][4], ci[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) cr[a][b] = ci[a][b] = 0.0;
    for (int kk = 0; kk < nb; kk += BK) {
        for (int idx = t; idx < BM * BK; idx += 256) {
            const int r = idx / BK, k = idx % BK;
            const int gi = m0 + r, gk = k0 + kk + k;
            const bool ok = gi < N && (kk + k) < nb;
            Asr[k][r] = ok ? Zr[(long)gi * ldZ + gk] : 0.0;
            Asi[k][r] = ok ? Zi[(long)gi * ldZ + gk] : 0.0;
        }
        for (int idx = t; idx < BK * BN; idx += 256) {
            const int k = idx / BN, c = idx % BN;
            const int gk = k0 + kk + k, gc = n0 + c;
            const bool ok = gc < N && (kk + k) < nb;
            Bsr[k][c] = ok ? Zr[(long)gk * ldZ + gc] : 0.0;
            Bsi[k][c] = ok ? Zi[(long)gk * ldZ + gc] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < BK; k++) {
            double ar[4], ai[4], br[4], bi[4];
#pragma unroll
            for (int a = 0; a < 4; a++) {
                ar[a] = Asr[k][tm + 16 * a];
                ai[a] = Asi[k][tm + 16 * a];
                br[a] = Bsr[k][tn + 16 * a];
                bi[a] = Bsi[k][tn + 16 * a];
            }
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    cr[a][b] += ar[a] * br[b] - ai[a] * bi[b];
                    ci[a][b] += ar[a] * bi[b] + ai[a] * br[b];
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int gi = m0 + tm + 16 * a, gc = n0 + tn + 16 * b;
            if (gi < N && gc < N) {
                Zr[(long)gi * ldZ + gc] -= cr[a][b];
                Zi[(long)gi * ldZ + gc] -= ci[a][b];
            }
        }
}

// trailing update on v_mfma_f64_16x16x4f64: C[r0:, r0:] -= A[r0:, k0:k0+nb] B[k0:k0+nb, r0:].
// 64x64 complex tile per workgroup, 32x32 per wave = 2x2 blocks of 16x16 per plane;
// Cr += Ar Br - Ai Bi and Ci += Ar Bi + Ai Br as four real MFMAs per block.  Lane
// l supplies A[l%16][k = l/16] and B[k = l/16][l%16]; accumulator v of lane l is
// C[4v + l/16][l%16].
typedef double dbl4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_gemm_f64mfma(int N, int nb, int k0, int r0,
                                                      double* __restrict__ Zr,
                                                      double* __restrict__ Zi, long ldZ)
{
    constexpr int BM = 64, BN = 64, BK = 16;
    __shared__ double Asr[BM][BK + 1], Asi[BM][BK + 1];
    __shared__ double Bsr[BK][BN + 1], Bsi[BK][BN + 1];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int m0 = r0 + blockIdx.y * BM, n0 = r0 + blockIdx.x * BN;
    dbl4 cr[2][2], ci[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int v = 0; v < 4; v++) cr[a][b][v] = ci[a][b][v] = 0.0;
    const int li = lane & 15, lk = lane >> 4;
    for (int kk = 0; kk < nb; kk += BK) {
        for (int idx = t; idx < BM * BK; idx += 256) {
            const int r = idx / BK, k = idx % BK, gi = m0 + r;
            const bool ok = gi < N && (kk + k) < nb;
            Asr[r][k] = ok ? Zr[(long)gi * ldZ + k0 + kk + k] : 0.0;
            Asi[r][k] = ok ? Zi[(long)gi * ldZ + k0 + kk + k] : 0.0;
        }
        for (int idx = t; idx < BK * BN; idx += 256) {
            const int k = idx / BN, c = idx % BN, gc = n0 + c;
            const bool ok = gc < N && (kk + k) < nb;
            Bsr[k][c] = ok ? Zr[(long)(k0 + kk + k) * ldZ + gc] : 0.0;
            Bsi[k][c] = ok ? Zi[(long)(k0 + kk + k) * ldZ + gc] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < BK; ks += 4) {
            double ar[2], ai[2], nai[2], br[2], bi[2];
#pragma unroll
            for (int a = 0; a < 2; a++) {
                ar[a] = Asr[wm + 16 * a + li][ks + lk];
                ai[a] = Asi[wm + 16 * a + li][ks + lk];
                nai[a] = -ai[a];
                br[a] = Bsr[ks + lk][wn + 16 * a + li];
                bi[a] = Bsi[ks + lk][wn + 16 * a + li];
            }
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int b = 0; b < 2; b++) {
                    cr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], br[b], cr[a][b], 0, 0, 0);
                    cr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai[a], bi[b], cr[a][b], 0, 0, 0);
                    ci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], bi[b], ci[a][b], 0, 0, 0);
                    ci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], br[b], ci[a][b], 0, 0, 0);
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int gi = m0 + wm + 16 * a + 4 * v + lk, gc = n0 + wn + 16 * b + li;
                if (gi < N && gc < N) {
                    Zr[(long)gi * ldZ + gc] -= cr[a][b][v];
                    Zi[(long)gi * ldZ + gc] -= ci[a][b][v];
                }
            }
}

// ---- triangular solves with the stored inverse diagonal blocks ------------
// In place: v_b <- Binv * v_b for one block of nb rows (split re/im vectors);
// Binv staged in LDS, 4 threads per row
__global__ __launch_bounds__(256) void k_block_apply(int nb, const double* __restrict__ Binv,
                                                     double* __restrict__ vr, double* __restrict__ vi,
                                                     int lower)
{
    __shared__ double Br[LU_NB][LU_NB + 1], Bi[LU_NB][LU_NB + 1];
    __shared__ double xr[LU_NB], xi[LU_NB];
    const int t = threadIdx.x;
    for (int idx = t; idx < LU_NB * LU_NB; idx += 256) {
        const int r = idx / LU_NB, c = idx % LU_NB;
        Br[r][c] = Binv[2 * idx];
        Bi[r][c] = Binv[2 * idx + 1];
    }
    if (t < nb) {
        xr[t] = vr[t];
        xi[t] = vi[t];
    }
    __syncthreads();
    const int r = t >> 2, part = t & 3;
    double sr = 0.0, si = 0.0;
    if (r < nb) {
        const int q0 = lower ? 0 : r, q1 = lower ? r : nb - 1;
        for (int q = q0 + part; q <= q1; q += 4) {
            sr += Br[r][q] * xr[q] - Bi[r][q] * xi[q];
            si += Br[r][q] * xi[q] + Bi[r][q] * xr[q];
        }
    }
    sr += __shfl_xor(sr, 1);
    si += __shfl_xor(si, 1);
    sr += __shfl_xor(sr, 2);
    si += __shfl_xor(si, 2);
    if (r < nb && part == 0) {
        vr[r] = sr;
        vi[r] = si;
    }
}

// x[rows] -= Z[rows, k0:k0+nb] * x[k0:k0+nb]  for rows in [ra, rb) (split vectors)
__global__ void k_panel_gemv(int ra, int rb, int nb, const double* __restrict__ Zr,
                             const double* __restrict__ Zi, long ldZ, int k0,
                             double* __restrict__ xr, double* __restrict__ xi)
{
    const int i = ra + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rb) return;
    double sr = 0.0, si = 0.0;
    for (int q = 0; q < nb; q++) {
        const double ar = Zr[(long)i * ldZ + k0 + q], ai = Zi[(long)i * ldZ + k0 + q];
        const double yr = xr[k0 + q], yi = xi[k0 + q];
        sr += ar * yr - ai * yi;
        si += ar * yi + ai * yr;
    }
    xr[i] -= sr;
    xi[i] -= si;
}

// ---- products with P -------------------------------------------------------
// y = Z x with Z = w .* P + diag(shift); x, y split (re block, im block) of length N.
// One wave per row, FP64.
__global__ __launch_bounds__(256) void k_zmatvec(int N, const double* __restrict__ P, long ldP,
                                                 const double* __restrict__ wr,
                                                 const double* __restrict__ wi,
                                                 const double* __restrict__ shift,
                                                 const double* __restrict__ x,
                                                 const double* __restrict__ b,
                                                 double* __restrict__ y)
{
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= N) return;
    const double* p = P + (long)row * ldP;
    double sr = 0.0, si = 0.0;
    for (int c = lane; c < N; c += 64) {
        const double v = p[c];
        sr += v * x[c];
        si += v * x[N + c];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        sr += __shfl_xor(sr, o);
        si += __shfl_xor(si, o);
    }
    if (lane == 0) {
        const double a = wr[row], bb = wi[row], s = shift[row];
        double zr = a * sr - bb * si + s * x[row];
        double zi = a * si + bb * sr + s * x[N + row];
        if (b) {   // residual b - Z x
            zr = b[row] - zr;
            zi = b[N + row] - zi;
        }
        y[row] = zr;
        y[N + row] = zi;
    }
}

}  // namespace

hipError_t launch_weight_fp64(dep_sys* s)
{
    dep_geom* g = s->g;
    hipLaunchKernelGGL(k_weight, dim3(dep_cdiv(g->N, 256 * 4), g->N), dim3(256), 0, g->ctx->stream,
                       g->N, g->P, g->ldP, s->wr, s->wi, s->shift, s->Zr, s->Zi, s->ldZ);
    return hipGetLastError();
}

// s->work must hold nblocks * 4 * LU_NB * LU_NB doubles for the inverses
hipError_t launch_lu_fp64(dep_sys* s, double* min_pivot)
{
    dep_geom* g = s->g;
    const int N = g->N;
    hipStream_t st = g->ctx->stream;
    const int nblk = dep_cdiv(N, LU_NB);
    double* inv = (double*)s->opA64;
    double* piv = s->dwork;
    // DEPSOLVER_FP64_GEMM=valu selects the vector-FMA trailing update (for comparison)
    const char* sel = getenv("DEPSOLVER_FP64_GEMM");
    const bool use_mfma = !(sel && strcmp(sel, "valu") == 0);
    for (int kb = 0; kb < nblk; kb++) {
        const int k0 = kb * LU_NB, nb = (N - k0 < LU_NB) ? N - k0 : LU_NB;
        double* invb = inv + (size_t)kb * 4 * LU_NB * LU_NB;
        hipLaunchKernelGGL(k_lu_diag, dim3(1), dim3(256), 0, st, nb, s->Zr, s->Zi, s->ldZ, k0, invb,
                           piv);
        const int r0 = k0 + nb;
        if (r0 < N) {
            hipLaunchKernelGGL(k_apply_uinv, dim3(dep_cdiv(N - r0, 16)), dim3(256), 0, st, N, nb, r0,
                               s->Zr, s->Zi, s->ldZ, k0, invb + 2 * LU_NB * LU_NB);
            hipLaunchKernelGGL(k_apply_linv, dim3(dep_cdiv(N - r0, 16)), dim3(256), 0, st, N, nb, r0,
                               s->Zr, s->Zi, s->ldZ, k0, invb);
            dim3 grid(dep_cdiv(N - r0, 64), dep_cdiv(N - r0, 64));
            hipEvent_t ev;
            prof_begin(g->ctx, &ev);
            if (use_mfma)
                hipLaunchKernelGGL(k_gemm_f64mfma, grid, dim3(256), 0, st, N, nb, k0, r0, s->Zr, s->Zi,
                                   s->ldZ);
            else
                hipLaunchKernelGGL(k_gemm_update, grid, dim3(256), 0, st, N, nb, k0, r0, s->Zr, s->Zi,
                                   s->ldZ);
            const double m = N - r0;
            prof_end(g->ctx, ev, "lu_gemm_fp64", 8.0 * m * m * nb, 32.0 * m * m + 32.0 * m * nb);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    std::vector<double> hp(nblk);
    hipError_t e = hipMemcpyAsync(hp.data(), piv, sizeof(double) * nblk, hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    double mn = 1e300;
    for (double v : hp) mn = v < mn ? v : mn;
    *min_pivot = mn;
    return hipSuccess;
}

// out = U^-1 L^-1 rhs; split (re block, im block) vectors of length N (may alias)
hipError_t launch_trsv_fp64(dep_sys* s, const double* rhs, double* out)
{
    dep_geom* g = s->g;
    const int N = g->N;
    hipStream_t st = g->ctx->stream;
    const int nblk = dep_cdiv(N, LU_NB);
    const double* inv = (const double*)s->opA64;
    double* xr = out;
    double* xi = out + N;
    hipError_t e;
    if (out != rhs) {
        e = hipMemcpyAsync(out, rhs, sizeof(double) * 2 * N, hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) return e;
    }
    for (int kb = 0; kb < nblk; kb++) {   // forward: L y = r (unit lower)
        const int k0 = kb * LU_NB, nb = (N - k0 < LU_NB) ? N - k0 : LU_NB;
        const double* Linv = inv + (size_t)kb * 4 * LU_NB * LU_NB;
        hipLaunchKernelGGL(k_block_apply, dim3(1), dim3(256), 0, st, nb, Linv, xr + k0, xi + k0, 1);
        if (k0 + nb < N)
            hipLaunchKernelGGL(k_panel_gemv, dim3(dep_cdiv(N - k0 - nb, 128)), dim3(128), 0, st,
                               k0 + nb, N, nb, s->Zr, s->Zi, s->ldZ, k0, xr, xi);
    }
    for (int kb = nblk - 1; kb >= 0; kb--) {   // backward: U x = y
        const int k0 = kb * LU_NB, nb = (N - k0 < LU_NB) ? N - k0 : LU_NB;
        const double* Uinv = inv + (size_t)kb * 4 * LU_NB * LU_NB + 2 * LU_NB * LU_NB;
        hipLaunchKernelGGL(k_block_apply, dim3(1), dim3(256), 0, st, nb, Uinv, xr + k0, xi + k0, 0);
        if (k0 > 0)
            hipLaunchKernelGGL(k_panel_gemv, dim3(dep_cdiv(k0, 128)), dim3(128), 0, st, 0, k0, nb,
                               s->Zr, s->Zi, s->ldZ, k0, xr, xi);
    }
    return hipGetLastError();
}

// r = b - Z x (b may be null: r = Z x); x, r split layout (re block, im block)
hipError_t launch_residual(dep_sys* s, const double* x, double* r)
{
    return launch_zmatvec_w(s, s->wr, s->wi, s->shift, x, s->b, r);
}

hipError_t launch_zmatvec_w(dep_sys* s, const double* wr, const double* wi, const double* shift,
                            const double* x, const double* b, double* y)
{
    dep_geom* g = s->g;
    hipEvent_t ev;
    prof_begin(g->ctx, &ev);
    hipLaunchKernelGGL(k_zmatvec, dim3(dep_cdiv(g->N, 4)), dim3(256), 0, g->ctx->stream, g->N, g->P,
                       g->ldP, wr, wi, shift, x, b, y);
    const double n = g->N;
    prof_end(g->ctx, ev, "zmatvec", 4.0 * n * n, 8.0 * n * n + 48.0 * n);
    return hipGetLastError();
}

hipError_t launch_matvec(dep_sys* s, const double* x, double* y)
{
    return launch_zmatvec_w(s, s->wr, s->wi, s->shift, x, nullptr, y);
}
