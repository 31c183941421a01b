// dep_solve.hip -- the complex system from the geometric blocks.
//
// Z = F - iG of the reference (src/assembly_el.c:565-573) is diag(w) P +
// diag(shift) with the per-row complex weights w: k_weight materialises it in
// FP64 for the FP64 LU (dep_lu.hip), and k_zmatvec evaluates Z x (or the
// residual b - Z x) straight from the real blocks P, streaming 8N^2 bytes per
// product instead of the 16N^2 of a complex FP64 Z.
#include "dep_common.h"
#include "dep_internal.h"

namespace {

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
