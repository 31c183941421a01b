// dep_gmres.hip -- device kernels of the FP64 GMRES (dep_gmres in dep_api.cpp).
//
// Replaces the vector work of solverGMRES_el/iterGMRES_el (src/gmres_el.c):
// the Krylov basis lives in device memory (max_iters + 1 vectors of 2N FP64)
// and the orthogonalisation of each new vector is classical Gram-Schmidt
// applied twice (CGS2) as two dot-product/update kernels per pass, instead of
// the reference's sequence of i+1 dependent modified-Gram-Schmidt steps
// (gmres_el.c:190-194).  Both give an orthonormal basis to working precision;
// CGS2 turns the latency-bound chain into bandwidth-bound GEMVs.  The small
// Hessenberg/Givens recurrence stays on the host, as in the reference.
#include "dep_common.h"
#include "dep_internal.h"

namespace {

// h[j] = dot(V[j], w) for j < nvec; one workgroup per j
__global__ __launch_bounds__(256) void k_dots(int nvec, long n2, const double* __restrict__ V,
                                              const double* __restrict__ w, double* __restrict__ h)
{
    __shared__ double sh[256];
    const int j = blockIdx.x;
    const double* v = V + (long)j * n2;
    double s = 0.0;
    for (long t = threadIdx.x; t < n2; t += 256) s += v[t] * w[t];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) h[j] = sh[0];
}

// w -= sum_j h[j] V[j]
__global__ void k_update(int nvec, long n2, const double* __restrict__ V, const double* __restrict__ h,
                         double* __restrict__ w)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n2) return;
    double s = w[t];
    for (int j = 0; j < nvec; j++) s -= h[j] * V[(long)j * n2 + t];
    w[t] = s;
}

// out = w * alpha
__global__ void k_scale(long n2, const double* __restrict__ w, double alpha, double* __restrict__ out)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n2) out[t] = w[t] * alpha;
}

// x += sum_j y[j] V[j]
__global__ void k_combine(int nvec, long n2, const double* __restrict__ V, const double* __restrict__ y,
                          double* __restrict__ x)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n2) return;
    double s = x[t];
    for (int j = 0; j < nvec; j++) s += y[j] * V[(long)j * n2 + t];
    x[t] = s;
}

// y = Z x for a dense complex Z (split planes), one wave per row; with b: y = b - Z x
__global__ __launch_bounds__(256) void k_dense_zmatvec(int N, const double* __restrict__ Zr,
                                                       const double* __restrict__ Zi, long ldZ,
                                                       const double* __restrict__ x,
                                                       const double* __restrict__ b,
                                                       double* __restrict__ y)
{
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= N) return;
    double sr = 0.0, si = 0.0;
    for (int c = lane; c < N; c += 64) {
        const double ar = Zr[(long)row * ldZ + c], ai = Zi[(long)row * ldZ + c];
        const double xr = x[c], xi = x[N + c];
        sr += ar * xr - ai * xi;
        si += ar * xi + ai * xr;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        sr += __shfl_xor(sr, o);
        si += __shfl_xor(si, o);
    }
    if (lane == 0) {
        if (b) {
            sr = b[row] - sr;
            si = b[N + row] - si;
        }
        y[row] = sr;
        y[N + row] = si;
    }
}

}  // namespace

hipError_t launch_dots(dep_ctx* c, int nvec, long n2, const double* V, const double* w, double* h)
{
    hipLaunchKernelGGL(k_dots, dim3(nvec), dim3(256), 0, c->stream, nvec, n2, V, w, h);
    return hipGetLastError();
}

hipError_t launch_update(dep_ctx* c, int nvec, long n2, const double* V, const double* h, double* w)
{
    hipLaunchKernelGGL(k_update, dim3(dep_cdiv(n2, 256)), dim3(256), 0, c->stream, nvec, n2, V, h, w);
    return hipGetLastError();
}

hipError_t launch_scale(dep_ctx* c, long n2, const double* w, double alpha, double* out)
{
    hipLaunchKernelGGL(k_scale, dim3(dep_cdiv(n2, 256)), dim3(256), 0, c->stream, n2, w, alpha, out);
    return hipGetLastError();
}

hipError_t launch_combine(dep_ctx* c, int nvec, long n2, const double* V, const double* y, double* x)
{
    hipLaunchKernelGGL(k_combine, dim3(dep_cdiv(n2, 256)), dim3(256), 0, c->stream, nvec, n2, V, y, x);
    return hipGetLastError();
}

hipError_t launch_dense_zmatvec(dep_ctx* c, int N, const double* Zr, const double* Zi, long ldZ,
                                const double* x, const double* b, double* y)
{
    hipLaunchKernelGGL(k_dense_zmatvec, dim3(dep_cdiv(N, 4)), dim3(256), 0, c->stream, N, Zr, Zi, ldZ, x,
                       b, y);
    return hipGetLastError();
}
