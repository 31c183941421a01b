// dep_mfma.h -- 16x16x4 matrix-core operations in FP64 and FP32 (gfx950).
//
// Lane l supplies A[l%16][k = l/16] and B[k = l/16][l%16]; accumulator v of
// lane l is row 4v + l/16 (FP64) or 4(l/16) + v (FP32) of the 16x16 block,
// column l%16.
#pragma once

#include <hip/hip_runtime.h>

template <class T>
struct mfma16;
template <>
struct mfma16<double> {
    typedef double acc __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc op(double a, double b, acc c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int v, int lane) { return 4 * v + (lane >> 4); }
};
template <>
struct mfma16<float> {
    typedef float acc __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc op(float a, float b, acc c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int v, int lane) { return 4 * (lane >> 4) + v; }
};
