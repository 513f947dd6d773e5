// sf_device.h -- shared device helpers for the gfx950 StokesFlow hot path.
//
// Layout convention (all kernels): arrays are column-major with x fastest, 0-based,
// A(i, j) = A[i + j * n1].  That is the memory layout of the reference's Julia
// Data.Array, so caller buffers are used as they are.
//
// Arithmetic convention: every expression keeps the reference's left-to-right
// evaluation order, and the library is compiled with -ffp-contract=off so that no
// a*b+c is fused into an FMA (Julia does not fuse without @fastmath/muladd).  Double
// division and floor() are IEEE-exact on gfx950, so FP64 results are bit-identical
// to a CPU evaluation of the same expressions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define SF_WAVE 64

namespace sf {

__host__ __device__ inline int64_t ix2(int64_t i, int64_t j, int64_t n1) { return i + j * n1; }

// Julia Base.min / Base.max on Float64 (NaN-propagating, -0.0 < 0.0).
__device__ inline double jl_min(double a, double b)
{
    if (__builtin_isnan(a) || __builtin_isnan(b)) return __builtin_nan("");
    if (a == b) return __builtin_signbit(a) ? a : b;
    return a < b ? a : b;
}
__device__ inline double jl_max(double a, double b)
{
    if (__builtin_isnan(a) || __builtin_isnan(b)) return __builtin_nan("");
    if (a == b) return __builtin_signbit(a) ? b : a;
    return a > b ? a : b;
}

// Value of lane (lane + d) within the 64-wide wave; lanes past the end keep their own value.
template <typename T>
__device__ inline T lane_down(T v, int d)
{
    return __shfl_down(v, (unsigned)d, SF_WAVE);
}

}  // namespace sf
