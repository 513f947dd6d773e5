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

// One-lane wave shifts as DPP moves (GFX9 DPP wave_shl:1 = 0x130, wave_shr:1 = 0x138): a VALU
// move instead of an LDS permute.  wave_shl:1 gives lane i the value of lane i+1 and
// wave_shr:1 that of lane i-1, as __shfl_down(v, 1) / __shfl_up(v, 1) do on a 64-wide wave; the
// lane without a source reads 0 (bound_ctrl), so the move needs no copy of its source first.
// Only halo lanes lack a source, and no kernel stores a halo lane's result.
template <int CTRL>
__device__ __forceinline__ int dpp_i32(int v)
{
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, true);
}
template <int CTRL>
__device__ __forceinline__ float dpp_shift(float v)
{
    return __int_as_float(dpp_i32<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ __forceinline__ double dpp_shift(double v)
{
    const int lo = dpp_i32<CTRL>(__double2loint(v));
    const int hi = dpp_i32<CTRL>(__double2hiint(v));
    return __hiloint2double(hi, lo);
}

// Value of lane (lane + d) within the 64-wide wave.  A shift by d > 1 repeats the one-lane
// shift; the last d lanes read 0, and no kernel stores a result of those halo lanes.
template <typename T>
__device__ __forceinline__ T lane_down(T v, int d)
{
    for (int k = 0; k < d; ++k) v = dpp_shift<0x130>(v);
    return v;
}

// Value of lane (lane - d); the first d lanes read 0.
template <typename T>
__device__ __forceinline__ T lane_up(T v, int d)
{
    for (int k = 0; k < d; ++k) v = dpp_shift<0x138>(v);
    return v;
}

}  // namespace sf
