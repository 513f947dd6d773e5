// sf_snapshot.hip -- device side of the snapshot files: FP64 -> FP32 casts (with the marker
// coordinate shift of saveMarkers, StokesFlow2D_multixpu.jl:233-234) into a staging buffer.
#include <algorithm>

#include "sf_device.h"
#include "sf_internal.h"

namespace sf {

// dst[k] = (float)(src[k] + shift): Julia's convert.(Float32, x .+ shift), round to nearest.
__global__ __launch_bounds__(256) void snap_cast_shift(const double* __restrict__ src, float* __restrict__ dst,
                                                       int64_t n, double shift)
{
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (int64_t)gridDim.x * 256)
        dst[k] = (float)(src[k] + shift);
}

hipError_t launch_snap_cast(const double* src, float* dst, int64_t n, double shift, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int64_t nb = std::min<int64_t>(4096, (n + 255) / 256);
    hipLaunchKernelGGL(snap_cast_shift, dim3((unsigned)nb), dim3(256), 0, s, src, dst, n, shift);
    return hipGetLastError();
}

}  // namespace sf
