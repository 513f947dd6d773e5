// sf_probe.hip -- DRAM bandwidth probe: a copy (1 read + 1 write stream) and a 3-read +
// 1-write kernel over arrays far larger than the caches, 16-byte accesses, grid-stride.
// It measures the ceiling that the stencil kernels' bandwidth is read against.
#include "sf_internal.h"

namespace sf {

__global__ __launch_bounds__(256) void probe_copy(const double4* __restrict__ a, double4* __restrict__ b, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) b[i] = a[i];
}

__global__ __launch_bounds__(256) void probe_r3w1(const double2* __restrict__ a, const double2* __restrict__ b,
                                                  const double2* __restrict__ c, double2* __restrict__ d, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const double2 x = a[i], y = b[i], z = c[i];
        d[i] = make_double2(x.x + y.x * z.x, x.y + y.y * z.y);
    }
}

hipError_t launch_probe_copy(const double* a, double* b, int64_t bytes, int grid, hipStream_t s)
{
    hipLaunchKernelGGL(probe_copy, dim3(grid), dim3(256), 0, s, (const double4*)a, (double4*)b, bytes / 32);
    return hipGetLastError();
}

hipError_t launch_probe_r3w1(const double* a, const double* b, const double* c, double* d, int64_t bytes, int grid,
                             hipStream_t s)
{
    hipLaunchKernelGGL(probe_r3w1, dim3(grid), dim3(256), 0, s, (const double2*)a, (const double2*)b,
                       (const double2*)c, (double2*)d, bytes / 16);
    return hipGetLastError();
}

}  // namespace sf
