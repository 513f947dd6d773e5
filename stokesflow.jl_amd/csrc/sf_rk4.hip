// sf_rk4.hip -- RK4 marker advection and seeded marker set-up for gfx950.
//
// moveMarkersRK4! (MoveMarkers.jl:15-69) with interpolateV (:76-174) and bilinearInterp
// (:181-187): one thread per marker, four stages of two bilinear gathers plus the
// continuity-based second-order correction, the failsafe index clamps, and the final
// clamp to [0, lx] x [0, ly].  Same IEEE operations in the same order as the reference,
// so positions are bit-identical to a CPU evaluation.  Marker arrays keep their order.
#include "sf_device.h"
#include "sf_internal.h"

namespace sf {

enum { RK_CLAMPS = 0, RK_NONFINITE = 1 };

__device__ inline bool key_ok_rk(double v) { return __builtin_isfinite(v) && v > -9.2e18 && v < 9.2e18; }

__device__ inline void topleft_rk(double xmin, double ymin, double x, double y, double dx, double dy, int64_t& ix,
                                  int64_t& iy, double& dxij, double& dyij)
{
    ix = (int64_t)floor((x - xmin) / dx) + 1;
    iy = (int64_t)floor((y - ymin) / dy) + 1;
    const double x_ix = xmin + (double)(ix - 1) * dx;
    const double y_iy = ymin + (double)(iy - 1) * dy;
    dxij = (x - x_ix) / dx;
    dyij = (y - y_iy) / dy;
}

__device__ inline double bilinear(double v1, double v2, double v3, double v4, double dxij, double dyij)
{
    const double s1 = ((1.0 - dxij) * (1.0 - dyij)) * v1;
    const double s2 = (dxij * (1.0 - dyij)) * v2;
    const double s3 = ((1.0 - dxij) * dyij) * v3;
    const double s4 = (dxij * dyij) * v4;
    return ((s1 + s2) + s3) + s4;
}

// interpolateV; returns false for a non-finite / unrepresentable position.
__device__ inline bool interp_v(double x, double y, const double* __restrict__ Vx, int n1x, int n2x,
                                const double* __restrict__ Vy, int n1y, int n2y, double x_vx_min, double y_vx_min,
                                double x_vy_min, double y_vy_min, double dx, double dy, double& vx, double& vy,
                                int& nclamp)
{
    if (!key_ok_rk((x - x_vx_min) / dx) || !key_ok_rk((y - y_vx_min) / dy) || !key_ok_rk((x - x_vy_min) / dx) ||
        !key_ok_rk((y - y_vy_min) / dy))
        return false;
    // 1-based accessors as in the reference
#define VX(i, j) Vx[(int64_t)((i) - 1) + (int64_t)((j) - 1) * n1x]
#define VY(i, j) Vy[(int64_t)((i) - 1) + (int64_t)((j) - 1) * n1y]
    int64_t ix, iy;
    double dxij, dyij;
    topleft_rk(x_vx_min, y_vx_min, x, y, dx, dy, ix, iy, dxij, dyij);
    if (ix < 2) { ix = 2; dxij = 0.0; ++nclamp; }
    if (iy < 1) { iy = 1; dyij = 0.0; ++nclamp; }
    if (ix >= n1x - 1) { ix = n1x - 2; dxij = 1.0; ++nclamp; }
    if (iy >= n2x) { iy = n2x - 1; dyij = 1.0; ++nclamp; }
    double v1 = VX(ix, iy), v2 = VX(ix + 1, iy), v3 = VX(ix, iy + 1), v4 = VX(ix + 1, iy + 1);
    double r = bilinear(v1, v2, v3, v4, dxij, dyij);
    double corr;
    if (dxij >= 0.5) {
        const double v5 = VX(ix + 2, iy), v6 = VX(ix + 2, iy + 1);
        corr = (0.5 * ((dxij - 0.5) * (dxij - 0.5))) * ((1.0 - dyij) * ((v1 - 2.0 * v2) + v5) + dyij * ((v3 - 2.0 * v4) + v6));
    } else {
        const double v5 = VX(ix - 1, iy), v6 = VX(ix - 1, iy + 1);
        corr = (0.5 * ((dxij - 0.5) * (dxij - 0.5))) * ((1.0 - dyij) * ((v5 - 2.0 * v1) + v2) + dyij * ((v6 - 2.0 * v3) + v4));
    }
    vx = r + corr;

    topleft_rk(x_vy_min, y_vy_min, x, y, dx, dy, ix, iy, dxij, dyij);
    if (ix < 1) { ix = 1; dxij = 0.0; ++nclamp; }
    if (iy < 2) { iy = 2; dyij = 0.0; ++nclamp; }
    if (ix >= n1y) { ix = n1y - 1; dxij = 1.0; ++nclamp; }
    if (iy >= n2y - 1) { iy = n2y - 2; dyij = 1.0; ++nclamp; }
    v1 = VY(ix, iy); v2 = VY(ix + 1, iy); v3 = VY(ix, iy + 1); v4 = VY(ix + 1, iy + 1);
    r = bilinear(v1, v2, v3, v4, dxij, dyij);
    if (dyij >= 0.5) {
        const double v5 = VY(ix, iy + 2), v6 = VY(ix + 1, iy + 2);
        corr = (0.5 * ((dyij - 0.5) * (dyij - 0.5))) * ((1.0 - dxij) * ((v1 - 2.0 * v3) + v5) + dxij * ((v2 - 2.0 * v4) + v6));
    } else {
        const double v5 = VY(ix, iy - 1), v6 = VY(ix + 1, iy - 1);
        corr = (0.5 * ((dyij - 0.5) * (dyij - 0.5))) * ((1.0 - dxij) * ((v5 - 2.0 * v1) + v3) + dxij * ((v6 - 2.0 * v2) + v4));
    }
    vy = r + corr;
#undef VX
#undef VY
    return true;
}

__global__ __launch_bounds__(256) void rk4_kernel(double* __restrict__ x_m, double* __restrict__ y_m, int64_t nm,
                                                  const double* __restrict__ Vx, int n1x, int n2x,
                                                  const double* __restrict__ Vy, int n1y, int n2y, double x_vx_min,
                                                  double y_vx_min, double x_vy_min, double y_vy_min, double dt_val,
                                                  const double* __restrict__ dt_ptr, double lx, double ly, double dx,
                                                  double dy, unsigned long long* __restrict__ counters)
{
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nm) return;
    const double dt = dt_ptr ? *dt_ptr : dt_val;  // a device dt: set by the solve earlier in a captured step
    const double rk_dt[4] = {0.0, 0.5 * dt, 0.5 * dt, dt};
    const double rk_wt[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
    const double x_old = x_m[m], y_old = y_m[m];
    double vx_eff = 0.0, vy_eff = 0.0, vx_rk = 0.0, vy_rk = 0.0;
    int nclamp = 0;
    for (int it = 0; it < 4; ++it) {
        const double x_rk = x_old + rk_dt[it] * vx_rk;
        const double y_rk = y_old + rk_dt[it] * vy_rk;
        if (!interp_v(x_rk, y_rk, Vx, n1x, n2x, Vy, n1y, n2y, x_vx_min, y_vx_min, x_vy_min, y_vy_min, dx, dy, vx_rk,
                      vy_rk, nclamp)) {
            atomicAdd(&counters[RK_NONFINITE], 1ull);
            return;
        }
        vx_eff = vx_eff + rk_wt[it] * vx_rk;
        vy_eff = vy_eff + rk_wt[it] * vy_rk;
    }
    const double x_new = x_old + vx_eff * dt;
    const double y_new = y_old + vy_eff * dt;
    x_m[m] = jl_min(jl_max(x_new, 0.0), lx);
    y_m[m] = jl_min(jl_max(y_new, 0.0), ly);
    if (nclamp) atomicAdd(&counters[RK_CLAMPS], (unsigned long long)nclamp);
}

// ---------------------------------------------------------------------------------------
// Seeded marker set-up (the reference builds its lattice on the host, common.jl:40-94; the
// benchmark workloads use md x md markers per cell with a seeded jitter instead).

__device__ inline uint64_t splitmix64(uint64_t seed, uint64_t counter)
{
    uint64_t z = seed + (counter + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ inline double uniform01(uint64_t seed, uint64_t counter)
{
    return (double)(splitmix64(seed, counter) >> 11) * 0x1.0p-53;
}

__global__ __launch_bounds__(256) void init_markers_kernel(double* __restrict__ x_m, double* __restrict__ y_m, int64_t nmx,
                                                           int64_t nmy, double dxm, double dym, double jx, double jy,
                                                           double lx, double ly, uint64_t seed)
{
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmx * nmy) return;
    const int64_t i = m / nmy, j = m - i * nmy;  // x outer, y inner (common.jl:76-83)
    const double x = ((double)i + 0.5) * dxm + (uniform01(seed, 2 * (uint64_t)m) - 0.5) * jx;
    const double y = ((double)j + 0.5) * dym + (uniform01(seed, 2 * (uint64_t)m + 1) - 0.5) * jy;
    x_m[m] = jl_min(jl_max(x, 0.0), lx);
    y_m[m] = jl_min(jl_max(y, 0.0), ly);
}

// Plume material: mat = {air_h, cx, cy, r, rho_air, rho_plume, rho_bg, mu_air, mu_plume, mu_bg}
__global__ __launch_bounds__(256) void plume_props_kernel(const double* __restrict__ x_m, const double* __restrict__ y_m,
                                                          int64_t nm, Mat10 mat,
                                                          double* __restrict__ rho_m, double* __restrict__ mu_m)
{
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nm) return;
    const double x = x_m[m], y = y_m[m];
    const double ddx = x - mat.v[1], ddy = y - mat.v[2];
    if (y < mat.v[0]) {
        rho_m[m] = mat.v[4];
        mu_m[m] = mat.v[7];
    } else if (ddx * ddx + ddy * ddy < mat.v[3] * mat.v[3]) {
        rho_m[m] = mat.v[5];
        mu_m[m] = mat.v[8];
    } else {
        rho_m[m] = mat.v[6];
        mu_m[m] = mat.v[9];
    }
}

// Two-layer Rayleigh-Taylor material: mat = {y0, amp, k, lx, rho_up, mu_up, rho_low, mu_low}.
// The upper (small-y) layer lies above y0 + amp * sin(2 pi k x / lx).
__global__ __launch_bounds__(256) void rt_props_kernel(const double* __restrict__ x_m, const double* __restrict__ y_m,
                                                       int64_t nm, Mat10 mat,
                                                       double* __restrict__ rho_m, double* __restrict__ mu_m)
{
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nm) return;
    const double yi = mat.v[0] + mat.v[1] * sin(2.0 * M_PI * mat.v[2] * x_m[m] / mat.v[3]);
    const bool up = y_m[m] < yi;
    rho_m[m] = up ? mat.v[4] : mat.v[6];
    mu_m[m] = up ? mat.v[5] : mat.v[7];
}

// The reference's deterministic marker lattice (initializeMarkersCPU, common.jl:40-83, one rank:
// all four boundary flags true): LinRange(dxm/2, lx - dxm/2, md (Nx-1)) in x, likewise in y,
// x outer / y inner.  A LinRange element is evaluated as Julia 1.8 does, (1-t) a + t b with
// t = j / (len - 1), unfused.  rand_pos adds (u - 0.5) dxm, (u - 0.5) dym (common.jl:85-87) with u
// from splitmix64(seed, index) (the reference's Julia RNG stream is not reproduced).
__device__ inline double linrange_at(double a, double b, int64_t len, int64_t j)
{
    if (len == 1) return a;
    const double t = (double)j / (double)(len - 1);
    return (1.0 - t) * a + t * b;
}

__global__ __launch_bounds__(256) void lattice_markers_kernel(double* __restrict__ x_m, double* __restrict__ y_m,
                                                              int64_t nmx, int64_t nmy, double xlo, double xhi,
                                                              double ylo, double yhi, double dxm, double dym,
                                                              int rand_pos, uint64_t seed)
{
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmx * nmy) return;
    const int64_t i = m / nmy, j = m - i * nmy;
    double x = linrange_at(xlo, xhi, nmx, i);
    double y = linrange_at(ylo, yhi, nmy, j);
    if (rand_pos) {
        x = x + (uniform01(seed, 2 * (uint64_t)m) - 0.5) * dxm;
        y = y + (uniform01(seed, 2 * (uint64_t)m + 1) - 0.5) * dym;
    }
    x_m[m] = x;
    y_m[m] = y;
}

// Material descriptor in place of the reference's density / viscosity callbacks
// (setInitialMarkerProperties!, common.jl:104-110): the first shape that contains the global
// position (x0 + x_m, y0 + y_m) gives rho and mu, else the background.  Each test is written as
// the reference's example callbacks write theirs (StokesFlow2D_multixpu.jl:387-404:
// y < air_height, (x - cx)^2 + (y - cy)^2 < r^2).
__global__ __launch_bounds__(256) void shape_props_kernel(const double* __restrict__ x_m, const double* __restrict__ y_m,
                                                          int64_t nm, double x0, double y0, ShapeSet set,
                                                          double* __restrict__ rho_m, double* __restrict__ mu_m)
{
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nm) return;
    const double x = x0 + x_m[m], y = y0 + y_m[m];
    double rho = set.rho_bg, mu = set.mu_bg;
    for (int k = 0; k < set.n; ++k) {
        const sf_shape& sh = set.s[k];
        bool in = false;
        switch (sh.kind) {
        case SF_SHAPE_Y_BELOW: in = y < sh.a; break;
        case SF_SHAPE_Y_ABOVE: in = y >= sh.a; break;
        case SF_SHAPE_CIRCLE: {
            const double ddx = x - sh.a, ddy = y - sh.b;
            in = ddx * ddx + ddy * ddy < sh.c * sh.c;
            break;
        }
        case SF_SHAPE_RECT: in = x >= sh.a && x < sh.b && y >= sh.c && y < sh.d; break;
        case SF_SHAPE_SINE_BELOW: in = y < sh.a + sh.b * sin(2.0 * M_PI * sh.c * x / sh.d); break;
        default: break;
        }
        if (in) {
            rho = sh.rho;
            mu = sh.mu;
            break;
        }
    }
    rho_m[m] = rho;
    mu_m[m] = mu;
}

static int g1(int64_t n) { return (int)((n + 255) / 256); }

hipError_t launch_lattice_markers(double* x_m, double* y_m, int nx, int ny, int md, double lx, double ly, double dx,
                                  double dy, int rand_pos, uint64_t seed, hipStream_t s)
{
    const double dxm = dx / md, dym = dy / md;
    const int64_t nmx = (int64_t)md * (nx - 1), nmy = (int64_t)md * (ny - 1);
    if (nmx * nmy <= 0) return hipSuccess;
    hipLaunchKernelGGL(lattice_markers_kernel, dim3(g1(nmx * nmy)), dim3(256), 0, s, x_m, y_m, nmx, nmy, dxm / 2,
                       lx - dxm / 2, dym / 2, ly - dym / 2, dxm, dym, rand_pos, seed);
    return hipGetLastError();
}

hipError_t launch_shape_properties(const double* x_m, const double* y_m, int64_t nm, double x0, double y0,
                                   const ShapeSet& set, double* rho_m, double* mu_m, hipStream_t s)
{
    if (nm <= 0) return hipSuccess;
    hipLaunchKernelGGL(shape_props_kernel, dim3(g1(nm)), dim3(256), 0, s, x_m, y_m, nm, x0, y0, set, rho_m, mu_m);
    return hipGetLastError();
}

hipError_t launch_rk4(double* x_m, double* y_m, int64_t nm, const double* vx, int vx_n1, int vx_n2, const double* vy,
                      int vy_n1, int vy_n2, double x_vx_min, double y_vx_min, double x_vy_min, double y_vy_min,
                      double dt, double lx, double ly, double dx, double dy, unsigned long long* counters, hipStream_t s,
                      const double* dt_ptr)
{
    if (nm <= 0) return hipSuccess;
    hipLaunchKernelGGL(rk4_kernel, dim3(g1(nm)), dim3(256), 0, s, x_m, y_m, nm, vx, vx_n1, vx_n2, vy, vy_n1, vy_n2,
                       x_vx_min, y_vx_min, x_vy_min, y_vy_min, dt, dt_ptr, lx, ly, dx, dy, counters);
    return hipGetLastError();
}

hipError_t launch_init_markers(double* x_m, double* y_m, int nx, int ny, int md, double dx, double dy,
                               double jitter_frac, uint64_t seed, hipStream_t s)
{
    const int64_t nmx = (int64_t)md * (nx - 1), nmy = (int64_t)md * (ny - 1);
    const double dxm = dx / md, dym = dy / md;
    const double lx = (double)(nx - 1) * dx, ly = (double)(ny - 1) * dy;
    const double jx = 2.0 * jitter_frac * dx, jy = 2.0 * jitter_frac * dy;
    hipLaunchKernelGGL(init_markers_kernel, dim3(g1(nmx * nmy)), dim3(256), 0, s, x_m, y_m, nmx, nmy, dxm, dym, jx, jy,
                       lx, ly, seed);
    return hipGetLastError();
}

hipError_t launch_plume_properties(const double* x_m, const double* y_m, int64_t nm, Mat10 mat, double* rho_m,
                                   double* mu_m, hipStream_t s)
{
    if (nm <= 0) return hipSuccess;
    hipLaunchKernelGGL(plume_props_kernel, dim3(g1(nm)), dim3(256), 0, s, x_m, y_m, nm, mat, rho_m, mu_m);
    return hipGetLastError();
}

hipError_t launch_rt_properties(const double* x_m, const double* y_m, int64_t nm, Mat10 mat, double* rho_m,
                                double* mu_m, hipStream_t s)
{
    if (nm <= 0) return hipSuccess;
    hipLaunchKernelGGL(rt_props_kernel, dim3(g1(nm)), dim3(256), 0, s, x_m, y_m, nm, mat, rho_m, mu_m);
    return hipGetLastError();
}

}  // namespace sf
