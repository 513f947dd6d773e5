// sf_pt.hip -- pseudo-transient Stokes iteration kernels for gfx950 (FP64 path).
//
// One PT iteration of the reference (StokesSolver_multixpu.jl:48-58: compute_P_tau!,
// compute_dV!, compute_V! and the four free-slip BC kernels) is ONE launch here:
//
//   * The state (Vx_s, Vy_s, P) is ping-ponged between the caller's arrays and a scratch
//     copy (A -> B), so a tile may read its neighbours' old values while other tiles
//     write new ones.  dVx/dVy are updated in place (each element is read and written
//     by the same lane).  Per cell the kernel moves exactly the algorithmic bytes:
//     read Vx, Vy, P, dVx, dVy, mu_p, mu_b, rho; write Vx, Vy, P, dVx, dVy (13 x 8 B).
//   * dtauVx, dtauVy, dtauP (compute_timesteps!, :128-133) and tau/divV are recomputed
//     in registers instead of being stored and re-read.
//   * Each 64-lane wave owns a strip of 64 columns and marches down the rows.  Row
//     neighbours stay in registers (sliding window); the x neighbours c+1, c+2 come from
//     lanes +1, +2.  The stencil only reaches right/down in this indexing, so lanes
//     0..61 produce outputs and lanes 62, 63 are the strip's halo (strips overlap by 2).
//   * Free-slip BCs (:53-56, order y-noflux then x-zero for Vx, y-zero then x-noflux for
//     Vy) are written by the lanes that own the adjacent interior values.
//   * The block -> tile map is XCD-aware: the 8 XCDs each receive a contiguous band of
//     tiles, so the strip/segment overlaps are shared through one L2.
//
// Check iterations (every ncheck) additionally store txx, tyy, txy, divV and Rx, then
// pt_resid_y computes Ry with post-update V and dt_check exactly as compute_Residuals!
// (:241-249).  Reductions use a fixed block count and a fixed combine order, so err and
// dt are deterministic.
#include <algorithm>

#include "sf_ptk_tile.h"

namespace sf {


// Launch: one wave per tile (max_waves < 0) or at most G waves grid-striding over the
// tiles.  In each round the waves of one XCD take a contiguous band of tiles (strips
// fastest), so the rows and columns two tiles share are read once through that XCD's L2.
template <int DEPTH, bool CHECK>
__global__ __launch_bounds__(64) void pt_fused_f64(PTParams p)
{
    const int G = gridDim.x;
    for (int base = 0; base < p.ntiles; base += G) {
        const int nround = min(G, p.ntiles - base);
        if ((int)blockIdx.x >= nround) break;
        const int logical = base + xcd_remap(blockIdx.x, nround);
        pt_tile<DEPTH, CHECK>(p, logical % p.nstrips, logical / p.nstrips, p.flags);
    }
}

// One wave per tile, no grid-stride loop (fewer live registers: 4 waves per SIMD).
template <int DEPTH, bool CHECK>
__global__ __launch_bounds__(64) void pt_tile_f64(PTParams p)
{
    const int logical = xcd_remap(blockIdx.x, gridDim.x);
    pt_tile<DEPTH, CHECK>(p, logical % p.nstrips, logical / p.nstrips, p.flags);
}

// Two PT iterations per pass, one wave per tile (sf_pt2_tile.h).
__global__ __launch_bounds__(64) void pt2_tile_f64(PTParams p)
{
    const int logical = xcd_remap(blockIdx.x, gridDim.x);
    pt2_tile(p, logical % p.nstrips, logical / p.nstrips);
}

// K PT iterations per pass, one wave per tile (sf_ptk_tile.h): the plain loop form for K = 3,
// the split form with one middle step per trip for K = 4.
template <int K>
__global__ __launch_bounds__(64) void ptk_tile_f64(PTParams p)
{
    const int logical = xcd_remap(blockIdx.x, gridDim.x);
    ptk_tile<K, K == 4 ? 1 : 0>(p, logical % p.nstrips, logical / p.nstrips);
}

// compute_ResY with post-update V and dt_check (compute_Residuals!, :241-249).
__global__ __launch_bounds__(256) void pt_resid_y(ResYParams p)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    const int nx = p.nx, ny = p.ny;
    if (i >= nx - 1 || j >= ny - 2) return;
    const int64_t lp = nx - 1, lvx = nx, lvy = nx + 1, lr = nx + 1;
    const double dt = *p.dt;
    const double rdx = p.rdx, rdy = p.rdy;
    const double av = 0.25 * (((p.vx[ix2(i, j + 1, lvx)] + p.vx[ix2(i + 1, j + 1, lvx)]) + p.vx[ix2(i, j + 2, lvx)]) +
                              p.vx[ix2(i + 1, j + 2, lvx)]);
    const double d_xi = p.rho[ix2(i + 2, j + 1, lr)] - p.rho[ix2(i, j + 1, lr)];
    const double d_yi = p.rho[ix2(i + 1, j + 2, lr)] - p.rho[ix2(i + 1, j, lr)];
    const double ry =
        (((p.tyy[ix2(i, j + 1, lp)] - p.tyy[ix2(i, j, lp)]) * rdy + (p.txy[ix2(i + 1, j + 1, lvx)] - p.txy[ix2(i, j + 1, lvx)]) * rdx) -
         (p.P[ix2(i, j + 1, lp)] - p.P[ix2(i, j, lp)]) * rdy) +
        p.g * (p.rho[ix2(i + 1, j + 1, lr)] -
               dt * (((av * d_xi) * 0.5) * rdx + ((p.vy[ix2(i + 1, j + 1, lvy)] * d_yi) * 0.5) * rdy));
    p.ry[ix2(i, j, nx - 1)] = ry;
}

// Per-block partial |a| sums or maxima over up to three arrays (blockIdx.y picks the
// array).  Fixed grid and a fixed tree order: the result depends only on the data.
__global__ __launch_bounds__(256) void reduce_abs_partials(RedParams p)
{
    __shared__ double sh[256];
    const int a = blockIdx.y;
    const double* __restrict__ src = p.a[a];
    const int64_t n = p.n[a];
    const int t = threadIdx.x;
    double acc = p.mode == RED_SUM ? 0.0 : -__builtin_inf();
    for (int64_t k = (int64_t)blockIdx.x * 256 + t; k < n; k += (int64_t)gridDim.x * 256) {
        const double v = fabs(src[k]);
        acc = (p.mode == RED_SUM) ? acc + v : jl_max(acc, v);
    }
    sh[t] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) sh[t] = (p.mode == RED_SUM) ? sh[t] + sh[t + s] : jl_max(sh[t], sh[t + s]);
        __syncthreads();
    }
    if (t == 0) p.part[(int64_t)a * gridDim.x + blockIdx.x] = sh[0];
}

// dt = maxdisp * min(dx / max|Vx|, dy / max|Vy|)  (compute_dt, :118-121)
// One wave: each lane takes the maximum of every 64th block partial, then the lanes combine.
// jl_max is commutative and associative (NaN wins, +0 beats -0 either way round), so the result
// does not depend on the order, and equals a serial pass over the partials.
__global__ void finalize_dt(const double* __restrict__ part, int nb, double dx, double dy, double maxdisp,
                            double* __restrict__ scal, int update_dt)
{
    const int lane = threadIdx.x;
    double mx = -__builtin_inf(), my = -__builtin_inf();
    for (int b = lane; b < nb; b += 64) {
        mx = jl_max(mx, part[b]);
        my = jl_max(my, part[nb + b]);
    }
    for (int d = 32; d >= 1; d >>= 1) {
        mx = jl_max(mx, __shfl_xor(mx, d, 64));
        my = jl_max(my, __shfl_xor(my, d, 64));
    }
    if (lane != 0) return;
    const double dtc = maxdisp * jl_min(dx / mx, dy / my);
    scal[SC_DT_CHECK] = dtc;
    if (update_dt) scal[SC_DT] = dtc;
}

// err = max(mean|Rx|, mean|Ry|, mean|divV|)  (compute_err, :103-109)
__global__ void finalize_err(const double* __restrict__ part, int nb, double n_rx, double n_ry, double n_dv,
                             double* __restrict__ scal)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int b = 0; b < nb; ++b) {
        s0 = s0 + part[b];
        s1 = s1 + part[nb + b];
        s2 = s2 + part[2 * nb + b];
    }
    scal[SC_ERR] = jl_max(jl_max(s0 / n_rx, s1 / n_ry), s2 / n_dv);
}

__global__ void set_scalar(double* __restrict__ dst, double v) { *dst = v; }

__global__ void zero_u64(unsigned long long* __restrict__ dst, int n)
{
    for (int k = threadIdx.x; k < n; k += blockDim.x) dst[k] = 0ull;
}

// Vx_pad[2:end-1,:] .= Vx_s; Vy_pad[:,2:end-1] .= Vy_s; bc_x_mirror!; bc_y_mirror!  (:89-92)
__global__ __launch_bounds__(256) void pad_mirror(int nx, int ny, const double* __restrict__ vx, const double* __restrict__ vy,
                                                  double* __restrict__ vx_pad, double* __restrict__ vy_pad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i < nx + 2 && j < ny + 1) {
        double v;
        if (i == 0) v = -vx[ix2(1, j, nx)];
        else if (i == nx + 1) v = -vx[ix2(nx - 2, j, nx)];
        else v = vx[ix2(i - 1, j, nx)];
        vx_pad[ix2(i, j, nx + 2)] = v;
    }
    if (i < nx + 1 && j < ny + 2) {
        double v;
        if (j == 0) v = -vy[ix2(i, 1, nx + 1)];
        else if (j == ny + 1) v = -vy[ix2(i, ny - 2, nx + 1)];
        else v = vy[ix2(i, j - 1, nx + 1)];
        vy_pad[ix2(i, j, nx + 1)] = v;
    }
}

// compute_timesteps! (:128-133) for callers that want the dtau arrays filled.
__global__ __launch_bounds__(256) void fill_timesteps(int nx, int ny, const double* __restrict__ mu_p, double* dtauvx,
                                                      double* dtauvy, double* dtaup, double vsc_mdxy2, double ptsc41,
                                                      double max_nxy)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    const int64_t lp = nx - 1;
    if (dtauvx && i < nx - 2 && j < ny - 1)
        dtauvx[ix2(i, j, nx - 2)] = (vsc_mdxy2 / ((mu_p[ix2(i, j, lp)] + mu_p[ix2(i + 1, j, lp)]) * 0.5)) / 4.1;
    if (dtauvy && i < nx - 1 && j < ny - 2)
        dtauvy[ix2(i, j, nx - 1)] = (vsc_mdxy2 / ((mu_p[ix2(i, j, lp)] + mu_p[ix2(i, j + 1, lp)]) * 0.5)) / 4.1;
    if (dtaup && i < nx - 1 && j < ny - 1) dtaup[ix2(i, j, lp)] = (ptsc41 * mu_p[ix2(i, j, lp)]) / max_nxy;
}

// ---------------------------------------------------------------------------------------
// host-side launchers

int pt_num_strips(int nx) { return (nx + PT_OUT_LANES - 1) / PT_OUT_LANES; }

// Waves of pt_fused_f64<depth> the current device holds at once (occupancy x CUs).
int pt_resident_waves(int depth)
{
    static int cached[64][2] = {};
    int dev = 0;
    const int d = depth > 1 ? 1 : 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 4096;
    if (cached[dev][d] == 0) {
        int per_cu = 0, cus = 0;
        hipError_t e = d ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pt_fused_f64<2, false>, 64, 0)
                         : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pt_fused_f64<1, false>, 64, 0);
        if (e != hipSuccess || per_cu < 1) per_cu = 4;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
        cached[dev][d] = per_cu * cus;
    }
    return cached[dev][d];
}

// Rows per tile.  Short tiles measured fastest on MI355X at every size from 512^2 to 4096^2
// (one wave per tile, all strips sweeping the same few rows at once; the one-row prologue
// a tile re-reads was just fetched into the same XCD's L2 by the tile above it).
// Default launch shape (measured on MI355X with plume data): 2-row tiles on grids up to
// 1024^2 (more waves), 4-row tiles above; one row of operands in flight (124 VGPRs, four
// waves per SIMD, beat two rows in flight at three waves per SIMD).
int pt_default_rows(int nx, int ny)
{
    const int r = (std::max(nx, ny) <= 1024) ? 2 : 4;
    return ny >= r ? r : ny;
}

int pt_default_depth(int nx, int ny)
{
    (void)nx;
    (void)ny;
    return 1;
}

hipError_t launch_pt_fused(const PTParams& p0, hipStream_t s)
{
    PTParams p = p0;
    if (p.seg <= 0) p.seg = pt_default_rows(p.nx, p.ny);
    const int depth = p.depth <= 0 ? pt_default_depth(p.nx, p.ny) : (p.depth > 1 ? 2 : 1);
    p.nstrips = pt_num_strips(p.nx);
    const int nseg = (p.ny + p.seg - 1) / p.seg;
    p.ntiles = p.nstrips * nseg;
    // max_waves: 0 = one wave per resident slot (persistent), < 0 = one wave per tile
    int grid = p.ntiles;
    if (p.max_waves == 0) grid = std::min(p.ntiles, pt_resident_waves(depth));
    else if (p.max_waves > 0) grid = std::min(p.ntiles, p.max_waves);
    const bool check = p.flags != 0;
    if (grid == p.ntiles) {  // one wave per tile
        if (depth == 2) {
            if (check) hipLaunchKernelGGL((pt_tile_f64<2, true>), dim3(grid), dim3(64), 0, s, p);
            else hipLaunchKernelGGL((pt_tile_f64<2, false>), dim3(grid), dim3(64), 0, s, p);
        } else {
            if (check) hipLaunchKernelGGL((pt_tile_f64<1, true>), dim3(grid), dim3(64), 0, s, p);
            else hipLaunchKernelGGL((pt_tile_f64<1, false>), dim3(grid), dim3(64), 0, s, p);
        }
    } else if (depth == 2) {
        if (check) hipLaunchKernelGGL((pt_fused_f64<2, true>), dim3(grid), dim3(64), 0, s, p);
        else hipLaunchKernelGGL((pt_fused_f64<2, false>), dim3(grid), dim3(64), 0, s, p);
    } else {
        if (check) hipLaunchKernelGGL((pt_fused_f64<1, true>), dim3(grid), dim3(64), 0, s, p);
        else hipLaunchKernelGGL((pt_fused_f64<1, false>), dim3(grid), dim3(64), 0, s, p);
    }
    return hipGetLastError();
}

// Rows per tile of the two-iteration pass: each tile recomputes 3 rows of the first
// iteration and 1 of the second around its segment, so taller tiles cut the re-reads, while
// smaller grids need short tiles for enough waves (2 per SIMD at 172 VGPRs).  Measured on
// MI355X (us per iteration, one-iteration kernel in brackets): 4096^2 16 rows 193-198 (356),
// 2048^2 12 rows 51.6 (90.7), 1024^2 12 rows 14.2 (19.8), 512^2 4 rows 6.6 (7.0).  Up to
// 256^2 a launch is latency-bound, and 2-row tiles (more waves, shorter row chains) gave 2.8-3.0
// us per iteration against 3.6-3.9 with 4 rows in the solver loop.
int pt2_default_rows(int nx, int ny)
{
    const int n = std::max(nx, ny);
    const int r = n > 2048 ? 16 : (n > 768 ? 12 : (n > 512 ? 8 : (n > 256 ? 4 : 2)));
    return ny >= r ? r : ny;
}

hipError_t launch_pt_pair(const PTParams& p0, hipStream_t s)
{
    PTParams p = p0;
    if (p.seg <= 0) p.seg = pt2_default_rows(p.nx, p.ny);
    p.nstrips = (p.nx - 1 + PT2_OUT_LANES - 1) / PT2_OUT_LANES;  // columns 0..nx-2 carry outputs
    const int nseg = (p.ny + p.seg - 1) / p.seg;
    p.ntiles = p.nstrips * nseg;
    hipLaunchKernelGGL(pt2_tile_f64, dim3(p.ntiles), dim3(64), 0, s, p);
    return hipGetLastError();
}

// Iterations per pass by grid size.  Pass timer, us per iteration at each K's best tile height,
// K = 4 / 3 / 2: 4096^2 108.7 / 129.1 / 190.4; 2048^2 30.0 / 34.3 / 50.7; 1024^2 11.2 / 11.5 / 12.7.
// Small grids are launch- and latency-bound, where the pass timer's own events weigh on the
// shorter passes, so there the solver's timing window decides (us per iteration, 2000 iterations,
// K = 4 / 3 / 2): 1024^2 9.70 / 9.80 / 11.90; 768^2 7.22 / 7.25 / 8.49; 512^2 4.96 / 5.03 / 5.53;
// 384^2 4.09 / 4.02 / 3.96; 256^2 3.18 / 3.04 / 3.00; 128^2 3.03 / 2.85 / 2.89; 64^2 3.05 / 2.87
// / 2.79.  Four iterations per pass from 449 nodes per side, two below.
int pt_default_fusion(int nx, int ny) { return std::max(nx, ny) > 448 ? 4 : 2; }

// Sweeps per pass of the FP32 correction solver when the context leaves it open (three from 769
// nodes per side; us per sweep, K=3 against K=2: 1024^2 7.82 (8-row tiles) against 8.8, 1280^2
// 9.54 against 11.2, 1536^2 12.45 against 14.45).  Above 1536, four (interleaved repetitions on
// one box, us per sweep, four against three: 4096^2 79.5 against 87.7, 3072^2 44.9 against 53.7,
// 2048^2 18.7 against 19.8); four sweeps move fewer bytes per sweep at two waves per SIMD and
// measured 78-80 us at 4096^2 on every box, where three ranged 76-92.  Above 3072, five (225
// VGPRs, still two waves per SIMD): 4096^2 74.6 against 79.4 for four; 2048^2 18.7 for both.
int corr_default_fusion(int nx, int ny)
{
    const int n = std::max(nx, ny);
    return n > 3072 ? 5 : (n > 1536 ? 4 : (n > 768 ? 3 : 2));
}

//
// The tile height of a K-iteration pass comes from a wave-round model.  A launch runs in rounds of
// resident waves (two per SIMD: 198 VGPRs at K = 3, 251 at K = 4) and a partly filled last round
// costs a whole one, so the height minimises rounds x (rows + 2K - 1 halo rows); a launch of at
// most one wave per SIMD runs each wave 1.38x faster (measured at 512^2, K = 4: 6-row tiles, 860
// waves, 1.70 us per row step; 4-row tiles, 1280 waves, 2.35 us), which is what small grids reach
// with short tiles.  Candidates: 2-12 rows in steps of 2, then 16-64 in steps of 4.  Measured (pass
// timer, us per iteration, model pick first), K = 4 in the plain loop form: 4096^2 52 rows 108.7,
// 44 111.7, 40 112.2, 60 110.1, 48 118.1; 3584^2 56 rows 80.2, 28 84.6; 3072^2 44 rows 63.5, 28
// 64.2, 24 66.5; 2560^2 60 rows 43.5, 32 47.7; 2048^2 40 rows 30.0, 16 34.7, 48 34.1; 1536^2 24
// rows 18.2, 16 19.8; 1280^2 16 rows 13.5, 8 16.0; 1024^2 10 rows 10.1, 12 11.2, 8 12.7; 768^2 6
// rows 7.75, 12 7.9, 8 8.9; 512^2 6 rows 5.5, 8 6.3, 4 6.5, 12 7.9; 64^2-256^2 2 rows.  K = 3 in
// the solver loop: 4096^2 48 rows 136.0, 32 141.8; 3072^2 40 rows 74.0, 28 76.7; 2048^2 40 rows
// 34.4, 20 36.6; 1536^2 24 rows 19.7, 12 21.2.
int ptk_default_rows(int k, int nx, int ny, int copies)
{
    static const int simds = [] {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        return 4 * cus;
    }();
    const int out = 64 - 2 * k;
    const int64_t strips = (nx - 1 + out - 1) / out;
    int best = 16;
    double best_cost = 1e300;
    for (int rows = 2; rows <= 64; rows += rows < 12 ? 2 : 4) {
        const int64_t waves = (int64_t)copies * strips * ((ny + rows - 1) / rows);
        const double steps = rows + 2 * k - 1;
        const double cost = waves <= simds ? steps : 1.38 * (double)((waves + 2 * simds - 1) / (2 * simds)) * steps;
        if (cost < best_cost) {
            best_cost = cost;
            best = rows;
        }
    }
    return ny >= best ? best : ny;
}

hipError_t launch_pt_multi(const PTParams& p0, int k, hipStream_t s)
{
    if (k == 2) return launch_pt_pair(p0, s);
    if (k != 3 && k != 4) return hipErrorInvalidValue;
    PTParams p = p0;
    if (p.seg <= 0) p.seg = ptk_default_rows(k, p.nx, p.ny);
    p.nstrips = (p.nx - 1 + (64 - 2 * k) - 1) / (64 - 2 * k);  // 64 - 2 K output columns per strip
    const int nseg = (p.ny + p.seg - 1) / p.seg;
    p.ntiles = p.nstrips * nseg;
    if (k == 3) hipLaunchKernelGGL(ptk_tile_f64<3>, dim3(p.ntiles), dim3(64), 0, s, p);
    else hipLaunchKernelGGL(ptk_tile_f64<4>, dim3(p.ntiles), dim3(64), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_pt_resid_y(const ResYParams& p, hipStream_t s)
{
    dim3 grid((p.nx - 1 + 255) / 256, p.ny - 2);
    hipLaunchKernelGGL(pt_resid_y, grid, dim3(256), 0, s, p);
    return hipGetLastError();
}

int red_num_blocks(int64_t nmax)
{
    int64_t nb = (nmax + 2047) / 2048;
    if (nb < 1) nb = 1;
    if (nb > RED_MAX_BLOCKS) nb = RED_MAX_BLOCKS;
    return (int)nb;
}

hipError_t launch_reduce(const RedParams& p, int narr, int nb, hipStream_t s)
{
    hipLaunchKernelGGL(reduce_abs_partials, dim3(nb, narr), dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_finalize_dt(const double* part, int nb, double dx, double dy, double maxdisp, double* scal,
                              int update, hipStream_t s)
{
    hipLaunchKernelGGL(finalize_dt, dim3(1), dim3(64), 0, s, part, nb, dx, dy, maxdisp, scal, update);
    return hipGetLastError();
}

hipError_t launch_finalize_err(const double* part, int nb, double nrx, double nry, double ndv, double* scal,
                               hipStream_t s)
{
    hipLaunchKernelGGL(finalize_err, dim3(1), dim3(64), 0, s, part, nb, nrx, nry, ndv, scal);
    return hipGetLastError();
}

hipError_t launch_set_scalar(double* dst, double v, hipStream_t s)
{
    hipLaunchKernelGGL(set_scalar, dim3(1), dim3(1), 0, s, dst, v);
    return hipGetLastError();
}

hipError_t launch_zero_u64(unsigned long long* dst, int n, hipStream_t s)
{
    hipLaunchKernelGGL(zero_u64, dim3(1), dim3(64), 0, s, dst, n);
    return hipGetLastError();
}

hipError_t launch_pad_mirror(int nx, int ny, const double* vx, const double* vy, double* vx_pad, double* vy_pad,
                             hipStream_t s)
{
    dim3 grid((nx + 2 + 255) / 256, ny + 2);
    hipLaunchKernelGGL(pad_mirror, grid, dim3(256), 0, s, nx, ny, vx, vy, vx_pad, vy_pad);
    return hipGetLastError();
}

hipError_t launch_fill_timesteps(int nx, int ny, const double* mu_p, double* dtauvx, double* dtauvy, double* dtaup,
                                 double vsc_mdxy2, double ptsc41, double max_nxy, hipStream_t s)
{
    dim3 grid((nx + 255) / 256, ny);
    hipLaunchKernelGGL(fill_timesteps, grid, dim3(256), 0, s, nx, ny, mu_p, dtauvx, dtauvy, dtaup, vsc_mdxy2, ptsc41,
                       max_nxy);
    return hipGetLastError();
}

}  // namespace sf
