// sf_batch.hip -- batched ensemble of independent Stokes solves in one persistent launch.
//
// Each workgroup (eight waves: one CU) takes an instance from a device work queue (largest
// first), runs that instance's whole PT solve -- the loop of solveStokes!
// (StokesSolver_multixpu.jl:43-77) with its ncheck / ndtupdate rules, the err and dt
// reductions and the padded/mirror epilogue (:89-95) -- and takes the next instance when
// it converged or hit iterMax.  Plain iterations run three per pass (ptk_tile<3>, the rule of
// sf_solve_stokes), the others one per pass; the barrier between passes is the workgroup
// barrier, so no workgroup ever waits for another and every instance exits on its own.  The
// waves of the workgroup share the instance's tiles; each tile is the same code as the single
// solver's pass kernels, so every iteration is bit-identical to a single solve.  Reductions are
// block-wide in a fixed order (deterministic).
//
// The instance descriptors (with the tile parameters of both ping-pong directions,
// prepared on the host) are read through the constant address space: scalar loads that the
// compiler re-issues where needed instead of holding ~70 uniform values in registers
// across the solve loop.
#include "sf_ptk_tile.h"

namespace sf {

// 8 waves per instance workgroup: the three-iteration tile needs ~200 VGPRs, two waves per SIMD
#define BATCH_MAX_THREADS 512
#define BATCH_RED_POW2 512  // power of two >= BATCH_MAX_THREADS for the reduction tree
#define SF_CONST __attribute__((address_space(4)))

typedef const BatchInstance SF_CONST CBatch;

// Fixed-order block reduction (sum or max) of one double per thread; result in every thread.
__device__ inline double block_reduce(double v, bool is_max, double* sh)
{
    const int t = threadIdx.x, nt = blockDim.x;
    sh[t] = v;
    __syncthreads();
    for (int s = BATCH_RED_POW2 / 2; s > 0; s >>= 1) {
        if (t < s && t + s < nt) sh[t] = is_max ? jl_max(sh[t], sh[t + s]) : sh[t] + sh[t + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

__device__ inline double block_absmax(const double* __restrict__ a, int64_t n, double* sh)
{
    double m = -__builtin_inf();
    for (int64_t k = threadIdx.x; k < n; k += blockDim.x) m = jl_max(m, fabs(a[k]));
    return block_reduce(m, true, sh);
}

__device__ inline double block_abssum(const double* __restrict__ a, int64_t n, double* sh)
{
    double s = 0.0;
    for (int64_t k = threadIdx.x; k < n; k += blockDim.x) s = s + fabs(a[k]);
    return block_reduce(s, false, sh);
}

// compute_Residuals! y part (StokesSolver_multixpu.jl:189-202) on the post-update state
// (P, V = the pass's outputs), with dt = dt_check.
__device__ inline void block_resid_y(const PTParams SF_CONST& p, double* __restrict__ ry, double dt)
{
    const int nx = p.nx, ny = p.ny;
    const int64_t lp = nx - 1, lvx = nx, lvy = nx + 1, lr = nx + 1;
    const int64_t n = (int64_t)(nx - 1) * (ny - 2);
    const double rdx = p.rdx, rdy = p.rdy, g = p.g;
    const double* __restrict__ P = p.p_b;
    const double* __restrict__ vx = p.vx_b;
    const double* __restrict__ vy = p.vy_b;
    const double* __restrict__ rho = p.rho;
    const double* __restrict__ tyy = p.tyy;
    const double* __restrict__ txy = p.txy;
    for (int64_t k = threadIdx.x; k < n; k += blockDim.x) {
        const int i = (int)(k % (nx - 1)), j = (int)(k / (nx - 1));
        const double av = 0.25 * (((vx[ix2(i, j + 1, lvx)] + vx[ix2(i + 1, j + 1, lvx)]) + vx[ix2(i, j + 2, lvx)]) +
                                  vx[ix2(i + 1, j + 2, lvx)]);
        const double d_xi = rho[ix2(i + 2, j + 1, lr)] - rho[ix2(i, j + 1, lr)];
        const double d_yi = rho[ix2(i + 1, j + 2, lr)] - rho[ix2(i + 1, j, lr)];
        ry[ix2(i, j, nx - 1)] =
            (((tyy[ix2(i, j + 1, lp)] - tyy[ix2(i, j, lp)]) * rdy + (txy[ix2(i + 1, j + 1, lvx)] - txy[ix2(i, j + 1, lvx)]) * rdx) -
             (P[ix2(i, j + 1, lp)] - P[ix2(i, j, lp)]) * rdy) +
            g * (rho[ix2(i + 1, j + 1, lr)] - dt * (((av * d_xi) * 0.5) * rdx + ((vy[ix2(i + 1, j + 1, lvy)] * d_yi) * 0.5) * rdy));
    }
}

__device__ inline void block_copy(double* __restrict__ dst, const double* __restrict__ src, int64_t n)
{
    for (int64_t k = threadIdx.x; k < n; k += blockDim.x) dst[k] = src[k];
}

__device__ void solve_instance(CBatch& I, BatchResult* __restrict__ R, double* sh)
{
    const int nx = I.nx, ny = I.ny;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), nwaves = blockDim.x >> 6;
    const int mn = nx < ny ? nx : ny;
    const int64_t ncheck = (5 * (int64_t)mn < 2000) ? 5 * (int64_t)mn : 2000;
    const int64_t ndtupdate = 10 * (int64_t)mn;
    const int64_t nvx = (int64_t)nx * (ny + 1), nvy = (int64_t)(nx + 1) * ny, np = (int64_t)(nx - 1) * (ny - 1);
    const double n_rx = (double)(nx - 2) * (ny - 1), n_ry = (double)(nx - 1) * (ny - 2), n_dv = (double)np;
    const double eps = I.eps;
    const int64_t iter_max = I.iter_max;
    const bool fss = I.use_fss != 0;

    if (threadIdx.x == 0) {
        const double* dd = I.dt_dev;
        I.scal[SC_DT] = fss ? (dd ? *dd : I.dt_in) : 0.0;
    }
    __syncthreads();

    double err = 2.0 * eps;
    int64_t iter = 1;
    int pf = 0;  // passes so far, mod 2: the state (V, P, dV) is in the caller's buffers when 0
    while (err > eps && iter <= iter_max) {
        // iterations iter..iter+2 in one pass when none is a check iteration, none but the last is
        // followed by a dt refresh and the last is not iter_max (the rule of sf_solve_stokes)
        bool fused = iter + 2 < iter_max;
        for (int64_t j = iter; fused && j < iter + 3; ++j) {
            if (j % ncheck == 0) fused = false;
            if (j < iter + 2 && fss && j % ndtupdate == 0) fused = false;
        }
        if (fused) {
            const PTParams SF_CONST& pp = I.wpt[4 + pf];
            const int ntiles = pp.ntiles, nstrips = pp.nstrips;
            for (int t = wave; t < ntiles; t += nwaves) ptk_tile<3, 0>(pp, t % nstrips, t / nstrips);
            pf ^= 1;
            __syncthreads();
            if (fss && (iter + 2) % ndtupdate == 0) {
                const double mx = block_absmax(pp.vx_b, nvx, sh);
                const double my = block_absmax(pp.vy_b, nvy, sh);
                if (threadIdx.x == 0) I.scal[SC_DT] = I.maxdisp * jl_min(I.dx / mx, I.dy / my);
                __syncthreads();
            }
            iter += 3;
            continue;
        }
        const PTParams SF_CONST& pp = I.wpt[pf];
        const bool check = (iter % ncheck) == 0;
        const bool last = iter == iter_max;
        const int flags = ((check || last) ? PT_WRITE_TAU : 0) | (check ? PT_WRITE_RX : 0);
        const int ntiles = pp.ntiles, nstrips = pp.nstrips;
        if (flags) {
            for (int t = wave; t < ntiles; t += nwaves) pt_tile<1, true>(pp, t % nstrips, t / nstrips, flags);
        } else {
            for (int t = wave; t < ntiles; t += nwaves) pt_tile<1, false>(pp, t % nstrips, t / nstrips, 0);
        }
        pf ^= 1;
        __syncthreads();

        const bool dt_upd = fss && (iter % ndtupdate) == 0;
        double dtc = 0.0;
        if (dt_upd || (check && fss)) {
            const double mx = block_absmax(pp.vx_b, nvx, sh);
            const double my = block_absmax(pp.vy_b, nvy, sh);
            dtc = I.maxdisp * jl_min(I.dx / mx, I.dy / my);
            if (dt_upd && threadIdx.x == 0) I.scal[SC_DT] = dtc;
        }
        if (check) {
            block_resid_y(pp, I.ry, fss ? dtc : 0.0);
            __syncthreads();
            const double s0 = block_abssum(pp.rx, (int64_t)n_rx, sh);
            const double s1 = block_abssum(I.ry, (int64_t)n_ry, sh);
            const double s2 = block_abssum(pp.divv, (int64_t)n_dv, sh);
            err = jl_max(jl_max(s0 / n_rx, s1 / n_ry), s2 / n_dv);
        }
        __syncthreads();
        iter += 1;
    }
    const int64_t niter = iter - 1;
    if (pf) {  // the last iterate (state and dV) is in the scratch buffers
        const PTParams SF_CONST& p0 = I.wpt[0];
        block_copy(I.Vx_s, p0.vx_b, nvx);
        block_copy(I.Vy_s, p0.vy_b, nvy);
        block_copy(I.P, p0.p_b, np);
        block_copy(p0.dvx, p0.dvx_o, (int64_t)(nx - 2) * (ny - 1));
        block_copy(p0.dvy, p0.dvy_o, (int64_t)(nx - 1) * (ny - 2));
    }
    __syncthreads();
    // Vx_pad[2:end-1,:] .= Vx_s; Vy_pad[:,2:end-1] .= Vy_s; bc_x_mirror!; bc_y_mirror!  (:89-92)
    const double* __restrict__ Vx_s = I.Vx_s;
    const double* __restrict__ Vy_s = I.Vy_s;
    double* __restrict__ Vx_pad = I.Vx_pad;
    double* __restrict__ Vy_pad = I.Vy_pad;
    const int64_t npx = (int64_t)(nx + 2) * (ny + 1), npy = (int64_t)(nx + 1) * (ny + 2);
    for (int64_t k = threadIdx.x; k < npx; k += blockDim.x) {
        const int i = (int)(k % (nx + 2)), j = (int)(k / (nx + 2));
        double v;
        if (i == 0) v = -Vx_s[ix2(1, j, nx)];
        else if (i == nx + 1) v = -Vx_s[ix2(nx - 2, j, nx)];
        else v = Vx_s[ix2(i - 1, j, nx)];
        Vx_pad[k] = v;
    }
    for (int64_t k = threadIdx.x; k < npy; k += blockDim.x) {
        const int i = (int)(k % (nx + 1)), j = (int)(k / (nx + 1));
        double v;
        if (j == 0) v = -Vy_s[ix2(i, 1, nx + 1)];
        else if (j == ny + 1) v = -Vy_s[ix2(i, ny - 2, nx + 1)];
        else v = Vy_s[ix2(i, j - 1, nx + 1)];
        Vy_pad[k] = v;
    }
    const double mx = block_absmax(Vx_s, nvx, sh);
    const double my = block_absmax(Vy_s, nvy, sh);
    if (threadIdx.x == 0) {
        R->dt_new = I.maxdisp * jl_min(I.dx / mx, I.dy / my);
        R->err = err;
        R->niter = niter;
        R->converged = err <= eps;
        R->dt_last = I.scal[SC_DT];
    }
    __syncthreads();
}

__global__ __launch_bounds__(BATCH_MAX_THREADS) void batch_solve_kernel(const BatchInstance* inst,
                                                                         const int* __restrict__ order, int n_inst,
                                                                         int* __restrict__ queue,
                                                                         BatchResult* __restrict__ res)
{
    __shared__ double sh[BATCH_RED_POW2];
    __shared__ int s_next;
    CBatch* cinst = (CBatch*)inst;
    for (;;) {
        if (threadIdx.x == 0) s_next = atomicAdd(queue, 1);
        __syncthreads();
        const int q = __builtin_amdgcn_readfirstlane(s_next);
        __syncthreads();
        if (q >= n_inst) return;  // every workgroup reaches this once the queue is drained
        const int id = __builtin_amdgcn_readfirstlane(order[q]);
        solve_instance(cinst[id], res + id, sh);
    }
}

// ---------------------------------------------------------------------------------------
// Wide mode: one launch runs one PT iteration of every instance in the active set.  Each
// 64-lane workgroup is one tile of one instance; the slot is found by a search over the
// tile prefix table in the kernel arguments.  The blockIdx -> tile map is XCD-banded, so an
// XCD works on a contiguous range of instances and their tiles share its L2.

template <int DEPTH>
__global__ __launch_bounds__(64) void wide_pt_kernel(WideArgs a)
{
    const int total = a.tile_start[a.n_active];
    const int t_glob = xcd_remap(blockIdx.x, gridDim.x);
    if (t_glob >= total) return;
    int lo = 0, hi = a.n_active - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.tile_start[mid] <= t_glob) lo = mid;
        else hi = mid - 1;
    }
    const int t = t_glob - a.tile_start[lo];
    CBatch* cinst = (CBatch*)a.desc;
    const int pf = a.slot_pf[lo];
    const PTParams SF_CONST& pp = cinst[a.slot_inst[lo]].wpt[pf & 1];
    const int flags = pf >> 1;
    const int nstrips = pp.nstrips;
    if (flags) pt_tile<DEPTH, true>(pp, t % nstrips, t / nstrips, flags);
    else pt_tile<DEPTH, false>(pp, t % nstrips, t / nstrips, 0);
}

// Two PT iterations of every active instance per launch (pt2_tile, sf_pt2_tile.h).
__global__ __launch_bounds__(64) void wide_pt2_kernel(WideArgs a)
{
    const int total = a.tile_start[a.n_active];
    const int t_glob = xcd_remap(blockIdx.x, gridDim.x);
    if (t_glob >= total) return;
    int lo = 0, hi = a.n_active - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.tile_start[mid] <= t_glob) lo = mid;
        else hi = mid - 1;
    }
    const int t = t_glob - a.tile_start[lo];
    CBatch* cinst = (CBatch*)a.desc;
    const PTParams SF_CONST& pp = cinst[a.slot_inst[lo]].wpt[2 + (a.slot_pf[lo] & 1)];
    const int nstrips = pp.nstrips;
    pt2_tile(pp, t % nstrips, t / nstrips);
}

hipError_t launch_wide_pt(const WideArgs& a, int depth, hipStream_t s)
{
    const int total = a.tile_start[a.n_active];
    if (total <= 0) return hipSuccess;
    if (depth > 1) hipLaunchKernelGGL(wide_pt_kernel<2>, dim3(total), dim3(64), 0, s, a);
    else hipLaunchKernelGGL(wide_pt_kernel<1>, dim3(total), dim3(64), 0, s, a);
    return hipGetLastError();
}

// K = 3 or 4 PT iterations of every active instance per launch (ptk_tile<K>; descriptor
// wpt[2K - 2 + parity]).  K = 3 runs the split loop form with two middle steps per trip (214
// VGPRs), K = 4 with one (252 VGPRs; two would need 260, past two waves per SIMD).
template <int K, int SPLIT>
__global__ __launch_bounds__(64) void wide_ptk_kernel(WideArgs a)
{
    const int total = a.tile_start[a.n_active];
    const int t_glob = xcd_remap(blockIdx.x, gridDim.x);
    if (t_glob >= total) return;
    int lo = 0, hi = a.n_active - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.tile_start[mid] <= t_glob) lo = mid;
        else hi = mid - 1;
    }
    const int t = t_glob - a.tile_start[lo];
    CBatch* cinst = (CBatch*)a.desc;
    const PTParams SF_CONST& pp = cinst[a.slot_inst[lo]].wpt[2 * K - 2 + (a.slot_pf[lo] & 1)];
    const int nstrips = pp.nstrips;
    ptk_tile<K, SPLIT>(pp, t % nstrips, t / nstrips);
}

hipError_t launch_wide_pt3(const WideArgs& a, hipStream_t s)
{
    const int total = a.tile_start[a.n_active];
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL((wide_ptk_kernel<3, 2>), dim3(total), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_wide_pt4(const WideArgs& a, hipStream_t s)
{
    const int total = a.tile_start[a.n_active];
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL((wide_ptk_kernel<4, 1>), dim3(total), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_wide_pt2(const WideArgs& a, hipStream_t s)
{
    const int total = a.tile_start[a.n_active];
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(wide_pt2_kernel, dim3(total), dim3(64), 0, s, a);
    return hipGetLastError();
}

__global__ void zero_int(int* p) { *p = 0; }

int batch_blocks_per_cu(int threads)
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_solve_kernel, threads, 0) != hipSuccess || n < 1) n = 1;
    return n;
}

hipError_t launch_batch_solve(const BatchInstance* inst, const int* order, int n_inst, int* queue,
                              BatchResult* res, int threads, int nblocks, hipStream_t s)
{
    hipLaunchKernelGGL(zero_int, dim3(1), dim3(1), 0, s, queue);
    hipLaunchKernelGGL(batch_solve_kernel, dim3(nblocks), dim3(threads), 0, s, inst, order, n_inst, queue, res);
    return hipGetLastError();
}

}  // namespace sf
