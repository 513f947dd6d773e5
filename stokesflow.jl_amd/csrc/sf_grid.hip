// sf_grid.hip -- one problem's whole PT solve in ONE launch that spans the chip.
//
// solveStokes! (StokesSolver_multixpu.jl:43-95) has a data-dependent loop (convergence is tested
// every ncheck iterations), which a hipGraph cannot hold.  Here the loop runs on the device: a
// launch of at most one workgroup per CU, of one to eight waves, walks the same pass schedule as
// sf_solve_stokes (passes of up to three PT iterations, the single iterations at checks, dt
// refreshes and iterMax), each wave taking tiles t = band, band + G, ... of every pass (the waves
// of one XCD hold consecutive bands), and a grid barrier separates the passes.  Small grids get one wave per CU (a wave alone on a SIMD runs its
// tile faster), larger ones up to eight (fewer tiles per wave).  The dt refresh and the err check
// reduce over the whole grid in two steps (a partial per wave, then every wave sums the partials
// in the same order, so every wave takes the same decision).  Each tile is the code of the single
// solver's pass kernels (sf_pt_tile.h, sf_pt2_tile.h, sf_ptk_tile.h), so every field is
// bit-identical to sf_solve_stokes; err is a sum of the same terms in another order (equal to
// rounding).
//
// The barrier is two-level, because an agent-scope release writes back the whole L2 of its XCD
// and agent-scope atomics on one word are serialised (measured with one-wave workgroups: 98 us per
// barrier at 2,048 waves with a release, an arrival and a poll per wave on one counter; 24 us
// with one counter per XCD and one release per XCD; 4 us at 256 arrivals).  A workgroup's waves
// meet at the workgroup barrier (their stores are then in the XCD's L2); its first thread counts
// the workgroup in on its XCD's counter; the XCD's last workgroup writes that L2 back (release,
// agent scope) and counts the XCD in on the global counter, which every workgroup's first thread
// polls before invalidating the caches (acquire, agent scope) and letting its waves go.  The
// counters are monotone within a launch and zeroed by a kernel before it; the workgroups per XCD
// are counted at the start.  The workgroups must all be resident, which the host ensures by
// sizing the launch from the kernel's occupancy; should one not be, nothing waits forever: a wait
// longer than the time limit (or a flag raised by another workgroup) sets the failure flag, every
// workgroup leaves the kernel, and the host reports the error.
#include <atomic>

#include "sf_ptk_tile.h"

namespace sf {

#define GRID_CONST __attribute__((address_space(4)))
typedef const BatchInstance GRID_CONST CGrid;

namespace {

typedef unsigned long long u64;

__device__ inline u64 ld_agent(const u64* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline u64 add_agent(u64* p) { return __hip_atomic_fetch_add(p, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Polls p until it reaches target; false after the time limit (the failure flag and its sticky
// copy then set) or when another workgroup has raised the flag.  s: the launch's sync words.
__device__ inline bool wait_agent(const u64* p, u64 target, u64* s, uint64_t limit)
{
    const uint64_t t0 = wall_clock64();
    while (ld_agent(p) < target) {
        if (ld_agent(s + GRID_SYNC_FAIL) != 0ull) return false;
        if (wall_clock64() - t0 > limit) {
            __hip_atomic_store(s + GRID_SYNC_FAIL, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(s + GRID_SYNC_STICKY, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(2);
    }
    return true;
}

__device__ inline int xcd_id()
{
    int v;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
    return v & (GRID_MAX_XCD - 1);
}

struct GridBarrier {
    u64* s;          // sync words (layout in sf_internal.h)
    u64 gtarget;     // global arrivals that complete the next barrier
    u64 xtarget;     // arrivals on this XCD's counter that complete it
    u64 n_xcd;       // XCDs with workgroups of this launch
    u64 w_here;      // workgroups of this launch on this XCD
    uint64_t limit;  // wall-clock ticks a workgroup may wait
    int xcd;
    int band;        // workgroups before this one when they are ordered by XCD, then by arrival

    // Counts the workgroups per XCD (once, at the start).
    __device__ bool init(u64* sync, uint64_t lim)
    {
        s = sync;
        limit = lim;
        xcd = xcd_id();
        gtarget = xtarget = 0ull;
        int ok = 1;
        u64 nx = 0, here = 0, before = 0;
        if (threadIdx.x == 0) {
            // the returned count is this workgroup's rank on its XCD; waiting for it also orders
            // this add before the start arrival
            const u64 rank = add_agent(s + GRID_SYNC_XCOUNT + 16 * xcd);
            if (rank >= (u64)gridDim.x) ok = 0;
            add_agent(s + GRID_SYNC_START);
            ok = wait_agent(s + GRID_SYNC_START, gridDim.x, s, limit) && ok;
            for (int x = 0; x < GRID_MAX_XCD; ++x) {
                const u64 c = ld_agent(s + GRID_SYNC_XCOUNT + 16 * x);
                nx += c > 0ull;
                if (x < xcd) before += c;
            }
            here = ld_agent(s + GRID_SYNC_XCOUNT + 16 * xcd);
            before += rank;
        }
        __shared__ u64 sh[3];
        if (threadIdx.x == 0) {
            sh[0] = nx;
            sh[1] = here;
            sh[2] = before;
        }
        const int bad = __syncthreads_or(threadIdx.x == 0 && !ok);
        n_xcd = sh[0];
        w_here = sh[1];
        band = (int)sh[2];
        return bad == 0;
    }

    __device__ bool sync()
    {
        gtarget += n_xcd;
        xtarget += w_here;
        __syncthreads();  // the workgroup's stores are in this XCD's L2
        int ok = 1;
        if (threadIdx.x == 0) {
            if (add_agent(s + GRID_SYNC_XARRIVE + 16 * xcd) + 1ull == xtarget) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // this XCD's L2 written back
                add_agent(s + GRID_SYNC_ARRIVE);
            }
            ok = wait_agent(s + GRID_SYNC_ARRIVE, gtarget, s, limit);
        }
        if (threadIdx.x == 0) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        return __syncthreads_or(threadIdx.x == 0 && !ok) == 0;
    }
};

// Sum over the wave in a fixed order; lane 0's value in every lane.
__device__ inline double wave_sum(double v)
{
    for (int d = 32; d > 0; d >>= 1) v = v + __shfl_down(v, d);
    return __shfl(v, 0);
}

__device__ inline double wave_max(double v)
{
    for (int d = 32; d > 0; d >>= 1) v = jl_max(v, __shfl_down(v, d));
    return __shfl(v, 0);
}

// Every wave reduces the nw partials part[0..nw) in the same order.
__device__ inline double partials_sum(const double* part, int nw)
{
    double s = 0.0;
    for (int k = threadIdx.x & 63; k < nw; k += 64) s = s + part[k];
    return wave_sum(s);
}

__device__ inline double partials_max(const double* part, int nw)
{
    double m = -__builtin_inf();
    for (int k = threadIdx.x & 63; k < nw; k += 64) m = jl_max(m, part[k]);
    return wave_max(m);
}

// compute_Residuals! y part (StokesSolver_multixpu.jl:189-202) on the post-update state, with
// dt = dt_check; grid-stride over the waves.
__device__ inline void grid_resid_y(const PTParams GRID_CONST& p, double* __restrict__ ry, double dt, int64_t gtid,
                                    int64_t gstride)
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
    for (int64_t k = gtid; k < n; k += gstride) {
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

__device__ inline double lane_absmax(const double* __restrict__ a, int64_t n, int64_t gtid, int64_t gstride)
{
    double m = -__builtin_inf();
    for (int64_t k = gtid; k < n; k += gstride) m = jl_max(m, fabs(a[k]));
    return m;
}

__device__ inline double lane_abssum(const double* __restrict__ a, int64_t n, int64_t gtid, int64_t gstride)
{
    double s = 0.0;
    for (int64_t k = gtid; k < n; k += gstride) s = s + fabs(a[k]);
    return s;
}

__device__ inline void grid_copy(double* __restrict__ dst, const double* __restrict__ src, int64_t n, int64_t gtid,
                                 int64_t gstride)
{
    for (int64_t k = gtid; k < n; k += gstride) dst[k] = src[k];
}

}  // namespace

// K = PT iterations per pass, 1 to 3 (the four-iteration tile spills beside the loop control).
template <int K>
__global__ __launch_bounds__(64 * GRID_WG_WAVES) void grid_solve_kernel(GridArgs a)
{
    CGrid& I = *(CGrid*)a.desc;
    // w is uniform in the wave (saying so keeps the tile indices in scalar registers)
    const int nw = gridDim.x * (blockDim.x >> 6),
              w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)));
    const int64_t gtid = (int64_t)w * 64 + (threadIdx.x & 63), gstride = (int64_t)nw * 64;
    GridBarrier bar;
    if (!bar.init(a.sync, a.limit)) return;
    // the tiles of a pass go out in bands: the waves of one XCD take consecutive tiles, so the
    // strips that share halo columns mostly share an L2
    const int wt = __builtin_amdgcn_readfirstlane(bar.band * (int)(blockDim.x >> 6) + (int)(threadIdx.x >> 6));
    double* part_a = a.part;           // nw partials (max of |Vx|, sum of |Rx|)
    double* part_b = a.part + nw;      // |Vy|, |Ry|
    double* part_c = a.part + 2 * nw;  // |divV|

    const int nx = I.nx, ny = I.ny;
    const int mn = nx < ny ? nx : ny;
    const int64_t ncheck = (5 * (int64_t)mn < 2000) ? 5 * (int64_t)mn : 2000;
    const int64_t ndtupdate = 10 * (int64_t)mn;
    const int64_t nvx = (int64_t)nx * (ny + 1), nvy = (int64_t)(nx + 1) * ny, np = (int64_t)(nx - 1) * (ny - 1);
    const double n_rx = (double)(nx - 2) * (ny - 1), n_ry = (double)(nx - 1) * (ny - 2), n_dv = (double)np;
    const double eps = I.eps;
    const int64_t iter_max = I.iter_max;
    const bool fss = I.use_fss != 0;

    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const double* dd = I.dt_dev;
        I.scal[SC_DT] = fss ? (dd ? *dd : I.dt_in) : 0.0;
    }
    if (!bar.sync()) return;

    // max(|Vx|), max(|Vy|) of the state in p -> maxdisp * min(dx / mx, dy / my) (compute_dt!)
    auto grid_dt = [&](const double* vx, const double* vy, double* dtc) -> bool {
        const double mx = wave_max(lane_absmax(vx, nvx, gtid, gstride));
        const double my = wave_max(lane_absmax(vy, nvy, gtid, gstride));
        if ((threadIdx.x & 63) == 0) {
            part_a[w] = mx;
            part_b[w] = my;
        }
        if (!bar.sync()) return false;
        const double gx = partials_max(part_a, nw), gy = partials_max(part_b, nw);
        *dtc = I.maxdisp * jl_min(I.dx / gx, I.dy / gy);
        return true;
    };

    double err = 2.0 * eps;
    int64_t iter = 1;
    int pf = 0;  // passes so far, mod 2: the state (V, P, dV) is in the caller's buffers when 0
    while (err > eps && iter <= iter_max) {
        // iterations iter..iter+K-1 in one pass when none is a check iteration, none but the last
        // is followed by a dt refresh and the last is not iter_max (the rule of sf_solve_stokes;
        // where K do not fit, single iterations follow, which gives the same bits)
        bool fused = K > 1 && iter + K - 1 < iter_max;
        for (int64_t j = iter; fused && j < iter + K; ++j) {
            if (j % ncheck == 0) fused = false;
            if (j < iter + K - 1 && fss && j % ndtupdate == 0) fused = false;
        }
        if (fused) {
            const PTParams GRID_CONST& pp = I.wpt[2 * K - 2 + pf];
            const int ntiles = pp.ntiles, nstrips = pp.nstrips;
            if constexpr (K == 3) {
                for (int t = wt; t < ntiles; t += nw) ptk_tile<3, 0>(pp, t % nstrips, t / nstrips);
            } else if constexpr (K == 2) {
                for (int t = wt; t < ntiles; t += nw) pt2_tile(pp, t % nstrips, t / nstrips);
            }
            pf ^= 1;
            if (!bar.sync()) return;
            if (fss && (iter + K - 1) % ndtupdate == 0) {
                double dtc;
                if (!grid_dt(pp.vx_b, pp.vy_b, &dtc)) return;
                if (blockIdx.x == 0 && threadIdx.x == 0) I.scal[SC_DT] = dtc;
                if (!bar.sync()) return;
            }
            iter += K;
            continue;
        }
        const PTParams GRID_CONST& pp = I.wpt[pf];
        const bool check = (iter % ncheck) == 0;
        const bool last = iter == iter_max;
        const int flags = ((check || last) ? PT_WRITE_TAU : 0) | (check ? PT_WRITE_RX : 0);
        const int ntiles = pp.ntiles, nstrips = pp.nstrips;
        if (flags) {
            for (int t = wt; t < ntiles; t += nw) pt_tile<1, true>(pp, t % nstrips, t / nstrips, flags);
        } else {
            for (int t = wt; t < ntiles; t += nw) pt_tile<1, false>(pp, t % nstrips, t / nstrips, 0);
        }
        pf ^= 1;
        if (!bar.sync()) return;

        const bool dt_upd = fss && (iter % ndtupdate) == 0;
        double dtc = 0.0;
        if (dt_upd || (check && fss)) {
            if (!grid_dt(pp.vx_b, pp.vy_b, &dtc)) return;
            if (dt_upd && w == 0 && threadIdx.x == 0) I.scal[SC_DT] = dtc;
        }
        if (check) {
            grid_resid_y(pp, I.ry, fss ? dtc : 0.0, gtid, gstride);
            if (!bar.sync()) return;
            const double s0 = wave_sum(lane_abssum(pp.rx, (int64_t)n_rx, gtid, gstride));
            const double s1 = wave_sum(lane_abssum(I.ry, (int64_t)n_ry, gtid, gstride));
            const double s2 = wave_sum(lane_abssum(pp.divv, (int64_t)n_dv, gtid, gstride));
            if ((threadIdx.x & 63) == 0) {
                part_a[w] = s0;
                part_b[w] = s1;
                part_c[w] = s2;
            }
            if (!bar.sync()) return;
            const double g0 = partials_sum(part_a, nw), g1 = partials_sum(part_b, nw), g2 = partials_sum(part_c, nw);
            err = jl_max(jl_max(g0 / n_rx, g1 / n_ry), g2 / n_dv);
        }
        if ((dt_upd || check) && !bar.sync()) return;  // dt written, partials read
        iter += 1;
    }
    const int64_t niter = iter - 1;
    if (pf) {  // the last iterate (state and dV) is in the scratch buffers
        const PTParams GRID_CONST& p0 = I.wpt[0];
        grid_copy(I.Vx_s, p0.vx_b, nvx, gtid, gstride);
        grid_copy(I.Vy_s, p0.vy_b, nvy, gtid, gstride);
        grid_copy(I.P, p0.p_b, np, gtid, gstride);
        grid_copy(p0.dvx, p0.dvx_o, (int64_t)(nx - 2) * (ny - 1), gtid, gstride);
        grid_copy(p0.dvy, p0.dvy_o, (int64_t)(nx - 1) * (ny - 2), gtid, gstride);
        if (!bar.sync()) return;
    }
    // Vx_pad[2:end-1,:] .= Vx_s; Vy_pad[:,2:end-1] .= Vy_s; bc_x_mirror!; bc_y_mirror!  (:89-92)
    const double* __restrict__ Vx_s = I.Vx_s;
    const double* __restrict__ Vy_s = I.Vy_s;
    double* __restrict__ Vx_pad = I.Vx_pad;
    double* __restrict__ Vy_pad = I.Vy_pad;
    const int64_t npx = (int64_t)(nx + 2) * (ny + 1), npy = (int64_t)(nx + 1) * (ny + 2);
    for (int64_t k = gtid; k < npx; k += gstride) {
        const int i = (int)(k % (nx + 2)), j = (int)(k / (nx + 2));
        double v;
        if (i == 0) v = -Vx_s[ix2(1, j, nx)];
        else if (i == nx + 1) v = -Vx_s[ix2(nx - 2, j, nx)];
        else v = Vx_s[ix2(i - 1, j, nx)];
        Vx_pad[k] = v;
    }
    for (int64_t k = gtid; k < npy; k += gstride) {
        const int i = (int)(k % (nx + 1)), j = (int)(k / (nx + 1));
        double v;
        if (j == 0) v = -Vy_s[ix2(i, 1, nx + 1)];
        else if (j == ny + 1) v = -Vy_s[ix2(i, ny - 2, nx + 1)];
        else v = Vy_s[ix2(i, j - 1, nx + 1)];
        Vy_pad[k] = v;
    }
    double dt_new;
    if (!grid_dt(Vx_s, Vy_s, &dt_new)) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        BatchResult* R = a.res;
        R->dt_new = dt_new;
        R->err = err;
        R->niter = niter;
        R->converged = err <= eps;
        R->dt_last = I.scal[SC_DT];
    }
}

// Every word but the sticky failure flag, which only the host clears.
__global__ void grid_zero_sync(unsigned long long* p)
{
    for (int k = threadIdx.x; k < GRID_SYNC_WORDS; k += blockDim.x)
        if (k != GRID_SYNC_STICKY) p[k] = 0ull;
}

// Resident workgroups of `waves` waves of grid_solve_kernel<k> on the current device, at most
// one per CU.
int grid_capacity(int k, int waves)
{
    static int cap[4][GRID_WG_WAVES + 1] = {};
    if (k < 1 || k > 3 || waves < 1 || waves > GRID_WG_WAVES) return 0;
    if (cap[k][waves] == 0) {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
            cus = 256;
        const void* f = k == 3 ? (const void*)grid_solve_kernel<3>
                               : (k == 2 ? (const void*)grid_solve_kernel<2> : (const void*)grid_solve_kernel<1>);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, 64 * waves, 0) != hipSuccess || per_cu < 1)
            cap[k][waves] = -1;
        else
            cap[k][waves] = cus;
    }
    return std::max(cap[k][waves], 0);
}

static std::atomic<double> grid_wait_s{(double)GRID_WAIT_SECONDS};

uint64_t grid_wait_limit()
{
    static const uint64_t khz = [] {
        int dev = 0, k = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&k, hipDeviceAttributeWallClockRate, dev) != hipSuccess || k < 1)
            k = 100000;
        return (uint64_t)k;
    }();
    return (uint64_t)(grid_wait_s.load() * 1000.0 * (double)khz);
}

// Launches already enqueued (or captured in a graph) keep the limit they were made with.
void grid_set_wait_seconds(double seconds) { grid_wait_s.store(seconds); }
double grid_wait_seconds() { return grid_wait_s.load(); }

hipError_t launch_grid_solve(const GridArgs& a, int k, hipStream_t s)
{
    if (a.nwg < 1 || a.nwg > grid_capacity(k, a.wg_waves)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(grid_zero_sync, dim3(1), dim3(256), 0, s, a.sync);
    const dim3 g(a.nwg), b(64 * a.wg_waves);
    if (k == 3) hipLaunchKernelGGL(grid_solve_kernel<3>, g, b, 0, s, a);
    else if (k == 2) hipLaunchKernelGGL(grid_solve_kernel<2>, g, b, 0, s, a);
    else hipLaunchKernelGGL(grid_solve_kernel<1>, g, b, 0, s, a);
    return hipGetLastError();
}

}  // namespace sf
