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

#include "sf_device.h"
#include "sf_internal.h"

namespace sf {

#define PT_OUT_LANES 62

// XCD-aware logical tile id: hardware hands workgroup k to XCD k % 8; give XCD x the
// contiguous logical range it would own in a blocked split.
__device__ inline int xcd_remap(int k, int total)
{
    const int nxcd = 8;
    int xcd = k % nxcd, idx = k / nxcd;
    int bpx = total / nxcd, rem = total % nxcd;
    return (xcd < rem) ? xcd * (bpx + 1) + idx : rem * (bpx + 1) + (xcd - rem) * bpx + idx;
}


// Operands one row step consumes (row r of the cell/basic quantities, row r+1 of V/rho,
// row r-1 of dVx/dVy).
struct RowIn {
    double vx1, vy1, rho1, pold, mup, mub, dvx, dvy;
};

// Per-lane column offsets, clamped into each array so that every load stays in bounds.
// Lanes whose column lies outside an array read a neighbouring in-array value; such
// values only reach lanes and rows whose results are never stored (every stored output
// reads in-array operands only), so no select is needed.
struct LaneCols {
    unsigned vx, vy, p, dvx, dvy;  // Vx/mu_b (nx), Vy/rho (nx+1), P/mu_p (nx-1), dVx (nx-2), dVy (nx-1)
};

// Uniform row index clamped into [0, nrows-1].
__device__ inline int crow(int r, int nrows) { return r < 0 ? 0 : (r >= nrows ? nrows - 1 : r); }

__device__ inline RowIn load_row(const PTParams& p, const LaneCols& lc, int r, bool need_dv)
{
    const int nx = p.nx, ny = p.ny;
    RowIn in;
    in.vx1 = (p.vx_a + (int64_t)crow(r + 1, ny + 1) * nx)[lc.vx];
    in.vy1 = (p.vy_a + (int64_t)crow(r + 1, ny) * (nx + 1))[lc.vy];
    in.rho1 = (p.rho + (int64_t)crow(r + 1, ny) * (nx + 1))[lc.vy];
    in.pold = (p.p_a + (int64_t)crow(r, ny - 1) * (nx - 1))[lc.p];
    in.mup = (p.mu_p + (int64_t)crow(r, ny - 1) * (nx - 1))[lc.p];
    in.mub = (p.mu_b + (int64_t)crow(r, ny) * nx)[lc.vx];
    if (need_dv) {
        in.dvx = (p.dvx + (int64_t)crow(r - 1, ny - 1) * (nx - 2))[lc.dvx];
        in.dvy = (p.dvy + (int64_t)crow(r - 1, ny - 2) * (nx - 1))[lc.dvy];
    } else {
        in.dvx = 0.0;
        in.dvy = 0.0;
    }
    return in;
}

template <int DEPTH, bool CHECK>
__device__ inline void pt_tile(const PTParams& p, int strip, int segid)
{
    const int lane = threadIdx.x;
    const int c = strip * PT_OUT_LANES + lane;
    const int nx = p.nx, ny = p.ny;
    const bool out_lane = lane < PT_OUT_LANES;
    const int r0 = segid * p.seg;
    const int r1 = min(r0 + p.seg, ny);
    const double dt = *p.dt;
    const int64_t lvx = nx, lvy = nx + 1, lp = nx - 1;
    const bool c_vx = c < nx;      // Vx_s, mu_b, txy columns
    const bool c_p = c < nx - 1;   // P, mu_p columns
    const double rdx = p.rdx, rdy = p.rdy;
    LaneCols lc;
    lc.vx = (unsigned)min(c, nx - 1);
    lc.vy = (unsigned)min(c, nx);
    lc.p = (unsigned)min(c, nx - 2);
    lc.dvx = (unsigned)min(c, nx - 3);
    lc.dvy = (unsigned)min(c, nx - 2);

    int r = (r0 == 0) ? 0 : r0 - 1;
    double vx_r = (p.vx_a + (int64_t)r * nx)[lc.vx];
    double vy_r = (p.vy_a + (int64_t)r * (nx + 1))[lc.vy];
    double rho_r = (p.rho + (int64_t)r * (nx + 1))[lc.vy];
    // the prologue row needs no dVx/dVy, and prefetch never reaches past the segment end
    RowIn nxt = load_row(p, lc, r, r >= r0);
    RowIn nxt2;
    if constexpr (DEPTH > 1)
        if (r + 1 < r1) nxt2 = load_row(p, lc, r + 1, r + 1 >= r0);
    double vx_r_p1 = lane_down(vx_r, 1);
    double vy_r_p1 = lane_down(vy_r, 1);
    double rho_r_p1 = lane_down(rho_r, 1);
    double rho_r_p2 = lane_down(rho_r, 2);
    double rho_rm1_p1 = 0.0;
    // row r-1 quantities carried across iterations (cell row / basic row)
    double pn_prev = 0.0, tyy_prev = 0.0, txx_prev = 0.0;
    double pn_prev_p1 = 0.0, txx_prev_p1 = 0.0, txy_prev_p1 = 0.0;
    double mup_prev = 0.0, mup_prev_p1 = 0.0;

    for (; r < r1; ++r) {
        const RowIn in = nxt;
        if constexpr (DEPTH > 1) {
            nxt = nxt2;
            if (r + 2 < r1) nxt2 = load_row(p, lc, r + 2, true);  // two rows of operands in flight
        } else {
            if (r + 1 < r1) nxt = load_row(p, lc, r + 1, true);   // next row's operands in flight
        }
        const bool emit = r >= r0;           // false on the prologue row of a segment
        const bool cell_row = r <= ny - 2;   // P-node rows
        const int j = r - 1;                 // output (dVx/dVy) row of this step
        const bool has_out = emit && j >= 0;
        const bool dvx_ok = has_out && c < nx - 2;
        const bool dvy_ok = has_out && c < nx - 1 && j <= ny - 3;

        const double vx_r1 = in.vx1, vy_r1 = in.vy1, rho_r1 = in.rho1;
        const double pold = in.pold, mup = in.mup, mub = in.mub;
        const double dvx_old = in.dvx, dvy_old = in.dvy;

        const double vx_r1_p1 = lane_down(vx_r1, 1);
        const double vy_r1_p1 = lane_down(vy_r1, 1);
        const double rho_r1_p1 = lane_down(rho_r1, 1);
        const double rho_r1_p2 = lane_down(rho_r1, 2);

        // compute_P_tau! at P node (c, r)  (:149-162)
        const double dvxdx = (vx_r1_p1 - vx_r1) * rdx;
        const double dvydy = (vy_r1_p1 - vy_r_p1) * rdy;
        const double div = dvxdx + dvydy;
        const double dtaup = (p.ptsc41 * mup) / p.max_nxy;
        const double pn = pold - dtaup * div;
        const double txx = (2.0 * mup) * (dvxdx - (1.0 / 3.0) * div);
        const double tyy = (2.0 * mup) * (dvydy - (1.0 / 3.0) * div);
        // compute_P_tau! at basic node (c, r)  (:164-166)
        const double txy = ((2.0 * mub) * 0.5) * ((vx_r1 - vx_r) * rdy + (vy_r_p1 - vy_r) * rdx);
        const double txy_p1 = lane_down(txy, 1);

        if (emit && out_lane) {
            if (c_p && cell_row) {
                p.p_b[ix2(c, r, lp)] = pn;
                if constexpr (CHECK) {
                    if (p.flags & PT_WRITE_TAU) {
                        p.divv[ix2(c, r, lp)] = div;
                        p.txx[ix2(c, r, lp)] = txx;
                        p.tyy[ix2(c, r, lp)] = tyy;
                    }
                }
            }
            if constexpr (CHECK)
                if ((p.flags & PT_WRITE_TAU) && c_vx) p.txy[ix2(c, r, lvx)] = txy;
        }

        if (has_out && out_lane) {
            // compute_ResX at dVx (c, j)  (:176-182)
            const double rx = ((txx_prev_p1 - txx_prev) * rdx + (txy_p1 - txy_prev_p1) * rdy) -
                              (pn_prev_p1 - pn_prev) * rdx;
            // compute_ResY at dVy (c, j)  (:189-202)
            const double av = 0.25 * (((vx_r + vx_r_p1) + vx_r1) + vx_r1_p1);
            const double d_xi = rho_r_p2 - rho_r;
            const double d_yi = rho_r1_p1 - rho_rm1_p1;
            const double ry = (((tyy - tyy_prev) * rdy + (txy_p1 - txy) * rdx) - (pn - pn_prev) * rdy) +
                              p.g * (rho_r_p1 - dt * (((av * d_xi) * 0.5) * rdx + ((vy_r_p1 * d_yi) * 0.5) * rdy));
            // compute_dV! (:209-219) and compute_V! (:226-234) with dtau from compute_timesteps!
            const double dvx_new = p.dampx * dvx_old + rx;
            const double dvy_new = p.dampy * dvy_old + ry;
            const double dtauvx = (p.vsc_mdxy2 / ((mup_prev + mup_prev_p1) * 0.5)) / 4.1;
            const double dtauvy = (p.vsc_mdxy2 / ((mup_prev + mup) * 0.5)) / 4.1;
            const double vx_new = vx_r_p1 + dtauvx * dvx_new;
            const double vy_new = vy_r_p1 + dtauvy * dvy_new;

            if (dvx_ok) {
                p.dvx[ix2(c, j, nx - 2)] = dvx_new;
                p.vx_b[ix2(c + 1, j + 1, lvx)] = vx_new;
                if (j == 0) p.vx_b[ix2(c + 1, 0, lvx)] = vx_new;        // bc_y_noflux!
                if (j == ny - 2) p.vx_b[ix2(c + 1, ny, lvx)] = vx_new;
                if constexpr (CHECK)
                    if (p.flags & PT_WRITE_RX) p.rx[ix2(c, j, nx - 2)] = rx;
            }
            if (c == 0 || c == nx - 2) {                                 // bc_x_zero!
                const int col = (c == 0) ? 0 : nx - 1;
                p.vx_b[ix2(col, j + 1, lvx)] = 0.0;
                if (j == 0) p.vx_b[ix2(col, 0, lvx)] = 0.0;
                if (j == ny - 2) p.vx_b[ix2(col, ny, lvx)] = 0.0;
            }
            if (dvy_ok) {
                p.dvy[ix2(c, j, nx - 1)] = dvy_new;
                p.vy_b[ix2(c + 1, j + 1, lvy)] = vy_new;
                if (c == 0) p.vy_b[ix2(0, j + 1, lvy)] = vy_new;         // bc_x_noflux!
                if (c == nx - 2) p.vy_b[ix2(nx, j + 1, lvy)] = vy_new;
            }
            if (c < nx - 1 && (j == 0 || j == ny - 2)) {                 // bc_y_zero!
                const int row = (j == 0) ? 0 : ny - 1;
                p.vy_b[ix2(c + 1, row, lvy)] = 0.0;
                if (c == 0) p.vy_b[ix2(0, row, lvy)] = 0.0;
                if (c == nx - 2) p.vy_b[ix2(nx, row, lvy)] = 0.0;
            }
        }

        // slide the window down one row
        pn_prev_p1 = lane_down(pn, 1);
        txx_prev_p1 = lane_down(txx, 1);
        mup_prev_p1 = lane_down(mup, 1);
        pn_prev = pn;
        txx_prev = txx;
        tyy_prev = tyy;
        txy_prev_p1 = txy_p1;
        mup_prev = mup;
        vx_r = vx_r1;
        vx_r_p1 = vx_r1_p1;
        vy_r = vy_r1;
        vy_r_p1 = vy_r1_p1;
        rho_rm1_p1 = rho_r_p1;
        rho_r = rho_r1;
        rho_r_p1 = rho_r1_p1;
        rho_r_p2 = rho_r1_p2;
    }
}

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
        pt_tile<DEPTH, CHECK>(p, logical % p.nstrips, logical / p.nstrips);
    }
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
__global__ void finalize_dt(const double* __restrict__ part, int nb, double dx, double dy, double maxdisp,
                            double* __restrict__ scal, int update_dt)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double mx = -__builtin_inf(), my = -__builtin_inf();
    for (int b = 0; b < nb; ++b) {
        mx = jl_max(mx, part[b]);
        my = jl_max(my, part[nb + b]);
    }
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
int pt_default_rows(int nx, int ny)
{
    (void)nx;
    return ny >= 4 ? 4 : ny;
}

hipError_t launch_pt_fused(const PTParams& p0, hipStream_t s)
{
    PTParams p = p0;
    if (p.seg <= 0) p.seg = pt_default_rows(p.nx, p.ny);
    const int depth = p.depth > 1 ? 2 : 1;
    p.nstrips = pt_num_strips(p.nx);
    const int nseg = (p.ny + p.seg - 1) / p.seg;
    p.ntiles = p.nstrips * nseg;
    // max_waves: 0 = one wave per resident slot (persistent), < 0 = one wave per tile
    int grid = p.ntiles;
    if (p.max_waves == 0) grid = std::min(p.ntiles, pt_resident_waves(depth));
    else if (p.max_waves > 0) grid = std::min(p.ntiles, p.max_waves);
    const bool check = p.flags != 0;
    if (depth == 2) {
        if (check) hipLaunchKernelGGL((pt_fused_f64<2, true>), dim3(grid), dim3(64), 0, s, p);
        else hipLaunchKernelGGL((pt_fused_f64<2, false>), dim3(grid), dim3(64), 0, s, p);
    } else {
        if (check) hipLaunchKernelGGL((pt_fused_f64<1, true>), dim3(grid), dim3(64), 0, s, p);
        else hipLaunchKernelGGL((pt_fused_f64<1, false>), dim3(grid), dim3(64), 0, s, p);
    }
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
