// sf_pt_tile.h -- one tile of the fused FP64 PT iteration (shared by the single-solve
// kernel in sf_pt.hip and the batched ensemble kernel in sf_batch.hip).
//
// A tile is a 64-lane wave over a strip of 64 columns (lanes 0..61 produce outputs,
// lanes 62, 63 are the halo) marching down a segment of rows; see sf_pt.hip for the
// scheme.  Every expression keeps the reference's evaluation order (compiled without FMA
// contraction), so a tile's outputs are bit-identical to a CPU evaluation.
#pragma once

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

template <class PP>
__device__ __forceinline__ RowIn load_row(const PP& p, const LaneCols& lc, int r, bool need_dv)
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

// PP is PTParams in any address space (a kernel argument, or constant memory in the
// batched kernel); the write flags (PT_WRITE_*) come as an argument.
template <int DEPTH, bool CHECK, class PP>
__device__ __forceinline__ void pt_tile(const PP& p, int strip, int segid, int flags)
{
    const int lane = threadIdx.x & 63;
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
    // dV is updated in place unless a second buffer is given (ping-pong in the wide batched mode)
    double* const dvx_w = p.dvx_o ? p.dvx_o : p.dvx;
    double* const dvy_w = p.dvy_o ? p.dvy_o : p.dvy;
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
                    if (flags & PT_WRITE_TAU) {
                        p.divv[ix2(c, r, lp)] = div;
                        p.txx[ix2(c, r, lp)] = txx;
                        p.tyy[ix2(c, r, lp)] = tyy;
                    }
                }
            }
            if constexpr (CHECK)
                if ((flags & PT_WRITE_TAU) && c_vx) p.txy[ix2(c, r, lvx)] = txy;
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
                dvx_w[ix2(c, j, nx - 2)] = dvx_new;
                p.vx_b[ix2(c + 1, j + 1, lvx)] = vx_new;
                if (j == 0) p.vx_b[ix2(c + 1, 0, lvx)] = vx_new;        // bc_y_noflux!
                if (j == ny - 2) p.vx_b[ix2(c + 1, ny, lvx)] = vx_new;
                if constexpr (CHECK)
                    if (flags & PT_WRITE_RX) p.rx[ix2(c, j, nx - 2)] = rx;
            }
            if (c == 0 || c == nx - 2) {                                 // bc_x_zero!
                const int col = (c == 0) ? 0 : nx - 1;
                p.vx_b[ix2(col, j + 1, lvx)] = 0.0;
                if (j == 0) p.vx_b[ix2(col, 0, lvx)] = 0.0;
                if (j == ny - 2) p.vx_b[ix2(col, ny, lvx)] = 0.0;
            }
            if (dvy_ok) {
                dvy_w[ix2(c, j, nx - 1)] = dvy_new;
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

}  // namespace sf
