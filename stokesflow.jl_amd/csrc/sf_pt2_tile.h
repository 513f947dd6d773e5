// sf_pt2_tile.h -- two PT iterations per tile pass (temporal blocking), FP64.
//
// pt_tile (sf_pt_tile.h) runs one PT iteration per launch and moves 13 x 8 B per node
// (read Vx, Vy, P, dVx, dVy, mu_p, mu_b, rho; write Vx, Vy, P, dVx, dVy).  At large grids
// it runs at the DRAM copy rate, so more iterations per second need fewer bytes per
// iteration.  pt2_tile runs iterations k+1 ("A") and k+2 ("B") in one pass: A marches
// down the rows exactly like pt_tile but keeps its results in registers, and B consumes
// them one row behind.  Per node a pass reads the state and the materials once and writes
// the state once, for two iterations.
//
// Lanes.  pt_tile's lane l (column c) produces P and dV at column c and V at column c+1
// from its inputs at columns c..c+2.  B needs iteration k+1's V at column c, i.e. A's
// output of lane l-1 (one lane up).  So B's output lanes are 1..60 (A's lanes 0..61 are
// the ones with valid outputs), strips advance by 60 columns and lane l holds column
// strip*60 + l - 1.
//
// Rows.  For B rows [r0, r1) the tile needs A's V rows r0-1..r1, P rows r0-1..r1-1 and
// dV rows r0-2..r1-2 (the prologue row of B included), so A runs rows max(r0-2, 0) ..
// min(r1, ny-1) and B runs rows max(r0-1, 0) .. r1-1; B's row r' runs right after A's
// row r'+1.  When r1 = ny, B's last row runs after A's last row with V row ny (Vx's copy
// of row ny-1).
//
// Boundaries.  Each A output handed to B is the value pt_tile would have stored,
// free-slip BCs included (Vx columns 0, nx-1 = 0; Vy column 0 = column 1 and column nx =
// column nx-1; Vx row 0 = row 1 and row ny = row ny-1; Vy rows 0, ny-1 = 0).
//
// Both iterations evaluate pt_tile's expressions in the same order (pt_row below), so
// the stored iteration k+2 is bit-identical to two pt_tile launches.  The pass writes dV
// to a second buffer (dvx_o, dvy_o): A reads iteration k's dV on rows that the tile
// above overwrites with iteration k+2's.
#pragma once

#include "sf_pt_tile.h"

namespace sf {

#define PT2_OUT_LANES 60

// The row window one PT iteration carries down a strip (pt_tile's register window).  Only the
// lane's own values are carried; their neighbours one or two lanes down are moved in again
// on every row (a DPP move is exact, so the values are the same), which keeps the window at
// 9 doubles instead of 16: a three-stage pass takes 198 VGPRs instead of 236, and a four-stage
// pass 238 (251 in the split loop form it runs), still two waves per SIMD.
struct PTWin {
    double vx_r, vy_r;
    double rho_r, rho_rm1_p1;
    double pn_prev, tyy_prev, txx_prev, txy_prev_p1, mup_prev;
};

// One row step's results: P at (c, r); dVx, dVy at (c, r-1); Vx, Vy at (c+1, r).
struct PTRow {
    double pn, dvx, dvy, vx, vy;
};

// The pseudo-time steps of a row (compute_timesteps!, StokesSolver_multixpu.jl:128-133): they
// depend on mu_p only, so a later iteration of the same row reuses the first one's.
struct PTTau {
    double p, vx, vy;
};

__device__ __forceinline__ void pt_win_init(PTWin& w, double vx_r, double vy_r, double rho_r)
{
    w.vx_r = vx_r;
    w.vy_r = vy_r;
    w.rho_r = rho_r;
    w.rho_rm1_p1 = 0.0;
    w.pn_prev = w.tyy_prev = w.txx_prev = 0.0;
    w.txy_prev_p1 = 0.0;
    w.mup_prev = 0.0;
}

// One row step of one PT iteration with pt_tile's expressions (compute_P_tau!, compute_ResX,
// compute_ResY, compute_dV!, compute_V!; StokesSolver_multixpu.jl:140-234), then the window
// slides down one row.  Results of lanes or rows without an output are computed but unused.
template <bool TAU_IN, class PP>
__device__ __forceinline__ PTRow pt_row(const PP& p, double dt, PTWin& w, const RowIn& in, PTTau& tau)
{
    const double rdx = p.rdx, rdy = p.rdy;
    const double w_vx_r_p1 = lane_down(w.vx_r, 1);
    const double w_vy_r_p1 = lane_down(w.vy_r, 1);
    const double w_rho_r_p1 = lane_down(w.rho_r, 1);
    const double w_rho_r_p2 = lane_down(w.rho_r, 2);
    const double w_pn_prev_p1 = lane_down(w.pn_prev, 1);
    const double w_txx_prev_p1 = lane_down(w.txx_prev, 1);
    const double w_mup_prev_p1 = lane_down(w.mup_prev, 1);
    const double vx_r1_p1 = lane_down(in.vx1, 1);
    const double vy_r1_p1 = lane_down(in.vy1, 1);
    const double rho_r1_p1 = lane_down(in.rho1, 1);

    const double dvxdx = (vx_r1_p1 - in.vx1) * rdx;
    const double dvydy = (vy_r1_p1 - w_vy_r_p1) * rdy;
    const double div = dvxdx + dvydy;
    if constexpr (!TAU_IN) tau.p = (p.ptsc41 * in.mup) / p.max_nxy;
    const double pn = in.pold - tau.p * div;
    const double txx = (2.0 * in.mup) * (dvxdx - (1.0 / 3.0) * div);
    const double tyy = (2.0 * in.mup) * (dvydy - (1.0 / 3.0) * div);
    const double txy = ((2.0 * in.mub) * 0.5) * ((in.vx1 - w.vx_r) * rdy + (w_vy_r_p1 - w.vy_r) * rdx);
    const double txy_p1 = lane_down(txy, 1);

    const double rx = ((w_txx_prev_p1 - w.txx_prev) * rdx + (txy_p1 - w.txy_prev_p1) * rdy) -
                      (w_pn_prev_p1 - w.pn_prev) * rdx;
    const double av = 0.25 * (((w.vx_r + w_vx_r_p1) + in.vx1) + vx_r1_p1);
    const double d_xi = w_rho_r_p2 - w.rho_r;
    const double d_yi = rho_r1_p1 - w.rho_rm1_p1;
    const double ry = (((tyy - w.tyy_prev) * rdy + (txy_p1 - txy) * rdx) - (pn - w.pn_prev) * rdy) +
                      p.g * (w_rho_r_p1 - dt * (((av * d_xi) * 0.5) * rdx + ((w_vy_r_p1 * d_yi) * 0.5) * rdy));
    PTRow o;
    o.pn = pn;
    o.dvx = p.dampx * in.dvx + rx;
    o.dvy = p.dampy * in.dvy + ry;
    if constexpr (!TAU_IN) {
        tau.vx = (p.vsc_mdxy2 / ((w.mup_prev + w_mup_prev_p1) * 0.5)) / 4.1;
        tau.vy = (p.vsc_mdxy2 / ((w.mup_prev + in.mup) * 0.5)) / 4.1;
    }
    o.vx = w_vx_r_p1 + tau.vx * o.dvx;
    o.vy = w_vy_r_p1 + tau.vy * o.dvy;

    w.pn_prev = pn;
    w.txx_prev = txx;
    w.tyy_prev = tyy;
    w.txy_prev_p1 = txy_p1;
    w.mup_prev = in.mup;
    w.vx_r = in.vx1;
    w.vy_r = in.vy1;
    w.rho_rm1_p1 = w_rho_r_p1;
    w.rho_r = in.rho1;
    return o;
}

// Stores of B's row step r (iteration k+2), with pt_tile's conditions and BC writes.
template <class PP>
__device__ __forceinline__ void pt2_store(const PP& p, const PTRow& o, int c, int r, int r0, bool out_lane)
{
    const int nx = p.nx, ny = p.ny;
    const int64_t lvx = nx, lvy = nx + 1, lp = nx - 1;
    if (!out_lane || r < r0) return;
    if (c < nx - 1 && r <= ny - 2) p.p_b[ix2(c, r, lp)] = o.pn;
    const int j = r - 1;
    if (j < 0) return;
    if (c < nx - 2) {
        p.dvx_o[ix2(c, j, nx - 2)] = o.dvx;
        p.vx_b[ix2(c + 1, j + 1, lvx)] = o.vx;
        if (j == 0) p.vx_b[ix2(c + 1, 0, lvx)] = o.vx;        // bc_y_noflux!
        if (j == ny - 2) p.vx_b[ix2(c + 1, ny, lvx)] = o.vx;
    }
    if (c == 0 || c == nx - 2) {                                 // bc_x_zero!
        const int col = (c == 0) ? 0 : nx - 1;
        p.vx_b[ix2(col, j + 1, lvx)] = 0.0;
        if (j == 0) p.vx_b[ix2(col, 0, lvx)] = 0.0;
        if (j == ny - 2) p.vx_b[ix2(col, ny, lvx)] = 0.0;
    }
    if (c < nx - 1 && j <= ny - 3) {
        p.dvy_o[ix2(c, j, nx - 1)] = o.dvy;
        p.vy_b[ix2(c + 1, j + 1, lvy)] = o.vy;
        if (c == 0) p.vy_b[ix2(0, j + 1, lvy)] = o.vy;           // bc_x_noflux!
        if (c == nx - 2) p.vy_b[ix2(nx, j + 1, lvy)] = o.vy;
    }
    if (c < nx - 1 && (j == 0 || j == ny - 2)) {                 // bc_y_zero!
        const int row = (j == 0) ? 0 : ny - 1;
        p.vy_b[ix2(c + 1, row, lvy)] = 0.0;
        if (c == 0) p.vy_b[ix2(0, row, lvy)] = 0.0;
        if (c == nx - 2) p.vy_b[ix2(nx, row, lvy)] = 0.0;
    }
}

// Two PT iterations over the tile (strip, segid): reads state k from (vx_a, vy_a, p_a, dvx,
// dvy), writes state k+2 to (vx_b, vy_b, p_b, dvx_o, dvy_o).  Needs nx, ny >= 3.
template <class PP>
__device__ __forceinline__ void pt2_tile(const PP& p, int strip, int segid)
{
    const int lane = threadIdx.x & 63;
    const int c = strip * PT2_OUT_LANES + lane - 1;
    const int nx = p.nx, ny = p.ny;
    const bool out_lane = lane >= 1 && lane <= PT2_OUT_LANES;
    const int r0 = segid * p.seg;
    const int r1 = min(r0 + p.seg, ny);
    const double dt = *p.dt;
    const int cc = c < 0 ? 0 : c;
    LaneCols lc;
    lc.vx = (unsigned)min(cc, nx - 1);
    lc.vy = (unsigned)min(cc, nx);
    lc.p = (unsigned)min(cc, nx - 2);
    lc.dvx = (unsigned)min(cc, nx - 3);
    lc.dvy = (unsigned)min(cc, nx - 2);
    // V^{k+1} at column c+1 from A's lane: free-slip column BCs
    const bool vx_zero = (c + 1 == 0) || (c + 1 == nx - 1);
    const bool vy_left = c + 1 == 0, vy_right = c + 1 == nx;

    const int ra0 = max(r0 - 2, 0);
    const int ra1 = min(r1 + 1, ny);
    const int rb0 = max(r0 - 1, 0);

    PTWin wa, wb;
    int r = ra0;
    pt_win_init(wa, (p.vx_a + (int64_t)r * nx)[lc.vx], (p.vy_a + (int64_t)r * (nx + 1))[lc.vy],
                (p.rho + (int64_t)r * (nx + 1))[lc.vy]);
    RowIn nxt = load_row(p, lc, r, r >= r0 - 1);
    // A's previous row step, as B consumes it
    double a_pn = 0.0, a_mup = 0.0, a_mub = 0.0, a_dvx = 0.0, a_dvy = 0.0, a_rho = 0.0, a_vx = 0.0, a_vy = 0.0;
    double a_vx_last = 0.0;
    PTTau a_tau{0.0, 0.0, 0.0};

    for (; r < ra1; ++r) {
        const RowIn in = nxt;
        if (r + 1 < ra1) nxt = load_row(p, lc, r + 1, true);
        const double rho_row = wa.rho_r;  // rho row r
        PTTau tau_a;
        const PTRow oa = pt_row<false>(p, dt, wa, in, tau_a);
        // iteration k+1's V row r at column c+1, as pt_tile stores it
        const double vx_a = vx_zero ? 0.0 : oa.vx;
        const double vy_dn = lane_down(oa.vy, 1), vy_up = lane_up(oa.vy, 1);
        const double vy_a = (r == ny - 1) ? 0.0 : (vy_left ? vy_dn : (vy_right ? vy_up : oa.vy));
        const double vx_in = lane_up(vx_a, 1);  // column c
        const double vy_in = lane_up(vy_a, 1);

        const int rb = r - 1;  // B's row step
        if (rb >= rb0) {
            if (rb == rb0) {
                // B's window at row rb0: V row rb0 of iteration k+1 (Vx row 0 = row 1, Vy row 0 = 0)
                const double vx0 = (rb0 == 0) ? vx_in : lane_up(a_vx, 1);
                const double vy0 = (rb0 == 0) ? 0.0 : lane_up(a_vy, 1);
                pt_win_init(wb, vx0, vy0, a_rho);
            }
            RowIn ib;
            ib.vx1 = vx_in;
            ib.vy1 = vy_in;
            ib.rho1 = rho_row;
            ib.pold = a_pn;
            ib.mup = a_mup;
            ib.mub = a_mub;
            ib.dvx = a_dvx;
            ib.dvy = a_dvy;
            const PTRow ob = pt_row<true>(p, dt, wb, ib, a_tau);
            pt2_store(p, ob, c, rb, r0, out_lane);
        }
        a_tau = tau_a;
        a_pn = oa.pn;
        a_mup = in.mup;
        a_mub = in.mub;
        a_dvx = oa.dvx;
        a_dvy = oa.dvy;
        a_rho = rho_row;
        a_vx = vx_a;
        a_vy = vy_a;
        a_vx_last = vx_in;
    }
    if (r1 == ny) {
        // B's row ny-1: Vx row ny is the copy of row ny-1; Vy row ny-1 is zero
        const int rb = ny - 1;
        if (rb == rb0) pt_win_init(wb, lane_up(a_vx, 1), 0.0, a_rho);
        RowIn ib;
        ib.vx1 = a_vx_last;
        ib.vy1 = 0.0;
        ib.rho1 = a_rho;
        ib.pold = a_pn;
        ib.mup = a_mup;
        ib.mub = a_mub;
        ib.dvx = a_dvx;
        ib.dvy = a_dvy;
        const PTRow ob = pt_row<true>(p, dt, wb, ib, a_tau);
        pt2_store(p, ob, c, rb, r0, out_lane);
    }
}

}  // namespace sf
