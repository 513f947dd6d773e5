// sf_ptk_tile.h -- K PT iterations per tile pass (temporal blocking of depth K), FP64.
//
// The K-stage generalisation of pt2_tile (sf_pt2_tile.h, which explains the scheme for K = 2).
// Stage s runs iteration k+1+s.  Stage 0 reads the state from memory; stage s >= 1 consumes
// stage s-1's rows one row behind, with V shifted by one lane (stage s-1's lane l produces V
// at column c+1).  Every stage keeps pt_tile's lane-to-column map, so stage s has valid
// results on lanes s..61-s and the last stage stores lanes K-1..64-K: strips advance by
// 64 - 2K columns.  For output rows [r0, r1) stage s runs rows r0-(K-s) .. r1-1+(K-1-s)
// (clamped to the grid); at the bottom edge the stages finish on a virtual row ny whose V is
// the free-slip copy (Vx row ny = row ny-1, Vy zero).  All stages evaluate pt_row, so the
// stored iteration k+K is bit-identical to K pt_tile launches.  dV goes to dvx_o / dvy_o.
//
// Three loop forms, with the same results.  The plain form (SPLIT = 0) runs every outer step
// through one body that checks each stage's row range, starts windows and runs the virtual row.
// The split forms run those checks only on the steps at the ends of the tile, and the steps in
// between, where every stage has a row of its range other than its first and no stage reaches
// row ny-1, through a body without them: one step per loop trip (SPLIT = 1) or two (SPLIT = 2,
// which also spares the copies between loop-carried registers at the back edge).  Measured
// (same box, us per iteration; 256-instance wide FP64 solve): four iterations per pass, one step
// per trip (251-252 VGPRs) against plain (238): 4096^2 106.5 / 108.7, 2048^2 28.2 / 30.0, 512^2
// 5.25 / 5.53, wide 12.84 / 13.74 s (two steps per trip need 260 VGPRs, past two waves per
// SIMD); three iterations per pass, two steps per trip (214 VGPRs): the wide solve 5% faster,
// the 2048^2-4096^2 single pass no faster, so the single three-iteration pass stays plain.
#pragma once

#include "sf_pt2_tile.h"

namespace sf {

// What a stage hands to the next one about its last computed row.
struct PTHand {
    double pn, mup, mub, dvx, dvy, rho, vx, vy;  // vx, vy: V at column c (shifted for the next stage)
    PTTau tau;                                   // the row's pseudo-time steps (stage 0's)
};

template <int K, int SPLIT, class PP>
__device__ __forceinline__ void ptk_tile(const PP& p, int strip, int segid)
{
    constexpr int OUT = 64 - 2 * K;
    const int lane = threadIdx.x & 63;
    const int c = strip * OUT + lane - (K - 1);
    const int nx = p.nx, ny = p.ny;
    const bool out_lane = lane >= K - 1 && lane < K - 1 + OUT;
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
    const bool vx_zero = (c + 1 == 0) || (c + 1 == nx - 1);
    const bool vy_left = c + 1 == 0, vy_right = c + 1 == nx;

    int rs0[K], rs1[K];  // row range of stage s (inclusive)
#pragma unroll
    for (int s = 0; s < K; ++s) {
        rs0[s] = max(r0 - (K - s), 0);
        rs1[s] = min(r1 - 1 + (K - 1 - s), ny - 1);
    }
    const int r_first = rs0[0];
    const int r_last = r1 - 1 + (K - 1);  // outer step at which the last stage ends

    PTWin w[K];
    PTHand h[K];  // stage s's hand-over from its previous row
    double vx_last[K];  // stage s's shifted V of row ny-1 (for the virtual row ny)
#pragma unroll
    for (int s = 0; s < K; ++s) {
        h[s].pn = h[s].mup = h[s].mub = h[s].dvx = h[s].dvy = h[s].rho = h[s].vx = h[s].vy = 0.0;
        h[s].tau = PTTau{0.0, 0.0, 0.0};
        vx_last[s] = 0.0;
    }
    int r = r_first;
    pt_win_init(w[0], (p.vx_a + (int64_t)r * nx)[lc.vx], (p.vy_a + (int64_t)r * (nx + 1))[lc.vy],
                (p.rho + (int64_t)r * (nx + 1))[lc.vy]);
    RowIn nxt = load_row(p, lc, r, r >= r0 - (K - 1));

    if constexpr (SPLIT == 0) {
        // The plain form: edge_step's body written out in the loop (called through the lambda
        // the same code ran 3% slower at 2048^2-4096^2, from a different scalar layout).
        for (; r <= r_last; ++r) {
            // stage s's output of this outer step, as stage s+1 consumes it in the same step; the
            // hand-over of its row (hn) replaces h[s] only after the step, since stage s+1 needs
            // stage s's previous row
            double cur_vx[K], cur_vy[K], cur_rho[K];
            PTHand hn[K];
            bool ran[K];
#pragma unroll
            for (int s = 0; s < K; ++s) {
                const int rs = r - s;
                ran[s] = false;
                cur_vx[s] = cur_vy[s] = cur_rho[s] = 0.0;
                hn[s] = h[s];
                if (rs < rs0[s] || rs > rs1[s]) {
                    if (s < K - 1 && rs == ny && rs - 1 >= rs0[s]) {
                        // virtual row ny of stage s: Vx row ny = row ny-1, Vy zero, rho clamped
                        cur_vx[s] = vx_last[s];
                        cur_vy[s] = 0.0;
                        cur_rho[s] = h[s].rho;
                    }
                    continue;
                }
                RowIn in;
                if (s == 0) {
                    in = nxt;
                    if (r + 1 <= rs1[0]) nxt = load_row(p, lc, r + 1, true);
                } else {
                    in.vx1 = cur_vx[s - 1];
                    in.vy1 = cur_vy[s - 1];
                    in.rho1 = cur_rho[s - 1];
                    in.pold = h[s - 1].pn;
                    in.mup = h[s - 1].mup;
                    in.mub = h[s - 1].mub;
                    in.dvx = h[s - 1].dvx;
                    in.dvy = h[s - 1].dvy;
                    if (rs == rs0[s]) {
                        // the window at row rs: V row rs of the previous stage (row 0: Vx = row 1, Vy = 0)
                        const double vx0 = (rs == 0) ? in.vx1 : h[s - 1].vx;
                        const double vy0 = (rs == 0) ? 0.0 : h[s - 1].vy;
                        pt_win_init(w[s], vx0, vy0, h[s - 1].rho);
                    }
                }
                const double rho_row = w[s].rho_r;
                PTTau tau;
                PTRow o;
                if (s == 0) {
                    o = pt_row<false>(p, dt, w[s], in, tau);
                } else {
                    tau = h[s - 1].tau;
                    o = pt_row<true>(p, dt, w[s], in, tau);
                }
                if (s == K - 1) {
                    pt2_store(p, o, c, rs, r0, out_lane);
                } else {
                    const double vx_a = vx_zero ? 0.0 : o.vx;
                    const double vy_dn = lane_down(o.vy, 1), vy_up = lane_up(o.vy, 1);
                    const double vy_a = (rs == ny - 1) ? 0.0 : (vy_left ? vy_dn : (vy_right ? vy_up : o.vy));
                    cur_vx[s] = lane_up(vx_a, 1);
                    cur_vy[s] = lane_up(vy_a, 1);
                    cur_rho[s] = rho_row;
                    hn[s].pn = o.pn;
                    hn[s].mup = in.mup;
                    hn[s].mub = in.mub;
                    hn[s].dvx = o.dvx;
                    hn[s].dvy = o.dvy;
                    hn[s].rho = rho_row;
                    hn[s].vx = cur_vx[s];
                    hn[s].vy = cur_vy[s];
                    hn[s].tau = tau;
                    ran[s] = true;
                }
            }
#pragma unroll
            for (int s = 0; s < K - 1; ++s) {
                if (ran[s]) vx_last[s] = cur_vx[s];
                h[s] = hn[s];
            }
        }
    } else {
        // One outer step with every check.  Stage s's output of the step is consumed by stage s+1
        // in the same step; the hand-over of its row (hn) replaces h[s] only after the step,
        // since stage s+1 needs stage s's previous row.
        auto edge_step = [&](int r) {
            double cur_vx[K], cur_vy[K], cur_rho[K];
            PTHand hn[K];
            bool ran[K];
#pragma unroll
            for (int s = 0; s < K; ++s) {
                const int rs = r - s;
                ran[s] = false;
                cur_vx[s] = cur_vy[s] = cur_rho[s] = 0.0;
                hn[s] = h[s];
                if (rs < rs0[s] || rs > rs1[s]) {
                    if (s < K - 1 && rs == ny && rs - 1 >= rs0[s]) {
                        // virtual row ny of stage s: Vx row ny = row ny-1, Vy zero, rho clamped
                        cur_vx[s] = vx_last[s];
                        cur_vy[s] = 0.0;
                        cur_rho[s] = h[s].rho;
                    }
                    continue;
                }
                RowIn in;
                if (s == 0) {
                    in = nxt;
                    if (r + 1 <= rs1[0]) nxt = load_row(p, lc, r + 1, true);
                } else {
                    in.vx1 = cur_vx[s - 1];
                    in.vy1 = cur_vy[s - 1];
                    in.rho1 = cur_rho[s - 1];
                    in.pold = h[s - 1].pn;
                    in.mup = h[s - 1].mup;
                    in.mub = h[s - 1].mub;
                    in.dvx = h[s - 1].dvx;
                    in.dvy = h[s - 1].dvy;
                    if (rs == rs0[s]) {
                        // the window at row rs: V row rs of the previous stage (row 0: Vx = row 1, Vy = 0)
                        const double vx0 = (rs == 0) ? in.vx1 : h[s - 1].vx;
                        const double vy0 = (rs == 0) ? 0.0 : h[s - 1].vy;
                        pt_win_init(w[s], vx0, vy0, h[s - 1].rho);
                    }
                }
                const double rho_row = w[s].rho_r;
                PTTau tau;
                PTRow o;
                if (s == 0) {
                    o = pt_row<false>(p, dt, w[s], in, tau);
                } else {
                    tau = h[s - 1].tau;
                    o = pt_row<true>(p, dt, w[s], in, tau);
                }
                if (s == K - 1) {
                    pt2_store(p, o, c, rs, r0, out_lane);
                } else {
                    const double vx_a = vx_zero ? 0.0 : o.vx;
                    const double vy_dn = lane_down(o.vy, 1), vy_up = lane_up(o.vy, 1);
                    const double vy_a = (rs == ny - 1) ? 0.0 : (vy_left ? vy_dn : (vy_right ? vy_up : o.vy));
                    cur_vx[s] = lane_up(vx_a, 1);
                    cur_vy[s] = lane_up(vy_a, 1);
                    cur_rho[s] = rho_row;
                    hn[s].pn = o.pn;
                    hn[s].mup = in.mup;
                    hn[s].mub = in.mub;
                    hn[s].dvx = o.dvx;
                    hn[s].dvy = o.dvy;
                    hn[s].rho = rho_row;
                    hn[s].vx = cur_vx[s];
                    hn[s].vy = cur_vy[s];
                    hn[s].tau = tau;
                    ran[s] = true;
                }
            }
#pragma unroll
            for (int s = 0; s < K - 1; ++s) {
                if (ran[s]) vx_last[s] = cur_vx[s];
                h[s] = hn[s];
            }
        };

        // An outer step in the middle of the tile: every stage runs a row of its range other
        // than its first, no stage reaches row ny-1 and stage 0's next row exists.  The
        // hand-over's rho and V are read only by a stage's first row and by the virtual row,
        // which are edge steps after an edge step has written them, so they are not carried.
        auto mid_step = [&](int r) {
            double cvx = 0.0, cvy = 0.0, crho = 0.0;
            PTHand hn[K];
#pragma unroll
            for (int s = 0; s < K; ++s) {
                const int rs = r - s;
                RowIn in;
                if (s == 0) {
                    in = nxt;
                    nxt = load_row(p, lc, r + 1, true);
                } else {
                    in = RowIn{cvx, cvy, crho, h[s - 1].pn, h[s - 1].mup, h[s - 1].mub, h[s - 1].dvx, h[s - 1].dvy};
                }
                const double rho_row = w[s].rho_r;
                PTTau tau;
                PTRow o;
                if (s == 0) {
                    o = pt_row<false>(p, dt, w[s], in, tau);
                } else {
                    tau = h[s - 1].tau;
                    o = pt_row<true>(p, dt, w[s], in, tau);
                }
                if (s == K - 1) {
                    pt2_store(p, o, c, rs, r0, out_lane);
                } else {
                    const double vx_a = vx_zero ? 0.0 : o.vx;
                    const double vy_dn = lane_down(o.vy, 1), vy_up = lane_up(o.vy, 1);
                    const double vy_a = vy_left ? vy_dn : (vy_right ? vy_up : o.vy);
                    cvx = lane_up(vx_a, 1);
                    cvy = lane_up(vy_a, 1);
                    crho = rho_row;
                    hn[s] = PTHand{o.pn, in.mup, in.mub, o.dvx, o.dvy, 0.0, 0.0, 0.0, tau};
                }
            }
#pragma unroll
            for (int s = 0; s < K - 1; ++s) h[s] = hn[s];
        };
        // middle steps: stage s's row r-s lies in (rs0[s], rs1[s]] and r+1 <= rs1[0]
        int r_mid0 = 0, r_mid1 = rs1[0] - 1;
#pragma unroll
        for (int s = 0; s < K; ++s) {
            r_mid0 = max(r_mid0, rs0[s] + s + 1);
            r_mid1 = min(r_mid1, rs1[s] + s);
        }
        for (; r < r_mid0 && r <= r_last; ++r) edge_step(r);
        if constexpr (SPLIT == 2) {
            for (; r + 1 <= r_mid1; r += 2) {
                mid_step(r);
                mid_step(r + 1);
            }
            if (r <= r_mid1) mid_step(r++);
        } else {
            for (; r <= r_mid1; ++r) mid_step(r);
        }
        for (; r <= r_last; ++r) edge_step(r);
    }
}

}  // namespace sf
