// sf_refine.hip -- FP32 pseudo-transient sweeps inside FP64 defect correction (gfx950).
//
// The PT operator of solveStokes! is affine in (V, P) for a fixed stabilisation dt:
//   Rx(u) = d(txx)/dx + d(txy)/dy - dP/dx,
//   Ry(u) = d(tyy)/dy + d(txy)/dx - dP/dy + g (rho - dt (avVx drho/dx + Vy drho/dy)),
//   C(u)  = div V,
// with tau linear in V (StokesSolver_multixpu.jl:140-202).  Defect correction:
//   1. FP64: r = (Rx, Ry, C)(u_k) at the current state (resid_f64), err_k = max mean |r|.
//   2. FP32: a window of PT iterations on the correction d with the same stencils,
//      homogeneous free-slip BCs and the FP64 residual as forcing (pt_corr_f32):
//        Pd  <- Pd - dtauP (div Vd + C_k)
//        Rdx  = Lx(d) + Rx_k,   Rdy = Ly(d) + Ry_k   (Ly holds the -g dt (...) term on Vd)
//        mdx <- dampX mdx + Rdx, Vdx <- Vdx + dtauVx mdx  (same for y)
//      The PT momentum md is carried from window to window, so the sequence of windows
//      is the PT iteration itself; only the FP32 rounding is reset at every window.
//   3. FP64: u_{k+1} = u_k + d (update_f64); dt refreshed every ndtupdate iterations.
// The FP32 kernel is the FP64 fused iteration's structure (one launch per iteration,
// register row-march, lane shifts for x neighbours) on 16 FP32 arrays per node: read
// Vd, Pd, md (x, y), mu_p, mu_b, rho, forcing (x, y, C); write Vd, Pd, md.
#include <algorithm>
#include <type_traits>

#include "sf_device.h"
#include "sf_internal.h"

namespace sf {

#define CORR_OUT_LANES 62

// The FP32 sweeps' pseudo-time steps (dtauP, dtauVx, dtauVy of compute_timesteps!,
// StokesSolver_multixpu.jl:128-133) take the hardware reciprocal (v_rcp_f32, 1 ulp) in place of
// IEEE division.  The steps only scale the correction sweeps and convergence is decided on FP64
// residuals, so their last bits do not matter; the five divisions per node were a tenth of the
// sweep's VALU work (2048^2: 21.9 -> 19.8 us per sweep).  Every FP32 kernel uses the same
// expressions, so all sweep groupings stay bit-identical to each other.
__device__ __forceinline__ float tau_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

__device__ inline int xcd_remap_r(int k, int total)
{
    const int nxcd = 8;
    int xcd = k % nxcd, idx = k / nxcd;
    int bpx = total / nxcd, rem = total % nxcd;
    return (xcd < rem) ? xcd * (bpx + 1) + idx : rem * (bpx + 1) + (xcd - rem) * bpx + idx;
}

__device__ inline int crow_r(int r, int nrows) { return r < 0 ? 0 : (r >= nrows ? nrows - 1 : r); }

struct CorrRowIn {
    float vx1, vy1, rho1, pold, mup, mub, fc, mx, my, fx, fy;
    // the next row's values one lane down (columns c + 1, c + 2), when the caller has them
    float vx1_p1, vy1_p1, rho1_p1, rho1_p2;
};

struct CorrCols {
    unsigned vx, vy, p, dvx, dvy;
};

// The FP32 sweep's arithmetic, shared by every FP32 kernel so that all sweep groupings give the
// same bits.  It is the PT iteration's (StokesSolver_multixpu.jl:140-234) on the correction, in an
// order of its own: products with a common factor are grouped, constants folded and a*b+c fused,
// which the FP64 kernels may not do (they reproduce the reference bit for bit) but the correction
// sweeps may: they only have to drive the FP64 residual down, and convergence is decided in FP64.
struct CorrC {
    float rdx, rdy;
    float kx, ky;  // g dt / 8 / dx and g dt / 2 / dy: the free-surface term's factors
    float dampx, dampy;
};

template <class PP>
__device__ __forceinline__ CorrC corr_consts(const PP& p, float dt)
{
    const float gdt = p.g * dt;
    return CorrC{p.rdx, p.rdy, (gdt * 0.125f) * p.rdx, (gdt * 0.5f) * p.rdy, p.dampx, p.dampy};
}

// The P update and the stresses at one node (compute_P_tau!, :140-169): m2 = 2 mu_p, taup = dtauP.
struct CorrPT {
    float dvxdx, dvydy, cres, pn, txx, tyy, txy;
};

__device__ __forceinline__ CorrPT corr_ptau(const CorrC& k, float vx1, float vx1_p1, float vx_r, float vy1_p1,
                                            float vy_r, float vy_r_p1, float m2, float mub, float fc, float pold,
                                            float taup)
{
    CorrPT o;
    o.dvxdx = (vx1_p1 - vx1) * k.rdx;
    o.dvydy = (vy1_p1 - vy_r_p1) * k.rdy;
    const float div = o.dvxdx + o.dvydy;
    o.cres = div + fc;  // continuity residual of the full state
    o.pn = __builtin_fmaf(-taup, o.cres, pold);
    const float d3 = div * (1.0f / 3.0f);
    o.txx = m2 * (o.dvxdx - d3);
    o.tyy = m2 * (o.dvydy - d3);
    o.txy = mub * __builtin_fmaf(vx1 - vx_r, k.rdy, (vy_r_p1 - vy_r) * k.rdx);
    return o;
}

// x momentum residual (compute_ResX, :176-182) plus the forcing.
__device__ __forceinline__ float corr_rx(const CorrC& k, float txx_p1, float txx, float pn_p1, float pn, float txy_p1,
                                         float txy_prev_p1, float fx)
{
    return __builtin_fmaf((txx_p1 - txx) - (pn_p1 - pn), k.rdx, __builtin_fmaf(txy_p1 - txy_prev_p1, k.rdy, fx));
}

// y momentum residual with the free-surface term (compute_ResY, :189-202) plus the forcing;
// av4 is four times the Vx average, d_xi and d_yi the two rho differences.
__device__ __forceinline__ float corr_ry(const CorrC& k, float tyy, float tyy_prev, float pn, float pn_prev,
                                         float txy_p1, float txy, float av4, float d_xi, float vy, float d_yi,
                                         float fy)
{
    const float stab = __builtin_fmaf(av4 * d_xi, k.kx, (vy * d_yi) * k.ky);
    return __builtin_fmaf((tyy - tyy_prev) - (pn - pn_prev), k.rdy, __builtin_fmaf(txy_p1 - txy, k.rdx, fy)) - stab;
}

// The pseudo-time steps (compute_timesteps!, :128-133) from mu_p.
template <class PP>
__device__ __forceinline__ float corr_taup(const PP& p, float mup)
{
    return (p.ptsc41 * mup) * tau_rcp(p.max_nxy);
}
template <class PP>
__device__ __forceinline__ float corr_tauv(const PP& p, float mup_a, float mup_b)
{
    return p.vsc_mdxy2 * tau_rcp((mup_a + mup_b) * 2.05f);
}

template <class PP>
__device__ __forceinline__ CorrRowIn corr_load_row(const PP& p, const CorrCols& lc, int r, bool need_j)
{
    const int nx = p.nx, ny = p.ny;
    CorrRowIn in;
    in.vx1 = (p.vx_a + (int64_t)crow_r(r + 1, ny + 1) * nx)[lc.vx];
    in.vy1 = (p.vy_a + (int64_t)crow_r(r + 1, ny) * (nx + 1))[lc.vy];
    in.rho1 = (p.rho + (int64_t)crow_r(r + 1, ny) * (nx + 1))[lc.vy];
    in.pold = (p.p_a + (int64_t)crow_r(r, ny - 1) * (nx - 1))[lc.p];
    in.mup = (p.mu_p + (int64_t)crow_r(r, ny - 1) * (nx - 1))[lc.p];
    in.fc = (p.fc + (int64_t)crow_r(r, ny - 1) * (nx - 1))[lc.p];
    in.mub = (p.mu_b + (int64_t)crow_r(r, ny) * nx)[lc.vx];
    if (need_j) {
        in.mx = (p.mx + (int64_t)crow_r(r - 1, ny - 1) * (nx - 2))[lc.dvx];
        in.fx = (p.fx + (int64_t)crow_r(r - 1, ny - 1) * (nx - 2))[lc.dvx];
        in.my = (p.my + (int64_t)crow_r(r - 1, ny - 2) * (nx - 1))[lc.dvy];
        in.fy = (p.fy + (int64_t)crow_r(r - 1, ny - 2) * (nx - 1))[lc.dvy];
    } else {
        in.mx = in.fx = in.my = in.fy = 0.f;
    }
    return in;
}

// PP is CorrParams in any address space (a kernel argument, or constant memory in the wide
// batched kernel).
template <int DEPTH, bool CHECK, class PP>
__device__ __forceinline__ void corr_tile(const PP& p, int strip, int segid)
{
    const int lane = threadIdx.x;
    const int c = strip * CORR_OUT_LANES + lane;
    const int nx = p.nx, ny = p.ny;
    const bool out_lane = lane < CORR_OUT_LANES;
    const int r0 = segid * p.seg;
    const int r1 = min(r0 + p.seg, ny);
    const float dt = p.dt;
    const int64_t lvx = nx, lvy = nx + 1, lp = nx - 1;
    const bool c_p = c < nx - 1;
    const CorrC k = corr_consts(p, dt);
    // momentum in place unless a second buffer is given (ping-pong in the wide batched mode)
    float* const mx_w = p.mx_o ? p.mx_o : p.mx;
    float* const my_w = p.my_o ? p.my_o : p.my;
    CorrCols lc;
    lc.vx = (unsigned)min(c, nx - 1);
    lc.vy = (unsigned)min(c, nx);
    lc.p = (unsigned)min(c, nx - 2);
    lc.dvx = (unsigned)min(c, nx - 3);
    lc.dvy = (unsigned)min(c, nx - 2);

    int r = (r0 == 0) ? 0 : r0 - 1;
    float vx_r = (p.vx_a + (int64_t)r * nx)[lc.vx];
    float vy_r = (p.vy_a + (int64_t)r * (nx + 1))[lc.vy];
    float rho_r = (p.rho + (int64_t)r * (nx + 1))[lc.vy];
    CorrRowIn nxt = corr_load_row(p, lc, r, r >= r0);
    CorrRowIn nxt2;
    if constexpr (DEPTH > 1)
        if (r + 1 < r1) nxt2 = corr_load_row(p, lc, r + 1, r + 1 >= r0);
    float vx_r_p1 = lane_down(vx_r, 1);
    float vy_r_p1 = lane_down(vy_r, 1);
    float rho_r_p2 = lane_down(rho_r, 2);
    float rho_rm1_p1 = 0.f, rho_r_ = rho_r;
    float pn_prev = 0.f, tyy_prev = 0.f, txx_prev = 0.f;
    float pn_prev_p1 = 0.f, txx_prev_p1 = 0.f, txy_prev_p1 = 0.f;
    float mup_prev = 0.f, mup_prev_p1 = 0.f;
    float rho_r_p1 = lane_down(rho_r, 1);

    for (; r < r1; ++r) {
        const CorrRowIn in = nxt;
        if constexpr (DEPTH > 1) {
            nxt = nxt2;
            if (r + 2 < r1) nxt2 = corr_load_row(p, lc, r + 2, true);  // two rows of operands in flight
        } else {
            if (r + 1 < r1) nxt = corr_load_row(p, lc, r + 1, true);
        }
        const bool emit = r >= r0;
        const bool cell_row = r <= ny - 2;
        const int j = r - 1;
        const bool has_out = emit && j >= 0;
        const bool dvx_ok = has_out && c < nx - 2;
        const bool dvy_ok = has_out && c < nx - 1 && j <= ny - 3;

        const float vx_r1_p1 = lane_down(in.vx1, 1);
        const float vy_r1_p1 = lane_down(in.vy1, 1);
        const float rho_r1_p1 = lane_down(in.rho1, 1);
        const float rho_r1_p2 = lane_down(in.rho1, 2);

        const CorrPT t = corr_ptau(k, in.vx1, vx_r1_p1, vx_r, vy_r1_p1, vy_r, vy_r_p1, 2.0f * in.mup, in.mub, in.fc,
                                   in.pold, corr_taup(p, in.mup));
        const float pn = t.pn, cres = t.cres, txx = t.txx, tyy = t.tyy, txy = t.txy;
        const float txy_p1 = lane_down(txy, 1);

        if (emit && out_lane && c_p && cell_row) {
            p.p_b[ix2(c, r, lp)] = pn;
            if constexpr (CHECK) p.rco[ix2(c, r, lp)] = cres;
        }
        if (has_out && out_lane) {
            const float rx = corr_rx(k, txx_prev_p1, txx_prev, pn_prev_p1, pn_prev, txy_p1, txy_prev_p1, in.fx);
            const float av4 = ((vx_r + vx_r_p1) + in.vx1) + vx_r1_p1;
            const float ry = corr_ry(k, tyy, tyy_prev, pn, pn_prev, txy_p1, txy, av4, rho_r_p2 - rho_r_, vy_r_p1,
                                     rho_r1_p1 - rho_rm1_p1, in.fy);
            const float mx_new = __builtin_fmaf(k.dampx, in.mx, rx);
            const float my_new = __builtin_fmaf(k.dampy, in.my, ry);
            const float vx_new = __builtin_fmaf(corr_tauv(p, mup_prev, mup_prev_p1), mx_new, vx_r_p1);
            const float vy_new = __builtin_fmaf(corr_tauv(p, mup_prev, in.mup), my_new, vy_r_p1);
            if (dvx_ok) {
                mx_w[ix2(c, j, nx - 2)] = mx_new;
                p.vx_b[ix2(c + 1, j + 1, lvx)] = vx_new;
                if (j == 0) p.vx_b[ix2(c + 1, 0, lvx)] = vx_new;
                if (j == ny - 2) p.vx_b[ix2(c + 1, ny, lvx)] = vx_new;
                if constexpr (CHECK) p.rxo[ix2(c, j, nx - 2)] = rx;
            }
            if (c == 0 || c == nx - 2) {
                const int col = (c == 0) ? 0 : nx - 1;
                p.vx_b[ix2(col, j + 1, lvx)] = 0.f;
                if (j == 0) p.vx_b[ix2(col, 0, lvx)] = 0.f;
                if (j == ny - 2) p.vx_b[ix2(col, ny, lvx)] = 0.f;
            }
            if (dvy_ok) {
                my_w[ix2(c, j, nx - 1)] = my_new;
                p.vy_b[ix2(c + 1, j + 1, lvy)] = vy_new;
                if (c == 0) p.vy_b[ix2(0, j + 1, lvy)] = vy_new;
                if (c == nx - 2) p.vy_b[ix2(nx, j + 1, lvy)] = vy_new;
                if constexpr (CHECK) p.ryo[ix2(c, j, nx - 1)] = ry;
            }
            if (c < nx - 1 && (j == 0 || j == ny - 2)) {
                const int row = (j == 0) ? 0 : ny - 1;
                p.vy_b[ix2(c + 1, row, lvy)] = 0.f;
                if (c == 0) p.vy_b[ix2(0, row, lvy)] = 0.f;
                if (c == nx - 2) p.vy_b[ix2(nx, row, lvy)] = 0.f;
            }
        }

        pn_prev_p1 = lane_down(pn, 1);
        txx_prev_p1 = lane_down(txx, 1);
        mup_prev_p1 = lane_down(in.mup, 1);
        pn_prev = pn;
        txx_prev = txx;
        tyy_prev = tyy;
        txy_prev_p1 = txy_p1;
        mup_prev = in.mup;
        vx_r = in.vx1;
        vx_r_p1 = vx_r1_p1;
        vy_r = in.vy1;
        vy_r_p1 = vy_r1_p1;
        rho_rm1_p1 = rho_r_p1;
        rho_r_ = in.rho1;
        rho_r_p1 = rho_r1_p1;
        rho_r_p2 = rho_r1_p2;
    }
}

// ---------------------------------------------------------------------------------------
// Two correction sweeps per pass (temporal blocking), the FP32 counterpart of pt2_tile
// (sf_pt2_tile.h, which explains the lane and row scheme): sweep A marches as corr_tile does
// and keeps its rows in registers, sweep B consumes them one row behind and stores.  Both
// sweeps evaluate corr_tile's expressions (corr_row), so the result is bit-identical to two
// pt_corr_f32 launches.  The momentum goes to a second buffer (mx_o, my_o).

#define CORR2_OUT_LANES 60

__device__ inline float lane_up_f(float v, int d) { return lane_up(v, d); }

struct CorrWin {
    float vx_r, vx_r_p1, vy_r, vy_r_p1;
    float rho_r, rho_r_p1, rho_r_p2, rho_rm1_p1;
    float pn_prev, tyy_prev, txx_prev, pn_prev_p1, txx_prev_p1, txy_prev_p1, mup_prev, mup_prev_p1;
};

struct CorrRow {
    float pn, mx, my, vx, vy;
};

__device__ __forceinline__ void corr_win_init(CorrWin& w, float vx_r, float vy_r, float rho_r)
{
    w.vx_r = vx_r;
    w.vy_r = vy_r;
    w.rho_r = rho_r;
    w.vx_r_p1 = lane_down(vx_r, 1);
    w.vy_r_p1 = lane_down(vy_r, 1);
    w.rho_r_p1 = lane_down(rho_r, 1);
    w.rho_r_p2 = lane_down(rho_r, 2);
    w.rho_rm1_p1 = 0.f;
    w.pn_prev = w.tyy_prev = w.txx_prev = 0.f;
    w.pn_prev_p1 = w.txx_prev_p1 = w.txy_prev_p1 = 0.f;
    w.mup_prev = w.mup_prev_p1 = 0.f;
}

// The pseudo-time steps of a row and 2 mu_p (they depend on mu_p only; later sweeps of a row
// reuse them).
struct CorrTau {
    float p, vx, vy, m2;
};

// One row step of one correction sweep with corr_tile's expressions, then the window slides.
// SHIFTED: in carries the next row's lane-shifted values (a later sweep of a multi-sweep pass
// has them from the sweep before it), else they are shifted here.
template <bool TAU_IN, bool SHIFTED, class PP>
__device__ __forceinline__ CorrRow corr_row(const PP& p, const CorrC& k, CorrWin& w, const CorrRowIn& in, CorrTau& tau)
{
    float vx_r1_p1, vy_r1_p1, rho_r1_p1, rho_r1_p2;
    if constexpr (SHIFTED) {
        vx_r1_p1 = in.vx1_p1;
        vy_r1_p1 = in.vy1_p1;
        rho_r1_p1 = in.rho1_p1;
        rho_r1_p2 = in.rho1_p2;
    } else {
        vx_r1_p1 = lane_down(in.vx1, 1);
        vy_r1_p1 = lane_down(in.vy1, 1);
        rho_r1_p1 = lane_down(in.rho1, 1);
        rho_r1_p2 = lane_down(in.rho1, 2);
    }
    if constexpr (!TAU_IN) {
        tau.p = corr_taup(p, in.mup);
        tau.m2 = 2.0f * in.mup;
    }
    const CorrPT t = corr_ptau(k, in.vx1, vx_r1_p1, w.vx_r, vy_r1_p1, w.vy_r, w.vy_r_p1, tau.m2, in.mub, in.fc,
                               in.pold, tau.p);
    const float txy_p1 = lane_down(t.txy, 1);
    const float rx = corr_rx(k, w.txx_prev_p1, w.txx_prev, w.pn_prev_p1, w.pn_prev, txy_p1, w.txy_prev_p1, in.fx);
    const float av4 = ((w.vx_r + w.vx_r_p1) + in.vx1) + vx_r1_p1;
    const float ry = corr_ry(k, t.tyy, w.tyy_prev, t.pn, w.pn_prev, txy_p1, t.txy, av4, w.rho_r_p2 - w.rho_r,
                             w.vy_r_p1, rho_r1_p1 - w.rho_rm1_p1, in.fy);
    CorrRow o;
    o.pn = t.pn;
    o.mx = __builtin_fmaf(k.dampx, in.mx, rx);
    o.my = __builtin_fmaf(k.dampy, in.my, ry);
    if constexpr (!TAU_IN) {
        tau.vx = corr_tauv(p, w.mup_prev, w.mup_prev_p1);
        tau.vy = corr_tauv(p, w.mup_prev, in.mup);
    }
    o.vx = __builtin_fmaf(tau.vx, o.mx, w.vx_r_p1);
    o.vy = __builtin_fmaf(tau.vy, o.my, w.vy_r_p1);

    w.pn_prev_p1 = lane_down(t.pn, 1);
    w.txx_prev_p1 = lane_down(t.txx, 1);
    if constexpr (!TAU_IN) {
        w.mup_prev_p1 = lane_down(in.mup, 1);
        w.mup_prev = in.mup;
    }
    w.pn_prev = t.pn;
    w.txx_prev = t.txx;
    w.tyy_prev = t.tyy;
    w.txy_prev_p1 = txy_p1;
    w.vx_r = in.vx1;
    w.vx_r_p1 = vx_r1_p1;
    w.vy_r = in.vy1;
    w.vy_r_p1 = vy_r1_p1;
    w.rho_rm1_p1 = w.rho_r_p1;
    w.rho_r = in.rho1;
    w.rho_r_p1 = rho_r1_p1;
    w.rho_r_p2 = rho_r1_p2;
    return o;
}

// Stores of sweep B's row step r, with corr_tile's conditions and (homogeneous) BC writes.
template <class PP>
__device__ __forceinline__ void corr2_store(const PP& p, const CorrRow& o, int c, int r, int r0, bool out_lane)
{
    const int nx = p.nx, ny = p.ny;
    const int64_t lvx = nx, lvy = nx + 1, lp = nx - 1;
    if (!out_lane || r < r0) return;
    if (c < nx - 1 && r <= ny - 2) p.p_b[ix2(c, r, lp)] = o.pn;
    const int j = r - 1;
    if (j < 0) return;
    if (c < nx - 2) {
        p.mx_o[ix2(c, j, nx - 2)] = o.mx;
        p.vx_b[ix2(c + 1, j + 1, lvx)] = o.vx;
        if (j == 0) p.vx_b[ix2(c + 1, 0, lvx)] = o.vx;
        if (j == ny - 2) p.vx_b[ix2(c + 1, ny, lvx)] = o.vx;
    }
    if (c == 0 || c == nx - 2) {
        const int col = (c == 0) ? 0 : nx - 1;
        p.vx_b[ix2(col, j + 1, lvx)] = 0.f;
        if (j == 0) p.vx_b[ix2(col, 0, lvx)] = 0.f;
        if (j == ny - 2) p.vx_b[ix2(col, ny, lvx)] = 0.f;
    }
    if (c < nx - 1 && j <= ny - 3) {
        p.my_o[ix2(c, j, nx - 1)] = o.my;
        p.vy_b[ix2(c + 1, j + 1, lvy)] = o.vy;
        if (c == 0) p.vy_b[ix2(0, j + 1, lvy)] = o.vy;
        if (c == nx - 2) p.vy_b[ix2(nx, j + 1, lvy)] = o.vy;
    }
    if (c < nx - 1 && (j == 0 || j == ny - 2)) {
        const int row = (j == 0) ? 0 : ny - 1;
        p.vy_b[ix2(c + 1, row, lvy)] = 0.f;
        if (c == 0) p.vy_b[ix2(0, row, lvy)] = 0.f;
        if (c == nx - 2) p.vy_b[ix2(nx, row, lvy)] = 0.f;
    }
}

// Two correction sweeps over the tile (strip, segid): state (vx_a, vy_a, p_a, mx, my) ->
// (vx_b, vy_b, p_b, mx_o, my_o).
template <class PP>
__device__ __forceinline__ void corr2_tile(const PP& p, int strip, int segid)
{
    const int lane = threadIdx.x & 63;
    const int c = strip * CORR2_OUT_LANES + lane - 1;
    const int nx = p.nx, ny = p.ny;
    const bool out_lane = lane >= 1 && lane <= CORR2_OUT_LANES;
    const int r0 = segid * p.seg;
    const int r1 = min(r0 + p.seg, ny);
    const CorrC k = corr_consts(p, p.dt);
    const int cc = c < 0 ? 0 : c;
    CorrCols lc;
    lc.vx = (unsigned)min(cc, nx - 1);
    lc.vy = (unsigned)min(cc, nx);
    lc.p = (unsigned)min(cc, nx - 2);
    lc.dvx = (unsigned)min(cc, nx - 3);
    lc.dvy = (unsigned)min(cc, nx - 2);
    const bool vx_zero = (c + 1 == 0) || (c + 1 == nx - 1);
    const bool vy_left = c + 1 == 0, vy_right = c + 1 == nx;

    const int ra0 = max(r0 - 2, 0);
    const int ra1 = min(r1 + 1, ny);
    const int rb0 = max(r0 - 1, 0);

    CorrWin wa, wb;
    int r = ra0;
    corr_win_init(wa, (p.vx_a + (int64_t)r * nx)[lc.vx], (p.vy_a + (int64_t)r * (nx + 1))[lc.vy],
                  (p.rho + (int64_t)r * (nx + 1))[lc.vy]);
    CorrRowIn nxt = corr_load_row(p, lc, r, r >= r0 - 1);
    float a_pn = 0.f, a_mup = 0.f, a_mub = 0.f, a_fc = 0.f, a_mx = 0.f, a_my = 0.f, a_fx = 0.f, a_fy = 0.f;
    float a_rho = 0.f, a_vx = 0.f, a_vy = 0.f, a_vx_last = 0.f;
    CorrTau a_tau{0.f, 0.f, 0.f, 0.f};

    for (; r < ra1; ++r) {
        const CorrRowIn in = nxt;
        if (r + 1 < ra1) nxt = corr_load_row(p, lc, r + 1, true);
        const float rho_row = wa.rho_r;
        CorrTau tau_a;
        const CorrRow oa = corr_row<false, false>(p, k, wa, in, tau_a);
        const float vx_a = vx_zero ? 0.f : oa.vx;
        const float vy_dn = lane_down(oa.vy, 1), vy_up = lane_up_f(oa.vy, 1);
        const float vy_a = (r == ny - 1) ? 0.f : (vy_left ? vy_dn : (vy_right ? vy_up : oa.vy));
        const float vx_in = lane_up_f(vx_a, 1);
        const float vy_in = lane_up_f(vy_a, 1);

        const int rb = r - 1;
        if (rb >= rb0) {
            if (rb == rb0) {
                const float vx0 = (rb0 == 0) ? vx_in : lane_up_f(a_vx, 1);
                const float vy0 = (rb0 == 0) ? 0.f : lane_up_f(a_vy, 1);
                corr_win_init(wb, vx0, vy0, a_rho);
            }
            CorrRowIn ib;
            ib.vx1 = vx_in;
            ib.vy1 = vy_in;
            ib.rho1 = rho_row;
            ib.pold = a_pn;
            ib.mup = a_mup;
            ib.mub = a_mub;
            ib.fc = a_fc;
            ib.mx = a_mx;
            ib.my = a_my;
            ib.fx = a_fx;
            ib.fy = a_fy;
            const CorrRow ob = corr_row<true, false>(p, k, wb, ib, a_tau);
            corr2_store(p, ob, c, rb, r0, out_lane);
        }
        a_tau = tau_a;
        a_pn = oa.pn;
        a_mup = in.mup;
        a_mub = in.mub;
        a_fc = in.fc;
        a_mx = oa.mx;
        a_my = oa.my;
        a_fx = in.fx;
        a_fy = in.fy;
        a_rho = rho_row;
        a_vx = vx_a;
        a_vy = vy_a;
        a_vx_last = vx_in;
    }
    if (r1 == ny) {
        CorrRowIn ib;
        ib.vx1 = a_vx_last;
        ib.vy1 = 0.f;
        ib.rho1 = a_rho;
        ib.pold = a_pn;
        ib.mup = a_mup;
        ib.mub = a_mub;
        ib.fc = a_fc;
        ib.mx = a_mx;
        ib.my = a_my;
        ib.fx = a_fx;
        ib.fy = a_fy;
        const CorrRow ob = corr_row<true, false>(p, k, wb, ib, a_tau);
        corr2_store(p, ob, c, ny - 1, r0, out_lane);
    }
}

// K correction sweeps per pass: the FP32 counterpart of ptk_tile (sf_ptk_tile.h, same stage
// scheme; the forcing and the materials pass down the stages with the rows).
struct CorrHand {
    float pn, mup, mub, fc, mx, my, fx, fy, rho, vx, vy;
    CorrTau tau;
};

// The row a stage hands to the next one within an outer step: V (lane-shifted for the next
// stage, and unshifted, which is the next stage's one-lane-down view), rho and its lane-down
// views (materials sit on the same columns in every stage).
struct CorrPass {
    float vx, vy, rho, vx_p1, vy_p1, rho_p1, rho_p2;
};

template <int K, class PP>
__device__ __forceinline__ void corrk_tile(const PP& p, int strip, int segid)
{
    constexpr int OUT = 64 - 2 * K;
    const int lane = threadIdx.x & 63;
    const int c = strip * OUT + lane - (K - 1);
    const int nx = p.nx, ny = p.ny;
    const bool out_lane = lane >= K - 1 && lane < K - 1 + OUT;
    const int r0 = segid * p.seg;
    const int r1 = min(r0 + p.seg, ny);
    const CorrC kc = corr_consts(p, p.dt);
    const int cc = c < 0 ? 0 : c;
    CorrCols lc;
    lc.vx = (unsigned)min(cc, nx - 1);
    lc.vy = (unsigned)min(cc, nx);
    lc.p = (unsigned)min(cc, nx - 2);
    lc.dvx = (unsigned)min(cc, nx - 3);
    lc.dvy = (unsigned)min(cc, nx - 2);
    const bool vx_zero = (c + 1 == 0) || (c + 1 == nx - 1);
    const bool vy_left = c + 1 == 0, vy_right = c + 1 == nx;

    int rs0[K], rs1[K];
#pragma unroll
    for (int s = 0; s < K; ++s) {
        rs0[s] = max(r0 - (K - s), 0);
        rs1[s] = min(r1 - 1 + (K - 1 - s), ny - 1);
    }
    const int r_last = r1 - 1 + (K - 1);

    CorrWin w[K];
    CorrHand h[K];
    float vx_last[K];
#pragma unroll
    for (int s = 0; s < K; ++s) {
        h[s] = CorrHand{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, CorrTau{0.f, 0.f, 0.f, 0.f}};
        vx_last[s] = 0.f;
    }
    int r = rs0[0];
    corr_win_init(w[0], (p.vx_a + (int64_t)r * nx)[lc.vx], (p.vy_a + (int64_t)r * (nx + 1))[lc.vy],
                  (p.rho + (int64_t)r * (nx + 1))[lc.vy]);
    CorrRowIn nxt = corr_load_row(p, lc, r, r >= r0 - (K - 1));

    // One outer step (stage s runs row r - s).  ALL: every stage runs a real row that is not its
    // first and stage 0 prefetches the next row -- the steady middle of the tile, compiled
    // without the per-stage range tests.
    auto step = [&](const int r, auto all_tag) {
        constexpr bool ALL = decltype(all_tag)::value;
        CorrPass cur[K];
        CorrHand hn[K];
        bool ran[K];
#pragma unroll
        for (int s = 0; s < K; ++s) {
            const int rs = r - s;
            ran[s] = false;
            cur[s] = CorrPass{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            hn[s] = h[s];
            if constexpr (!ALL) {
                if (rs < rs0[s] || rs > rs1[s]) {
                    if (s < K - 1 && rs == ny && rs - 1 >= rs0[s]) {  // the virtual row ny
                        cur[s].vx = vx_last[s];
                        cur[s].vx_p1 = lane_down(vx_last[s], 1);
                        cur[s].rho = h[s].rho;
                        cur[s].rho_p1 = lane_down(h[s].rho, 1);
                        cur[s].rho_p2 = lane_down(h[s].rho, 2);
                    }
                    continue;
                }
            }
            CorrRowIn in;
            if (s == 0) {
                in = nxt;
                if (ALL || r + 1 <= rs1[0]) nxt = corr_load_row(p, lc, r + 1, true);
            } else {
                in.vx1 = cur[s - 1].vx;
                in.vy1 = cur[s - 1].vy;
                in.rho1 = cur[s - 1].rho;
                in.vx1_p1 = cur[s - 1].vx_p1;
                in.vy1_p1 = cur[s - 1].vy_p1;
                in.rho1_p1 = cur[s - 1].rho_p1;
                in.rho1_p2 = cur[s - 1].rho_p2;
                in.pold = h[s - 1].pn;
                in.mup = h[s - 1].mup;
                in.mub = h[s - 1].mub;
                in.fc = h[s - 1].fc;
                in.mx = h[s - 1].mx;
                in.my = h[s - 1].my;
                in.fx = h[s - 1].fx;
                in.fy = h[s - 1].fy;
                if (!ALL && rs == rs0[s]) {
                    const float vx0 = (rs == 0) ? in.vx1 : h[s - 1].vx;
                    const float vy0 = (rs == 0) ? 0.f : h[s - 1].vy;
                    corr_win_init(w[s], vx0, vy0, h[s - 1].rho);
                }
            }
            const float rho_row = w[s].rho_r, rho_row_p1 = w[s].rho_r_p1, rho_row_p2 = w[s].rho_r_p2;
            CorrTau tau;
            CorrRow o;
            if (s == 0) {
                o = corr_row<false, false>(p, kc, w[s], in, tau);
            } else {
                tau = h[s - 1].tau;
                o = corr_row<true, true>(p, kc, w[s], in, tau);
            }
            if (s == K - 1) {
                corr2_store(p, o, c, rs, r0, out_lane);
            } else {
                const float vx_a = vx_zero ? 0.f : o.vx;
                const float vy_dn = lane_down(o.vy, 1), vy_up = lane_up_f(o.vy, 1);
                const float vy_a = (rs == ny - 1) ? 0.f : (vy_left ? vy_dn : (vy_right ? vy_up : o.vy));
                cur[s] = CorrPass{lane_up_f(vx_a, 1), lane_up_f(vy_a, 1), rho_row, vx_a, vy_a, rho_row_p1, rho_row_p2};
                hn[s].pn = o.pn;
                hn[s].mup = in.mup;
                hn[s].mub = in.mub;
                hn[s].fc = in.fc;
                hn[s].mx = o.mx;
                hn[s].my = o.my;
                hn[s].fx = in.fx;
                hn[s].fy = in.fy;
                hn[s].rho = rho_row;
                hn[s].vx = cur[s].vx;
                hn[s].vy = cur[s].vy;
                hn[s].tau = tau;
                ran[s] = true;
            }
        }
#pragma unroll
        for (int s = 0; s < K - 1; ++s) {
            if (ALL || ran[s]) vx_last[s] = cur[s].vx;
            h[s] = hn[s];
        }
    };
    int lo = 0, hi = rs1[0] - 1;
#pragma unroll
    for (int s = 0; s < K; ++s) {
        lo = max(lo, rs0[s] + s + 1);
        hi = min(hi, rs1[s] + s);
    }
    for (; r < lo && r <= r_last; ++r) step(r, std::false_type{});
    // the steady middle two rows per trip: a value carried from one row step to the next then
    // alternates between two registers instead of being copied back each step
    for (; r + 1 <= hi; r += 2) {
        step(r, std::true_type{});
        step(r + 1, std::true_type{});
    }
    if (r <= hi) {  // at most one steady row left
        step(r, std::true_type{});
        ++r;
    }
    for (; r <= r_last; ++r) step(r, std::false_type{});
}

template <int K>
__global__ __launch_bounds__(64) void pt_corrk_f32(CorrParams p)
{
    const int logical = xcd_remap_r(blockIdx.x, gridDim.x);
    corrk_tile<K>(p, logical % p.nstrips, logical / p.nstrips);
}

__global__ __launch_bounds__(64) void pt_corr2_f32(CorrParams p)
{
    const int logical = xcd_remap_r(blockIdx.x, gridDim.x);
    corr2_tile(p, logical % p.nstrips, logical / p.nstrips);
}

// One wave per tile, XCD-banded tile order (as the FP64 kernel).
template <int DEPTH, bool CHECK>
__global__ __launch_bounds__(64) void pt_corr_f32(CorrParams p)
{
    const int logical = xcd_remap_r(blockIdx.x, gridDim.x);
    corr_tile<DEPTH, CHECK>(p, logical % p.nstrips, logical / p.nstrips);
}

// Wide batched mode: one launch runs one FP32 correction iteration of every instance in the
// active set (see wide_pt_kernel in sf_batch.hip for the slot search and the tile order).
__global__ __launch_bounds__(64) void wide_corr_kernel(WideCorrArgs a)
{
    const int total = a.tile_start[a.n_active];
    const int t_glob = xcd_remap_r(blockIdx.x, gridDim.x);
    if (t_glob >= total) return;
    int lo = 0, hi = a.n_active - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.tile_start[mid] <= t_glob) lo = mid;
        else hi = mid - 1;
    }
    const int t = t_glob - a.tile_start[lo];
    typedef const CorrParams __attribute__((address_space(4))) CCorr;
    CCorr& cp = ((CCorr*)a.desc)[a.slot_desc[lo]];
    const int nstrips = cp.nstrips;
    corr_tile<1, false>(cp, t % nstrips, t / nstrips);
}

// Two correction sweeps of every active instance per launch (corr2_tile).
__global__ __launch_bounds__(64) void wide_corr2_kernel(WideCorrArgs a)
{
    const int total = a.tile_start[a.n_active];
    const int t_glob = xcd_remap_r(blockIdx.x, gridDim.x);
    if (t_glob >= total) return;
    int lo = 0, hi = a.n_active - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.tile_start[mid] <= t_glob) lo = mid;
        else hi = mid - 1;
    }
    const int t = t_glob - a.tile_start[lo];
    typedef const CorrParams __attribute__((address_space(4))) CCorr;
    CCorr& cp = ((CCorr*)a.desc)[a.slot_desc[lo]];
    const int nstrips = cp.nstrips;
    corr2_tile(cp, t % nstrips, t / nstrips);
}

hipError_t launch_wide_corr2(const WideCorrArgs& a, hipStream_t s)
{
    const int total = a.tile_start[a.n_active];
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(wide_corr2_kernel, dim3(total), dim3(64), 0, s, a);
    return hipGetLastError();
}

// Three correction sweeps of every active instance per launch (corrk_tile<3>).
__global__ __launch_bounds__(64) void wide_corr3_kernel(WideCorrArgs a)
{
    const int total = a.tile_start[a.n_active];
    const int t_glob = xcd_remap_r(blockIdx.x, gridDim.x);
    if (t_glob >= total) return;
    int lo = 0, hi = a.n_active - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.tile_start[mid] <= t_glob) lo = mid;
        else hi = mid - 1;
    }
    const int t = t_glob - a.tile_start[lo];
    typedef const CorrParams __attribute__((address_space(4))) CCorr;
    CCorr& cp = ((CCorr*)a.desc)[a.slot_desc[lo]];
    const int nstrips = cp.nstrips;
    corrk_tile<3>(cp, t % nstrips, t / nstrips);
}

hipError_t launch_wide_corr3(const WideCorrArgs& a, hipStream_t s)
{
    const int total = a.tile_start[a.n_active];
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(wide_corr3_kernel, dim3(total), dim3(64), 0, s, a);
    return hipGetLastError();
}

int corr2_num_strips(int nx) { return (nx - 1 + CORR2_OUT_LANES - 1) / CORR2_OUT_LANES; }

int corr_num_strips(int nx) { return (nx + CORR_OUT_LANES - 1) / CORR_OUT_LANES; }

hipError_t launch_wide_corr(const WideCorrArgs& a, hipStream_t s)
{
    const int total = a.tile_start[a.n_active];
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(wide_corr_kernel, dim3(total), dim3(64), 0, s, a);
    return hipGetLastError();
}

// FP64 residual of the state u = (V, P) without a pressure update, with stabilisation dt.
// Writes the FP32 forcing of the correction solve and per-block sums of |Rx|, |Ry|, |C|.
__global__ __launch_bounds__(256) void resid_f64(ResidParams p)
{
    __shared__ double sh[3][256];
    const int nx = p.nx, ny = p.ny;
    const int64_t lvx = nx, lvy = nx + 1, lp = nx - 1, lr = nx + 1;
    const int t = threadIdx.x;
    double s_rx = 0.0, s_ry = 0.0, s_c = 0.0;
    const int64_t ncell = (int64_t)(nx - 1) * (ny - 1);
    const double rdx = p.rdx, rdy = p.rdy;
    const double* Vx = p.vx;
    const double* Vy = p.vy;
    auto txx_at = [&](int i, int j) {
        const double dvxdx = (Vx[ix2(i + 1, j + 1, lvx)] - Vx[ix2(i, j + 1, lvx)]) * rdx;
        const double dvydy = (Vy[ix2(i + 1, j + 1, lvy)] - Vy[ix2(i + 1, j, lvy)]) * rdy;
        return (2.0 * p.mu_p[ix2(i, j, lp)]) * (dvxdx - (1.0 / 3.0) * (dvxdx + dvydy));
    };
    auto tyy_at = [&](int i, int j) {
        const double dvxdx = (Vx[ix2(i + 1, j + 1, lvx)] - Vx[ix2(i, j + 1, lvx)]) * rdx;
        const double dvydy = (Vy[ix2(i + 1, j + 1, lvy)] - Vy[ix2(i + 1, j, lvy)]) * rdy;
        return (2.0 * p.mu_p[ix2(i, j, lp)]) * (dvydy - (1.0 / 3.0) * (dvxdx + dvydy));
    };
    auto txy_at = [&](int a, int b) {
        return ((2.0 * p.mu_b[ix2(a, b, lvx)]) * 0.5) *
               ((Vx[ix2(a, b + 1, lvx)] - Vx[ix2(a, b, lvx)]) * rdy + (Vy[ix2(a + 1, b, lvy)] - Vy[ix2(a, b, lvy)]) * rdx);
    };
    for (int64_t k = (int64_t)blockIdx.x * 256 + t; k < ncell; k += (int64_t)gridDim.x * 256) {
        const int i = (int)(k % (nx - 1)), j = (int)(k / (nx - 1));
        // continuity at P node (i, j)
        const double cdiv = (Vx[ix2(i + 1, j + 1, lvx)] - Vx[ix2(i, j + 1, lvx)]) * rdx +
                            (Vy[ix2(i + 1, j + 1, lvy)] - Vy[ix2(i + 1, j, lvy)]) * rdy;
        p.fc[ix2(i, j, lp)] = (float)cdiv;
        s_c += fabs(cdiv);
        if (i < nx - 2) {  // x-momentum at dVx node (i, j)
            const double rx = ((txx_at(i + 1, j) - txx_at(i, j)) * rdx + (txy_at(i + 1, j + 1) - txy_at(i + 1, j)) * rdy) -
                              (p.P[ix2(i + 1, j, lp)] - p.P[ix2(i, j, lp)]) * rdx;
            p.fx[ix2(i, j, nx - 2)] = (float)rx;
            s_rx += fabs(rx);
        }
        if (j < ny - 2) {  // y-momentum at dVy node (i, j)
            const double av = 0.25 * (((Vx[ix2(i, j + 1, lvx)] + Vx[ix2(i + 1, j + 1, lvx)]) + Vx[ix2(i, j + 2, lvx)]) +
                                      Vx[ix2(i + 1, j + 2, lvx)]);
            const double d_xi = p.rho[ix2(i + 2, j + 1, lr)] - p.rho[ix2(i, j + 1, lr)];
            const double d_yi = p.rho[ix2(i + 1, j + 2, lr)] - p.rho[ix2(i + 1, j, lr)];
            const double ry = (((tyy_at(i, j + 1) - tyy_at(i, j)) * rdy + (txy_at(i + 1, j + 1) - txy_at(i, j + 1)) * rdx) -
                               (p.P[ix2(i, j + 1, lp)] - p.P[ix2(i, j, lp)]) * rdy) +
                              p.g * (p.rho[ix2(i + 1, j + 1, lr)] -
                                     p.dt * (((av * d_xi) * 0.5) * rdx + ((Vy[ix2(i + 1, j + 1, lvy)] * d_yi) * 0.5) * rdy));
            p.fy[ix2(i, j, nx - 1)] = (float)ry;
            s_ry += fabs(ry);
        }
    }
    sh[0][t] = s_rx;
    sh[1][t] = s_ry;
    sh[2][t] = s_c;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s)
            for (int q = 0; q < 3; ++q) sh[q][t] += sh[q][t + s];
        __syncthreads();
    }
    if (t == 0)
        for (int q = 0; q < 3; ++q) p.part[(int64_t)q * gridDim.x + blockIdx.x] = sh[q][0];
}

// u += d in FP64 (the corrections satisfy the homogeneous free-slip BCs, so u keeps its BCs)
__global__ __launch_bounds__(256) void update_f64(double* __restrict__ a, const float* __restrict__ d, int64_t n)
{
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (int64_t)gridDim.x * 256)
        a[k] = a[k] + (double)d[k];
}

__global__ __launch_bounds__(256) void cast_f64(double* __restrict__ dst, const float* __restrict__ src, int64_t n)
{
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (int64_t)gridDim.x * 256)
        dst[k] = (double)src[k];
}

__global__ __launch_bounds__(256) void cast_f32(float* __restrict__ dst, const double* __restrict__ src, int64_t n)
{
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (int64_t)gridDim.x * 256)
        dst[k] = (float)src[k];
}

__global__ __launch_bounds__(256) void fill_f32(float* __restrict__ dst, int64_t n, float v)
{
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (int64_t)gridDim.x * 256) dst[k] = v;
}

__global__ __launch_bounds__(256) void fill_f64(double* __restrict__ dst, int64_t n, double v)
{
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (int64_t)gridDim.x * 256) dst[k] = v;
}

// Per-block sums of |a| over up to three FP32 arrays (FP64 accumulation, fixed order).
__global__ __launch_bounds__(256) void reduce_abs_f32(RedParamsF p)
{
    __shared__ double sh[256];
    const int a = blockIdx.y;
    const float* __restrict__ src = p.a[a];
    const int64_t n = p.n[a];
    const int t = threadIdx.x;
    double acc = 0.0;
    for (int64_t k = (int64_t)blockIdx.x * 256 + t; k < n; k += (int64_t)gridDim.x * 256) acc += fabs((double)src[k]);
    sh[t] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) sh[t] += sh[t + s];
        __syncthreads();
    }
    if (t == 0) p.part[(int64_t)a * gridDim.x + blockIdx.x] = sh[0];
}

static int grid_for(int64_t n) { return (int)std::min<int64_t>(4096, std::max<int64_t>(1, (n + 255) / 256)); }

hipError_t launch_pt_corr(const CorrParams& p0, bool check, hipStream_t s)
{
    CorrParams p = p0;
    if (p.seg <= 0) p.seg = 4;
    p.nstrips = (p.nx + CORR_OUT_LANES - 1) / CORR_OUT_LANES;
    const int nseg = (p.ny + p.seg - 1) / p.seg;
    p.ntiles = p.nstrips * nseg;
    if (p.depth > 1) {
        if (check) hipLaunchKernelGGL((pt_corr_f32<2, true>), dim3(p.ntiles), dim3(64), 0, s, p);
        else hipLaunchKernelGGL((pt_corr_f32<2, false>), dim3(p.ntiles), dim3(64), 0, s, p);
    } else {
        if (check) hipLaunchKernelGGL((pt_corr_f32<1, true>), dim3(p.ntiles), dim3(64), 0, s, p);
        else hipLaunchKernelGGL((pt_corr_f32<1, false>), dim3(p.ntiles), dim3(64), 0, s, p);
    }
    return hipGetLastError();
}

// Rows per tile of the two-sweep pass (as the FP64 pass: taller tiles cut the halo re-reads,
// short ones give small grids enough waves; us per sweep in the refined solver: 256^2 2.66 with
// 2 rows against 3.44 with 4, 512^2 3.93 against 4.12).
int corr2_default_rows(int nx, int ny)
{
    const int n = std::max(nx, ny);
    const int r = n > 2048 ? 16 : (n > 768 ? 12 : (n > 512 ? 8 : 2));
    return ny >= r ? r : ny;
}

hipError_t launch_pt_corr2(const CorrParams& p0, hipStream_t s)
{
    CorrParams p = p0;
    if (p.seg <= 0) p.seg = corr2_default_rows(p.nx, p.ny);
    p.nstrips = (p.nx - 1 + CORR2_OUT_LANES - 1) / CORR2_OUT_LANES;
    const int nseg = (p.ny + p.seg - 1) / p.seg;
    p.ntiles = p.nstrips * nseg;
    hipLaunchKernelGGL(pt_corr2_f32, dim3(p.ntiles), dim3(64), 0, s, p);
    return hipGetLastError();
}

// Tile rows of the four- and five-sweep passes: the launch runs in rounds of resident waves (two
// per SIMD at 217 / 253 VGPRs), and a last round that is only partly filled costs a whole one, so
// the height is chosen to minimise rounds x (rows + 2K-1 halo rows) over 24-96 rows.  Measured,
// us per sweep: K=5 4096^2 80 rows 59.8-60.2 (the pick, a two-round launch), 40 62.3-63.2, 32 69.2,
// 48 70.1, 56 66.3, 64 72.2, 96 70.1, 128 87.3; K=4 3072^2 84 rows 39.0 (the pick, one round), 28
// 40.0, 44 40.4; 2560^2 60 rows 28.3 (the pick), 40 30.7; 2048^2 40 rows 15.1 (the pick), 80 21.4.
static int corrk_rows(int nx, int ny, int k)
{
    static const int slots = [] {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        return 2 * 4 * cus;
    }();
    const int out = 64 - 2 * k;
    const int64_t strips = (nx - 1 + out - 1) / out;
    int best = 40;
    double best_cost = 1e300;
    for (int rows = 24; rows <= 96; rows += 4) {
        const int64_t waves = strips * ((ny + rows - 1) / rows);
        const double cost = (double)((waves + slots - 1) / slots) * (rows + 2 * k - 1);
        if (cost < best_cost) {
            best_cost = cost;
            best = rows;
        }
    }
    return best;
}

hipError_t launch_pt_corrk(const CorrParams& p0, int k, hipStream_t s)
{
    if (k < 3 || k > 5) return hipErrorInvalidValue;
    CorrParams p = p0;
    if (p.seg <= 0) {  // three sweeps, 1024^2: 7.82 us per sweep with 8 rows, 9.48 with 16
        const int n = std::max(p.nx, p.ny);
        const int r = k >= 4 ? corrk_rows(p.nx, p.ny, k) : (n > 3072 ? 48 : (n > 1024 ? 16 : 8));
        p.seg = p.ny >= r ? r : p.ny;
    }
    const int out = 64 - 2 * k;
    p.nstrips = (p.nx - 1 + out - 1) / out;
    const int nseg = (p.ny + p.seg - 1) / p.seg;
    p.ntiles = p.nstrips * nseg;
    if (k == 3) hipLaunchKernelGGL(pt_corrk_f32<3>, dim3(p.ntiles), dim3(64), 0, s, p);
    else if (k == 4) hipLaunchKernelGGL(pt_corrk_f32<4>, dim3(p.ntiles), dim3(64), 0, s, p);
    else hipLaunchKernelGGL(pt_corrk_f32<5>, dim3(p.ntiles), dim3(64), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_resid_f64(const ResidParams& p, int nb, hipStream_t s)
{
    hipLaunchKernelGGL(resid_f64, dim3(nb), dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_update_f64(double* a, const float* d, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(update_f64, dim3(grid_for(n)), dim3(256), 0, s, a, d, n);
    return hipGetLastError();
}

hipError_t launch_cast_f32(float* dst, const double* src, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(cast_f32, dim3(grid_for(n)), dim3(256), 0, s, dst, src, n);
    return hipGetLastError();
}

hipError_t launch_cast_f64(double* dst, const float* src, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(cast_f64, dim3(grid_for(n)), dim3(256), 0, s, dst, src, n);
    return hipGetLastError();
}

hipError_t launch_fill_f32(float* dst, int64_t n, float v, hipStream_t s)
{
    hipLaunchKernelGGL(fill_f32, dim3(grid_for(n)), dim3(256), 0, s, dst, n, v);
    return hipGetLastError();
}

hipError_t launch_fill_f64(double* dst, int64_t n, double v, hipStream_t s)
{
    hipLaunchKernelGGL(fill_f64, dim3(grid_for(n)), dim3(256), 0, s, dst, n, v);
    return hipGetLastError();
}

hipError_t launch_reduce_f32(const RedParamsF& p, int narr, int nb, hipStream_t s)
{
    hipLaunchKernelGGL(reduce_abs_f32, dim3(nb, narr), dim3(256), 0, s, p);
    return hipGetLastError();
}

}  // namespace sf
