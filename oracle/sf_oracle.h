/*
 * sf_oracle.h -- CPU restatement of the StokesFlow.jl FP64 hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP product is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never links or calls it.
 *
 * Arrays are column-major with the first (x) index fastest and 0-based, i.e.
 * A(i, j) = A[i + j * size1], which is Julia's layout of the reference's
 * Data.Array with 1-based (ix, iy) = (i + 1, j + 1).
 *
 * Nx, Ny are node counts.  Array shapes (reference StokesFlow2D_multixpu.jl:89-114):
 *   P, txx, tyy, divV, dtauP, mu_p : (Nx-1, Ny-1)
 *   txy, mu_b                      : (Nx,   Ny)
 *   Vx_s : (Nx, Ny+1)     Vx_pad : (Nx+2, Ny+1)
 *   Vy_s : (Nx+1, Ny)     Vy_pad : (Nx+1, Ny+2)
 *   rho_vy : (Nx+1, Ny)
 *   Rx, dVx, dtauVx : (Nx-2, Ny-1)     Ry, dVy, dtauVy : (Nx-1, Ny-2)
 *
 * Pinning status of every function is recorded in the header of
 * sf_oracle.c next to each definition.
 */
#ifndef SF_ORACLE_H
#define SF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double  dt_new;    /* compute_dt(Vx_s, Vy_s) after the loop (StokesSolver_multixpu.jl:95) */
    double  err;       /* last err from compute_err (:71) */
    double  t_eff;     /* reference T_eff in GB/s (:79-81) */
    double  runtime;   /* seconds from PT iteration 11 to the end of the loop (:44-45,:78) */
    int64_t niter;     /* PT iterations done (:76) */
    int32_t converged; /* err <= eps at exit */
    int32_t reserved;
    double  dt_last;   /* dt used by the stabilisation term in the last iteration */
} sfo_solve_out;

/* StokesSolver_multixpu.jl:128-133 */
void sfo_compute_timesteps(int nx, int ny, const double *mu_p, double *dtauVx, double *dtauVy,
                           double *dtauP, double Vsc, double Ptsc, double min_dxy2, int max_nxy);

/* StokesSolver_multixpu.jl:140-169 */
void sfo_compute_P_tau(int nx, int ny, double *divV, double *P, const double *Vx, const double *Vy,
                       const double *dtauP, double *txx, double *tyy, double *txy,
                       const double *mu_p, const double *mu_b, double _dx, double _dy);

/* StokesSolver_multixpu.jl:209-219 (with compute_ResX :176-182 and compute_ResY :189-202) */
void sfo_compute_dV(int nx, int ny, double *dVx, double *dVy, const double *P, const double *rho,
                    const double *txx, const double *tyy, const double *txy, const double *Vx,
                    const double *Vy, double g_y, double dampX, double dampY, double _dx, double _dy,
                    double dt);

/* StokesSolver_multixpu.jl:226-234 */
void sfo_compute_V(int nx, int ny, double *Vx, double *Vy, const double *dVx, const double *dVy,
                   const double *dtauVx, const double *dtauVy);

/* bc_y_noflux!(Vx), bc_x_zero!(Vx), bc_y_zero!(Vy), bc_x_noflux!(Vy) in the call order of
 * StokesSolver_multixpu.jl:53-56 (definitions :256-271, :289-303). */
void sfo_apply_free_slip(int nx, int ny, double *Vx, double *Vy);

/* StokesSolver_multixpu.jl:241-249 */
void sfo_compute_residuals(int nx, int ny, double *Rx, double *Ry, const double *P, const double *rho,
                           const double *txx, const double *tyy, const double *txy, const double *Vx,
                           const double *Vy, double g_y, double _dx, double _dy, double dt);

/* StokesSolver_multixpu.jl:103-109 at one rank: max(mean|Rx|, mean|Ry|, mean|divV|) */
double sfo_compute_err(int nx, int ny, const double *Rx, const double *Ry, const double *divV);

/* StokesSolver_multixpu.jl:118-121 at one rank */
double sfo_compute_dt(int nx, int ny, const double *Vx, const double *Vy, double maxdisp,
                      double dx, double dy);

/* Mean of |a| with Julia's pairwise summation structure (blocksize 1024). */
double sfo_mean_abs(const double *a, int64_t n);

/* Full solver, StokesSolver_multixpu.jl:14-96 at one rank (nx_g = Nx, ny_g = Ny).
 * Every scratch argument must be a caller-owned array of the documented shape. */
int sfo_solve_stokes(double *P, double *Vx_pad, double *Vy_pad, const double *rho_vy,
                     const double *mu_b, const double *mu_p, double *txx, double *tyy, double *txy,
                     double *divV, double *dtauP, double *Rx, double *Ry, double *dVx, double *dVy,
                     double *dtauVx, double *dtauVy, double *Vx_s, double *Vy_s, double g_y,
                     double dx, double dy, int nx, int ny, double dt, double maxdisp, int use_fss,
                     double eps, int64_t iter_max, sfo_solve_out *out);

/* n_iter plain PT iterations at a fixed stabilisation dt (the body of the loop at
 * StokesSolver_multixpu.jl:48-58, no dt refresh and no check).  dtau* must already hold
 * compute_timesteps! output.  Used to pin the GPU iteration bit for bit. */
void sfo_pt_iterations(int nx, int ny, int64_t n_iter, double *P, double *Vx_s, double *Vy_s,
                       double *dVx, double *dVy, const double *rho, const double *mu_b,
                       const double *mu_p, double *txx, double *tyy, double *txy, double *divV,
                       const double *dtauP, const double *dtauVx, const double *dtauVy,
                       double g_y, double dx, double dy, double dt);

/* common.jl:22-33.  ix, iy are 1-based like the reference. */
void sfo_topleft_index(double x_grid_min, double y_grid_min, double x, double y, double dx,
                       double dy, int64_t *ix, int64_t *iy, double *dxij, double *dyij);

/* MarkerToGrid.jl:18-38 with the CPU path :110-155 and safeDivision :164-172 at one rank.
 * val_wt_sum / wt_sum have shape (ws_nx, ws_ny) >= (gnx, gny) and are zeroed whole.
 * Returns 0, or -1 for a non-finite marker coordinate (floor(Int, NaN) throws in Julia).
 * n_bad counts markers outside the grid range, n_empty nodes left unchanged. */
int sfo_marker_to_grid(double x_grid_min, double y_grid_min, double *val_grid, int gnx, int gny,
                       const double *x_m, const double *y_m, const double *val_m, int64_t nm,
                       double dx, double dy, double *val_wt_sum, double *wt_sum, int ws_nx,
                       int ws_ny, int64_t *n_bad, int64_t *n_empty);

/* Cell keys of bilinearMarkerToGrid!'s CPU path (MarkerToGrid.jl:116-122 with
 * topleftIndexRelDist, common.jl:22-33): for the 1-based top-left node (ix, iy) of marker m,
 * keys[m] = iy*(gnx+1) + ix when 0 <= ix <= gnx and 0 <= iy <= gny, else -1 (the markers the
 * reference reports as erroneous).  counts[(gnx+1)*(gny+1)] (may be NULL) gets the number of
 * markers per key.  Returns -1 for a non-finite coordinate. */
int sfo_cell_keys(double x_grid_min, double y_grid_min, int gnx, int gny, const double *x_m,
                  const double *y_m, int64_t nm, double dx, double dy, int32_t *keys, int32_t *counts,
                  int64_t *n_bad);

/* MoveMarkers.jl:76-174 (with bilinearInterp :181-187).  n_clamp (may be NULL) is incremented
 * once per failsafe clamp taken.  Returns -1 for a non-finite position. */
int sfo_interpolate_v(double x, double y, const double *Vx, int vx_n1, int vx_n2, const double *Vy,
                      int vy_n1, int vy_n2, double x_vx_min, double y_vx_min, double x_vy_min,
                      double y_vy_min, double dx, double dy, double *vx, double *vy,
                      int64_t *n_clamp);

/* MoveMarkers.jl:15-69 */
int sfo_move_markers_rk4(double *x_m, double *y_m, int64_t nm, const double *Vx, int vx_n1,
                         int vx_n2, const double *Vy, int vy_n1, int vy_n2, double x_vx_min,
                         double y_vx_min, double x_vy_min, double y_vy_min, double dt, double lx,
                         double ly, double dx, double dy, int64_t *n_clamp);

/* common.jl:40-94 at one rank with RAND_MARKER_POS=false: the LinRange lattice, x outer,
 * y inner.  Returns the number of markers written (nmx * nmy). */
int64_t sfo_lattice_markers(int nx, int ny, int marker_density, double lx, double ly, double dx,
                            double dy, double *x_m, double *y_m);

/* Seeded marker generator shared bit for bit with the device generator (sf_init_markers in the
 * product library): md x md sub-lattice per cell, x outer / y inner, uniform jitter of
 * +-jitter_frac cells from splitmix64(seed, index), clamped to [0, lx] x [0, ly]. */
void sfo_seeded_markers(int nx, int ny, int md, double dx, double dy, double jitter_frac,
                        uint64_t seed, double *x_m, double *y_m);

/* splitmix64 of (seed, counter) mapped to [0, 1) with 53 bits. */
double sfo_uniform(uint64_t seed, uint64_t counter);

/* Material rule of the plume setups: air layer y < air_h first, then the disc
 * (x-cx)^2 + (y-cy)^2 < r^2, else background (test_StokesFlow2D.jl:66-83). */
void sfo_plume_properties(const double *x_m, const double *y_m, int64_t nm, double air_h,
                          double cx, double cy, double r, double rho_air, double rho_plume,
                          double rho_bg, double mu_air, double mu_plume, double mu_bg,
                          double *rho_m, double *mu_m);

#ifdef __cplusplus
}
#endif
#endif
