/*
 * sf_hip.h -- C ABI of libsfhip, the gfx950 (MI355X) implementation of the
 * StokesFlow.jl per-timestep hot path: pseudo-transient Stokes solve, marker -> grid
 * interpolation and RK4 marker advection, on one GPU.
 *
 * Conventions
 *   - Every array pointer is a DEVICE pointer to float64 data in column-major order with
 *     the first (x) index fastest -- the memory layout of the reference's Julia
 *     Data.Array / ROCArray.  A(i, j) = A[i + j * size1] with 0-based (i, j).
 *   - Node counts Nx, Ny and array shapes follow StokesFlow2D_multixpu.jl:89-114:
 *       P, mu_p, txx, tyy, divV, dtauP : (Nx-1, Ny-1)    txy, mu_b : (Nx, Ny)
 *       Vx_s : (Nx, Ny+1)   Vx_pad : (Nx+2, Ny+1)        rho_vy : (Nx+1, Ny)
 *       Vy_s : (Nx+1, Ny)   Vy_pad : (Nx+1, Ny+2)
 *       Rx, dVxdtau, dtauVx : (Nx-2, Ny-1)    Ry, dVydtau, dtauVy : (Nx-1, Ny-2)
 *   - Work is enqueued on the context's HIP stream; calls that return results on the
 *     host (solve outputs, counters) synchronise that stream before returning.
 *   - Every function returns SF_OK (0) or a negative status; sf_last_error() returns a
 *     message for the calling thread's last failure.  One context per stream/thread; a
 *     context is not re-entrant.
 *   - Data problems the reference only prints about ("erroneous marker position",
 *     "no markers close enough", the velocity-interpolation failsafe clamps) come back as
 *     counters.  Inputs the reference would throw on (floor(Int, NaN)) return
 *     SF_ERR_NONFINITE.
 */
#ifndef SF_HIP_H
#define SF_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SF_OK 0
#define SF_ERR_ARG (-1)       /* invalid argument (shape, NULL pointer, size) */
#define SF_ERR_HIP (-2)       /* HIP runtime error */
#define SF_ERR_NONFINITE (-3) /* non-finite / unrepresentable marker coordinate */
#define SF_ERR_SIZE (-4)      /* problem larger than the 32-bit index limits */

typedef struct sf_ctx sf_ctx;

typedef struct {
    double dt_new;      /* compute_dt(Vx_s, Vy_s) after the solve: 1st return value */
    double err;         /* last err (max of mean |Rx|, |Ry|, |divV|) */
    double t_eff;       /* reference T_eff = 6*Nx*Ny*8 B / t_it, GB/s: 2nd return value */
    double runtime;     /* seconds from PT iteration 11 to the end: 3rd return value */
    int64_t niter;      /* PT iterations performed */
    int32_t converged;  /* err <= eps at exit */
    int32_t reserved;
    double dt_last;     /* stabilisation dt in effect during the last iteration */
    double t_it;        /* seconds per PT iteration over the timed window */
    double gbs_alg;     /* algorithmic GB/s = 13-array bytes per iteration / t_it */
} sf_solve_out;

typedef struct {
    int64_t n_bad_markers; /* markers outside the target grid (skipped) */
    int64_t n_empty_nodes; /* nodes with zero weight (value left unchanged) */
} sf_m2g_out;

typedef struct {
    int64_t n_clamped; /* failsafe index clamps taken in interpolateV */
} sf_rk4_out;

/* Library version (major*10000 + minor*100 + patch). */
int sf_version(void);
/* Message describing the calling thread's last failed call. */
const char *sf_last_error(void);

/* Context: owns device scratch for an Nx x Ny node grid and the HIP stream to use
 * (hip_stream may be NULL for the default stream).  Scratch for markers grows on demand.
 * nx = ny = 0 creates a marker-only context (marker -> grid, RK4, set-up; no solve). */
int sf_ctx_create(sf_ctx **ctx, int nx, int ny, void *hip_stream);
int sf_ctx_destroy(sf_ctx *ctx);
/* Replace the stream work is enqueued on. */
int sf_ctx_set_stream(sf_ctx *ctx, void *hip_stream);
/* Launch tuning of the PT iteration kernel; results do not depend on it.  rows: grid rows
 * each tile marches through (0 = size default); depth: rows of operands prefetched (1 or 2;
 * 0 = size default); max_waves: 0 = one wave per resident slot (persistent), > 0 = at most
 * that many waves, < 0 = one wave per tile.  No reference counterpart (ParallelStencil picks launch
 * shapes itself). */
int sf_ctx_set_pt_tuning(sf_ctx *ctx, int rows, int depth, int max_waves);
/* Temporal blocking of the PT iteration: iters_per_launch = 2, 3 or 4 runs that many plain
 * iterations (no check iteration, no dt refresh before the last of them, not the last
 * iteration or the first timed one) as one pass that reads and writes the state once (where
 * they do not fit, shorter passes); 1 runs one iteration per launch; 0 is the size default (4
 * above 448 nodes per side, else 2).  The FP32 sweeps of sf_solve_stokes_refined take the same
 * setting with 5 also allowed (the FP64 passes cap it at 4); their default is 5 above 3072
 * nodes per side, 4 above 1536, 3 above 768, else 2.
 * rows: grid rows per tile of the pass (0 = size default).  Results are bit-identical either
 * way; the pass needs a second dVx/dVy buffer, allocated in the context on first use.  No
 * reference counterpart (the reference launches each of its kernels once per iteration). */
int sf_ctx_set_pt_fusion(sf_ctx *ctx, int iters_per_launch, int rows);

/* Autotuner of the launch shapes: times `iters` PT iterations of every candidate on a copy of
 * the caller's problem (P ... mu_p as for sf_pt_iterations; the caller's arrays are not
 * modified) or, when any of them is NULL, on a synthetic one (materials 1, state 0), in scratch
 * it allocates and frees.  Candidates: the one-iteration kernel (rows 1, 2, 4, 8, 16 x depth
 * 1, 2 x one wave per tile or persistent; iters_per_launch = 1), then passes of 2, 3 and 4
 * iterations (rows 2, 4, 8, 12, 16, 24, 32, 48, 64 and the pass kind's size default; at most 50
 * candidates in all).  The fastest one-iteration shape is set on the
 * context (check iterations use it), and the fastest overall decides the iterations per
 * launch and the pass rows of the FP64 solver (sf_solve_stokes, sf_pt_iterations); the FP32
 * sweeps of sf_solve_stokes_refined keep their own size defaults (or the explicit
 * sf_ctx_set_pt_fusion setting, which also clears the tuned FP64 choice).  Timing can depend
 * on the data, so tune on the real problem.
 * best (may be NULL) receives the fastest overall; table (may be NULL, table_cap entries)
 * every candidate's time, n_table how many were timed.  No reference counterpart. */
typedef struct {
    int rows, depth, max_waves, iters_per_launch;
    double us_per_iter;
} sf_pt_tuning;
int sf_ctx_autotune_pt(sf_ctx *ctx, int iters, const double *P, const double *Vx_s, const double *Vy_s,
                       const double *dVxdtau, const double *dVydtau, const double *rho_vy, const double *mu_b,
                       const double *mu_p, double g_y, double dx, double dy, sf_pt_tuning *best, sf_pt_tuning *table,
                       int table_cap, int *n_table);

/* Per-launch timing of the pass kernels, for roofline figures.  enable = 1 brackets every later
 * launch of the FP64 multi-iteration passes (sf_solve_stokes, sf_pt_iterations) and of the FP32
 * correction passes (sf_solve_stokes_refined) on this context with two HIP events on its stream
 * and clears the totals; enable = 0 stops.  Results do not change.  sf_ctx_pass_times returns the
 * totals since the last enable, one entry per (precision, iterations per launch) launched; it
 * waits for the timed launches.  No reference counterpart (instrumentation). */
typedef struct {
    int32_t precision;        /* 0: FP64 PT iterations, 1: FP32 correction sweeps */
    int32_t iters_per_launch; /* PT iterations (FP32 sweeps) per launch */
    int64_t launches;
    double seconds;           /* summed durations of those launches */
} sf_pass_time;
int sf_ctx_time_passes(sf_ctx *ctx, int enable);
int sf_ctx_pass_times(sf_ctx *ctx, sf_pass_time *out, int cap, int *n_out);

/* DRAM probe: best-of-8 bandwidth of a copy (1 read + 1 write) and of a 3-read + 1-write
 * kernel over four `bytes`-sized scratch arrays (allocated and freed here), in GB/s.  The
 * ceiling the stencil kernels' bandwidth is read against.  No reference counterpart. */
int sf_probe_dram(sf_ctx *ctx, int64_t bytes, double *copy_gbs, double *r3w1_gbs);

/* Snapshot files of the timestep loop, the files viz.jl reads (viz.jl:11-161), as MAT-file
 * level 5 (MAT.jl reads it; the reference's MAT.jl writes the HDF5-based v7.3 layout).
 * Replaces saveStats (StokesFlow2D_multixpu.jl:212-223): "dims" (3 x 1 Int64), "nt" (1 x 1
 * Int64); host only. */
int sf_save_stats(const char *path, const int64_t dims[3], int64_t nt);
/* Replaces saveMarkers (:230-240): Float32 "x_m" = x_m + x_shift, "y_m" = y_m + y_shift,
 * "rho_m" (nm x 1 each) from device arrays (x_shift = (lx - dx) * coords[1] in the reference,
 * 0 at one rank).  Streams through pinned memory in 32 MiB chunks on the context's stream. */
int sf_save_markers(sf_ctx *ctx, const char *path, const double *x_m, const double *y_m, const double *rho_m,
                    int64_t nm, double x_shift, double y_shift);
/* Replaces saveGrid (:247-299): Float32 node coordinates x, y, x_p, y_p, x_vx, y_vx, x_vy, y_vy
 * (:117-124) and the device grids mu_b (Nx x Ny), P (Nx-1 x Ny-1), Vx (padded, Nx+2 x Ny+1),
 * Vy (padded, Nx+1 x Ny+2). */
int sf_save_grid(sf_ctx *ctx, const char *path, int nx, int ny, double dx, double dy, const double *mu_b,
                 const double *P, const double *Vx_pad, const double *Vy_pad);

/* solveStokes! -- replaces scripts/StokesSolver_multixpu.jl:14-96
 * (solveStokes!(P, Vx_pad, Vy_pad, rho_vy, mu_b, mu_p, txx, tyy, txy, divV, dtauPt, Rx, Ry,
 *  dVxdtau, dVydtau, dtauVx, dtauVy, Vx_s, Vy_s, g_y, dx, dy, Nx, Ny, dt, maxdisp, comm;
 *  use_free_surface_stabilization, eps, print_info, iterMax) -> (dt, T_eff, runtime)).
 * In/out: P, Vx_s, Vy_s, dVxdtau, dVydtau (state), Vx_pad, Vy_pad (output).
 * Scratch arguments txx, tyy, txy, divV, dtauP, Rx, Ry, dtauVx, dtauVy may be NULL; when
 * given they are left holding what the reference leaves in them.  comm is absent: one
 * rank.  The caller's rho_vy, mu_b, mu_p are read only.  print_info: 0 silent, 1 the
 * reference's summary line (:82-83), 2 also one line per convergence check. */
int sf_solve_stokes(sf_ctx *ctx, double *P, double *Vx_pad, double *Vy_pad, const double *rho_vy,
                    const double *mu_b, const double *mu_p, double *txx, double *tyy, double *txy,
                    double *divV, double *dtauP, double *Rx, double *Ry, double *dVxdtau, double *dVydtau,
                    double *dtauVx, double *dtauVy, double *Vx_s, double *Vy_s, double g_y, double dx,
                    double dy, int nx, int ny, double dt, double maxdisp, int use_fss, double eps,
                    int64_t iter_max, int print_info, sf_solve_out *out);

/* sf_solve_stokes without a context and without allocating: the same solve (size-default launch
 * shapes, results bit-identical to sf_solve_stokes on a new context), with its scratch -- the PT
 * ping-pong state, the check-iteration arrays the caller passes as NULL, reduction partials and
 * device scalars -- carved from one static device array of SF_NOALLOC_SCRATCH_BYTES per GPU,
 * held until the process ends.  It calls no hipMalloc, hipFree, hipHostMalloc or hipEventCreate;
 * the timing window (runtime, T_eff) is taken on the host clock.  A grid whose scratch does not
 * fit returns SF_ERR_SIZE (about 88 B per node with every scratch argument NULL, 40 B with
 * txx, tyy, txy, divV, Rx, Ry passed: up to about 860^2 or 1290^2 nodes).  Work runs on
 * hip_stream (NULL: the default stream).  Every call on one GPU shares that GPU's array, whose
 * contents persist from call to call: calls that overlap on one GPU race.  No reference
 * counterpart: for callers that cannot allocate device memory (tight serving loops, harnesses
 * that forbid hipMalloc). */
#define SF_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int sf_solve_stokes_noalloc(void *hip_stream, double *P, double *Vx_pad, double *Vy_pad, const double *rho_vy,
                            const double *mu_b, const double *mu_p, double *txx, double *tyy, double *txy,
                            double *divV, double *dtauP, double *Rx, double *Ry, double *dVxdtau, double *dVydtau,
                            double *dtauVx, double *dtauVy, double *Vx_s, double *Vy_s, double g_y, double dx,
                            double dy, int nx, int ny, double dt, double maxdisp, int use_fss, double eps,
                            int64_t iter_max, int print_info, sf_solve_out *out);

/* FP32 PT sweeps inside FP64 defect correction.  Same arguments, shapes and outputs as
 * sf_solve_stokes.  The PT operator is affine in (V, P) for a fixed stabilisation dt.
 * Every rebase_interval iterations the FP64 residual of the current state is computed
 * exactly; the next window of PT iterations runs in FP32 on the correction with that
 * residual as forcing, and the correction is added to the state in FP64.  The PT momentum
 * is carried from window to window (and starts from, and returns to, dVxdtau/dVydtau),
 * so the windows continue one PT iteration; dt is refreshed every ndtupdate iterations as
 * the reference does.  Convergence is declared by one FP64 PT iteration whose err
 * (computed exactly as compute_err) is <= eps, so the returned state meets the
 * reference's own test.  No reference counterpart: an added solver mode beside
 * solveStokes!. */
typedef struct {
    int64_t rebase_interval; /* FP32 iterations between FP64 re-bases (<= 0: the reference's ncheck) */
    int64_t iter_max;        /* cap on FP32 iterations (<= 0: 100000, the reference's default iterMax) */
    int32_t verbose;         /* 1: one line per re-base on stderr */
    int32_t reserved;
} sf_refine_params;

typedef struct {
    int32_t n_outer;     /* re-bases (FP32 windows) */
    int32_t reserved;
    int64_t n_fp32_iter; /* FP32 PT iterations in total */
    int64_t n_fp64_iter; /* FP64 PT (check) iterations */
    double err_fp64;     /* err of the returned state, as compute_err computes it */
    double seconds;      /* wall time of the solve */
} sf_refine_out;

int sf_solve_stokes_refined(sf_ctx *ctx, double *P, double *Vx_pad, double *Vy_pad, const double *rho_vy,
                            const double *mu_b, const double *mu_p, double *txx, double *tyy, double *txy,
                            double *divV, double *dtauP, double *Rx, double *Ry, double *dVxdtau,
                            double *dVydtau, double *dtauVx, double *dtauVy, double *Vx_s, double *Vy_s,
                            double g_y, double dx, double dy, int nx, int ny, double dt, double maxdisp,
                            int use_fss, double eps, const sf_refine_params *prm, sf_solve_out *out,
                            sf_refine_out *refine_out);

/* Batched ensemble: many independent solveStokes! problems (different sizes allowed).
 * PERSISTENT: one launch; each workgroup takes an instance from a device work queue (largest
 * first), runs its whole solve with the reference's ncheck/ndtupdate rules, exits it on its
 * own convergence, and takes the next (no host synchronisation; capturable).
 * WIDE: one launch per PT iteration covers every instance of an active set whose working
 * sets fit budget_mb; each instance is checked at its own ncheck iterations and retired on
 * convergence, and the set is refilled from the largest-first queue.
 * GRID: the instances one after another, each solved by ONE launch that spans the chip (at most
 * one workgroup of one to eight waves per CU): the whole PT loop with its checks, dt refreshes
 * and convergence test runs on the device, the passes (up to three PT iterations each, or
 * prm->depth) separated by a grid barrier; no host synchronisation, so it can be captured
 * (sf_graph_*).  A barrier wait longer than the limit (5 s unless sf_set_grid_wait_limit sets
 * another; workgroups not all resident) ends the launch and the call returns SF_ERR_HIP.
 * With precision 1 (WIDE only) every instance is solved as sf_solve_stokes_refined solves it,
 * with one launch per FP32 sweep of all active instances.
 * In FP64 every PT iteration is bit-identical to sf_solve_stokes.  outs[i] gets instance i's dt_new,
 * err, niter, converged, dt_last (runtime = wall time of the batch).  No reference
 * counterpart (the reference solves one problem per process). */
typedef struct {
    double *P, *Vx_pad, *Vy_pad, *dVxdtau, *dVydtau, *Vx_s, *Vy_s;
    const double *rho_vy, *mu_b, *mu_p;
    int nx, ny;
    double dx, dy, g_y, dt, maxdisp, eps;
    int64_t iter_max;
    int use_fss;
    int reserved;
} sf_batch_instance;

enum { SF_BATCH_AUTO = 0, SF_BATCH_PERSISTENT = 1, SF_BATCH_WIDE = 2, SF_BATCH_GRID = 3 };
typedef struct {
    int rows;         /* grid rows per tile (<= 0: per instance, as the single solver's default) */
    int depth;        /* wide: rows of operands prefetched, 1 or 2 (<= 0: 1); grid: PT iterations per
                         pass, 1 to 3 (<= 0: 3) */
    int threads;      /* persistent: threads per instance workgroup, multiple of 64, <= 512 (<= 0: 512) */
    int max_blocks;   /* persistent: cap on concurrently resident instances (<= 0: as many as fit) */
    int mode;         /* SF_BATCH_*; AUTO = WIDE */
    int precision;    /* wide: 0 FP64 PT iterations; 1 FP32 PT sweeps inside FP64 defect correction (the
                         algorithm of sf_solve_stokes_refined, FP64 convergence test) */
    double budget_mb; /* wide: working-set budget of the active set in MiB (<= 0: 1536; 1024 with
                         precision 1) */
} sf_batch_params;

int sf_solve_stokes_batched(sf_ctx *ctx, const sf_batch_instance *inst, int n, const sf_batch_params *prm,
                            sf_solve_out *outs);

/* Autotuner of the batched solver's wide FP64 passes.  For every grid size in inst it times the
 * wide two-, three- and four-iteration passes with tile heights of 4..64 rows over an active set of that
 * size's instances, as large as prm->budget_mb holds, `iters` PT iterations per candidate, and
 * takes the fastest (iterations per pass, rows) per size.  Each size's choice is then checked on
 * the whole batch -- every instance capped at 240 PT iterations, with the size defaults and with
 * the choice added -- and kept on the context only where the capped run is faster (one launch
 * mixes the sizes' tiles).  Copies of the caller's states are used (allocated and freed here;
 * the caller's arrays are not modified).  Later sf_solve_stokes_batched calls (wide mode) and
 * ensembles created from the context use the kept shapes; results do not depend on them.
 * table (may be NULL, table_cap entries) gets every candidate's time, n_table how many were
 * timed.  sf_batch_tuning_clear returns to the size defaults.  No reference counterpart. */
typedef struct {
    int nx, ny;
    int iters_per_launch, rows;
    int instances;           /* active-set size the candidate was timed on */
    int kept;                /* 1: this size's choice, kept after the check on the whole batch */
    double ns_per_node_iter; /* time per grid node and PT iteration */
} sf_batch_tuning;
int sf_batch_autotune(sf_ctx *ctx, const sf_batch_instance *inst, int n, const sf_batch_params *prm, int iters,
                      sf_batch_tuning *table, int table_cap, int *n_table);
int sf_batch_tuning_clear(sf_ctx *ctx);

/* n_iter iterations of the PT loop body (StokesSolver_multixpu.jl:48-58) at a fixed
 * stabilisation dt, with no dt refresh and no convergence check.  Building block for
 * fixed-iteration benchmarks and for bit-exact per-iteration parity. */
int sf_pt_iterations(sf_ctx *ctx, int64_t n_iter, double *P, double *Vx_s, double *Vy_s, double *dVxdtau,
                     double *dVydtau, const double *rho_vy, const double *mu_b, const double *mu_p,
                     double g_y, double dx, double dy, int nx, int ny, double dt);

/* bilinearMarkerToGrid! -- replaces scripts/MarkerToGrid.jl:18-38
 * (bilinearMarkerToGrid!(x_grid_min, y_grid_min, val_grid, x_m, y_m, val_m, dx, dy,
 *  val_wt_sum, wt_sum, grid)).  Deterministic: bit-identical to the reference's sequential
 * CPU path (:110-155).  val_wt_sum / wt_sum (shape ws_nx x ws_ny >= gnx x gny) may be
 * NULL; when given they are zeroed and hold the weight sums like the reference's helpers.
 * Marker arrays are read only and keep their order. */
int sf_marker_to_grid(sf_ctx *ctx, double x_grid_min, double y_grid_min, double *val_grid, int gnx, int gny,
                      const double *x_m, const double *y_m, const double *val_m, int64_t nm, double dx,
                      double dy, double *val_wt_sum, double *wt_sum, int ws_nx, int ws_ny, sf_m2g_out *out);

/* The three bilinearMarkerToGrid! calls of one timestep (StokesFlow2D_multixpu.jl:164-166) in
 * one call: rho_vy (nx+1, ny) from rho_m, mu_b (nx, ny) and mu_p (nx-1, ny-1) from mu_m, at the
 * reference's grid origins.  out[0..2] (may be NULL) get each pass's counters.  Each grid keeps
 * its own sort: its cell keys come from its own origin in floating point, so one sort cannot
 * serve the three grids bit for bit. */
int sf_markers_to_grids3(sf_ctx *ctx, double *rho_vy, double *mu_b, double *mu_p, int nx, int ny, const double *x_m,
                         const double *y_m, const double *rho_m, const double *mu_m, int64_t nm, double dx, double dy,
                         sf_m2g_out out[3]);

/* The integer cell keys of bilinearMarkerToGrid! (MarkerToGrid.jl:116-122 with
 * topleftIndexRelDist, common.jl:22-33), for checking them bit for bit: keys[m] =
 * iy*(gnx+1) + ix for marker m's 1-based top-left node (ix, iy) when 0 <= ix <= gnx and
 * 0 <= iy <= gny, else -1 (a marker the reference reports as erroneous; counted in
 * out->n_bad_markers).  counts[(gnx+1)*(gny+1)] gets the markers per key.  These are the keys
 * sf_marker_to_grid sorts by.  No reference entry point: the reference computes them inline. */
int sf_marker_cell_keys(sf_ctx *ctx, double x_grid_min, double y_grid_min, int gnx, int gny, const double *x_m,
                        const double *y_m, int64_t nm, double dx, double dy, int32_t *keys, int32_t *counts,
                        sf_m2g_out *out);

/* moveMarkersRK4! -- replaces scripts/MoveMarkers.jl:15-23
 * (moveMarkersRK4!(x_m, y_m, Vx, Vy, x_vx_min, y_vx_min, x_vy_min, y_vy_min, dt, lx, ly,
 *  dx, dy)).  Vx is (vx_n1, vx_n2) = size(Vx), normally the padded (Nx+2, Ny+1) array;
 * Vy likewise (Nx+1, Ny+2).  x_m / y_m are updated in place, order preserved. */
int sf_move_markers_rk4(sf_ctx *ctx, double *x_m, double *y_m, int64_t nm, const double *Vx, int vx_n1,
                        int vx_n2, const double *Vy, int vy_n1, int vy_n2, double x_vx_min, double y_vx_min,
                        double x_vy_min, double y_vy_min, double dt, double lx, double ly, double dx,
                        double dy, sf_rk4_out *out);

/* One timestep of StokesFlow2D at one rank -- replaces the loop body
 * scripts/StokesFlow2D_multixpu.jl:164-181 (three bilinearMarkerToGrid! calls on the
 * rho_vy, mu_b and mu_p grids, solveStokes!, moveMarkersRK4!; exchangeMarkers is the
 * identity at one rank).  dt on input is the previous step's dt (0 on the first step)
 * and on output the new one, as the reference's loop variable. */
typedef struct {
    int nx, ny;
    double dx, dy;
    double g_y, maxdisp, eps;
    int64_t iter_max;
    int use_fss;
    double dt;
    double *P, *Vx_pad, *Vy_pad, *rho_vy, *mu_b, *mu_p, *dVxdtau, *dVydtau, *Vx_s, *Vy_s;
    double *x_m, *y_m, *rho_m, *mu_m;
    int64_t nm;
} sf_step_args;

int sf_step(sf_ctx *ctx, sf_step_args *args, sf_solve_out *solve_out, sf_m2g_out m2g_out[3],
            sf_rk4_out *rk4_out);

/* Ensemble timestep with hipGraphs: n independent problems (grid sizes may differ), each
 * described like sf_step (its dt field is the initial dt, 0 on a first step).  One step of
 * every member = the three marker -> grid passes per member, ONE batched solve of all members
 * and RK4 per member with the dt that the solve produced, read on the device -- the loop body
 * of scripts/StokesFlow2D_multixpu.jl:164-181 for each member.
 * prm->mode PERSISTENT: the whole step is one hipGraph (the persistent batched kernel in the
 * middle), replayed with no host synchronisation; best for small members.  GRID: the same
 * one-graph step with one chip-wide grid-mode solve per member (sf_solve_stokes_batched's GRID
 * mode, prm->depth PT iterations per pass).  WIDE (= AUTO):
 * the marker -> grid passes and the RK4 passes of all members are two hipGraphs, and the
 * wide batched solve runs between them; best when members reach 256^2 and beyond.
 * Without use_graph the same kernels are enqueued every step (identical results).  The
 * ensemble keeps the members' array pointers: the arrays must outlive it.  It runs on a
 * stream of its own, ordered after the context's stream at each call.  outs[i] / dt[i]
 * (either may be NULL) receive member i's solve results and dt after the last step.  No
 * reference counterpart (the reference runs one problem per process). */
typedef struct sf_ensemble sf_ensemble;
int sf_ensemble_create(sf_ctx *ctx, const sf_step_args *members, int n, const sf_batch_params *prm, int use_graph,
                       sf_ensemble **ens);
int sf_ensemble_step(sf_ensemble *ens, int n_steps, sf_solve_out *outs, double *dt);
int sf_ensemble_destroy(sf_ensemble *ens);

/* Captured timestep of ONE problem -- the loop body of scripts/StokesFlow2D_multixpu.jl:164-181
 * (the three bilinearMarkerToGrid! calls, solveStokes!, moveMarkersRK4!) as one hipGraph: the
 * solve runs in the GRID mode of sf_solve_stokes_batched (one launch over the whole chip with the
 * PT loop, its checks and its convergence test on the device), and RK4 reads the new dt on the
 * device, so sf_graph_step replays n_steps timesteps with no host synchronisation in between.
 * args as for sf_step (its dt is the first step's input dt); iters_per_pass 1 to 3 (0: 3).
 * Fields, markers and dt are bit-identical to n_steps sf_step calls; err agrees to rounding
 * (its sum runs in another order).  The arrays must outlive the graph; it runs on a stream of its
 * own, ordered after the context's stream at each call, and sf_graph_step returns once the steps
 * are done.  out / dt (may be NULL) receive the last step's solve results and the dt after it.
 * Other work on the GPU while a step runs can keep the solve's workgroups from being resident
 * together; the step then returns SF_ERR_HIP after the barrier wait limit (5 s by default,
 * fixed when the graph is created).  Replaces sf_step's host loop; the
 * reference has no counterpart (its loop is host-driven, StokesFlow2D_multixpu.jl:157-218). */
/* Time limit of one grid-barrier wait in the GRID mode (default 5 s): a launch whose wait exceeds
 * it ends and its call returns SF_ERR_HIP.  For grids whose passes take longer, or to test the
 * error path (0: every wait that does not find its barrier complete fails at once).  Launches
 * already enqueued or captured keep the limit they were made with.  No reference counterpart. */
int sf_set_grid_wait_limit(double seconds);

typedef struct sf_graph sf_graph;
int sf_graph_create(sf_ctx *ctx, const sf_step_args *args, int iters_per_pass, sf_graph **g);
int sf_graph_step(sf_graph *g, int n_steps, sf_solve_out *out, double *dt);
int sf_graph_destroy(sf_graph *g);

/* Seeded marker set-up: md x md markers per cell, x outer / y inner (the order of
 * common.jl:76-83), uniform jitter of +-jitter_frac cells from splitmix64(seed, index),
 * clamped to the domain.  x_m, y_m hold md^2 (nx-1)(ny-1) values.  No reference
 * counterpart (the reference seeds Julia's RNG, common.jl:85-87). */
int sf_init_markers(sf_ctx *ctx, double *x_m, double *y_m, int nx, int ny, int md, double dx, double dy,
                    double jitter_frac, uint64_t seed);

/* Material descriptors replacing the reference's density/viscosity callbacks
 * (setInitialMarkerProperties!, common.jl:104-110).
 * plume: mat = {air_h, cx, cy, r, rho_air, rho_plume, rho_bg, mu_air, mu_plume, mu_bg}:
 *   y < air_h -> air, else (x-cx)^2+(y-cy)^2 < r^2 -> plume, else background.
 * rt:    mat = {y0, amp, k, lx, rho_up, mu_up, rho_low, mu_low}: y < y0 + amp sin(2 pi k x/lx)
 *   -> upper layer, else lower layer. */
int sf_plume_properties(sf_ctx *ctx, const double *x_m, const double *y_m, int64_t nm, const double mat[10],
                        double *rho_m, double *mu_m);
int sf_rt_properties(sf_ctx *ctx, const double *x_m, const double *y_m, int64_t nm, const double mat[8],
                     double *rho_m, double *mu_m);

/* initializeMarkersCPU -- replaces scripts/common.jl:40-87 at one rank: the deterministic marker
 * lattice LinRange(dxm/2, lx - dxm/2, md (Nx-1)) x LinRange(dym/2, ly - dym/2, md (Ny-1)),
 * dxm = dx / md, x outer / y inner, each element evaluated as Julia 1.8's LinRange does (bit for
 * bit).  rand_pos = 1 adds (u - 0.5) dxm and (u - 0.5) dym as RAND_MARKER_POS does, with u from
 * the build's splitmix64(seed, index) (Julia's RNG stream is not reproduced).  x_m, y_m hold
 * md^2 (nx-1)(ny-1) values. */
int sf_lattice_markers(sf_ctx *ctx, double *x_m, double *y_m, int nx, int ny, int md, double lx, double ly,
                       double dx, double dy, int rand_pos, uint64_t seed);

/* setInitialMarkerProperties! -- replaces scripts/common.jl:104-110 with a material descriptor
 * in place of the density / viscosity callbacks: rho_m, mu_m of the first shape (in order) that
 * contains the global position (x0 + x_m, y0 + y_m), else rho_bg, mu_bg.  Shapes (a..e their
 * parameters):
 *   SF_SHAPE_Y_BELOW    y < a                      SF_SHAPE_Y_ABOVE   y >= a
 *   SF_SHAPE_CIRCLE     (x - a)^2 + (y - b)^2 < c^2
 *   SF_SHAPE_RECT       a <= x < b and c <= y < d
 *   SF_SHAPE_SINE_BELOW y < a + b sin(2 pi c x / d)
 * The plume callbacks of StokesFlow2D_multixpu.jl:387-404 are {Y_BELOW air_height, CIRCLE}. */
enum { SF_SHAPE_Y_BELOW = 1, SF_SHAPE_Y_ABOVE = 2, SF_SHAPE_CIRCLE = 3, SF_SHAPE_RECT = 4, SF_SHAPE_SINE_BELOW = 5 };
#define SF_MAX_SHAPES 16
typedef struct {
    int32_t kind; /* SF_SHAPE_* */
    int32_t reserved;
    double a, b, c, d, e;
    double rho, mu; /* material inside the shape */
} sf_shape;
int sf_marker_properties(sf_ctx *ctx, const double *x_m, const double *y_m, int64_t nm, double x0, double y0,
                         const sf_shape *shapes, int n_shapes, double rho_bg, double mu_bg, double *rho_m,
                         double *mu_m);

#ifdef __cplusplus
}
#endif
#endif /* SF_HIP_H */
