// sf_internal.h -- kernel parameter blocks and host launchers shared inside libsfhip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sf_hip.h"

namespace sf {

enum { PT_WRITE_TAU = 1, PT_WRITE_RX = 2 };
enum { RED_SUM = 0, RED_MAX = 1 };
enum { RED_MAX_BLOCKS = 1024 };
// device scalar slots
enum { SC_DT = 0, SC_DT_CHECK = 1, SC_ERR = 2, SC_COUNT = 8 };

struct PTParams {
    int nx, ny;           // node counts
    int seg;              // basic-node rows per tile (<= 0: size heuristic)
    int flags;            // PT_WRITE_TAU | PT_WRITE_RX
    int depth;            // rows of operands prefetched (1 or 2)
    int max_waves;        // 0: one wave per resident slot, > 0: cap, < 0: one wave per tile
    int nstrips, ntiles;  // filled by the launcher
    const double* vx_a;   // Vx_s (nx, ny+1), state in
    const double* vy_a;   // Vy_s (nx+1, ny)
    const double* p_a;    // P (nx-1, ny-1)
    double* vx_b;         // state out
    double* vy_b;
    double* p_b;
    double* dvx;          // (nx-2, ny-1), in place
    double* dvy;          // (nx-1, ny-2), in place
    double* dvx_o;        // two-iteration pass: dVx, dVy out (dvx, dvy are then read only)
    double* dvy_o;
    const double* rho;    // rho_vy (nx+1, ny)
    const double* mu_b;   // (nx, ny)
    const double* mu_p;   // (nx-1, ny-1)
    const double* dt;     // device scalar: stabilisation dt
    double* txx;          // check-iteration outputs
    double* tyy;
    double* txy;
    double* divv;
    double* rx;
    double g, dampx, dampy, rdx, rdy, vsc_mdxy2, ptsc41, max_nxy;
};

struct ResYParams {
    int nx, ny;
    const double* tyy;
    const double* txy;
    const double* P;
    const double* rho;
    const double* vx;
    const double* vy;
    const double* dt;
    double* ry;
    double g, rdx, rdy;
};

struct RedParams {
    const double* a[3];
    int64_t n[3];
    double* part;
    int mode;
};

int pt_num_strips(int nx);
int pt_default_rows(int nx, int ny);
int pt_default_depth(int nx, int ny);
int pt_resident_waves(int depth);
int red_num_blocks(int64_t nmax);
hipError_t launch_pt_fused(const PTParams& p, hipStream_t s);
// Two PT iterations per pass (pt2_tile): state (vx_a, vy_a, p_a, dvx, dvy) -> (vx_b, vy_b, p_b,
// dvx_o, dvy_o); p.seg = rows per tile (<= 0: default).  Bit-identical to two launch_pt_fused.
hipError_t launch_pt_pair(const PTParams& p, hipStream_t s);
int pt2_default_rows(int nx, int ny);
// k = 2, 3 or 4 PT iterations per pass (k >= 3: ptk_tile<k>, 64 - 2k output columns per strip;
// p.seg <= 0: ptk_default_rows).
hipError_t launch_pt_multi(const PTParams& p, int k, hipStream_t s);
// tile height of a k-iteration pass (k >= 3) when `copies` grids of this size share the launch
int ptk_default_rows(int k, int nx, int ny, int copies = 1);
// Snapshot staging: dst = (float)(src + shift) (sf_snapshot.hip).
hipError_t launch_snap_cast(const double* src, float* dst, int64_t n, double shift, hipStream_t s);
int pt_default_fusion(int nx, int ny);  // iterations per pass when the context leaves it open
int corr_default_fusion(int nx, int ny);  // FP32 correction sweeps per pass, likewise
hipError_t launch_pt_resid_y(const ResYParams& p, hipStream_t s);
hipError_t launch_reduce(const RedParams& p, int narr, int nb, hipStream_t s);
hipError_t launch_finalize_dt(const double* part, int nb, double dx, double dy, double maxdisp, double* scal,
                              int update, hipStream_t s);
hipError_t launch_finalize_err(const double* part, int nb, double nrx, double nry, double ndv, double* scal,
                               hipStream_t s);
hipError_t launch_set_scalar(double* dst, double v, hipStream_t s);
hipError_t launch_zero_u64(unsigned long long* dst, int n, hipStream_t s);
hipError_t launch_pad_mirror(int nx, int ny, const double* vx, const double* vy, double* vx_pad, double* vy_pad,
                             hipStream_t s);
hipError_t launch_fill_timesteps(int nx, int ny, const double* mu_p, double* dtauvx, double* dtauvy, double* dtaup,
                                 double vsc_mdxy2, double ptsc41, double max_nxy, hipStream_t s);

// ---- FP32 sweeps in FP64 defect correction ----
struct CorrParams {
    int nx, ny, seg, nstrips, ntiles, depth;
    const float* vx_a;  // correction state in (Vd_x, Vd_y, Pd)
    const float* vy_a;
    const float* p_a;
    float* vx_b;  // correction state out
    float* vy_b;
    float* p_b;
    float* mx;  // PT momentum of the correction (in place)
    float* my;
    float* mx_o;  // two-sweep pass: momentum out (mx, my are then read only)
    float* my_o;
    const float* rho;  // FP32 materials
    const float* mu_b;
    const float* mu_p;
    const float* fx;  // FP64 residual of the outer state, as forcing
    const float* fy;
    const float* fc;
    float* rxo;  // check outputs: correction residuals
    float* ryo;
    float* rco;
    float dt, g, dampx, dampy, rdx, rdy, vsc_mdxy2, ptsc41, max_nxy;
};

struct ResidParams {
    int nx, ny;
    const double* vx;
    const double* vy;
    const double* P;
    const double* rho;
    const double* mu_b;
    const double* mu_p;
    float* fx;
    float* fy;
    float* fc;
    double* part;  // 3 x nblocks
    double dt, g, rdx, rdy;
};

struct RedParamsF {
    const float* a[3];
    int64_t n[3];
    double* part;
};

hipError_t launch_pt_corr(const CorrParams& p, bool check, hipStream_t s);
// Two correction sweeps per pass: (vx_a, vy_a, p_a, mx, my) -> (vx_b, vy_b, p_b, mx_o, my_o);
// p.seg = rows per tile (<= 0: default).  Bit-identical to two launch_pt_corr.
hipError_t launch_pt_corr2(const CorrParams& p, hipStream_t s);
// k = 3 or 4 correction sweeps per pass (corrk_tile<k>, 64 - 2k output columns per strip).
hipError_t launch_pt_corrk(const CorrParams& p, int k, hipStream_t s);
hipError_t launch_resid_f64(const ResidParams& p, int nb, hipStream_t s);
hipError_t launch_update_f64(double* a, const float* d, int64_t n, hipStream_t s);
hipError_t launch_cast_f32(float* dst, const double* src, int64_t n, hipStream_t s);
hipError_t launch_fill_f32(float* dst, int64_t n, float v, hipStream_t s);
hipError_t launch_fill_f64(double* dst, int64_t n, double v, hipStream_t s);
hipError_t launch_cast_f64(double* dst, const float* src, int64_t n, hipStream_t s);
hipError_t launch_reduce_f32(const RedParamsF& p, int narr, int nb, hipStream_t s);

// ---- DRAM probe ----
hipError_t launch_probe_copy(const double* a, double* b, int64_t bytes, int grid, hipStream_t s);
hipError_t launch_probe_r3w1(const double* a, const double* b, const double* c, double* d, int64_t bytes, int grid,
                             hipStream_t s);

// ---- batched ensemble ----
// One ensemble instance.  pt[0] runs the caller's state into the scratch buffers, pt[1]
// runs it back (tile parameters complete, flags unused); the rest is what the solve loop
// and the epilogue need.
struct BatchInstance {
    PTParams pt[2];   // one iteration, state ping-pong by parity, dV in place (the host's template of wpt)
    PTParams wpt[8];  // every mode, dV ping-pong too: [parity] one iteration, [2 + parity] two, [4 + parity]
                      // three, [6 + parity] four
    double *P, *Vx_s, *Vy_s, *Vx_pad, *Vy_pad;  // caller state (in/out)
    double *ry, *scal;                          // library scratch (scal[SC_DT] = stabilisation dt)
    const double* dt_dev;                       // if set, the input dt is read here (device) instead of dt_in
    int nx, ny, use_fss, kmax;  // kmax: most PT iterations per wide pass (2, 3 or 4)
    double dx, dy, dt_in, maxdisp, eps;
    int64_t iter_max;
};

struct BatchResult {
    double err, dt_new, dt_last;
    int64_t niter;
    int converged, pad;
};

// Grid mode (sf_grid.hip): one instance's whole solve in one launch of nwg workgroups of wg_waves
// waves (at most one workgroup per CU, grid_capacity(k, wg_waves), all resident), the passes
// separated by a grid barrier.  sync holds GRID_SYNC_WORDS words, zeroed by launch_grid_solve:
// [GRID_SYNC_ARRIVE] XCDs arrived, [GRID_SYNC_FAIL] failure flag (a barrier wait exceeded the
// limit), [GRID_SYNC_STICKY] its copy, which launch_grid_solve leaves alone (the host clears it),
// [GRID_SYNC_START] workgroups started, [GRID_SYNC_XARRIVE + 16 x] workgroups of XCD x
// arrived, [GRID_SYNC_XCOUNT + 16 x] workgroups on XCD x (one 128-byte line each); part holds
// 3 * wg_waves * nwg partials.
enum {
    GRID_WG_WAVES = 8,
    GRID_MAX_WAVES = 4096,
    GRID_WAIT_SECONDS = 5,
    GRID_MAX_XCD = 16,
    GRID_SYNC_ARRIVE = 0,
    GRID_SYNC_FAIL = 1,
    GRID_SYNC_START = 2,
    GRID_SYNC_STICKY = 3,
    GRID_SYNC_XARRIVE = 16,
    GRID_SYNC_XCOUNT = 16 + 16 * GRID_MAX_XCD,
    GRID_SYNC_WORDS = 16 + 32 * GRID_MAX_XCD,
};
struct GridArgs {
    const BatchInstance* desc;
    BatchResult* res;
    unsigned long long* sync;
    double* part;
    int nwg, wg_waves;
    uint64_t limit;  // wall-clock ticks a workgroup may wait at a barrier
};
// k = PT iterations per pass (1, 2 or 3; four need more than 256 VGPRs in this kernel); waves =
// waves per workgroup (1 to GRID_WG_WAVES).
int grid_capacity(int k, int waves);
uint64_t grid_wait_limit();  // wall-clock ticks of the current limit (sf_set_grid_wait_limit)
void grid_set_wait_seconds(double seconds);
double grid_wait_seconds();  // the current limit in seconds
hipError_t launch_grid_solve(const GridArgs& a, int k, hipStream_t s);

// Wide mode of the batched solver: the active set of one PT-iteration launch.
#define WIDE_MAX 64
struct WideArgs {
    const BatchInstance* desc;
    int n_active;
    int tile_start[WIDE_MAX + 1];  // prefix of the slots' tile counts
    int slot_inst[WIDE_MAX];       // descriptor index
    int slot_pf[WIDE_MAX];         // ping-pong parity | PT_WRITE_* flags << 1
};
// One PT iteration of every slot (descriptor wpt[parity]).
hipError_t launch_wide_pt(const WideArgs& a, int depth, hipStream_t s);
// Two PT iterations of every slot (descriptor wpt[2 + parity]; flags must be 0).
hipError_t launch_wide_pt2(const WideArgs& a, hipStream_t s);
// Three PT iterations of every slot (descriptor wpt[4 + parity]).
hipError_t launch_wide_pt3(const WideArgs& a, hipStream_t s);
// Four PT iterations of every slot (descriptor wpt[6 + parity]).
hipError_t launch_wide_pt4(const WideArgs& a, hipStream_t s);

// Wide mode of the FP32 correction sweeps: slot k runs descriptor slot_desc[k] (two
// descriptors per instance, one per ping-pong direction).
struct WideCorrArgs {
    const CorrParams* desc;
    int n_active;
    int tile_start[WIDE_MAX + 1];
    int slot_desc[WIDE_MAX];
};
hipError_t launch_wide_corr(const WideCorrArgs& a, hipStream_t s);
hipError_t launch_wide_corr2(const WideCorrArgs& a, hipStream_t s);  // two sweeps per slot
hipError_t launch_wide_corr3(const WideCorrArgs& a, hipStream_t s);  // three sweeps per slot
int corr_num_strips(int nx);
int corr2_num_strips(int nx);
int corr2_default_rows(int nx, int ny);
// Correction descriptors per instance in the wide refined mode: [parity] one sweep,
// [2 + parity] two sweeps, [4 + parity] three; the state and the momentum ping-pong together.
#define CDESC_PER_INST 6

hipError_t launch_batch_solve(const BatchInstance* inst, const int* order, int n_inst, int* queue,
                              BatchResult* res, int threads, int nblocks, hipStream_t s);
int batch_blocks_per_cu(int threads);

// ---- marker -> grid ----
// One marker's contribution in cell order: the relative position in its cell, its value and
// its index (the merge key).  32 bytes, so a merge step reads one record from one cache line.
struct M2GRec {
    double dx, dy, val;
    int idx, pad;
};

struct M2GScratch {
    int64_t cap_markers = 0;
    int64_t cap_cells = 0;
    int* key = nullptr;      // [nm] cell key or -1
    int* count = nullptr;    // [cells]
    int* offset = nullptr;   // [cells + 1]
    int* fill = nullptr;     // [cells]
    int* rec_idx = nullptr;  // [nm] marker index in cell order
    M2GRec* rec = nullptr;   // [nm] records in cell order
    int* blocksums = nullptr;  // scan scratch
    int64_t cap_blocks = 0;
    unsigned long long* counters = nullptr;  // [4]: bad, empty, nonfinite, spare
};

hipError_t m2g_run(M2GScratch& s, double xmin, double ymin, double* val_grid, int gnx, int gny, const double* x_m,
                   const double* y_m, const double* val_m, int64_t nm, double dx, double dy, double* val_wt_sum,
                   double* wt_sum, int ws_nx, int ws_ny, hipStream_t st);

hipError_t m2g_cell_keys(M2GScratch& s, double xmin, double ymin, int gnx, int gny, const double* x_m,
                         const double* y_m, int64_t nm, double dx, double dy, int* keys, int* counts, hipStream_t st);

// ---- RK4 / markers ----
hipError_t launch_rk4(double* x_m, double* y_m, int64_t nm, const double* vx, int vx_n1, int vx_n2, const double* vy,
                      int vy_n1, int vy_n2, double x_vx_min, double y_vx_min, double x_vy_min, double y_vy_min,
                      double dt, double lx, double ly, double dx, double dy, unsigned long long* counters,
                      hipStream_t s, const double* dt_ptr = nullptr);
hipError_t launch_init_markers(double* x_m, double* y_m, int nx, int ny, int md, double dx, double dy,
                               double jitter_frac, uint64_t seed, hipStream_t s);
struct Mat10 {
    double v[10];
};
hipError_t launch_plume_properties(const double* x_m, const double* y_m, int64_t nm, Mat10 mat, double* rho_m,
                                   double* mu_m, hipStream_t s);
hipError_t launch_rt_properties(const double* x_m, const double* y_m, int64_t nm, Mat10 mat, double* rho_m,
                                double* mu_m, hipStream_t s);
hipError_t launch_lattice_markers(double* x_m, double* y_m, int nx, int ny, int md, double lx, double ly, double dx,
                                  double dy, int rand_pos, uint64_t seed, hipStream_t s);
// A material descriptor passed by value to the kernel.
struct ShapeSet {
    int n;
    double rho_bg, mu_bg;
    sf_shape s[SF_MAX_SHAPES];
};
hipError_t launch_shape_properties(const double* x_m, const double* y_m, int64_t nm, double x0, double y0,
                                   const ShapeSet& set, double* rho_m, double* mu_m, hipStream_t s);

}  // namespace sf
