// sf_host.h -- private declarations shared by the host files of libsfhip (sf_capi.cpp,
// sf_batch_host.cpp, sf_ensemble.cpp).  Not part of the C ABI.
#pragma once

#include "../../include/sf_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <utility>
#include <vector>

#include "sf_internal.h"

namespace sf {

// Records the message for sf_last_error() and returns code.
int set_err(int code, const char* fmt, ...);

template <typename T>
inline hipError_t dalloc(T** p, size_t n)
{
    return hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T));
}

struct SolveConsts {
    double vsc_mdxy2, ptsc41, max_nxy, dampx, dampy, rdx, rdy;
};
SolveConsts solve_consts(int nx, int ny, double dx, double dy);

// compute_dt (StokesSolver_multixpu.jl:118-121) into scal[SC_DT_CHECK] (and scal[SC_DT] if update).
int enqueue_dt_raw(hipStream_t s, int nx, int ny, const double* vx, const double* vy, double dx, double dy,
                   double maxdisp, int update, double* part, double* scal);

void m2g_free(M2GScratch& S);
int m2g_reserve(M2GScratch& S, int64_t nm, int64_t ncells);

// ---- batched solver (sf_batch_host.cpp), also used by the ensembles ----

// Wide mode: the host keeps an active set of instances whose working sets fit a byte budget
// (the Infinity Cache), launches one PT iteration of all of them per kernel launch, runs each
// instance's dt refresh and convergence check at its own iteration numbers (every check reads
// err back, as the reference does), retires converged instances and refills the set from the
// largest-first queue.  Every instance runs exactly the iterations of solveStokes!.
// Buffers a wide-mode run uses: its stream, device descriptors, reduction partials
// (3 * RED_MAX_BLOCKS) and pinned per-instance scalars (SC_COUNT per instance).
struct WideEnv {
    hipStream_t s;
    const BatchInstance* desc_dev;
    double* part;
    double* hscal;
};

// Per-instance FP32 buffers of the refined wide mode (the layout of sf_solve_stokes_refined's).
struct RefineBufs {
    float *vx[2], *vy[2], *p[2], *mx, *my, *rho, *mub, *mup, *fx, *fy, *fc;
    float *mx2, *my2;  // second momentum (ping-pong with the state in the wide mode)
    double* part;  // 3 * RED_MAX_BLOCKS residual partial sums
};

struct RefineEnv {
    WideEnv w;
    const std::vector<RefineBufs>* rb;
    CorrParams* cdesc_dev;   // 2 per instance
    double* hpart;           // pinned, 3 * RED_MAX_BLOCKS per instance
    float* hdt;              // pinned, 1 per instance
    int64_t rebase;          // FP32 window length (<= 0: ncheck)
};

size_t batch_scratch_doubles(int nx, int ny);
int batch_check(const sf_batch_instance& a, int i, const char* who);

// Launch shape of an instance's wide FP64 passes (results do not depend on it): the most PT
// iterations per pass (2, 3 or 4) and the tile rows of the two-, three- and four-iteration
// passes (0: size default).  sf_batch_autotune keeps one per grid size on the context.
struct BatchShape {
    int kmax = 0, rows2 = 0, rows3 = 0, rows4 = 0;
    // instances of the batch: below WIDE_MAX the default three- and four-iteration tile heights
    // come from the wave-round model over that many copies of the grid (0: the fixed table)
    int copies = 0;
};
struct BatchTune {
    int nx, ny;
    BatchShape sh;
};
BatchShape batch_shape(const sf_ctx* c, int nx, int ny);
BatchShape grid_shape(int nx, int ny);
struct GridShape {
    int nwg, waves;  // workgroups, waves per workgroup
};
GridShape grid_launch_shape(const BatchInstance& d, int k);
BatchInstance batch_descriptor(const sf_batch_instance& a, int rows, double* q, const double* dt_dev,
                               const BatchShape& sh = BatchShape());
std::vector<int> batch_order(const std::vector<std::pair<int, int>>& dims);
int batch_grid(sf_ctx* c, int n, int threads, int max_blocks);
int batch_threads(const sf_batch_params* prm, int* threads, int* rows, const char* who);
int batch_run_wide(const WideEnv& env, const sf_batch_instance* in, int n, const std::vector<BatchInstance>& desc,
                   const std::vector<int>& order, double budget, int depth, std::vector<BatchResult>& res);
size_t refine_floats(int nx, int ny);
RefineBufs refine_carve(int nx, int ny, float* q, double* part);
void refine_descriptors(const sf_batch_instance& a, const RefineBufs& b, int rows, CorrParams out[CDESC_PER_INST]);
int batch_run_wide_refined(const RefineEnv& env, const sf_batch_instance* in, int n,
                           const std::vector<BatchInstance>& desc, std::vector<CorrParams>& cdesc,
                           const std::vector<int>& order, double budget, std::vector<BatchResult>& res);

}  // namespace sf

#define HIPCHK(expr)                                                                                      \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess) return sf::set_err(SF_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                                 __FILE__, __LINE__);                                     \
    } while (0)

namespace sf {

// Ping-pong buffers of the PT loop: a second copy of (Vx_s, Vy_s, P) and, for the
// multi-iteration passes, of (dVx, dVy).
struct PTBuffers {
    double *vx_b, *vy_b, *p_b, *dvx_b, *dvy_b;
};

// Per-launch timing of the pass kernels on one stream: two HIP events around each launch, kept
// in a ring and resolved lazily (a slot is resolved when the ring wraps, long after its launch
// ended), summed per (precision, iterations per launch).  Instrumentation only.
struct PassTimer {
    enum { RING = 256, KINDS = 16 };
    bool on = false;
    hipEvent_t ev[2 * RING] = {};
    int kind[RING] = {};
    int head = 0, count = 0;
    double sec[KINDS] = {};
    int64_t n[KINDS] = {};
    // precision 0: FP64 PT iterations, 1: FP32 correction sweeps; k iterations per launch
    int begin(hipStream_t s, int precision, int k);
    int end(hipStream_t s);
    int flush_one();  // resolves the oldest pending launch (waits for it)
    int flush();      // resolves every pending launch
    void reset();
    void release();
};

// The PT loop state shared by the solve and the fixed-count iteration entry points: the
// ping-pong pair of (Vx_s, Vy_s, P) and, for multi-iteration passes, of (dVx, dVy).  pp holds
// the constants, materials and dt pointer; the state pointers are managed here.
struct PTLoop {
    PTParams pp;
    int nx, ny;
    double *cur_vx, *cur_vy, *cur_p, *nxt_vx, *nxt_vy, *nxt_p;
    double *cur_dvx, *cur_dvy, *nxt_dvx, *nxt_dvy;
    double *vx0, *vy0, *p0, *dvx0, *dvy0;  // the caller's arrays
    PTBuffers b;
    bool pairs;  // multi-iteration passes enabled (b.dvx_b, b.dvy_b must then be set)
    int kmax;    // iterations per pass (2, 3 or 4)
    int rows;    // tile rows of the passes (0: size default)
    PassTimer* tm = nullptr;  // times the pass launches when set and on

    // fuse: iterations per pass (0 size default, 1 one per launch, 2 to 4; larger values cap at 4)
    void init(const PTBuffers& buf, int fuse, int pass_rows, int nx_, int ny_, double* Vx_s, double* Vy_s, double* P,
              double* dVx, double* dVy)
    {
        nx = nx_;
        ny = ny_;
        b = buf;
        cur_vx = vx0 = Vx_s;
        cur_vy = vy0 = Vy_s;
        cur_p = p0 = P;
        cur_dvx = dvx0 = dVx;
        cur_dvy = dvy0 = dVy;
        nxt_vx = b.vx_b;
        nxt_vy = b.vy_b;
        nxt_p = b.p_b;
        nxt_dvx = b.dvx_b;
        nxt_dvy = b.dvy_b;
        rows = pass_rows;
        pairs = fuse != 1;
        kmax = fuse == 0 ? pt_default_fusion(nx, ny) : std::min(fuse, 4);
    }
    // One iteration (flags: check-iteration outputs); dV in place.
    int single(int flags, hipStream_t s);
    // k plain iterations in one pass.
    int pair(hipStream_t s, int k = 2);
    // State back into the caller's arrays.
    int finish(hipStream_t s);
    // n_iter plain iterations (passes of kmax while they fit, then smaller ones, then one), then finish.
    int run_fixed(int64_t n_iter, hipStream_t s);
};

// Scratch of one FP64 solve: the context's (sf_solve_stokes) or carved from a static device
// array (sf_solve_stokes_noalloc).
struct SolveScratch {
    hipStream_t s;
    PTBuffers pt;                              // ping-pong state (dvx_b, dvy_b: NULL when fuse == 1)
    double *txx, *tyy, *txy, *divv, *rx, *ry;  // used where the caller passes NULL
    double* part;                              // 3 * RED_MAX_BLOCKS reduction partials
    double* scal;                              // SC_COUNT device scalars
    double* hscal;                             // host mirror of scal
    hipEvent_t ev0, ev1;                       // timing window; NULL: taken on the host clock
    int seg, depth, max_waves;                 // one-iteration kernel shape
    int fuse, rows;                            // iterations per pass and pass tile rows (0: defaults)
    PassTimer* tm;                             // may be NULL
};

// The body of solveStokes! (StokesSolver_multixpu.jl:14-96) on the given scratch; `who` names the
// entry point in error messages.
int solve_core(const SolveScratch& w, const char* who, double* P, double* Vx_pad, double* Vy_pad,
               const double* rho_vy, const double* mu_b, const double* mu_p, double* txx, double* tyy, double* txy,
               double* divV, double* dtauP, double* Rx, double* Ry, double* dVx, double* dVy, double* dtauVx,
               double* dtauVy, double* Vx_s, double* Vy_s, double g_y, double dx, double dy, int nx, int ny,
               double dt, double maxdisp, int use_fss, double eps, int64_t iter_max, int print_info,
               sf_solve_out* out);

}  // namespace sf

// The context: device scratch for one grid size and the stream work is enqueued on.
struct sf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int nx = 0, ny = 0;
    // PT ping-pong state and check-iteration scratch
    double *vx_b = nullptr, *vy_b = nullptr, *p_b = nullptr;
    double *txx = nullptr, *tyy = nullptr, *txy = nullptr, *divv = nullptr, *rx = nullptr, *ry = nullptr;
    double *part = nullptr;   // 3 * RED_MAX_BLOCKS
    double *scal = nullptr;   // SC_COUNT device scalars
    double *hscal = nullptr;  // pinned host mirror
    unsigned long long *counters = nullptr;   // device counters (RK4)
    unsigned long long *hcounters = nullptr;  // pinned host mirror
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    sf::M2GScratch m2g;
    // FP32 correction scratch (allocated on first refined solve)
    float *r_vx[2] = {nullptr, nullptr}, *r_vy[2] = {nullptr, nullptr}, *r_p[2] = {nullptr, nullptr};
    float *r_mx = nullptr, *r_my = nullptr, *r_rho = nullptr, *r_mub = nullptr, *r_mup = nullptr;
    float *r_fx = nullptr, *r_fy = nullptr, *r_fc = nullptr, *r_rxo = nullptr, *r_ryo = nullptr, *r_rco = nullptr;
    float *r_mx2 = nullptr, *r_my2 = nullptr;  // second momentum (two-sweep pass)
    double* hpart = nullptr;  // pinned host copy of reduction partials
    // batched ensemble scratch
    double* b_pool = nullptr;
    size_t b_pool_cap = 0;
    sf::BatchInstance* b_inst = nullptr;
    int* b_order = nullptr;
    int* b_queue = nullptr;
    sf::BatchResult* b_res = nullptr;
    int b_cap = 0;
    double* b_hscal = nullptr;  // pinned per-instance scalars (wide mode)
    int b_hscal_cap = 0;
    // refined wide mode
    float* b_pool32 = nullptr;
    size_t b_pool32_cap = 0;
    double* b_part = nullptr;      // 3 * RED_MAX_BLOCKS per instance
    sf::CorrParams* b_cdesc = nullptr;  // CDESC_PER_INST per instance
    double* b_hpart = nullptr;     // pinned
    float* b_hdt = nullptr;        // pinned
    int b_rcap = 0;
    // PT launch tuning (results do not depend on it)
    int seg = 0;        // rows per tile; 0 = size heuristic
    int depth = 0;      // rows of operands prefetched; 0 = size default
    int max_waves = -1; // 0 persistent (resident slots), > 0 cap, < 0 one wave per tile
    int fuse = 0;       // sf_ctx_set_pt_fusion: iterations / FP32 sweeps per pass (0 size default, 1..5)
    int pair_rows = 0;  // rows per tile of the multi-iteration passes; 0 = size default
    int fuse64 = 0;     // autotuner's choice for the FP64 passes only (0: none; then fuse / pair_rows apply)
    int rows64 = 0;
    double *dvx_b = nullptr, *dvy_b = nullptr;  // second dVx/dVy (two-iteration pass)
    sf::PassTimer timer;  // sf_ctx_time_passes
    std::vector<sf::BatchTune> btune;  // sf_batch_autotune: wide-pass shapes per grid size
    // grid mode of the batched solver
    unsigned long long* g_sync = nullptr;  // GRID_SYNC_WORDS barrier words per instance
    int g_cap = 0;
    double* g_part = nullptr;  // 3 * GRID_MAX_WAVES partials
};
