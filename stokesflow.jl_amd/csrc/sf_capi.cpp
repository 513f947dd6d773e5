// sf_capi.cpp -- C ABI of libsfhip (include/sf_hip.h): error state, context, the PT solve
// loop (FP64 and FP32-in-FP64 defect correction), the autotuner, the marker entry points and
// the one-rank timestep.  The batched solver is in sf_batch_host.cpp, ensembles in
// sf_ensemble.cpp.
#include <cstdarg>

#include "sf_host.h"

using namespace sf;

#define SF_VERSION_NUM 100  // 0.1.0

thread_local char g_err[512] = "";

int sf::set_err(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

extern "C" int sf_version(void) { return SF_VERSION_NUM; }
extern "C" const char* sf_last_error(void) { return g_err; }

static void free_ctx(sf_ctx* c)
{
    if (!c) return;
    void* bufs[] = {c->dvx_b, c->dvy_b, c->vx_b, c->vy_b, c->p_b, c->txx, c->tyy, c->txy, c->divv, c->rx, c->ry, c->part, c->scal,
                    c->counters, c->m2g.key, c->m2g.count, c->m2g.offset, c->m2g.fill, c->m2g.rec_idx,
                    c->m2g.rec, c->m2g.blocksums, c->m2g.counters};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    void* rbufs[] = {c->r_mx2, c->r_my2, c->r_vx[0], c->r_vx[1], c->r_vy[0], c->r_vy[1], c->r_p[0], c->r_p[1], c->r_mx, c->r_my,
                     c->r_rho, c->r_mub, c->r_mup, c->r_fx, c->r_fy, c->r_fc, c->r_rxo, c->r_ryo, c->r_rco};
    for (void* b : rbufs)
        if (b) (void)hipFree(b);
    if (c->hpart) (void)hipHostFree(c->hpart);
    if (c->b_hscal) (void)hipHostFree(c->b_hscal);
    for (void* b : {(void*)c->b_pool32, (void*)c->b_part, (void*)c->b_cdesc})
        if (b) (void)hipFree(b);
    for (void* b : {(void*)c->b_hpart, (void*)c->b_hdt})
        if (b) (void)hipHostFree(b);
    for (void* b : {(void*)c->b_pool, (void*)c->b_inst, (void*)c->b_order, (void*)c->b_queue, (void*)c->b_res,
                    (void*)c->g_sync, (void*)c->g_part})
        if (b) (void)hipFree(b);
    if (c->hscal) (void)hipHostFree(c->hscal);
    if (c->hcounters) (void)hipHostFree(c->hcounters);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    c->timer.release();
    delete c;
}

extern "C" int sf_ctx_create(sf_ctx** out, int nx, int ny, void* stream)
{
    if (!out) return set_err(SF_ERR_ARG, "sf_ctx_create: ctx pointer is NULL");
    *out = nullptr;
    const bool markers_only = (nx == 0 && ny == 0);
    if (!markers_only && (nx < 3 || ny < 3))
        return set_err(SF_ERR_ARG, "sf_ctx_create: need Nx, Ny >= 3, or 0 x 0 for a marker-only context (got %d, %d)",
                       nx, ny);
    if ((int64_t)(nx + 2) * (ny + 2) >= (int64_t)1 << 31)
        return set_err(SF_ERR_SIZE, "sf_ctx_create: grid %d x %d exceeds the 32-bit cell index", nx, ny);
    sf_ctx* c = new sf_ctx();
    c->nx = nx;
    c->ny = ny;
    c->stream = (hipStream_t)stream;
    hipError_t e = hipGetDevice(&c->device);
    // a marker-only context keeps 1-element placeholders for the grid scratch
    const size_t gx = markers_only ? 3 : (size_t)nx, gy = markers_only ? 3 : (size_t)ny;
    const size_t nvx = gx * (gy + 1), nvy = (gx + 1) * gy, np = (gx - 1) * (gy - 1);
    if (e == hipSuccess) e = dalloc(&c->vx_b, nvx);
    if (e == hipSuccess) e = dalloc(&c->vy_b, nvy);
    if (e == hipSuccess) e = dalloc(&c->p_b, np);
    if (e == hipSuccess) e = dalloc(&c->txx, np);
    if (e == hipSuccess) e = dalloc(&c->tyy, np);
    if (e == hipSuccess) e = dalloc(&c->txy, gx * gy);
    if (e == hipSuccess) e = dalloc(&c->divv, np);
    if (e == hipSuccess) e = dalloc(&c->rx, (gx - 2) * (gy - 1));
    if (e == hipSuccess) e = dalloc(&c->ry, (gx - 1) * (gy - 2));
    if (e == hipSuccess) e = dalloc(&c->part, 3 * (size_t)RED_MAX_BLOCKS);
    if (e == hipSuccess) e = dalloc(&c->scal, (size_t)SC_COUNT);
    if (e == hipSuccess) e = dalloc(&c->counters, 4);
    if (e == hipSuccess) e = dalloc(&c->m2g.counters, 4);
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->hscal, SC_COUNT * sizeof(double), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->hcounters, 8 * sizeof(unsigned long long), hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e != hipSuccess) {
        free_ctx(c);
        return set_err(SF_ERR_HIP, "sf_ctx_create: %s", hipGetErrorString(e));
    }
    *out = c;
    return SF_OK;
}

extern "C" int sf_ctx_destroy(sf_ctx* c)
{
    if (!c) return SF_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    else (void)hipDeviceSynchronize();
    free_ctx(c);
    return SF_OK;
}

extern "C" int sf_ctx_set_pt_tuning(sf_ctx* c, int rows, int depth, int max_waves)
{
    if (!c) return set_err(SF_ERR_ARG, "sf_ctx_set_pt_tuning: NULL ctx");
    if (rows < 0 || rows > 4096) return set_err(SF_ERR_ARG, "sf_ctx_set_pt_tuning: rows %d outside [0, 4096]", rows);
    if (depth < 0 || depth > 2) return set_err(SF_ERR_ARG, "sf_ctx_set_pt_tuning: depth %d not 0, 1 or 2", depth);
    c->seg = rows;
    c->depth = depth;
    c->max_waves = max_waves;
    return SF_OK;
}

extern "C" int sf_ctx_set_pt_fusion(sf_ctx* c, int iters_per_launch, int rows)
{
    if (!c) return set_err(SF_ERR_ARG, "sf_ctx_set_pt_fusion: NULL ctx");
    if (iters_per_launch < 0 || iters_per_launch > 5)
        return set_err(SF_ERR_ARG, "sf_ctx_set_pt_fusion: iters_per_launch %d not in 0..5", iters_per_launch);
    if (rows < 0 || rows > 4096) return set_err(SF_ERR_ARG, "sf_ctx_set_pt_fusion: rows %d outside [0, 4096]", rows);
    c->fuse = iters_per_launch;
    c->pair_rows = rows;
    c->fuse64 = c->rows64 = 0;  // an explicit setting replaces the autotuner's FP64 choice
    return SF_OK;
}

int PTLoop::single(int flags, hipStream_t s)
{
    pp.flags = flags;
    pp.vx_a = cur_vx;
    pp.vy_a = cur_vy;
    pp.p_a = cur_p;
    pp.vx_b = nxt_vx;
    pp.vy_b = nxt_vy;
    pp.p_b = nxt_p;
    pp.dvx = cur_dvx;
    pp.dvy = cur_dvy;
    pp.dvx_o = nullptr;  // in place
    pp.dvy_o = nullptr;
    HIPCHK(launch_pt_fused(pp, s));
    std::swap(cur_vx, nxt_vx);
    std::swap(cur_vy, nxt_vy);
    std::swap(cur_p, nxt_p);
    return SF_OK;
}

int PTLoop::pair(hipStream_t s, int k)
{
    const int seg1 = pp.seg;
    pp.flags = 0;
    pp.seg = rows;
    pp.vx_a = cur_vx;
    pp.vy_a = cur_vy;
    pp.p_a = cur_p;
    pp.vx_b = nxt_vx;
    pp.vy_b = nxt_vy;
    pp.p_b = nxt_p;
    pp.dvx = cur_dvx;
    pp.dvy = cur_dvy;
    pp.dvx_o = nxt_dvx;
    pp.dvy_o = nxt_dvy;
    if (tm && tm->on) {
        const int r = tm->begin(s, 0, k);
        if (r) return r;
    }
    const hipError_t e = launch_pt_multi(pp, k, s);
    pp.seg = seg1;  // the one-iteration kernel keeps its own tile rows
    HIPCHK(e);
    if (tm && tm->on) {
        const int r = tm->end(s);
        if (r) return r;
    }
    std::swap(cur_vx, nxt_vx);
    std::swap(cur_vy, nxt_vy);
    std::swap(cur_p, nxt_p);
    std::swap(cur_dvx, nxt_dvx);
    std::swap(cur_dvy, nxt_dvy);
    return SF_OK;
}

int PTLoop::finish(hipStream_t s)
{
    const size_t gx = nx, gy = ny;
    if (cur_vx != vx0) {
        HIPCHK(hipMemcpyAsync(vx0, cur_vx, sizeof(double) * gx * (gy + 1), hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(vy0, cur_vy, sizeof(double) * (gx + 1) * gy, hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(p0, cur_p, sizeof(double) * (gx - 1) * (gy - 1), hipMemcpyDeviceToDevice, s));
    }
    if (cur_dvx != dvx0) {
        HIPCHK(hipMemcpyAsync(dvx0, cur_dvx, sizeof(double) * (gx - 2) * (gy - 1), hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(dvy0, cur_dvy, sizeof(double) * (gx - 1) * (gy - 2), hipMemcpyDeviceToDevice, s));
    }
    cur_vx = vx0;
    cur_vy = vy0;
    cur_p = p0;
    cur_dvx = dvx0;
    cur_dvy = dvy0;
    nxt_vx = b.vx_b;
    nxt_vy = b.vy_b;
    nxt_p = b.p_b;
    nxt_dvx = b.dvx_b;
    nxt_dvy = b.dvy_b;
    return SF_OK;
}

int PTLoop::run_fixed(int64_t n_iter, hipStream_t s)
{
    int64_t it = 0;
    int rc = SF_OK;
    for (int k = kmax; pairs && k >= 2; --k)
        for (; it + k <= n_iter && rc == SF_OK; it += k) rc = pair(s, k);
    for (; it < n_iter && rc == SF_OK; ++it) rc = single(0, s);
    if (rc) return rc;
    return finish(s);
}

// The context's ping-pong buffers of the PT loop; the second dVx/dVy (multi-iteration passes)
// is allocated on first use.
static int ctx_pt_buffers(sf_ctx* c, bool pairs, PTBuffers* b)
{
    if (pairs && !c->dvx_b) {
        const size_t nx = c->nx, ny = c->ny;
        hipError_t e = dalloc(&c->dvx_b, (nx - 2) * (ny - 1));
        if (e == hipSuccess) e = dalloc(&c->dvy_b, (nx - 1) * (ny - 2));
        if (e != hipSuccess) {
            if (c->dvx_b) (void)hipFree(c->dvx_b);
            c->dvx_b = c->dvy_b = nullptr;
            return set_err(SF_ERR_HIP, "PT loop: dV buffer allocation failed: %s", hipGetErrorString(e));
        }
    }
    *b = PTBuffers{c->vx_b, c->vy_b, c->p_b, c->dvx_b, c->dvy_b};
    return SF_OK;
}

// The PT loop of a context: the autotuner's FP64 choice, else the explicit setting, else the
// size default.
static int ctx_pt_loop(sf_ctx* c, PTLoop* L, double* Vx_s, double* Vy_s, double* P, double* dVx, double* dVy)
{
    const int fuse = c->fuse64 > 0 ? c->fuse64 : c->fuse;
    const int rows = c->fuse64 > 0 ? c->rows64 : c->pair_rows;
    PTBuffers b{};
    const int rc = ctx_pt_buffers(c, fuse != 1, &b);
    if (rc) return rc;
    L->init(b, fuse, rows, c->nx, c->ny, Vx_s, Vy_s, P, dVx, dVy);
    L->tm = &c->timer;
    return SF_OK;
}

// ---------------------------------------------------------------------------------------
// per-launch timing of the pass kernels

int PassTimer::begin(hipStream_t s, int precision, int k)
{
    if (count == RING) {  // resolve the oldest launch (recorded RING launches ago)
        const int r = flush_one();
        if (r) return r;
    }
    const int slot = (head + count) % RING;
    for (int q = 0; q < 2; ++q)
        if (!ev[2 * slot + q]) HIPCHK(hipEventCreate(&ev[2 * slot + q]));
    kind[slot] = (precision ? 8 : 0) + std::min(std::max(k, 0), 7);
    HIPCHK(hipEventRecord(ev[2 * slot], s));
    return SF_OK;
}

int PassTimer::end(hipStream_t s)
{
    const int slot = (head + count) % RING;
    HIPCHK(hipEventRecord(ev[2 * slot + 1], s));
    ++count;
    return SF_OK;
}

int PassTimer::flush_one()
{
    float ms = 0.f;
    HIPCHK(hipEventSynchronize(ev[2 * head + 1]));
    HIPCHK(hipEventElapsedTime(&ms, ev[2 * head], ev[2 * head + 1]));
    sec[kind[head]] += 1e-3 * (double)ms;
    n[kind[head]] += 1;
    head = (head + 1) % RING;
    --count;
    return SF_OK;
}

int PassTimer::flush()
{
    while (count > 0) {
        const int r = flush_one();
        if (r) return r;
    }
    return SF_OK;
}

void PassTimer::reset()
{
    head = count = 0;
    for (int q = 0; q < KINDS; ++q) {
        sec[q] = 0.0;
        n[q] = 0;
    }
}

void PassTimer::release()
{
    for (hipEvent_t& e : ev)
        if (e) {
            (void)hipEventDestroy(e);
            e = nullptr;
        }
    reset();
    on = false;
}

extern "C" int sf_ctx_time_passes(sf_ctx* c, int enable)
{
    if (!c) return set_err(SF_ERR_ARG, "sf_ctx_time_passes: NULL ctx");
    if (c->timer.count) {  // drop what an earlier timing left pending
        if (c->stream) HIPCHK(hipStreamSynchronize(c->stream));
        else HIPCHK(hipDeviceSynchronize());
    }
    c->timer.reset();
    c->timer.on = enable != 0;
    return SF_OK;
}

extern "C" int sf_ctx_pass_times(sf_ctx* c, sf_pass_time* out, int cap, int* n_out)
{
    if (!c) return set_err(SF_ERR_ARG, "sf_ctx_pass_times: NULL ctx");
    if (cap > 0 && !out) return set_err(SF_ERR_ARG, "sf_ctx_pass_times: NULL output with cap > 0");
    const int rc = c->timer.flush();
    if (rc) return rc;
    int m = 0;
    for (int q = 0; q < PassTimer::KINDS; ++q) {
        if (!c->timer.n[q]) continue;
        if (m < cap) out[m] = sf_pass_time{q >= 8 ? 1 : 0, q % 8, c->timer.n[q], c->timer.sec[q]};
        ++m;
    }
    if (n_out) *n_out = m;
    return SF_OK;
}

extern "C" int sf_ctx_set_stream(sf_ctx* c, void* stream)
{
    if (!c) return set_err(SF_ERR_ARG, "sf_ctx_set_stream: NULL ctx");
    c->stream = (hipStream_t)stream;
    return SF_OK;
}

// ---------------------------------------------------------------------------------------
// PT solve

// Constants of StokesSolver_multixpu.jl:17-27 with nx_g() = Nx, ny_g() = Ny.
SolveConsts sf::solve_consts(int nx, int ny, double dx, double dy)
{
    const double Vdmp = 4.0, Vsc = 0.45, Ptsc = 1.0 / 2.0;
    const double m = std::fmin(dx, dy);
    SolveConsts k;
    k.vsc_mdxy2 = Vsc * (m * m);
    k.ptsc41 = Ptsc * 4.1;
    k.max_nxy = (double)std::max(nx, ny);
    k.dampx = 1.0 - Vdmp / (double)nx;
    k.dampy = 1.0 - Vdmp / (double)ny;
    k.rdx = 1.0 / dx;
    k.rdy = 1.0 / dy;
    return k;
}

namespace {

double bytes_per_iter(int nx, int ny)
{
    const double n_vx = (double)nx * (ny + 1), n_vy = (double)(nx + 1) * ny, n_p = (double)(nx - 1) * (ny - 1);
    const double n_dvx = (double)(nx - 2) * (ny - 1), n_dvy = (double)(nx - 1) * (ny - 2);
    const double n_mub = (double)nx * ny, n_rho = (double)(nx + 1) * ny;
    return 8.0 * (2.0 * (n_vx + n_vy + n_p + n_dvx + n_dvy) + n_p + n_mub + n_rho);
}

}  // namespace

static int check_ctx(sf_ctx* c, int nx, int ny, const char* fn)
{
    if (!c) return set_err(SF_ERR_ARG, "%s: NULL ctx", fn);
    if (c->nx == 0) return set_err(SF_ERR_ARG, "%s: marker-only context has no grid scratch", fn);
    if (nx != c->nx || ny != c->ny)
        return set_err(SF_ERR_ARG, "%s: grid %d x %d does not match the context's %d x %d", fn, nx, ny, c->nx, c->ny);
    return SF_OK;
}

// max|Vx|, max|Vy| -> dt_check (and dt if update) on the device
// compute_dt (StokesSolver_multixpu.jl:118-121) into scal[SC_DT_CHECK] (and scal[SC_DT] if update).
int sf::enqueue_dt_raw(hipStream_t s, int nx, int ny, const double* vx, const double* vy, double dx, double dy,
                           double maxdisp, int update, double* part, double* scal)
{
    RedParams rp{};
    rp.a[0] = vx;
    rp.n[0] = (int64_t)nx * (ny + 1);
    rp.a[1] = vy;
    rp.n[1] = (int64_t)(nx + 1) * ny;
    rp.part = part;
    rp.mode = RED_MAX;
    const int nb = red_num_blocks(std::max(rp.n[0], rp.n[1]));
    HIPCHK(launch_reduce(rp, 2, nb, s));
    HIPCHK(launch_finalize_dt(part, nb, dx, dy, maxdisp, scal, update, s));
    return SF_OK;
}

static int enqueue_dt(sf_ctx* c, const double* vx, const double* vy, double dx, double dy, double maxdisp, int update)
{
    return enqueue_dt_raw(c->stream, c->nx, c->ny, vx, vy, dx, dy, maxdisp, update, c->part, c->scal);
}

int sf::solve_core(const SolveScratch& w, const char* who, double* P, double* Vx_pad, double* Vy_pad,
                   const double* rho_vy, const double* mu_b, const double* mu_p, double* txx, double* tyy, double* txy,
                   double* divV, double* dtauP, double* Rx, double* Ry, double* dVx, double* dVy, double* dtauVx,
                   double* dtauVy, double* Vx_s, double* Vy_s, double g_y, double dx, double dy, int nx, int ny,
                   double dt, double maxdisp, int use_fss, double eps, int64_t iter_max, int print_info,
                   sf_solve_out* out)
{
    if (!P || !Vx_pad || !Vy_pad || !rho_vy || !mu_b || !mu_p || !dVx || !dVy || !Vx_s || !Vy_s)
        return set_err(SF_ERR_ARG, "%s: a state or material array is NULL", who);
    if (!(dx > 0.0) || !(dy > 0.0)) return set_err(SF_ERR_ARG, "%s: dx, dy must be > 0", who);
    hipStream_t s = w.s;
    const SolveConsts k = solve_consts(nx, ny, dx, dy);
    if (!use_fss) dt = 0.0;

    double* t_xx = txx ? txx : w.txx;
    double* t_yy = tyy ? tyy : w.tyy;
    double* t_xy = txy ? txy : w.txy;
    double* d_v = divV ? divV : w.divv;
    double* r_x = Rx ? Rx : w.rx;
    double* r_y = Ry ? Ry : w.ry;
    if (dtauP || dtauVx || dtauVy)
        HIPCHK(launch_fill_timesteps(nx, ny, mu_p, dtauVx, dtauVy, dtauP, k.vsc_mdxy2, k.ptsc41, k.max_nxy, s));
    HIPCHK(launch_set_scalar(w.scal + SC_DT, dt, s));

    const int mn = std::min(nx, ny);
    const int64_t ncheck = std::min<int64_t>(5 * (int64_t)mn, 2000);
    const int64_t ndtupdate = 10 * (int64_t)mn;
    const int64_t itert1 = 11;
    const double n_rx = (double)(nx - 2) * (ny - 1), n_ry = (double)(nx - 1) * (ny - 2), n_dv = (double)(nx - 1) * (ny - 1);
    auto dt_of = [&](const double* vx, const double* vy, int update) {
        return enqueue_dt_raw(s, nx, ny, vx, vy, dx, dy, maxdisp, update, w.part, w.scal);
    };

    PTParams pp{};
    pp.nx = nx;
    pp.ny = ny;
    pp.seg = w.seg;
    pp.depth = w.depth;
    pp.max_waves = w.max_waves;
    pp.dvx = dVx;
    pp.dvy = dVy;
    pp.rho = rho_vy;
    pp.mu_b = mu_b;
    pp.mu_p = mu_p;
    pp.dt = w.scal + SC_DT;
    pp.txx = t_xx;
    pp.tyy = t_yy;
    pp.txy = t_xy;
    pp.divv = d_v;
    pp.rx = r_x;
    pp.g = g_y;
    pp.dampx = k.dampx;
    pp.dampy = k.dampy;
    pp.rdx = k.rdx;
    pp.rdy = k.rdy;
    pp.vsc_mdxy2 = k.vsc_mdxy2;
    pp.ptsc41 = k.ptsc41;
    pp.max_nxy = k.max_nxy;

    PTLoop L;
    L.init(w.pt, w.fuse, w.rows, nx, ny, Vx_s, Vy_s, P, dVx, dVy);
    L.tm = w.tm;
    L.pp = pp;

    double err = 2.0 * eps;
    int64_t iter = 1, niter = 0;
    bool timed = false;
    auto t_host0 = std::chrono::steady_clock::now();
    int rc = SF_OK;
    while (err > eps && iter <= iter_max) {
        if (iter == itert1) {  // the reference's timing window starts here (:44)
            if (w.ev0) {
                HIPCHK(hipEventRecord(w.ev0, s));
            } else {
                HIPCHK(hipStreamSynchronize(s));
                t_host0 = std::chrono::steady_clock::now();
            }
            timed = true;
        }
        // iterations iter..iter+k-1 in one pass when none is a check iteration, none but the
        // last is followed by a dt refresh, the last is not iter_max and the timing window does
        // not start inside the pass
        auto fusable = [&](int64_t k) {
            if (iter + k - 1 >= iter_max) return false;
            for (int64_t j = iter; j < iter + k; ++j) {
                if (j % ncheck == 0) return false;
                if (j < iter + k - 1 && use_fss && j % ndtupdate == 0) return false;
                if (j > iter && j == itert1) return false;
            }
            return true;
        };
        int kf = 1;
        for (int k = L.pairs ? L.kmax : 1; k >= 2 && kf == 1; --k)
            if (fusable(k)) kf = k;
        if (kf > 1) {
            rc = L.pair(s, kf);
            if (rc) return rc;
            if (use_fss && (iter + kf - 1) % ndtupdate == 0) {
                rc = dt_of(L.cur_vx, L.cur_vy, 1);
                if (rc) return rc;
            }
            iter += kf;
            niter += kf;
            continue;
        }
        const bool check = (iter % ncheck) == 0;
        const bool last = iter == iter_max;
        rc = L.single(((check || last) ? PT_WRITE_TAU : 0) | (check ? PT_WRITE_RX : 0), s);
        if (rc) return rc;
        double* cur_vx = L.cur_vx;
        double* cur_vy = L.cur_vy;
        double* cur_p = L.cur_p;

        const bool dt_upd = use_fss && (iter % ndtupdate) == 0;
        if (dt_upd) {
            rc = dt_of(cur_vx, cur_vy, 1);
            if (rc) return rc;
        }
        if (check) {
            if (use_fss) {
                if (!dt_upd) {
                    rc = dt_of(cur_vx, cur_vy, 0);
                    if (rc) return rc;
                }
            } else {
                HIPCHK(launch_set_scalar(w.scal + SC_DT_CHECK, 0.0, s));
            }
            ResYParams rq{};
            rq.nx = nx;
            rq.ny = ny;
            rq.tyy = t_yy;
            rq.txy = t_xy;
            rq.P = cur_p;
            rq.rho = rho_vy;
            rq.vx = cur_vx;
            rq.vy = cur_vy;
            rq.dt = w.scal + SC_DT_CHECK;
            rq.ry = r_y;
            rq.g = g_y;
            rq.rdx = k.rdx;
            rq.rdy = k.rdy;
            HIPCHK(launch_pt_resid_y(rq, s));
            RedParams rp{};
            rp.a[0] = r_x;
            rp.n[0] = (int64_t)n_rx;
            rp.a[1] = r_y;
            rp.n[1] = (int64_t)n_ry;
            rp.a[2] = d_v;
            rp.n[2] = (int64_t)n_dv;
            rp.part = w.part;
            rp.mode = RED_SUM;
            const int nb = red_num_blocks((int64_t)n_dv);
            HIPCHK(launch_reduce(rp, 3, nb, s));
            HIPCHK(launch_finalize_err(w.part, nb, n_rx, n_ry, n_dv, w.scal, s));
            HIPCHK(hipMemcpyAsync(w.hscal, w.scal, SC_COUNT * sizeof(double), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            err = w.hscal[SC_ERR];
            if (print_info >= 2) {
                printf("  iter = %lld, err = %1.3e\n", (long long)iter, err);
                fflush(stdout);
            }
        }
        iter += 1;
        niter += 1;
    }
    double runtime = 0.0;
    if (timed) {
        if (w.ev0) {
            HIPCHK(hipEventRecord(w.ev1, s));
        } else {
            HIPCHK(hipStreamSynchronize(s));
            runtime = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_host0).count();
        }
    }

    rc = L.finish(s);  // state back into the caller's arrays
    if (rc) return rc;
    HIPCHK(launch_pad_mirror(nx, ny, Vx_s, Vy_s, Vx_pad, Vy_pad, s));
    rc = dt_of(Vx_s, Vy_s, 0);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(w.hscal, w.scal, SC_COUNT * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));

    if (timed && w.ev0) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, w.ev0, w.ev1));
        runtime = 1e-3 * (double)ms;
    }
    const double t_it = runtime / (double)(niter - itert1 + 1);
    const double a_eff = (3.0 * 2.0) / 1e9 * (double)nx * (double)ny * 8.0;
    if (print_info)
        printf("Total steps = %lld, err = %1.3e, time = %1.3e sec (@ T_eff = %1.2f GB/s) \n", (long long)niter, err,
               runtime, timed ? a_eff / t_it : 0.0);
    if (out) {
        out->dt_new = w.hscal[SC_DT_CHECK];
        out->err = err;
        out->t_eff = timed ? a_eff / t_it : 0.0;
        out->runtime = runtime;
        out->niter = niter;
        out->converged = err <= eps;
        out->reserved = 0;
        out->dt_last = w.hscal[SC_DT];
        out->t_it = timed ? t_it : 0.0;
        out->gbs_alg = timed ? bytes_per_iter(nx, ny) / t_it / 1e9 : 0.0;
    }
    return SF_OK;
}

extern "C" int sf_solve_stokes(sf_ctx* c, double* P, double* Vx_pad, double* Vy_pad, const double* rho_vy,
                               const double* mu_b, const double* mu_p, double* txx, double* tyy, double* txy,
                               double* divV, double* dtauP, double* Rx, double* Ry, double* dVx, double* dVy,
                               double* dtauVx, double* dtauVy, double* Vx_s, double* Vy_s, double g_y, double dx,
                               double dy, int nx, int ny, double dt, double maxdisp, int use_fss, double eps,
                               int64_t iter_max, int print_info, sf_solve_out* out)
{
    int rc = check_ctx(c, nx, ny, "sf_solve_stokes");
    if (rc) return rc;
    // the autotuner's FP64 choice, else the explicit setting, else the size default
    const int fuse = c->fuse64 > 0 ? c->fuse64 : c->fuse;
    const int rows = c->fuse64 > 0 ? c->rows64 : c->pair_rows;
    PTBuffers b{};
    rc = ctx_pt_buffers(c, fuse != 1, &b);
    if (rc) return rc;
    const SolveScratch w{c->stream, b,       c->txx,    c->tyy,       c->txy, c->divv, c->rx,
                         c->ry,     c->part, c->scal,   c->hscal,     c->ev0, c->ev1,  c->seg,
                         c->depth,  c->max_waves, fuse, rows,         &c->timer};
    return solve_core(w, "sf_solve_stokes", P, Vx_pad, Vy_pad, rho_vy, mu_b, mu_p, txx, tyy, txy, divV, dtauP, Rx,
                      Ry, dVx, dVy, dtauVx, dtauVy, Vx_s, Vy_s, g_y, dx, dy, nx, ny, dt, maxdisp, use_fss, eps,
                      iter_max, print_info, out);
}

extern "C" int sf_pt_iterations(sf_ctx* c, int64_t n_iter, double* P, double* Vx_s, double* Vy_s, double* dVx,
                                double* dVy, const double* rho_vy, const double* mu_b, const double* mu_p, double g_y,
                                double dx, double dy, int nx, int ny, double dt)
{
    int rc = check_ctx(c, nx, ny, "sf_pt_iterations");
    if (rc) return rc;
    if (!P || !Vx_s || !Vy_s || !dVx || !dVy || !rho_vy || !mu_b || !mu_p)
        return set_err(SF_ERR_ARG, "sf_pt_iterations: NULL array");
    if (n_iter < 0) return set_err(SF_ERR_ARG, "sf_pt_iterations: n_iter < 0");
    hipStream_t s = c->stream;
    const SolveConsts k = solve_consts(nx, ny, dx, dy);
    HIPCHK(launch_set_scalar(c->scal + SC_DT, dt, s));
    PTParams pp{};
    pp.nx = nx;
    pp.ny = ny;
    pp.seg = c->seg;
    pp.depth = c->depth;
    pp.max_waves = c->max_waves;
    pp.flags = 0;
    pp.dvx = dVx;
    pp.dvy = dVy;
    pp.rho = rho_vy;
    pp.mu_b = mu_b;
    pp.mu_p = mu_p;
    pp.dt = c->scal + SC_DT;
    pp.txx = c->txx;
    pp.tyy = c->tyy;
    pp.txy = c->txy;
    pp.divv = c->divv;
    pp.rx = c->rx;
    pp.g = g_y;
    pp.dampx = k.dampx;
    pp.dampy = k.dampy;
    pp.rdx = k.rdx;
    pp.rdy = k.rdy;
    pp.vsc_mdxy2 = k.vsc_mdxy2;
    pp.ptsc41 = k.ptsc41;
    pp.max_nxy = k.max_nxy;
    PTLoop L;
    rc = ctx_pt_loop(c, &L, Vx_s, Vy_s, P, dVx, dVy);
    if (rc) return rc;
    L.pp = pp;
    return L.run_fixed(n_iter, s);
}

// ---------------------------------------------------------------------------------------
// autotuner of the PT kernel's launch shape: times each candidate on a copy of the caller's
// problem, or a synthetic one of the context's size (results do not depend on the shape, but
// timings depend on the data)

extern "C" int sf_ctx_autotune_pt(sf_ctx* c, int iters, const double* P_in, const double* Vx_in, const double* Vy_in,
                                  const double* dVx_in, const double* dVy_in, const double* rho_in,
                                  const double* mub_in, const double* mup_in, double g_y, double dx_in, double dy_in,
                                  sf_pt_tuning* best, sf_pt_tuning* table, int table_cap, int* n_table)
{
    const char* who = "sf_ctx_autotune_pt";
    if (!c) return set_err(SF_ERR_ARG, "%s: NULL ctx", who);
    if (c->nx < 3 || c->ny < 3) return set_err(SF_ERR_ARG, "%s: marker-only context", who);
    if (iters < 1) iters = 20;
    const int nx = c->nx, ny = c->ny;
    const size_t nvx = (size_t)nx * (ny + 1), nvy = (size_t)(nx + 1) * ny, np = (size_t)(nx - 1) * (ny - 1);
    const size_t ndx = (size_t)(nx - 2) * (ny - 1), ndy = (size_t)(nx - 1) * (ny - 2);
    double *P = nullptr, *vx = nullptr, *vy = nullptr, *dvx = nullptr, *dvy = nullptr, *rho = nullptr, *mub = nullptr,
           *mup = nullptr;
    hipError_t e = dalloc(&P, np);
    if (e == hipSuccess) e = dalloc(&vx, nvx);
    if (e == hipSuccess) e = dalloc(&vy, nvy);
    if (e == hipSuccess) e = dalloc(&dvx, ndx);
    if (e == hipSuccess) e = dalloc(&dvy, ndy);
    if (e == hipSuccess) e = dalloc(&rho, nvy);
    if (e == hipSuccess) e = dalloc(&mub, (size_t)nx * ny);
    if (e == hipSuccess) e = dalloc(&mup, np);
    auto release = [&]() {
        for (double* p : {P, vx, vy, dvx, dvy, rho, mub, mup})
            if (p) (void)hipFree(p);
    };
    if (e != hipSuccess) {
        release();
        return set_err(SF_ERR_HIP, "%s: %s", who, hipGetErrorString(e));
    }
    hipStream_t s = c->stream;
    const double one = 1.0;
    // the caller's problem (copied: its arrays are not touched), or a synthetic one
    const bool given = P_in && Vx_in && Vy_in && dVx_in && dVy_in && rho_in && mub_in && mup_in;
    const double gy = given ? g_y : 1.0;
    const double dx = given ? dx_in : 1.0 / (nx - 1), dy = given ? dy_in : 1.0 / (ny - 1);
    if (given) {
        const std::pair<double*, const double*> cp[] = {{P, P_in},   {vx, Vx_in},   {vy, Vy_in},   {dvx, dVx_in},
                                                        {dvy, dVy_in}, {rho, rho_in}, {mub, mub_in}, {mup, mup_in}};
        const size_t len[] = {np, nvx, nvy, ndx, ndy, nvy, (size_t)nx * ny, np};
        for (int q = 0; q < 8 && e == hipSuccess; ++q)
            e = hipMemcpyAsync(cp[q].first, cp[q].second, len[q] * sizeof(double), hipMemcpyDeviceToDevice, s);
    } else {
        e = launch_fill_f64(rho, (int64_t)nvy, one, s);
        if (e == hipSuccess) e = launch_fill_f64(mub, (int64_t)nx * ny, one, s);
        if (e == hipSuccess) e = launch_fill_f64(mup, (int64_t)np, one, s);
        for (double* p : {P, vx, vy, dvx, dvy}) {
            const size_t n = p == P ? np : p == vx ? nvx : p == vy ? nvy : p == dvx ? ndx : ndy;
            if (e == hipSuccess) e = launch_fill_f64(p, (int64_t)n, 0.0, s);
        }
    }
    if (e != hipSuccess) {
        release();
        return set_err(SF_ERR_HIP, "%s: %s", who, hipGetErrorString(e));
    }
    const int saved_seg = c->seg, saved_depth = c->depth, saved_waves = c->max_waves, saved_fuse64 = c->fuse64,
              saved_rows64 = c->rows64;
    int nt = 0;
    // times `n` iterations of the context's current shape (after a warm-up of `warm`)
    auto time_shape = [&](int warm, int n, double* us) -> int {
        int r = sf_pt_iterations(c, warm, P, vx, vy, dvx, dvy, rho, mub, mup, gy, dx, dy, nx, ny, 0.0);
        if (r) return r;
        if (hipEventRecord(c->ev0, s) != hipSuccess) return set_err(SF_ERR_HIP, "%s: event record failed", who);
        r = sf_pt_iterations(c, n, P, vx, vy, dvx, dvy, rho, mub, mup, gy, dx, dy, nx, ny, 0.0);
        if (r) return r;
        float ms = 0.f;
        if (hipEventRecord(c->ev1, s) != hipSuccess || hipEventSynchronize(c->ev1) != hipSuccess ||
            hipEventElapsedTime(&ms, c->ev0, c->ev1) != hipSuccess)
            return set_err(SF_ERR_HIP, "%s: event timing failed", who);
        *us = 1e3 * (double)ms / n;
        return SF_OK;
    };
    auto record = [&](const sf_pt_tuning& t) {
        if (table && nt < table_cap) table[nt] = t;
        ++nt;
    };
    sf_pt_tuning bestv{}, best_all{};
    bestv.us_per_iter = 1e300;
    // first the shapes of the one-iteration kernel; every loop ends at the first error
    const int rc = [&]() -> int {
        c->fuse64 = 1;
        for (int rows : {1, 2, 4, 8, 16})
            for (int depth : {1, 2})
                for (int waves : {-1, 0}) {
                    c->seg = rows;
                    c->depth = depth;
                    c->max_waves = waves;
                    sf_pt_tuning t{rows, depth, waves, 1, 0.0};
                    const int r = time_shape(3, iters, &t.us_per_iter);
                    if (r) return r;
                    record(t);
                    if (t.us_per_iter < bestv.us_per_iter) bestv = t;
                }
        // then passes of 2 to 4 iterations (the check iterations keep the best one-iteration shape)
        best_all = bestv;
        c->seg = bestv.rows;
        c->depth = bestv.depth;
        c->max_waves = bestv.max_waves;
        // fixed tile heights plus each pass kind's size default (the wave-round model's pick)
        const int n_it = std::max(12, (iters + 11) / 12 * 12);  // whole passes of every kind
        for (int k : {2, 3, 4}) {
            std::vector<int> prows_c = {2, 4, 8, 12, 16, 24, 32, 48, 64};
            const int dflt = k == 2 ? pt2_default_rows(nx, ny) : ptk_default_rows(k, nx, ny);
            if (std::find(prows_c.begin(), prows_c.end(), dflt) == prows_c.end()) prows_c.push_back(dflt);
            for (int rows : prows_c) {
                if (rows > ny && rows != prows_c[0]) continue;
                c->fuse64 = k;
                c->rows64 = rows;
                sf_pt_tuning t{rows, 1, -1, k, 0.0};
                const int r = time_shape(k, n_it, &t.us_per_iter);
                if (r) return r;
                record(t);
                if (t.us_per_iter < best_all.us_per_iter) best_all = t;
            }
        }
        return SF_OK;
    }();
    (void)hipStreamSynchronize(s);
    release();
    if (rc) {
        c->seg = saved_seg;
        c->depth = saved_depth;
        c->max_waves = saved_waves;
        c->fuse64 = saved_fuse64;
        c->rows64 = saved_rows64;
        return rc;
    }
    c->seg = bestv.rows;
    c->depth = bestv.depth;
    c->max_waves = bestv.max_waves;
    // the FP64 passes only: the FP32 sweeps of the refined solver keep their own defaults
    c->fuse64 = best_all.iters_per_launch;
    c->rows64 = best_all.iters_per_launch > 1 ? best_all.rows : 0;
    if (best) *best = best_all;
    if (n_table) *n_table = nt;
    return SF_OK;
}

// ---------------------------------------------------------------------------------------
// DRAM probe

extern "C" int sf_probe_dram(sf_ctx* c, int64_t bytes, double* copy_gbs, double* r3w1_gbs)
{
    if (!c) return set_err(SF_ERR_ARG, "sf_probe_dram: NULL ctx");
    if (bytes < ((int64_t)1 << 24) || bytes % 64 != 0)
        return set_err(SF_ERR_ARG, "sf_probe_dram: bytes must be a multiple of 64 and >= 16 MiB");
    double* buf[4] = {nullptr, nullptr, nullptr, nullptr};
    hipError_t e = hipSuccess;
    for (int k = 0; k < 4 && e == hipSuccess; ++k) e = dalloc(&buf[k], (size_t)bytes / sizeof(double));
    hipStream_t s = c->stream;
    for (int k = 0; k < 4 && e == hipSuccess; ++k) e = launch_fill_f64(buf[k], bytes / 8, 1.0, s);
    float best_c = 1e30f, best_t = 1e30f;
    for (int grid : {16384, 65536}) {
        for (int rep = 0; rep < 4 && e == hipSuccess; ++rep) {
            float ms = 0.f;
            e = hipEventRecord(c->ev0, s);
            if (e == hipSuccess) e = launch_probe_copy(buf[0], buf[1], bytes, grid, s);
            if (e == hipSuccess) e = hipEventRecord(c->ev1, s);
            if (e == hipSuccess) e = hipEventSynchronize(c->ev1);
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->ev0, c->ev1);
            if (e == hipSuccess && ms < best_c) best_c = ms;
            if (e == hipSuccess) e = hipEventRecord(c->ev0, s);
            if (e == hipSuccess) e = launch_probe_r3w1(buf[0], buf[1], buf[2], buf[3], bytes, grid, s);
            if (e == hipSuccess) e = hipEventRecord(c->ev1, s);
            if (e == hipSuccess) e = hipEventSynchronize(c->ev1);
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->ev0, c->ev1);
            if (e == hipSuccess && ms < best_t) best_t = ms;
        }
    }
    (void)hipStreamSynchronize(s);
    for (double* p : buf)
        if (p) (void)hipFree(p);
    if (e != hipSuccess) return set_err(SF_ERR_HIP, "sf_probe_dram: %s", hipGetErrorString(e));
    if (copy_gbs) *copy_gbs = 2.0 * (double)bytes / (1e6 * (double)best_c);
    if (r3w1_gbs) *r3w1_gbs = 4.0 * (double)bytes / (1e6 * (double)best_t);
    return SF_OK;
}

// ---------------------------------------------------------------------------------------
// marker -> grid

void sf::m2g_free(M2GScratch& S)
{
    for (void* p : {(void*)S.key, (void*)S.rec_idx, (void*)S.rec, (void*)S.count,
                    (void*)S.offset, (void*)S.fill, (void*)S.blocksums, (void*)S.counters})
        if (p) (void)hipFree(p);
    S = M2GScratch();
}

int sf::m2g_reserve(M2GScratch& S, int64_t nm, int64_t ncells)
{
    if (nm > S.cap_markers) {
        for (void* p : {(void*)S.key, (void*)S.rec_idx, (void*)S.rec})
            if (p) HIPCHK(hipFree(p));
        S.key = S.rec_idx = nullptr;
        S.rec = nullptr;
        S.cap_markers = 0;
        const int64_t cap = nm + nm / 8;
        HIPCHK(dalloc(&S.key, (size_t)cap));
        HIPCHK(dalloc(&S.rec_idx, (size_t)cap));
        HIPCHK(dalloc(&S.rec, (size_t)cap));
        S.cap_markers = cap;
    }
    if (ncells > S.cap_cells) {
        for (void* p : {(void*)S.count, (void*)S.offset, (void*)S.fill, (void*)S.blocksums})
            if (p) HIPCHK(hipFree(p));
        S.count = S.offset = S.fill = S.blocksums = nullptr;
        S.cap_cells = 0;
        HIPCHK(dalloc(&S.count, (size_t)ncells));
        HIPCHK(dalloc(&S.offset, (size_t)ncells + 1));
        HIPCHK(dalloc(&S.fill, (size_t)ncells));
        HIPCHK(dalloc(&S.blocksums, (size_t)(ncells / 2048 + 2)));
        S.cap_cells = ncells;
    }
    return SF_OK;
}

static int m2g_call(sf_ctx* c, double xmin, double ymin, double* val_grid, int gnx, int gny, const double* x_m,
                    const double* y_m, const double* val_m, int64_t nm, double dx, double dy, double* vws, double* ws,
                    int ws_nx, int ws_ny, sf_m2g_out* out)
{
    if (!val_grid || (nm > 0 && (!x_m || !y_m || !val_m)))
        return set_err(SF_ERR_ARG, "sf_marker_to_grid: NULL array");
    if (gnx < 1 || gny < 1) return set_err(SF_ERR_ARG, "sf_marker_to_grid: empty grid %d x %d", gnx, gny);
    if (nm < 0) return set_err(SF_ERR_ARG, "sf_marker_to_grid: nm < 0");
    if ((vws == nullptr) != (ws == nullptr))
        return set_err(SF_ERR_ARG, "sf_marker_to_grid: pass both helper arrays or neither");
    if (ws && (ws_nx < gnx || ws_ny < gny))
        return set_err(SF_ERR_ARG, "sf_marker_to_grid: helper arrays %d x %d smaller than grid %d x %d", ws_nx, ws_ny,
                       gnx, gny);
    if (!(dx > 0.0) || !(dy > 0.0)) return set_err(SF_ERR_ARG, "sf_marker_to_grid: dx, dy must be > 0");
    const int64_t ncells = (int64_t)(gnx + 1) * (gny + 1);
    if (nm >= ((int64_t)1 << 31) - 1 || ncells >= ((int64_t)1 << 31) - 1)
        return set_err(SF_ERR_SIZE, "sf_marker_to_grid: %lld markers / %lld cells exceed the 32-bit index limit",
                       (long long)nm, (long long)ncells);
    int rc = m2g_reserve(c->m2g, nm, ncells);
    if (rc) return rc;
    HIPCHK(m2g_run(c->m2g, xmin, ymin, val_grid, gnx, gny, x_m, y_m, val_m, nm, dx, dy, vws, ws, ws_nx, ws_ny,
                   c->stream));
    HIPCHK(hipMemcpyAsync(c->hcounters, c->m2g.counters, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->hcounters[2] > 0)
        return set_err(SF_ERR_NONFINITE, "sf_marker_to_grid: %llu markers have non-finite coordinates",
                       (unsigned long long)c->hcounters[2]);
    if (out) {
        out->n_bad_markers = (int64_t)c->hcounters[0];
        out->n_empty_nodes = (int64_t)c->hcounters[1];
    }
    return SF_OK;
}

extern "C" int sf_marker_to_grid(sf_ctx* c, double xmin, double ymin, double* val_grid, int gnx, int gny,
                                 const double* x_m, const double* y_m, const double* val_m, int64_t nm, double dx,
                                 double dy, double* val_wt_sum, double* wt_sum, int ws_nx, int ws_ny, sf_m2g_out* out)
{
    if (!c) return set_err(SF_ERR_ARG, "sf_marker_to_grid: NULL ctx");
    return m2g_call(c, xmin, ymin, val_grid, gnx, gny, x_m, y_m, val_m, nm, dx, dy, val_wt_sum, wt_sum, ws_nx, ws_ny,
                    out);
}

extern "C" int sf_marker_cell_keys(sf_ctx* c, double xmin, double ymin, int gnx, int gny, const double* x_m,
                                   const double* y_m, int64_t nm, double dx, double dy, int32_t* keys,
                                   int32_t* counts, sf_m2g_out* out)
{
    if (!c) return set_err(SF_ERR_ARG, "sf_marker_cell_keys: NULL ctx");
    if (nm < 0) return set_err(SF_ERR_ARG, "sf_marker_cell_keys: nm < 0");
    if (!counts || (nm > 0 && (!x_m || !y_m || !keys))) return set_err(SF_ERR_ARG, "sf_marker_cell_keys: NULL array");
    if (gnx < 1 || gny < 1) return set_err(SF_ERR_ARG, "sf_marker_cell_keys: empty grid %d x %d", gnx, gny);
    if (!(dx > 0.0) || !(dy > 0.0)) return set_err(SF_ERR_ARG, "sf_marker_cell_keys: dx, dy must be > 0");
    const int64_t ncells = (int64_t)(gnx + 1) * (gny + 1);
    if (nm >= ((int64_t)1 << 31) - 1 || ncells >= ((int64_t)1 << 31) - 1)
        return set_err(SF_ERR_SIZE, "sf_marker_cell_keys: %lld markers / %lld cells exceed the 32-bit index limit",
                       (long long)nm, (long long)ncells);
    HIPCHK(m2g_cell_keys(c->m2g, xmin, ymin, gnx, gny, x_m, y_m, nm, dx, dy, keys, counts, c->stream));
    HIPCHK(hipMemcpyAsync(c->hcounters, c->m2g.counters, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->hcounters[2] > 0)
        return set_err(SF_ERR_NONFINITE, "sf_marker_cell_keys: %llu markers have non-finite coordinates",
                       (unsigned long long)c->hcounters[2]);
    if (out) {
        out->n_bad_markers = (int64_t)c->hcounters[0];
        out->n_empty_nodes = 0;
    }
    return SF_OK;
}

// ---------------------------------------------------------------------------------------
// RK4

extern "C" int sf_move_markers_rk4(sf_ctx* c, double* x_m, double* y_m, int64_t nm, const double* Vx, int vx_n1,
                                   int vx_n2, const double* Vy, int vy_n1, int vy_n2, double x_vx_min, double y_vx_min,
                                   double x_vy_min, double y_vy_min, double dt, double lx, double ly, double dx,
                                   double dy, sf_rk4_out* out)
{
    if (!c) return set_err(SF_ERR_ARG, "sf_move_markers_rk4: NULL ctx");
    if (nm < 0) return set_err(SF_ERR_ARG, "sf_move_markers_rk4: nm < 0");
    if (nm > 0 && (!x_m || !y_m || !Vx || !Vy)) return set_err(SF_ERR_ARG, "sf_move_markers_rk4: NULL array");
    // the failsafe clamps keep every access inside arrays of at least 3 x 2 / 2 x 3 (+ stencil)
    if (vx_n1 < 4 || vx_n2 < 2 || vy_n1 < 2 || vy_n2 < 4)
        return set_err(SF_ERR_ARG, "sf_move_markers_rk4: velocity arrays too small (%d x %d, %d x %d)", vx_n1, vx_n2,
                       vy_n1, vy_n2);
    if (!(dx > 0.0) || !(dy > 0.0)) return set_err(SF_ERR_ARG, "sf_move_markers_rk4: dx, dy must be > 0");
    hipStream_t s = c->stream;
    HIPCHK(launch_zero_u64(c->counters, 4, s));
    HIPCHK(launch_rk4(x_m, y_m, nm, Vx, vx_n1, vx_n2, Vy, vy_n1, vy_n2, x_vx_min, y_vx_min, x_vy_min, y_vy_min, dt, lx,
                      ly, dx, dy, c->counters, s));
    HIPCHK(hipMemcpyAsync(c->hcounters, c->counters, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (c->hcounters[1] > 0)
        return set_err(SF_ERR_NONFINITE, "sf_move_markers_rk4: %llu markers have non-finite positions",
                       (unsigned long long)c->hcounters[1]);
    if (out) out->n_clamped = (int64_t)c->hcounters[0];
    return SF_OK;
}

// ---------------------------------------------------------------------------------------
// timestep (StokesFlow2D_multixpu.jl:164-181 at one rank)

extern "C" int sf_markers_to_grids3(sf_ctx* c, double* rho_vy, double* mu_b, double* mu_p, int nx, int ny,
                                    const double* x_m, const double* y_m, const double* rho_m, const double* mu_m,
                                    int64_t nm, double dx, double dy, sf_m2g_out out[3])
{
    if (!c) return set_err(SF_ERR_ARG, "sf_markers_to_grids3: NULL ctx");
    if (nx < 3 || ny < 3) return set_err(SF_ERR_ARG, "sf_markers_to_grids3: need Nx, Ny >= 3");
    // grid origins, StokesFlow2D_multixpu.jl:117-126
    const double x_rho = 0.0 * dx - 0.5 * dx, y_rho = 0.0 * dy;
    const double x_p = 0.0 * dx + 0.5 * dx, y_p = 0.0 * dy + 0.5 * dy;
    sf_m2g_out mo[3] = {};
    int rc = m2g_call(c, x_rho, y_rho, rho_vy, nx + 1, ny, x_m, y_m, rho_m, nm, dx, dy, nullptr, nullptr, 0, 0, &mo[0]);
    if (rc) return rc;
    rc = m2g_call(c, 0.0, 0.0, mu_b, nx, ny, x_m, y_m, mu_m, nm, dx, dy, nullptr, nullptr, 0, 0, &mo[1]);
    if (rc) return rc;
    rc = m2g_call(c, x_p, y_p, mu_p, nx - 1, ny - 1, x_m, y_m, mu_m, nm, dx, dy, nullptr, nullptr, 0, 0, &mo[2]);
    if (rc) return rc;
    if (out) std::memcpy(out, mo, sizeof(mo));
    return SF_OK;
}

extern "C" int sf_step(sf_ctx* c, sf_step_args* a, sf_solve_out* so, sf_m2g_out m2g_out[3], sf_rk4_out* ro)
{
    if (!c || !a) return set_err(SF_ERR_ARG, "sf_step: NULL argument");
    int rc = check_ctx(c, a->nx, a->ny, "sf_step");
    if (rc) return rc;
    const int nx = a->nx, ny = a->ny;
    const double dx = a->dx, dy = a->dy;
    const double lx = (double)(nx - 1) * dx, ly = (double)(ny - 1) * dy;
    // velocity grid origins, StokesFlow2D_multixpu.jl:121-124
    const double x_vx = -1.0 * dx, y_vx = 0.0 * dy - 0.5 * dy;
    const double x_vy = 0.0 * dx - 0.5 * dx, y_vy = -1.0 * dy;
    sf_m2g_out mo[3] = {};
    rc = sf_markers_to_grids3(c, a->rho_vy, a->mu_b, a->mu_p, nx, ny, a->x_m, a->y_m, a->rho_m, a->mu_m, a->nm, dx, dy,
                              mo);
    if (rc) return rc;
    sf_solve_out sout{};
    rc = sf_solve_stokes(c, a->P, a->Vx_pad, a->Vy_pad, a->rho_vy, a->mu_b, a->mu_p, nullptr, nullptr, nullptr,
                         nullptr, nullptr, nullptr, nullptr, a->dVxdtau, a->dVydtau, nullptr, nullptr, a->Vx_s,
                         a->Vy_s, a->g_y, dx, dy, nx, ny, a->dt, a->maxdisp, a->use_fss, a->eps, a->iter_max, 0, &sout);
    if (rc) return rc;
    a->dt = sout.dt_new;
    sf_rk4_out rout{};
    rc = sf_move_markers_rk4(c, a->x_m, a->y_m, a->nm, a->Vx_pad, nx + 2, ny + 1, a->Vy_pad, nx + 1, ny + 2, x_vx, y_vx,
                             x_vy, y_vy, a->dt, lx, ly, dx, dy, &rout);
    if (rc) return rc;
    if (so) *so = sout;
    if (m2g_out) std::memcpy(m2g_out, mo, sizeof(mo));
    if (ro) *ro = rout;
    return SF_OK;
}

// ---------------------------------------------------------------------------------------
// marker set-up

extern "C" int sf_init_markers(sf_ctx* c, double* x_m, double* y_m, int nx, int ny, int md, double dx, double dy,
                               double jitter_frac, uint64_t seed)
{
    if (!c || !x_m || !y_m) return set_err(SF_ERR_ARG, "sf_init_markers: NULL argument");
    if (nx < 2 || ny < 2 || md < 1) return set_err(SF_ERR_ARG, "sf_init_markers: bad sizes");
    HIPCHK(launch_init_markers(x_m, y_m, nx, ny, md, dx, dy, jitter_frac, seed, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return SF_OK;
}

extern "C" int sf_plume_properties(sf_ctx* c, const double* x_m, const double* y_m, int64_t nm, const double mat[10],
                                   double* rho_m, double* mu_m)
{
    if (!c || !mat || (nm > 0 && (!x_m || !y_m || !rho_m || !mu_m)))
        return set_err(SF_ERR_ARG, "sf_plume_properties: NULL argument");
    Mat10 m{};
    std::memcpy(m.v, mat, 10 * sizeof(double));
    HIPCHK(launch_plume_properties(x_m, y_m, nm, m, rho_m, mu_m, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return SF_OK;
}

extern "C" int sf_lattice_markers(sf_ctx* c, double* x_m, double* y_m, int nx, int ny, int md, double lx, double ly,
                                  double dx, double dy, int rand_pos, uint64_t seed)
{
    if (!c || !x_m || !y_m) return set_err(SF_ERR_ARG, "sf_lattice_markers: NULL argument");
    if (nx < 2 || ny < 2 || md < 1) return set_err(SF_ERR_ARG, "sf_lattice_markers: bad sizes");
    if (!(dx > 0.0) || !(dy > 0.0)) return set_err(SF_ERR_ARG, "sf_lattice_markers: dx, dy must be > 0");
    HIPCHK(launch_lattice_markers(x_m, y_m, nx, ny, md, lx, ly, dx, dy, rand_pos, seed, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return SF_OK;
}

extern "C" int sf_marker_properties(sf_ctx* c, const double* x_m, const double* y_m, int64_t nm, double x0, double y0,
                                    const sf_shape* shapes, int n_shapes, double rho_bg, double mu_bg, double* rho_m,
                                    double* mu_m)
{
    if (!c || (nm > 0 && (!x_m || !y_m || !rho_m || !mu_m)) || (n_shapes > 0 && !shapes))
        return set_err(SF_ERR_ARG, "sf_marker_properties: NULL argument");
    if (n_shapes < 0 || n_shapes > SF_MAX_SHAPES)
        return set_err(SF_ERR_ARG, "sf_marker_properties: %d shapes (0..%d)", n_shapes, SF_MAX_SHAPES);
    ShapeSet set{};
    set.n = n_shapes;
    set.rho_bg = rho_bg;
    set.mu_bg = mu_bg;
    for (int k = 0; k < n_shapes; ++k) {
        if (shapes[k].kind < SF_SHAPE_Y_BELOW || shapes[k].kind > SF_SHAPE_SINE_BELOW)
            return set_err(SF_ERR_ARG, "sf_marker_properties: shape %d has unknown kind %d", k, shapes[k].kind);
        set.s[k] = shapes[k];
    }
    HIPCHK(launch_shape_properties(x_m, y_m, nm, x0, y0, set, rho_m, mu_m, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return SF_OK;
}

extern "C" int sf_rt_properties(sf_ctx* c, const double* x_m, const double* y_m, int64_t nm, const double mat[8],
                                double* rho_m, double* mu_m)
{
    if (!c || !mat || (nm > 0 && (!x_m || !y_m || !rho_m || !mu_m)))
        return set_err(SF_ERR_ARG, "sf_rt_properties: NULL argument");
    Mat10 m{};
    std::memcpy(m.v, mat, 8 * sizeof(double));
    HIPCHK(launch_rt_properties(x_m, y_m, nm, m, rho_m, mu_m, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return SF_OK;
}

// ---------------------------------------------------------------------------------------
// FP32 PT sweeps inside FP64 defect correction

// Allocates the refined solver's scratch (all but the second momentum) once per context.
// hpart is allocated last and is the "reserved" mark: after any failed allocation every
// buffer is freed and reset, so the next call starts from a consistent (empty) context.
static int refine_reserve(sf_ctx* c)
{
    if (c->hpart) return SF_OK;
    const size_t nx = c->nx, ny = c->ny;
    const size_t nvx = nx * (ny + 1), nvy = (nx + 1) * ny, np = (nx - 1) * (ny - 1);
    const size_t nmx = (nx - 2) * (ny - 1), nmy = (nx - 1) * (ny - 2);
    const std::pair<float**, size_t> bufs[] = {
        {&c->r_vx[0], nvx}, {&c->r_vx[1], nvx}, {&c->r_vy[0], nvy},        {&c->r_vy[1], nvy}, {&c->r_p[0], np},
        {&c->r_p[1], np},   {&c->r_mx, nmx},    {&c->r_my, nmy},           {&c->r_rho, (nx + 1) * ny},
        {&c->r_mub, nx * ny}, {&c->r_mup, np},  {&c->r_fx, nmx},           {&c->r_fy, nmy},    {&c->r_fc, np},
        {&c->r_rxo, nmx},   {&c->r_ryo, nmy},   {&c->r_rco, np}};
    hipError_t e = hipSuccess;
    for (const auto& b : bufs)
        if (e == hipSuccess && !*b.first) e = dalloc(b.first, b.second);
    if (e == hipSuccess)
        e = hipHostMalloc((void**)&c->hpart, 3 * (size_t)RED_MAX_BLOCKS * sizeof(double), hipHostMallocDefault);
    if (e != hipSuccess) {
        for (const auto& b : bufs) {
            if (*b.first) (void)hipFree(*b.first);
            *b.first = nullptr;
        }
        c->hpart = nullptr;
        return set_err(SF_ERR_HIP, "sf_solve_stokes_refined: scratch allocation failed: %s", hipGetErrorString(e));
    }
    return SF_OK;
}

// The second PT momentum of the multi-sweep passes: both buffers or neither.
static int refine_reserve_m2(sf_ctx* c)
{
    if (c->r_mx2 && c->r_my2) return SF_OK;
    hipError_t e = c->r_mx2 ? hipSuccess : dalloc(&c->r_mx2, (size_t)(c->nx - 2) * (c->ny - 1));
    if (e == hipSuccess && !c->r_my2) e = dalloc(&c->r_my2, (size_t)(c->nx - 1) * (c->ny - 2));
    if (e != hipSuccess) {
        for (float** b : {&c->r_mx2, &c->r_my2}) {
            if (*b) (void)hipFree(*b);
            *b = nullptr;
        }
        return set_err(SF_ERR_HIP, "sf_solve_stokes_refined: momentum buffer allocation failed: %s",
                       hipGetErrorString(e));
    }
    return SF_OK;
}

// One FP64 PT iteration in check mode from the caller's arrays (state copied back), giving
// err exactly as compute_err does at a check iteration (StokesSolver_multixpu.jl:64-73).
static int fp64_check_iteration(sf_ctx* c, PTParams pp, double* P, double* Vx_s, double* Vy_s, const double* rho_vy,
                                double* r_y, double g_y, double dx, double dy, double maxdisp, int use_fss,
                                double rdx, double rdy, double* err_out)
{
    const int nx = c->nx, ny = c->ny;
    hipStream_t s = c->stream;
    pp.flags = PT_WRITE_TAU | PT_WRITE_RX;
    pp.vx_a = Vx_s;
    pp.vy_a = Vy_s;
    pp.p_a = P;
    pp.vx_b = c->vx_b;
    pp.vy_b = c->vy_b;
    pp.p_b = c->p_b;
    HIPCHK(launch_pt_fused(pp, s));
    if (use_fss) {
        int rc = enqueue_dt(c, c->vx_b, c->vy_b, dx, dy, maxdisp, 0);
        if (rc) return rc;
    } else {
        HIPCHK(launch_set_scalar(c->scal + SC_DT_CHECK, 0.0, s));
    }
    ResYParams rq{};
    rq.nx = nx;
    rq.ny = ny;
    rq.tyy = pp.tyy;
    rq.txy = pp.txy;
    rq.P = c->p_b;
    rq.rho = rho_vy;
    rq.vx = c->vx_b;
    rq.vy = c->vy_b;
    rq.dt = c->scal + SC_DT_CHECK;
    rq.ry = r_y;
    rq.g = g_y;
    rq.rdx = rdx;
    rq.rdy = rdy;
    HIPCHK(launch_pt_resid_y(rq, s));
    const double n_rx = (double)(nx - 2) * (ny - 1), n_ry = (double)(nx - 1) * (ny - 2), n_dv = (double)(nx - 1) * (ny - 1);
    RedParams rp{};
    rp.a[0] = pp.rx;
    rp.n[0] = (int64_t)n_rx;
    rp.a[1] = r_y;
    rp.n[1] = (int64_t)n_ry;
    rp.a[2] = pp.divv;
    rp.n[2] = (int64_t)n_dv;
    rp.part = c->part;
    rp.mode = RED_SUM;
    const int nb = red_num_blocks((int64_t)n_dv);
    HIPCHK(launch_reduce(rp, 3, nb, s));
    HIPCHK(launch_finalize_err(c->part, nb, n_rx, n_ry, n_dv, c->scal, s));
    HIPCHK(hipMemcpyAsync(Vx_s, c->vx_b, sizeof(double) * (size_t)nx * (ny + 1), hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(Vy_s, c->vy_b, sizeof(double) * (size_t)(nx + 1) * ny, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(P, c->p_b, sizeof(double) * (size_t)(nx - 1) * (ny - 1), hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(c->hscal, c->scal, SC_COUNT * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    *err_out = c->hscal[SC_ERR];
    return SF_OK;
}

extern "C" int sf_solve_stokes_refined(sf_ctx* c, double* P, double* Vx_pad, double* Vy_pad, const double* rho_vy,
                                       const double* mu_b, const double* mu_p, double* txx, double* tyy, double* txy,
                                       double* divV, double* dtauP, double* Rx, double* Ry, double* dVx, double* dVy,
                                       double* dtauVx, double* dtauVy, double* Vx_s, double* Vy_s, double g_y,
                                       double dx, double dy, int nx, int ny, double dt, double maxdisp, int use_fss,
                                       double eps, const sf_refine_params* prm, sf_solve_out* out,
                                       sf_refine_out* rout)
{
    int rc = check_ctx(c, nx, ny, "sf_solve_stokes_refined");
    if (rc) return rc;
    if (!P || !Vx_pad || !Vy_pad || !rho_vy || !mu_b || !mu_p || !dVx || !dVy || !Vx_s || !Vy_s)
        return set_err(SF_ERR_ARG, "sf_solve_stokes_refined: a state or material array is NULL");
    if (!(dx > 0.0) || !(dy > 0.0)) return set_err(SF_ERR_ARG, "sf_solve_stokes_refined: dx, dy must be > 0");
    const auto t_start = std::chrono::steady_clock::now();
    hipStream_t s = c->stream;
    const SolveConsts k = solve_consts(nx, ny, dx, dy);
    if (!use_fss) dt = 0.0;
    const int mn = std::min(nx, ny);
    const int64_t ncheck = std::min<int64_t>(5 * (int64_t)mn, 2000);
    const int64_t ndtupdate = 10 * (int64_t)mn;
    const int64_t rebase = (prm && prm->rebase_interval > 0) ? prm->rebase_interval : ncheck;
    const int64_t iter_max = (prm && prm->iter_max > 0) ? prm->iter_max : (int64_t)100000;
    const int verbose = prm ? prm->verbose : 0;

    rc = refine_reserve(c);
    if (rc) return rc;
    double* t_xx = txx ? txx : c->txx;
    double* t_yy = tyy ? tyy : c->tyy;
    double* t_xy = txy ? txy : c->txy;
    double* d_v = divV ? divV : c->divv;
    double* r_x = Rx ? Rx : c->rx;
    double* r_y = Ry ? Ry : c->ry;
    if (dtauP || dtauVx || dtauVy)
        HIPCHK(launch_fill_timesteps(nx, ny, mu_p, dtauVx, dtauVy, dtauP, k.vsc_mdxy2, k.ptsc41, k.max_nxy, s));

    const int64_t nvx = (int64_t)nx * (ny + 1), nvy = (int64_t)(nx + 1) * ny, np = (int64_t)(nx - 1) * (ny - 1);
    const int64_t nmx = (int64_t)(nx - 2) * (ny - 1), nmy = (int64_t)(nx - 1) * (ny - 2);
    HIPCHK(launch_cast_f32(c->r_rho, rho_vy, (int64_t)(nx + 1) * ny, s));
    HIPCHK(launch_cast_f32(c->r_mub, mu_b, (int64_t)nx * ny, s));
    HIPCHK(launch_cast_f32(c->r_mup, mu_p, np, s));
    // the PT momentum continues from the caller's dVx/dVy (warm start across timesteps)
    HIPCHK(launch_cast_f32(c->r_mx, dVx, nmx, s));
    HIPCHK(launch_cast_f32(c->r_my, dVy, nmy, s));

    PTParams pp{};
    pp.nx = nx;
    pp.ny = ny;
    pp.seg = c->seg;
    pp.depth = c->depth;
    pp.max_waves = c->max_waves;
    pp.dvx = dVx;
    pp.dvy = dVy;
    pp.rho = rho_vy;
    pp.mu_b = mu_b;
    pp.mu_p = mu_p;
    pp.dt = c->scal + SC_DT;
    pp.txx = t_xx;
    pp.tyy = t_yy;
    pp.txy = t_xy;
    pp.divv = d_v;
    pp.rx = r_x;
    pp.g = g_y;
    pp.dampx = k.dampx;
    pp.dampy = k.dampy;
    pp.rdx = k.rdx;
    pp.rdy = k.rdy;
    pp.vsc_mdxy2 = k.vsc_mdxy2;
    pp.ptsc41 = k.ptsc41;
    pp.max_nxy = k.max_nxy;

    CorrParams cp{};
    cp.nx = nx;
    cp.ny = ny;
    const int corr_seg = c->seg > 0 ? c->seg : 4;  // the context's tile rows, when set
    cp.depth = c->depth > 1 ? 2 : 1;
    // two sweeps per pass unless the context asks for one iteration per launch; the pass
    // writes the momentum to a second buffer, and the context's pointers are swapped after
    // each pass so that r_mx / r_my always hold the current momentum.  The setting is the
    // explicit one of sf_ctx_set_pt_fusion (the autotuner's FP64 choice does not apply here).
    const bool corr_pairs = c->fuse != 1;
    // sweeps per pass: the context's setting, or the size default of the FP32 sweeps
    const int corr_k = c->fuse == 0 ? corr_default_fusion(nx, ny) : c->fuse;
    if (corr_pairs) {
        rc = refine_reserve_m2(c);
        if (rc) return rc;
    }
    cp.rho = c->r_rho;
    cp.mu_b = c->r_mub;
    cp.mu_p = c->r_mup;
    cp.fx = c->r_fx;
    cp.fy = c->r_fy;
    cp.fc = c->r_fc;
    cp.rxo = c->r_rxo;
    cp.ryo = c->r_ryo;
    cp.rco = c->r_rco;
    cp.g = (float)g_y;
    cp.dampx = (float)k.dampx;
    cp.dampy = (float)k.dampy;
    cp.rdx = (float)k.rdx;
    cp.rdy = (float)k.rdy;
    cp.vsc_mdxy2 = (float)k.vsc_mdxy2;
    cp.ptsc41 = (float)k.ptsc41;
    cp.max_nxy = (float)k.max_nxy;

    const double n_rx = (double)nmx, n_ry = (double)nmy, n_dv = (double)np;
    const int nbr = red_num_blocks(np);
    // one FP32 pass of kk sweeps, bracketed by the context's pass timer when it is on
    auto timed_corr = [&](int kk, auto&& launch) -> int {
        if (c->timer.on) {
            const int r = c->timer.begin(s, 1, kk);
            if (r) return r;
        }
        HIPCHK(launch());
        return c->timer.on ? c->timer.end(s) : SF_OK;
    };
    double dt_k = dt;
    int64_t n32 = 0, n64 = 0;
    int rebases = 0;
    double err_ref = 2.0 * eps, err_k = 0.0;
    bool converged = false;
    int64_t checked_at = -1;
    while (n32 < iter_max) {
        // FP64 residual of the current state: the exact forcing of the next FP32 window
        ResidParams rp{};
        rp.nx = nx;
        rp.ny = ny;
        rp.vx = Vx_s;
        rp.vy = Vy_s;
        rp.P = P;
        rp.rho = rho_vy;
        rp.mu_b = mu_b;
        rp.mu_p = mu_p;
        rp.fx = c->r_fx;
        rp.fy = c->r_fy;
        rp.fc = c->r_fc;
        rp.part = c->part;
        rp.dt = dt_k;
        rp.g = g_y;
        rp.rdx = k.rdx;
        rp.rdy = k.rdy;
        HIPCHK(launch_resid_f64(rp, nbr, s));
        HIPCHK(hipMemcpyAsync(c->hpart, c->part, 3 * (size_t)nbr * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        double sx = 0.0, sy = 0.0, sc = 0.0;
        for (int b = 0; b < nbr; ++b) {
            sx += c->hpart[b];
            sy += c->hpart[nbr + b];
            sc += c->hpart[2 * nbr + b];
        }
        err_k = std::max(std::max(sx / n_rx, sy / n_ry), sc / n_dv);
        if (verbose > 0)
            fprintf(stderr, "FP32 iterations %lld: FP64 residual err %.3e, dt %.4e\n", (long long)n32, err_k, dt_k);
        if (!std::isfinite(err_k)) break;
        // convergence is tested where the reference tests it, at multiples of ncheck (:63)
        if (n32 > 0 && n32 % ncheck == 0 && n32 != checked_at && err_k <= eps) {
            checked_at = n32;
            // confirm with the reference's own check: one FP64 PT iteration and compute_err
            HIPCHK(launch_cast_f64(dVx, c->r_mx, nmx, s));
            HIPCHK(launch_cast_f64(dVy, c->r_my, nmy, s));
            HIPCHK(launch_set_scalar(c->scal + SC_DT, dt_k, s));
            rc = fp64_check_iteration(c, pp, P, Vx_s, Vy_s, rho_vy, r_y, g_y, dx, dy, maxdisp, use_fss, k.rdx, k.rdy,
                                      &err_ref);
            if (rc) return rc;
            n64 += 1;
            HIPCHK(launch_cast_f32(c->r_mx, dVx, nmx, s));
            HIPCHK(launch_cast_f32(c->r_my, dVy, nmy, s));
            if (verbose > 0) fprintf(stderr, "  FP64 check iteration: err %.3e\n", err_ref);
            if (err_ref <= eps) {
                converged = true;
                break;
            }
            continue;
        }
        // an FP32 window on the correction, the momentum carried over from the previous one
        for (int b = 0; b < 2; ++b) {
            HIPCHK(launch_fill_f32(c->r_vx[b], nvx, 0.f, s));
            HIPCHK(launch_fill_f32(c->r_vy[b], nvy, 0.f, s));
            HIPCHK(launch_fill_f32(c->r_p[b], np, 0.f, s));
        }
        cp.dt = (float)dt_k;
        int cur = 0;
        // windows end at every multiple of ncheck and (with stabilisation) of ndtupdate, so the
        // convergence test and the dt refresh happen at the reference's iteration numbers
        int64_t steps = std::min(rebase, iter_max - n32);
        steps = std::min(steps, ncheck - n32 % ncheck);
        if (use_fss) steps = std::min(steps, ndtupdate - n32 % ndtupdate);
        int64_t it = 0;
        for (; corr_k >= 3 && it + corr_k <= steps; it += corr_k) {
            cp.vx_a = c->r_vx[cur];
            cp.vy_a = c->r_vy[cur];
            cp.p_a = c->r_p[cur];
            cp.vx_b = c->r_vx[1 - cur];
            cp.vy_b = c->r_vy[1 - cur];
            cp.p_b = c->r_p[1 - cur];
            cp.mx = c->r_mx;
            cp.my = c->r_my;
            cp.mx_o = c->r_mx2;
            cp.my_o = c->r_my2;
            cp.seg = c->pair_rows;  // 0: size default
            rc = timed_corr(corr_k, [&] { return launch_pt_corrk(cp, corr_k, s); });
            if (rc) return rc;
            cur = 1 - cur;
            std::swap(c->r_mx, c->r_mx2);
            std::swap(c->r_my, c->r_my2);
        }
        for (; corr_pairs && it + 2 <= steps; it += 2) {
            cp.vx_a = c->r_vx[cur];
            cp.vy_a = c->r_vy[cur];
            cp.p_a = c->r_p[cur];
            cp.vx_b = c->r_vx[1 - cur];
            cp.vy_b = c->r_vy[1 - cur];
            cp.p_b = c->r_p[1 - cur];
            cp.mx = c->r_mx;
            cp.my = c->r_my;
            cp.mx_o = c->r_mx2;
            cp.my_o = c->r_my2;
            cp.seg = c->pair_rows;  // 0: size default
            rc = timed_corr(2, [&] { return launch_pt_corr2(cp, s); });
            if (rc) return rc;
            cur = 1 - cur;
            std::swap(c->r_mx, c->r_mx2);
            std::swap(c->r_my, c->r_my2);
        }
        for (; it < steps; ++it) {
            cp.vx_a = c->r_vx[cur];
            cp.vy_a = c->r_vy[cur];
            cp.p_a = c->r_p[cur];
            cp.vx_b = c->r_vx[1 - cur];
            cp.vy_b = c->r_vy[1 - cur];
            cp.p_b = c->r_p[1 - cur];
            cp.mx = c->r_mx;
            cp.my = c->r_my;
            cp.mx_o = nullptr;  // in place
            cp.my_o = nullptr;
            cp.seg = corr_seg;
            rc = timed_corr(1, [&] { return launch_pt_corr(cp, false, s); });
            if (rc) return rc;
            cur = 1 - cur;
        }
        const int64_t before = n32;
        n32 += steps;
        // fold the correction into the FP64 state
        HIPCHK(launch_update_f64(Vx_s, c->r_vx[cur], nvx, s));
        HIPCHK(launch_update_f64(Vy_s, c->r_vy[cur], nvy, s));
        HIPCHK(launch_update_f64(P, c->r_p[cur], np, s));
        rebases += 1;
        // the reference refreshes the stabilisation dt every ndtupdate iterations (:60-62)
        if (use_fss && (n32 / ndtupdate) > (before / ndtupdate)) {
            rc = enqueue_dt(c, Vx_s, Vy_s, dx, dy, maxdisp, 0);
            if (rc) return rc;
            HIPCHK(hipMemcpyAsync(c->hscal, c->scal, SC_COUNT * sizeof(double), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            dt_k = c->hscal[SC_DT_CHECK];
        }
    }
    // hand the PT momentum back in FP64
    if (!converged) {
        HIPCHK(launch_cast_f64(dVx, c->r_mx, nmx, s));
        HIPCHK(launch_cast_f64(dVy, c->r_my, nmy, s));
    }

    HIPCHK(launch_pad_mirror(nx, ny, Vx_s, Vy_s, Vx_pad, Vy_pad, s));
    rc = enqueue_dt(c, Vx_s, Vy_s, dx, dy, maxdisp, 0);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(c->hscal, c->scal, SC_COUNT * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    if (!converged && std::isfinite(err_k)) err_ref = err_k;
    if (out) {
        out->dt_new = c->hscal[SC_DT_CHECK];
        out->err = err_ref;
        out->t_eff = 0.0;
        out->runtime = secs;
        out->niter = n32 + n64;
        out->converged = converged;
        out->reserved = 0;
        out->dt_last = dt_k;
        out->t_it = 0.0;
        out->gbs_alg = 0.0;
    }
    if (rout) {
        rout->n_outer = rebases;
        rout->reserved = 0;
        rout->n_fp32_iter = n32;
        rout->n_fp64_iter = n64;
        rout->err_fp64 = err_ref;
        rout->seconds = secs;
    }
    return SF_OK;
}

// ---------------------------------------------------------------------------------------
// batched ensemble: many independent solves in one persistent launch

