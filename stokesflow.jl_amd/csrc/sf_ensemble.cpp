// sf_ensemble.cpp -- ensemble timesteps (sf_ensemble_*): the StokesFlow2D loop body for many
// independent problems, with the marker passes and RK4 captured in hipGraphs around one
// batched solve (or the whole step in one graph with the persistent batched kernel, or with one
// chip-wide grid-mode solve per member), and the captured timestep of one problem (sf_graph_*).
#include "sf_host.h"

using namespace sf;

struct sf_ensemble {
    sf_ctx* c = nullptr;
    hipStream_t stream = nullptr;  // own non-blocking stream: capture needs one
    hipEvent_t ev = nullptr;
    int n = 0, threads = 512, nblocks = 0, use_graph = 1, wide = 1, depth = 1;
    int grid = 0, gk = 3;  // grid mode: one chip-wide launch per member, gk PT iterations per pass
    double budget = 0.0;
    std::vector<sf_step_args> mem;
    std::vector<sf_batch_instance> bi;     // the members as batched-solver instances (dt updated per step)
    std::vector<BatchInstance> desc;       // host copies of the device descriptors
    std::vector<int> order;
    std::vector<double> dt;                // members' dt between steps (wide form)
    double* pool = nullptr;
    BatchInstance* inst = nullptr;
    int* order_dev = nullptr;
    int* queue = nullptr;
    BatchResult* res = nullptr;          // persistent form: res[i].dt_new is member i's dt between steps
    double* part = nullptr;              // wide form: reduction partials
    double* hscal = nullptr;             // wide form: pinned per-member scalars
    unsigned long long* cnt = nullptr;   // per member: 3 x 4 marker -> grid counters, 4 RK4 counters
    unsigned long long* hcnt = nullptr;  // pinned
    BatchResult* hres = nullptr;         // pinned
    unsigned long long* gsync = nullptr;  // grid mode: GRID_SYNC_WORDS barrier words per member
    unsigned long long* hsync = nullptr;  // pinned
    double* gpart = nullptr;              // grid mode: 3 * GRID_MAX_WAVES partials
    std::vector<GridShape> gshape;        // grid mode: each member's launch
    hipGraph_t graph[2] = {nullptr, nullptr};  // persistent: [0] whole step; wide: [0] marker -> grid, [1] RK4
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    M2GScratch m2g;  // own sort scratch: the captured graphs hold its addresses
    // wide form with FP32 sweeps inside FP64 defect correction
    int refined = 0;
    float* pool32 = nullptr;
    double* rpart = nullptr;
    CorrParams* cdesc_dev = nullptr;
    double* hpart = nullptr;
    float* hdt = nullptr;
    std::vector<RefineBufs> rb;
    std::vector<CorrParams> cdesc;
};

enum { ENS_CNT = 16 };

static void free_ensemble(sf_ensemble* e)
{
    if (!e) return;
    for (int k = 0; k < 2; ++k) {
        if (e->exec[k]) (void)hipGraphExecDestroy(e->exec[k]);
        if (e->graph[k]) (void)hipGraphDestroy(e->graph[k]);
    }
    for (void* p : {(void*)e->pool, (void*)e->inst, (void*)e->order_dev, (void*)e->queue, (void*)e->res,
                    (void*)e->cnt, (void*)e->part, (void*)e->pool32, (void*)e->rpart, (void*)e->cdesc_dev,
                    (void*)e->gsync, (void*)e->gpart})
        if (p) (void)hipFree(p);
    for (void* p : {(void*)e->hcnt, (void*)e->hres, (void*)e->hscal, (void*)e->hpart, (void*)e->hdt, (void*)e->hsync})
        if (p) (void)hipHostFree(p);
    if (e->ev) (void)hipEventDestroy(e->ev);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    m2g_free(e->m2g);
    delete e;
}

// The three marker -> grid passes of every member.
static hipError_t ensemble_enqueue_m2g(sf_ensemble* e)
{
    hipStream_t s = e->stream;
    for (int i = 0; i < e->n; ++i) {
        const sf_step_args& a = e->mem[(size_t)i];
        const double dx = a.dx, dy = a.dy;
        // grid origins, StokesFlow2D_multixpu.jl:117-126
        const double x_rho = 0.0 * dx - 0.5 * dx, y_rho = 0.0 * dy;
        const double x_p = 0.0 * dx + 0.5 * dx, y_p = 0.0 * dy + 0.5 * dy;
        M2GScratch S = e->m2g;
        S.counters = e->cnt + (size_t)ENS_CNT * i;
        hipError_t r = m2g_run(S, x_rho, y_rho, a.rho_vy, a.nx + 1, a.ny, a.x_m, a.y_m, a.rho_m, a.nm, dx, dy, nullptr,
                               nullptr, 0, 0, s);
        if (r != hipSuccess) return r;
        S.counters = e->cnt + (size_t)ENS_CNT * i + 4;
        r = m2g_run(S, 0.0, 0.0, a.mu_b, a.nx, a.ny, a.x_m, a.y_m, a.mu_m, a.nm, dx, dy, nullptr, nullptr, 0, 0, s);
        if (r != hipSuccess) return r;
        S.counters = e->cnt + (size_t)ENS_CNT * i + 8;
        r = m2g_run(S, x_p, y_p, a.mu_p, a.nx - 1, a.ny - 1, a.x_m, a.y_m, a.mu_m, a.nm, dx, dy, nullptr, nullptr, 0,
                    0, s);
        if (r != hipSuccess) return r;
    }
    return hipSuccess;
}

// RK4 of every member with the dt the solve produced, read on the device.
static hipError_t ensemble_enqueue_rk4(sf_ensemble* e)
{
    hipStream_t s = e->stream;
    for (int i = 0; i < e->n; ++i) {
        const sf_step_args& a = e->mem[(size_t)i];
        const double dx = a.dx, dy = a.dy;
        const double lx = (double)(a.nx - 1) * dx, ly = (double)(a.ny - 1) * dy;
        const double x_vx = -1.0 * dx, y_vx = 0.0 * dy - 0.5 * dy;
        const double x_vy = 0.0 * dx - 0.5 * dx, y_vy = -1.0 * dy;
        unsigned long long* rc = e->cnt + (size_t)ENS_CNT * i + 12;
        const double* dt_dev = e->wide ? e->desc[(size_t)i].scal + SC_DT_CHECK : &e->res[i].dt_new;
        hipError_t r = launch_zero_u64(rc, 4, s);
        if (r != hipSuccess) return r;
        r = launch_rk4(a.x_m, a.y_m, a.nm, a.Vx_pad, a.nx + 2, a.ny + 1, a.Vy_pad, a.nx + 1, a.ny + 2, x_vx, y_vx, x_vy,
                       y_vy, 0.0, lx, ly, dx, dy, rc, s, dt_dev);
        if (r != hipSuccess) return r;
    }
    return hipSuccess;
}

// Persistent and grid forms: the whole step, with the persistent batched kernel (or one grid-mode
// launch per member, largest first) in the middle.
static hipError_t ensemble_enqueue_persistent(sf_ensemble* e)
{
    hipError_t r = ensemble_enqueue_m2g(e);
    if (r != hipSuccess) return r;
    if (e->grid) {
        for (int q = 0; q < e->n && r == hipSuccess; ++q) {
            const int i = e->order[(size_t)q];
            const GridShape& gs = e->gshape[(size_t)i];
            const GridArgs ga{e->inst + i, e->res + i, e->gsync + (size_t)GRID_SYNC_WORDS * i, e->gpart, gs.nwg, gs.waves,
                              grid_wait_limit()};
            r = launch_grid_solve(ga, e->gk, e->stream);
        }
    } else {
        r = launch_batch_solve(e->inst, e->order_dev, e->n, e->queue, e->res, e->threads, e->nblocks, e->stream);
    }
    if (r != hipSuccess) return r;
    return ensemble_enqueue_rk4(e);
}

static hipError_t ensemble_capture(sf_ensemble* e, int k, hipError_t (*fn)(sf_ensemble*))
{
    hipError_t er = hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal);
    if (er != hipSuccess) return er;
    const hipError_t eq = fn(e);
    er = hipStreamEndCapture(e->stream, &e->graph[k]);
    if (eq != hipSuccess) return eq;
    if (er != hipSuccess) return er;
    return hipGraphInstantiate(&e->exec[k], e->graph[k], nullptr, nullptr, 0);
}

extern "C" int sf_ensemble_create(sf_ctx* c, const sf_step_args* members, int n, const sf_batch_params* prm,
                                  int use_graph, sf_ensemble** out)
{
    const char* who = "sf_ensemble_create";
    if (!c || !out || (n > 0 && !members)) return set_err(SF_ERR_ARG, "%s: NULL argument", who);
    *out = nullptr;
    if (n < 1) return set_err(SF_ERR_ARG, "%s: n < 1", who);
    int threads, rows;
    int rc = batch_threads(prm, &threads, &rows, who);
    if (rc) return rc;
    const int mode = prm ? prm->mode : SF_BATCH_AUTO;
    if (mode != SF_BATCH_AUTO && mode != SF_BATCH_WIDE && mode != SF_BATCH_PERSISTENT && mode != SF_BATCH_GRID)
        return set_err(SF_ERR_ARG, "%s: unknown mode %d", who, mode);
    std::vector<sf_batch_instance> bi((size_t)n);
    int64_t max_nm = 0, max_cells = 0;
    size_t need = 0;
    std::vector<size_t> off((size_t)n);
    for (int i = 0; i < n; ++i) {
        const sf_step_args& a = members[i];
        sf_batch_instance& b = bi[(size_t)i];
        std::memset(&b, 0, sizeof(b));
        b.P = a.P;
        b.Vx_pad = a.Vx_pad;
        b.Vy_pad = a.Vy_pad;
        b.dVxdtau = a.dVxdtau;
        b.dVydtau = a.dVydtau;
        b.Vx_s = a.Vx_s;
        b.Vy_s = a.Vy_s;
        b.rho_vy = a.rho_vy;
        b.mu_b = a.mu_b;
        b.mu_p = a.mu_p;
        b.nx = a.nx;
        b.ny = a.ny;
        b.dx = a.dx;
        b.dy = a.dy;
        b.g_y = a.g_y;
        b.dt = a.dt;
        b.maxdisp = a.maxdisp;
        b.eps = a.eps;
        b.iter_max = a.iter_max;
        b.use_fss = a.use_fss;
        rc = batch_check(b, i, who);
        if (rc) return rc;
        if (a.nm < 0 || (a.nm > 0 && (!a.x_m || !a.y_m || !a.rho_m || !a.mu_m)))
            return set_err(SF_ERR_ARG, "%s: member %d: marker arrays", who, i);
        const int64_t cells = (int64_t)(a.nx + 2) * (a.ny + 1);  // the largest of the three grids' cell counts
        if (a.nm >= ((int64_t)1 << 31) - 1) return set_err(SF_ERR_SIZE, "%s: member %d: too many markers", who, i);
        max_nm = std::max(max_nm, a.nm);
        max_cells = std::max(max_cells, cells);
        off[(size_t)i] = need;
        need += batch_scratch_doubles(a.nx, a.ny);
    }
    sf_ensemble* e = new sf_ensemble();
    rc = m2g_reserve(e->m2g, max_nm, max_cells);
    if (rc == SF_OK && dalloc(&e->m2g.counters, 4) != hipSuccess) rc = set_err(SF_ERR_HIP, "%s: hipMalloc failed", who);
    if (rc) {
        free_ensemble(e);
        return rc;
    }
    e->c = c;
    e->n = n;
    e->threads = threads;
    e->use_graph = use_graph ? 1 : 0;
    e->wide = (mode != SF_BATCH_PERSISTENT && mode != SF_BATCH_GRID);
    e->grid = (mode == SF_BATCH_GRID);
    e->gk = (prm && prm->depth >= 1 && prm->depth <= 3) ? prm->depth : 3;
    e->refined = prm ? prm->precision : 0;
    if (e->refined != 0 && (e->refined != 1 || !e->wide)) {
        free_ensemble(e);
        return set_err(SF_ERR_ARG, "%s: FP32 refinement needs the wide form (precision %d)", who, prm->precision);
    }
    e->depth = (prm && prm->depth > 1) ? 2 : 1;
    e->budget = (prm && prm->budget_mb > 0.0) ? prm->budget_mb * 1048576.0
                                              : (e->refined ? 1024.0 : 1536.0) * 1048576.0;
    e->mem.assign(members, members + n);
    e->bi = bi;
    hipError_t er = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (er == hipSuccess) er = hipEventCreateWithFlags(&e->ev, hipEventDisableTiming);
    if (er == hipSuccess) er = dalloc(&e->pool, need);
    if (er == hipSuccess) er = dalloc(&e->inst, (size_t)n);
    if (er == hipSuccess) er = dalloc(&e->order_dev, (size_t)n);
    if (er == hipSuccess) er = dalloc(&e->queue, 1);
    if (er == hipSuccess) er = dalloc(&e->res, (size_t)n);
    if (er == hipSuccess) er = dalloc(&e->cnt, (size_t)ENS_CNT * n);
    if (er == hipSuccess) er = dalloc(&e->part, 3 * (size_t)RED_MAX_BLOCKS);
    if (er == hipSuccess) er = hipHostMalloc((void**)&e->hcnt, sizeof(unsigned long long) * ENS_CNT * n, hipHostMallocDefault);
    if (er == hipSuccess) er = hipHostMalloc((void**)&e->hres, sizeof(BatchResult) * n, hipHostMallocDefault);
    if (er == hipSuccess) er = hipHostMalloc((void**)&e->hscal, sizeof(double) * SC_COUNT * n, hipHostMallocDefault);
    if (er == hipSuccess && e->grid) er = dalloc(&e->gsync, (size_t)GRID_SYNC_WORDS * n);
    if (er == hipSuccess && e->grid) er = dalloc(&e->gpart, 3 * (size_t)GRID_MAX_WAVES);
    if (er == hipSuccess && e->grid)
        er = hipHostMalloc((void**)&e->hsync, sizeof(unsigned long long) * GRID_SYNC_WORDS * n, hipHostMallocDefault);
    if (er != hipSuccess) {
        free_ensemble(e);
        return set_err(SF_ERR_HIP, "%s: %s", who, hipGetErrorString(er));
    }
    e->desc.resize((size_t)n);
    e->dt.resize((size_t)n);
    std::vector<std::pair<int, int>> dims((size_t)n);
    for (int i = 0; i < n; ++i) {
        BatchShape sh = e->grid ? grid_shape(members[i].nx, members[i].ny) : batch_shape(c, members[i].nx, members[i].ny);
        if (!e->grid) sh.copies = n;
        e->desc[(size_t)i] = batch_descriptor(bi[(size_t)i], rows, e->pool + off[(size_t)i], e->wide ? nullptr : &e->res[i].dt_new,
                                              sh);
        dims[(size_t)i] = {members[i].nx, members[i].ny};
        std::memset(&e->hres[i], 0, sizeof(BatchResult));
        e->hres[i].dt_new = members[i].use_fss ? members[i].dt : 0.0;
        e->dt[(size_t)i] = members[i].dt;
    }
    e->order = batch_order(dims);
    if (e->grid) {
        e->gshape.resize((size_t)n);
        for (int i = 0; i < n; ++i) e->gshape[(size_t)i] = grid_launch_shape(e->desc[(size_t)i], e->gk);
    }
    e->nblocks = batch_grid(c, n, threads, prm ? prm->max_blocks : 0);
    if (e->refined) {
        size_t nf = 0;
        std::vector<size_t> foff((size_t)n);
        for (int i = 0; i < n; ++i) {
            foff[(size_t)i] = nf;
            nf += refine_floats(members[i].nx, members[i].ny);
        }
        er = dalloc(&e->pool32, nf);
        if (er == hipSuccess) er = dalloc(&e->rpart, (size_t)3 * RED_MAX_BLOCKS * n);
        if (er == hipSuccess) er = dalloc(&e->cdesc_dev, CDESC_PER_INST * (size_t)n);
        if (er == hipSuccess)
            er = hipHostMalloc((void**)&e->hpart, sizeof(double) * 3 * RED_MAX_BLOCKS * (size_t)n, hipHostMallocDefault);
        if (er == hipSuccess) er = hipHostMalloc((void**)&e->hdt, sizeof(float) * (size_t)n, hipHostMallocDefault);
        if (er != hipSuccess) {
            free_ensemble(e);
            return set_err(SF_ERR_HIP, "%s: %s", who, hipGetErrorString(er));
        }
        e->rb.resize((size_t)n);
        e->cdesc.resize(CDESC_PER_INST * (size_t)n);
        for (int i = 0; i < n; ++i) {
            e->rb[(size_t)i] = refine_carve(members[i].nx, members[i].ny, e->pool32 + foff[(size_t)i],
                                            e->rpart + (size_t)3 * RED_MAX_BLOCKS * i);
            refine_descriptors(bi[(size_t)i], e->rb[(size_t)i], prm ? prm->rows : 0, &e->cdesc[CDESC_PER_INST * (size_t)i]);
        }
        er = hipMemcpyAsync(e->cdesc_dev, e->cdesc.data(), sizeof(CorrParams) * CDESC_PER_INST * (size_t)n, hipMemcpyHostToDevice,
                            e->stream);
        if (er != hipSuccess) {
            free_ensemble(e);
            return set_err(SF_ERR_HIP, "%s: %s", who, hipGetErrorString(er));
        }
    }
    // make the caller's earlier work on its stream visible, then upload on our stream
    er = hipEventRecord(e->ev, c->stream);
    if (er == hipSuccess) er = hipStreamWaitEvent(e->stream, e->ev, 0);
    if (er == hipSuccess && e->grid)  // the sticky failure flags start clear (outside any graph)
        er = hipMemsetAsync(e->gsync, 0, sizeof(unsigned long long) * GRID_SYNC_WORDS * (size_t)n, e->stream);
    if (er == hipSuccess)
        er = hipMemcpyAsync(e->inst, e->desc.data(), sizeof(BatchInstance) * (size_t)n, hipMemcpyHostToDevice, e->stream);
    if (er == hipSuccess)
        er = hipMemcpyAsync(e->order_dev, e->order.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, e->stream);
    if (er == hipSuccess)
        er = hipMemcpyAsync(e->res, e->hres, sizeof(BatchResult) * (size_t)n, hipMemcpyHostToDevice, e->stream);
    if (er == hipSuccess) er = hipStreamSynchronize(e->stream);
    if (er == hipSuccess && e->use_graph) {
        if (e->wide) {
            er = ensemble_capture(e, 0, ensemble_enqueue_m2g);
            if (er == hipSuccess) er = ensemble_capture(e, 1, ensemble_enqueue_rk4);
        } else {
            er = ensemble_capture(e, 0, ensemble_enqueue_persistent);
        }
    }
    if (er != hipSuccess) {
        free_ensemble(e);
        return set_err(SF_ERR_HIP, "%s: %s", who, hipGetErrorString(er));
    }
    *out = e;
    return SF_OK;
}

extern "C" int sf_ensemble_step(sf_ensemble* e, int n_steps, sf_solve_out* outs, double* dt)
{
    const char* who = "sf_ensemble_step";
    if (!e) return set_err(SF_ERR_ARG, "%s: NULL ensemble", who);
    if (n_steps < 0) return set_err(SF_ERR_ARG, "%s: n_steps < 0", who);
    const auto t_start = std::chrono::steady_clock::now();
    HIPCHK(hipEventRecord(e->ev, e->c->stream));
    HIPCHK(hipStreamWaitEvent(e->stream, e->ev, 0));
    std::vector<BatchResult> res;
    for (int k = 0; k < n_steps; ++k) {
        if (!e->wide) {
            if (e->use_graph)
                HIPCHK(hipGraphLaunch(e->exec[0], e->stream));
            else
                HIPCHK(ensemble_enqueue_persistent(e));
            continue;
        }
        if (e->use_graph)
            HIPCHK(hipGraphLaunch(e->exec[0], e->stream));
        else
            HIPCHK(ensemble_enqueue_m2g(e));
        for (int i = 0; i < e->n; ++i) e->bi[(size_t)i].dt = e->dt[(size_t)i];
        const WideEnv env{e->stream, e->inst, e->part, e->hscal};
        int rc;
        if (e->refined) {
            const RefineEnv renv{env, &e->rb, e->cdesc_dev, e->hpart, e->hdt, 0};
            rc = batch_run_wide_refined(renv, e->bi.data(), e->n, e->desc, e->cdesc, e->order, e->budget, res);
        } else {
            rc = batch_run_wide(env, e->bi.data(), e->n, e->desc, e->order, e->budget, e->depth, res);
        }
        if (rc) return rc;
        for (int i = 0; i < e->n; ++i) e->dt[(size_t)i] = res[(size_t)i].dt_new;
        if (e->use_graph)
            HIPCHK(hipGraphLaunch(e->exec[1], e->stream));
        else
            HIPCHK(ensemble_enqueue_rk4(e));
    }
    if (!e->wide)
        HIPCHK(hipMemcpyAsync(e->hres, e->res, sizeof(BatchResult) * (size_t)e->n, hipMemcpyDeviceToHost, e->stream));
    if (e->grid)
        HIPCHK(hipMemcpyAsync(e->hsync, e->gsync, sizeof(unsigned long long) * GRID_SYNC_WORDS * (size_t)e->n,
                              hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemcpyAsync(e->hcnt, e->cnt, sizeof(unsigned long long) * ENS_CNT * (size_t)e->n, hipMemcpyDeviceToHost,
                          e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    if (n_steps > 0 && e->grid)
        for (int i = 0; i < e->n; ++i)
            if (e->hsync[(size_t)GRID_SYNC_WORDS * i + GRID_SYNC_STICKY])
                return set_err(SF_ERR_HIP,
                               "%s: member %d: a grid barrier wait exceeded the limit its launch was made with "
                               "(workgroups not all resident?)",
                               who, i);
    if (n_steps > 0)
        for (int i = 0; i < e->n; ++i) {
            const unsigned long long* q = e->hcnt + (size_t)ENS_CNT * i;
            if (q[2] + q[6] + q[10] > 0 || q[13] > 0)
                return set_err(SF_ERR_NONFINITE, "%s: member %d has markers with non-finite coordinates", who, i);
        }
    for (int i = 0; i < e->n; ++i) {
        if (e->wide && n_steps > 0) e->hres[i] = res[(size_t)i];
        const BatchResult& r = e->hres[i];
        if (outs) {
            sf_solve_out& o = outs[i];
            std::memset(&o, 0, sizeof(o));
            o.dt_new = r.dt_new;
            o.err = r.err;
            o.niter = r.niter;
            o.converged = r.converged;
            o.dt_last = r.dt_last;
            o.runtime = secs;
        }
        if (dt) dt[i] = e->wide ? e->dt[(size_t)i] : r.dt_new;
    }
    return SF_OK;
}

extern "C" int sf_ensemble_destroy(sf_ensemble* e)
{
    if (!e) return SF_OK;
    (void)hipStreamSynchronize(e->stream);
    free_ensemble(e);
    return SF_OK;
}

// ---------------------------------------------------------------------------------------
// The captured timestep of one problem: a one-member ensemble in the grid form, so the whole
// step (three marker -> grid passes, the solve with its loop on the device, RK4 with the new
// dt read on the device) is one hipGraph.

struct sf_graph {
    sf_ensemble* e = nullptr;
};

extern "C" int sf_graph_create(sf_ctx* c, const sf_step_args* args, int iters_per_pass, sf_graph** out)
{
    const char* who = "sf_graph_create";
    if (!c || !args || !out) return set_err(SF_ERR_ARG, "%s: NULL argument", who);
    *out = nullptr;
    if (iters_per_pass < 0 || iters_per_pass > 3) return set_err(SF_ERR_ARG, "%s: iters_per_pass %d", who, iters_per_pass);
    sf_batch_params prm;
    std::memset(&prm, 0, sizeof(prm));
    prm.mode = SF_BATCH_GRID;
    prm.depth = iters_per_pass;
    sf_graph* g = new sf_graph();
    const int rc = sf_ensemble_create(c, args, 1, &prm, 1, &g->e);
    if (rc) {
        delete g;
        return rc;
    }
    *out = g;
    return SF_OK;
}

extern "C" int sf_graph_step(sf_graph* g, int n_steps, sf_solve_out* out, double* dt)
{
    if (!g || !g->e) return set_err(SF_ERR_ARG, "sf_graph_step: NULL graph");
    return sf_ensemble_step(g->e, n_steps, out, dt);
}

extern "C" int sf_graph_destroy(sf_graph* g)
{
    if (!g) return SF_OK;
    const int rc = sf_ensemble_destroy(g->e);
    delete g;
    return rc;
}
