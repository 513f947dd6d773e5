// sf_batch_host.cpp -- host side of the batched solver (sf_solve_stokes_batched): instance
// descriptors and scratch, the wide-mode schedulers (FP64, and FP32 sweeps inside FP64
// defect correction) and the persistent launch.  The kernels are in sf_batch.hip and
// sf_refine.hip.
#include "sf_host.h"

using namespace sf;

namespace sf {

// Per-instance scratch (doubles) of the batched solver: ping-pong state, tau/divV, Rx/Ry, scalars.
size_t batch_scratch_doubles(int nx_, int ny_)
{
    const size_t nx = nx_, ny = ny_, np = (nx - 1) * (ny - 1);
    const size_t ndv = (nx - 2) * (ny - 1) + (nx - 1) * (ny - 2);
    return nx * (ny + 1) + (nx + 1) * ny + np + 3 * np + nx * ny + ndv + SC_COUNT + ndv;
}

int batch_check(const sf_batch_instance& a, int i, const char* who)
{
    if (a.nx < 3 || a.ny < 3) return set_err(SF_ERR_ARG, "%s: instance %d: need Nx, Ny >= 3", who, i);
    if ((int64_t)(a.nx + 2) * (a.ny + 2) >= (int64_t)1 << 31)
        return set_err(SF_ERR_SIZE, "%s: instance %d: grid too large", who, i);
    if (!a.P || !a.Vx_pad || !a.Vy_pad || !a.rho_vy || !a.mu_b || !a.mu_p || !a.dVxdtau || !a.dVydtau || !a.Vx_s ||
        !a.Vy_s)
        return set_err(SF_ERR_ARG, "%s: instance %d: a state or material array is NULL", who, i);
    if (!(a.dx > 0.0) || !(a.dy > 0.0)) return set_err(SF_ERR_ARG, "%s: instance %d: dx, dy must be > 0", who, i);
    return SF_OK;
}

// The device descriptor of one instance; scratch carved from q (batch_scratch_doubles long).
BatchShape batch_shape(const sf_ctx* c, int nx, int ny)
{
    if (c)
        for (const BatchTune& t : c->btune)
            if (t.nx == nx && t.ny == ny) return t.sh;
    return BatchShape();
}

BatchInstance batch_descriptor(const sf_batch_instance& a, int rows, double* q, const double* dt_dev,
                               const BatchShape& sh)
{
    BatchInstance b;
    std::memset(&b, 0, sizeof(b));
    const int nx = a.nx, ny = a.ny;
    const size_t np = (size_t)(nx - 1) * (ny - 1);
    double* vx_b = q; q += (size_t)nx * (ny + 1);
    double* vy_b = q; q += (size_t)(nx + 1) * ny;
    double* p_b = q; q += np;
    double* t_xx = q; q += np;
    double* t_yy = q; q += np;
    double* d_v = q; q += np;
    double* t_xy = q; q += (size_t)nx * ny;
    double* r_x = q; q += (size_t)(nx - 2) * (ny - 1);
    double* r_y = q; q += (size_t)(nx - 1) * (ny - 2);
    double* scal = q; q += SC_COUNT;
    double* dvx2 = q; q += (size_t)(nx - 2) * (ny - 1);  // second dV (wide mode)
    double* dvy2 = q;
    const SolveConsts k = solve_consts(nx, ny, a.dx, a.dy);
    if (rows <= 0) rows = pt_default_rows(nx, ny);
    rows = std::min(rows, ny);
    PTParams pp{};
    pp.nx = nx;
    pp.ny = ny;
    pp.seg = rows;
    pp.depth = 1;
    pp.nstrips = pt_num_strips(nx);
    pp.ntiles = pp.nstrips * ((ny + rows - 1) / rows);
    pp.dvx = a.dVxdtau;
    pp.dvy = a.dVydtau;
    pp.rho = a.rho_vy;
    pp.mu_b = a.mu_b;
    pp.mu_p = a.mu_p;
    pp.dt = scal + SC_DT;
    pp.txx = t_xx;
    pp.tyy = t_yy;
    pp.txy = t_xy;
    pp.divv = d_v;
    pp.rx = r_x;
    pp.g = a.g_y;
    pp.dampx = k.dampx;
    pp.dampy = k.dampy;
    pp.rdx = k.rdx;
    pp.rdy = k.rdy;
    pp.vsc_mdxy2 = k.vsc_mdxy2;
    pp.ptsc41 = k.ptsc41;
    pp.max_nxy = k.max_nxy;
    b.pt[0] = pp;
    b.pt[0].vx_a = a.Vx_s;
    b.pt[0].vy_a = a.Vy_s;
    b.pt[0].p_a = a.P;
    b.pt[0].vx_b = vx_b;
    b.pt[0].vy_b = vy_b;
    b.pt[0].p_b = p_b;
    b.pt[1] = pp;
    b.pt[1].vx_a = vx_b;
    b.pt[1].vy_a = vy_b;
    b.pt[1].p_a = p_b;
    b.pt[1].vx_b = a.Vx_s;
    b.pt[1].vy_b = a.Vy_s;
    b.pt[1].p_b = a.P;
    // wide mode: dV ping-pongs with the state, so one parity describes both; the two-iteration
    // descriptors carry the pass's own tile geometry
    for (int par = 0; par < 2; ++par) {
        PTParams w = b.pt[par];
        w.dvx = par ? dvx2 : a.dVxdtau;
        w.dvy = par ? dvy2 : a.dVydtau;
        w.dvx_o = par ? a.dVxdtau : dvx2;
        w.dvy_o = par ? a.dVydtau : dvy2;
        b.wpt[par] = w;
        // at least 4 rows: the wide mode fills the chip with many instances, so the short tiles
        // that suit one small grid alone would only add halo rows
        const int rows2 = sh.rows2 > 0 ? std::min(sh.rows2, ny) : std::max(pt2_default_rows(nx, ny), std::min(4, ny));
        w.seg = rows2;
        w.nstrips = (nx - 1 + 59) / 60;
        w.ntiles = w.nstrips * ((ny + rows2 - 1) / rows2);
        b.wpt[2 + par] = w;
        // few instances fill the chip only with shorter tiles: below WIDE_MAX instances the height
        // comes from the wave-round model over that many copies of the grid (16 x 256^2, one box:
        // 48-row four-iteration tiles ran 480 waves per launch)
        const bool model = sh.copies > 0 && sh.copies < WIDE_MAX;
        const int rows3 = std::min(ny, sh.rows3 > 0 ? sh.rows3
                                                    : (model ? ptk_default_rows(3, nx, ny, sh.copies)
                                                             : (std::max(nx, ny) > 256 ? 16 : 8)));
        w.seg = rows3;
        w.nstrips = (nx - 1 + 57) / 58;
        w.ntiles = w.nstrips * ((ny + rows3 - 1) / rows3);
        b.wpt[4 + par] = w;
        const int n = std::max(nx, ny);
        const int rows4 = std::min(ny, sh.rows4 > 0 ? sh.rows4
                                                    : (model ? ptk_default_rows(4, nx, ny, sh.copies)
                                                             : (n > 384 ? 64 : (n > 192 ? 48 : 8))));
        w.seg = rows4;
        w.nstrips = (nx - 1 + 55) / 56;
        w.ntiles = w.nstrips * ((ny + rows4 - 1) / rows4);
        b.wpt[6 + par] = w;
    }
    b.P = a.P;
    b.Vx_s = a.Vx_s;
    b.Vy_s = a.Vy_s;
    b.Vx_pad = a.Vx_pad;
    b.Vy_pad = a.Vy_pad;
    b.ry = r_y;
    b.scal = scal;
    b.dt_dev = dt_dev;
    b.nx = nx;
    b.ny = ny;
    b.use_fss = a.use_fss;
    // size default: four iterations per pass above 192 nodes per side, three below (wide passes,
    // ns per node and iteration on the 256-instance ensemble's active sets, best tile heights,
    // K = 4 / 3: 512^2 0.00691 / 0.00783; 256^2 0.00833 / 0.00905; 128^2 0.01306 / 0.01219;
    // 64^2 0.03259 / 0.03056)
    b.kmax = sh.kmax > 0 ? std::min(sh.kmax, 4) : (std::max(nx, ny) > 192 ? 4 : 3);
    b.dx = a.dx;
    b.dy = a.dy;
    b.dt_in = a.dt;
    b.maxdisp = a.maxdisp;
    b.eps = a.eps;
    b.iter_max = a.iter_max;
    return b;
}

// Largest expected work first: nodes x an iteration count that grows with the grid size.
std::vector<int> batch_order(const std::vector<std::pair<int, int>>& dims)
{
    const int n = (int)dims.size();
    std::vector<int> order((size_t)n);
    for (int i = 0; i < n; ++i) order[(size_t)i] = i;
    auto work = [&](int i) {
        const double a = dims[(size_t)i].first, b = dims[(size_t)i].second;
        return a * b * std::max(a, b);
    };
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return work(x) > work(y); });
    return order;
}

int batch_grid(sf_ctx* c, int n, int threads, int max_blocks)
{
    int cus = 256;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || cus < 1) cus = 256;
    int resident = cus * batch_blocks_per_cu(threads);
    if (max_blocks > 0) resident = std::min(resident, max_blocks);
    return std::min(n, resident);
}

int batch_threads(const sf_batch_params* prm, int* threads, int* rows, const char* who)
{
    *threads = (prm && prm->threads > 0) ? prm->threads : 512;
    *rows = (prm && prm->rows > 0) ? prm->rows : 0;  // 0: per instance, pt_default_rows
    if (*threads % 64 != 0 || *threads > 512) return set_err(SF_ERR_ARG, "%s: threads %d", who, *threads);
    return SF_OK;
}

// Grid mode: the single solver's tile heights for every pass kind (one instance fills the chip).
BatchShape grid_shape(int nx, int ny)
{
    BatchShape sh;
    sh.kmax = 3;
    sh.rows2 = std::min(pt2_default_rows(nx, ny), ny);
    sh.rows3 = std::min(ptk_default_rows(3, nx, ny), ny);
    sh.rows4 = std::min(ptk_default_rows(4, nx, ny), ny);
    return sh;
}

// Grid mode: the launch of an instance, one wave per tile of its largest pass spread one wave per
// CU first, then up to GRID_WG_WAVES waves per workgroup (at most one workgroup per CU).
GridShape grid_launch_shape(const BatchInstance& d, int k)
{
    int t = d.wpt[0].ntiles;
    if (k >= 2) t = std::max(t, d.wpt[2 * k - 2].ntiles);
    int cus = 256;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        cus = 256;
    const int waves = std::max(1, std::min((t + cus - 1) / cus, (int)GRID_WG_WAVES));
    const int nwg = std::max(1, std::min((t + waves - 1) / waves, grid_capacity(k, waves)));
    return GridShape{nwg, waves};
}

}  // namespace sf

extern "C" int sf_set_grid_wait_limit(double seconds)
{
    if (!(seconds >= 0.0) || seconds > 3600.0)
        return set_err(SF_ERR_ARG, "sf_set_grid_wait_limit: %g s is not in [0, 3600]", seconds);
    grid_set_wait_seconds(seconds);
    return SF_OK;
}

namespace sf {

// Grid mode: the barrier words and the partials of the context.
static int grid_reserve(sf_ctx* c, int n)
{
    if (!c->g_part) HIPCHK(dalloc(&c->g_part, 3 * (size_t)GRID_MAX_WAVES));
    if (n > c->g_cap) {
        if (c->g_sync) HIPCHK(hipFree(c->g_sync));
        c->g_sync = nullptr;
        c->g_cap = 0;
        HIPCHK(dalloc(&c->g_sync, (size_t)GRID_SYNC_WORDS * n));
        c->g_cap = n;
    }
    return SF_OK;
}

// Bytes an instance keeps live while it iterates: both ping-pong copies of Vx, Vy, P, plus
// dVx, dVy and the three material grids.
static double batch_working_set(int nx_, int ny_)
{
    const double nx = nx_, ny = ny_, np = (nx - 1) * (ny - 1);
    return 8.0 * (2.0 * (nx * (ny + 1) + (nx + 1) * ny + np) + (nx - 2) * (ny - 1) + (nx - 1) * (ny - 2) +
                  (nx + 1) * ny + nx * ny + np);
}

int batch_run_wide(const WideEnv& env, const sf_batch_instance* in, int n,
                          const std::vector<BatchInstance>& desc, const std::vector<int>& order, double budget,
                          int depth, std::vector<BatchResult>& res)
{
    hipStream_t s = env.s;
    struct Run {
        int id;
        int64_t iter, ncheck, ndtupdate;
        double err, ws;
        bool checking;
        int k;            // iterations of this step's launch (1 to 4)
        int64_t nlaunch;  // launches so far: their parity says where the state (V, P, dV) is
        // A check's err is read one step later: the host waits for the step before the one it
        // has just enqueued, so the GPU never drains at a check.  Meanwhile the instance may run
        // one plain pass (no check, no last iteration, no dt refresh); it reads the checked
        // state and writes the other ping-pong buffers, so if the check converged, the checked
        // state is still intact and the pass is dropped.
        bool pend, skip;
        int64_t chk_iter, chk_nlaunch;  // iter and nlaunch right after the pending check
    };
    std::vector<Run> act;
    size_t q = 0;
    double ws_active = 0.0;
    res.assign((size_t)n, BatchResult{});
    auto start = [&](int id) -> int {
        const sf_batch_instance& a = in[id];
        const int mn = std::min(a.nx, a.ny);
        Run r{id, 1, std::min<int64_t>(5 * (int64_t)mn, 2000), 10 * (int64_t)mn, 2.0 * a.eps,
              batch_working_set(a.nx, a.ny), false, 1, 0, false, false, 0, 0};
        HIPCHK(launch_set_scalar(desc[(size_t)id].scal + SC_DT, a.use_fss ? a.dt : 0.0, s));
        act.push_back(r);
        ws_active += r.ws;
        return SF_OK;
    };
    auto finish = [&](const Run& r) -> int {
        const sf_batch_instance& a = in[r.id];
        const BatchInstance& d = desc[(size_t)r.id];
        const int64_t niter = r.iter - 1;
        if (r.nlaunch & 1) {  // the last iterate is in the scratch buffers
            const PTParams& p0 = d.wpt[0];
            HIPCHK(hipMemcpyAsync(a.Vx_s, p0.vx_b, sizeof(double) * (size_t)a.nx * (a.ny + 1), hipMemcpyDeviceToDevice, s));
            HIPCHK(hipMemcpyAsync(a.Vy_s, p0.vy_b, sizeof(double) * (size_t)(a.nx + 1) * a.ny, hipMemcpyDeviceToDevice, s));
            HIPCHK(hipMemcpyAsync(a.P, p0.p_b, sizeof(double) * (size_t)(a.nx - 1) * (a.ny - 1), hipMemcpyDeviceToDevice,
                                  s));
            HIPCHK(hipMemcpyAsync(a.dVxdtau, p0.dvx_o, sizeof(double) * (size_t)(a.nx - 2) * (a.ny - 1),
                                  hipMemcpyDeviceToDevice, s));
            HIPCHK(hipMemcpyAsync(a.dVydtau, p0.dvy_o, sizeof(double) * (size_t)(a.nx - 1) * (a.ny - 2),
                                  hipMemcpyDeviceToDevice, s));
        }
        HIPCHK(launch_pad_mirror(a.nx, a.ny, a.Vx_s, a.Vy_s, a.Vx_pad, a.Vy_pad, s));
        int rc = enqueue_dt_raw(s, a.nx, a.ny, a.Vx_s, a.Vy_s, a.dx, a.dy, a.maxdisp, 0, env.part, d.scal);
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(env.hscal + (size_t)SC_COUNT * r.id, d.scal, SC_COUNT * sizeof(double),
                              hipMemcpyDeviceToHost, s));
        BatchResult& o = res[(size_t)r.id];
        o.err = r.err;
        o.niter = niter;
        o.converged = r.err <= a.eps;
        return SF_OK;
    };
    int rc = SF_OK;
    struct Events {  // one per step parity: recorded after a step's check reads
        hipEvent_t e[2] = {nullptr, nullptr};
        ~Events()
        {
            for (hipEvent_t x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } ev;
    for (int k = 0; k < 2; ++k) HIPCHK(hipEventCreateWithFlags(&ev.e[k], hipEventDisableTiming));
    bool step_checked[2] = {false, false};
    int64_t step = 0;
    WideArgs wa, w2, w3, w4;
    std::memset(&wa, 0, sizeof(wa));
    std::memset(&w2, 0, sizeof(w2));
    std::memset(&w3, 0, sizeof(w3));
    std::memset(&w4, 0, sizeof(w4));
    wa.desc = env.desc_dev;
    w2.desc = env.desc_dev;
    w3.desc = env.desc_dev;
    w4.desc = env.desc_dev;
    while (!act.empty() || q < order.size()) {
        while (q < order.size() && (int)act.size() < WIDE_MAX &&
               (act.empty() || ws_active + batch_working_set(in[order[q]].nx, in[order[q]].ny) <= budget)) {
            rc = start(order[q++]);
            if (rc) return rc;
        }
        // every active instance advances by three or two PT iterations in one pass where its next
        // iterations are plain (the rule of sf_solve_stokes), else by one; each kind is a launch
        wa.n_active = 0;
        w2.n_active = 0;
        w3.n_active = 0;
        w4.n_active = 0;
        int tiles = 0, tiles2 = 0, tiles3 = 0, tiles4 = 0;
        for (size_t k = 0; k < act.size(); ++k) {
            Run& r = act[k];
            const sf_batch_instance& a = in[r.id];
            const bool check = (r.iter % r.ncheck) == 0;
            const bool last = r.iter == a.iter_max;
            auto fusable = [&](int64_t kk) {
                if (r.iter + kk - 1 >= a.iter_max) return false;
                for (int64_t j = r.iter; j < r.iter + kk; ++j) {
                    if (j % r.ncheck == 0) return false;
                    if (j < r.iter + kk - 1 && a.use_fss && j % r.ndtupdate == 0) return false;
                }
                return true;
            };
            const int kmax = desc[(size_t)r.id].kmax;
            r.k = 1;
            for (int kk = kmax; kk >= 2 && r.k == 1; --kk)
                if (fusable(kk)) r.k = kk;
            r.checking = check;
            // while its check is pending an instance runs only a plain pass
            r.skip = r.pend && (check || last || r.iter > a.iter_max ||
                                (a.use_fss && (r.iter + r.k - 1) % r.ndtupdate == 0));
            if (r.skip) continue;
            const int par = (int)(r.nlaunch & 1);
            if (r.k == 4) {
                w4.tile_start[w4.n_active] = tiles4;
                w4.slot_inst[w4.n_active] = r.id;
                w4.slot_pf[w4.n_active] = par;
                tiles4 += desc[(size_t)r.id].wpt[6].ntiles;
                w4.n_active += 1;
            } else if (r.k == 3) {
                w3.tile_start[w3.n_active] = tiles3;
                w3.slot_inst[w3.n_active] = r.id;
                w3.slot_pf[w3.n_active] = par;
                tiles3 += desc[(size_t)r.id].wpt[4].ntiles;
                w3.n_active += 1;
            } else if (r.k == 2) {
                w2.tile_start[w2.n_active] = tiles2;
                w2.slot_inst[w2.n_active] = r.id;
                w2.slot_pf[w2.n_active] = par;
                tiles2 += desc[(size_t)r.id].wpt[2].ntiles;
                w2.n_active += 1;
            } else {
                const int flags = ((check || last) ? PT_WRITE_TAU : 0) | (check ? PT_WRITE_RX : 0);
                wa.tile_start[wa.n_active] = tiles;
                wa.slot_inst[wa.n_active] = r.id;
                wa.slot_pf[wa.n_active] = par | (flags << 1);
                tiles += desc[(size_t)r.id].wpt[0].ntiles;
                wa.n_active += 1;
            }
        }
        wa.tile_start[wa.n_active] = tiles;
        w2.tile_start[w2.n_active] = tiles2;
        w3.tile_start[w3.n_active] = tiles3;
        w4.tile_start[w4.n_active] = tiles4;
        if (w4.n_active) HIPCHK(launch_wide_pt4(w4, s));
        if (w3.n_active) HIPCHK(launch_wide_pt3(w3, s));
        if (w2.n_active) HIPCHK(launch_wide_pt2(w2, s));
        if (wa.n_active) HIPCHK(launch_wide_pt(wa, depth, s));
        bool any_check = false;
        for (Run& r : act) {
            if (r.skip) continue;
            const sf_batch_instance& a = in[r.id];
            const BatchInstance& d = desc[(size_t)r.id];
            const PTParams& p = d.wpt[r.nlaunch & 1];  // its outputs are now the current state
            const int64_t it_done = r.iter + r.k - 1;  // the last iteration just run
            const bool dt_upd = a.use_fss && (it_done % r.ndtupdate) == 0;
            if (dt_upd) {
                rc = enqueue_dt_raw(s, a.nx, a.ny, p.vx_b, p.vy_b, a.dx, a.dy, a.maxdisp, 1, env.part, d.scal);
                if (rc) return rc;
            }
            if (!r.checking) continue;
            any_check = true;
            if (a.use_fss) {
                if (!dt_upd) {
                    rc = enqueue_dt_raw(s, a.nx, a.ny, p.vx_b, p.vy_b, a.dx, a.dy, a.maxdisp, 0, env.part, d.scal);
                    if (rc) return rc;
                }
            } else {
                HIPCHK(launch_set_scalar(d.scal + SC_DT_CHECK, 0.0, s));
            }
            const SolveConsts k = solve_consts(a.nx, a.ny, a.dx, a.dy);
            ResYParams rq{};
            rq.nx = a.nx;
            rq.ny = a.ny;
            rq.tyy = p.tyy;
            rq.txy = p.txy;
            rq.P = p.p_b;
            rq.rho = a.rho_vy;
            rq.vx = p.vx_b;
            rq.vy = p.vy_b;
            rq.dt = d.scal + SC_DT_CHECK;
            rq.ry = d.ry;
            rq.g = a.g_y;
            rq.rdx = k.rdx;
            rq.rdy = k.rdy;
            HIPCHK(launch_pt_resid_y(rq, s));
            const double n_rx = (double)(a.nx - 2) * (a.ny - 1), n_ry = (double)(a.nx - 1) * (a.ny - 2),
                         n_dv = (double)(a.nx - 1) * (a.ny - 1);
            RedParams rp{};
            rp.a[0] = p.rx;
            rp.n[0] = (int64_t)n_rx;
            rp.a[1] = d.ry;
            rp.n[1] = (int64_t)n_ry;
            rp.a[2] = p.divv;
            rp.n[2] = (int64_t)n_dv;
            rp.part = env.part;
            rp.mode = RED_SUM;
            const int nb = red_num_blocks((int64_t)n_dv);
            HIPCHK(launch_reduce(rp, 3, nb, s));
            HIPCHK(launch_finalize_err(env.part, nb, n_rx, n_ry, n_dv, d.scal, s));
            HIPCHK(hipMemcpyAsync(env.hscal + (size_t)SC_COUNT * r.id, d.scal, SC_COUNT * sizeof(double),
                                  hipMemcpyDeviceToHost, s));
        }
        const int sp = (int)(step & 1);
        step_checked[sp] = any_check;
        if (any_check) HIPCHK(hipEventRecord(ev.e[sp], s));
        // the previous step's checks: wait for that step only (this one stays queued)
        if (step_checked[sp ^ 1]) HIPCHK(hipEventSynchronize(ev.e[sp ^ 1]));
        step_checked[sp ^ 1] = false;
        // advance, resolve the previous step's checks, retire converged / capped instances
        std::vector<Run> keep;
        keep.reserve(act.size());
        for (Run& r : act) {
            const sf_batch_instance& a = in[r.id];
            const bool was_pending = r.pend;
            if (!r.skip) {
                r.iter += r.k;
                r.nlaunch += 1;
            }
            if (was_pending) {
                r.err = env.hscal[(size_t)SC_COUNT * r.id + SC_ERR];
                r.pend = false;
                if (r.err <= a.eps || r.chk_iter > a.iter_max) {
                    Run f = r;  // the checked state; a plain pass after it is dropped
                    f.iter = r.chk_iter;
                    f.nlaunch = r.chk_nlaunch;
                    rc = finish(f);
                    if (rc) return rc;
                    ws_active -= r.ws;
                    continue;
                }
            }
            if (!r.skip && r.checking) {
                r.pend = true;
                r.chk_iter = r.iter;
                r.chk_nlaunch = r.nlaunch;
                keep.push_back(r);
            } else if (r.iter <= a.iter_max) {
                keep.push_back(r);
            } else {
                rc = finish(r);
                if (rc) return rc;
                ws_active -= r.ws;
            }
        }
        act.swap(keep);
        ++step;
    }
    HIPCHK(hipStreamSynchronize(s));
    for (int i = 0; i < n; ++i) {
        res[(size_t)i].dt_new = env.hscal[(size_t)SC_COUNT * i + SC_DT_CHECK];
        res[(size_t)i].dt_last = env.hscal[(size_t)SC_COUNT * i + SC_DT];
    }
    return SF_OK;
}

// ---- wide mode with FP32 sweeps inside FP64 defect correction ----

size_t refine_floats(int nx_, int ny_)
{
    const size_t nx = nx_, ny = ny_, nvx = nx * (ny + 1), nvy = (nx + 1) * ny, np = (nx - 1) * (ny - 1);
    const size_t nmx = (nx - 2) * (ny - 1), nmy = (nx - 1) * (ny - 2);
    return 2 * (nvx + nvy + np) + 2 * (nmx + nmy) + nvy + nx * ny + 2 * np + nmx + nmy;
}

RefineBufs refine_carve(int nx_, int ny_, float* q, double* part)
{
    const size_t nx = nx_, ny = ny_, nvx = nx * (ny + 1), nvy = (nx + 1) * ny, np = (nx - 1) * (ny - 1);
    const size_t nmx = (nx - 2) * (ny - 1), nmy = (nx - 1) * (ny - 2);
    RefineBufs b;
    for (int k = 0; k < 2; ++k) {
        b.vx[k] = q; q += nvx;
        b.vy[k] = q; q += nvy;
        b.p[k] = q; q += np;
    }
    b.mx = q; q += nmx;
    b.my = q; q += nmy;
    b.rho = q; q += nvy;
    b.mub = q; q += nx * ny;
    b.mup = q; q += np;
    b.fx = q; q += nmx;
    b.fy = q; q += nmy;
    b.fc = q; q += np;
    b.mx2 = q; q += nmx;
    b.my2 = q; q += nmy;
    b.part = part;
    return b;
}

// FP32 working set of a refined instance while it iterates.
static double refine_working_set(int nx, int ny) { return 4.0 * (double)refine_floats(nx, ny); }

// The FP32 correction descriptors of one instance: [0] runs vx[0] -> vx[1], [1] back.
void refine_descriptors(const sf_batch_instance& a, const RefineBufs& b, int rows, CorrParams out[CDESC_PER_INST])
{
    const SolveConsts k = solve_consts(a.nx, a.ny, a.dx, a.dy);
    CorrParams cp{};
    cp.nx = a.nx;
    cp.ny = a.ny;
    cp.seg = std::min(rows > 0 ? rows : 4, a.ny);
    cp.nstrips = corr_num_strips(a.nx);
    cp.ntiles = cp.nstrips * ((a.ny + cp.seg - 1) / cp.seg);
    cp.depth = 1;
    cp.mx = b.mx;
    cp.my = b.my;
    cp.rho = b.rho;
    cp.mu_b = b.mub;
    cp.mu_p = b.mup;
    cp.fx = b.fx;
    cp.fy = b.fy;
    cp.fc = b.fc;
    cp.dt = (float)(a.use_fss ? a.dt : 0.0);
    cp.g = (float)a.g_y;
    cp.dampx = (float)k.dampx;
    cp.dampy = (float)k.dampy;
    cp.rdx = (float)k.rdx;
    cp.rdy = (float)k.rdy;
    cp.vsc_mdxy2 = (float)k.vsc_mdxy2;
    cp.ptsc41 = (float)k.ptsc41;
    cp.max_nxy = (float)k.max_nxy;
    float* const mx[2] = {b.mx, b.mx2};
    float* const my[2] = {b.my, b.my2};
    // at least 4 rows, as the FP64 wide pass: many instances fill the chip
    const int rows2 = std::max(corr2_default_rows(a.nx, a.ny), std::min(4, a.ny));
    for (int d = 0; d < 2; ++d) {
        out[d] = cp;
        out[d].vx_a = b.vx[d];
        out[d].vy_a = b.vy[d];
        out[d].p_a = b.p[d];
        out[d].vx_b = b.vx[1 - d];
        out[d].vy_b = b.vy[1 - d];
        out[d].p_b = b.p[1 - d];
        out[d].mx = mx[d];
        out[d].my = my[d];
        out[d].mx_o = mx[1 - d];
        out[d].my_o = my[1 - d];
        out[2 + d] = out[d];
        out[2 + d].seg = rows2;
        out[2 + d].nstrips = corr2_num_strips(a.nx);
        out[2 + d].ntiles = out[2 + d].nstrips * ((a.ny + rows2 - 1) / rows2);
        const int rows3 = std::min(a.ny, std::max(a.nx, a.ny) > 256 ? 16 : 8);
        out[4 + d] = out[d];
        out[4 + d].seg = rows3;
        out[4 + d].nstrips = (a.nx - 1 + 57) / 58;
        out[4 + d].ntiles = out[4 + d].nstrips * ((a.ny + rows3 - 1) / rows3);
    }
}

static int batch_refine_reserve(sf_ctx* c, int n, size_t nfloats)
{
    if (nfloats > c->b_pool32_cap) {
        if (c->b_pool32) HIPCHK(hipFree(c->b_pool32));
        c->b_pool32 = nullptr;
        c->b_pool32_cap = 0;
        HIPCHK(dalloc(&c->b_pool32, nfloats));
        c->b_pool32_cap = nfloats;
    }
    if (n > c->b_rcap) {
        for (void* p : {(void*)c->b_part, (void*)c->b_cdesc})
            if (p) HIPCHK(hipFree(p));
        for (void* p : {(void*)c->b_hpart, (void*)c->b_hdt})
            if (p) HIPCHK(hipHostFree(p));
        c->b_part = nullptr;
        c->b_cdesc = nullptr;
        c->b_hpart = nullptr;
        c->b_hdt = nullptr;
        c->b_rcap = 0;
        HIPCHK(dalloc(&c->b_part, (size_t)3 * RED_MAX_BLOCKS * n));
        HIPCHK(dalloc(&c->b_cdesc, CDESC_PER_INST * (size_t)n));
        HIPCHK(hipHostMalloc((void**)&c->b_hpart, sizeof(double) * 3 * RED_MAX_BLOCKS * (size_t)n, hipHostMallocDefault));
        HIPCHK(hipHostMalloc((void**)&c->b_hdt, sizeof(float) * (size_t)n, hipHostMallocDefault));
        c->b_rcap = n;
    }
    return SF_OK;
}

// One FP64 PT iteration of instance `a` with the reference's check (compute_Residuals! and
// compute_err, StokesSolver_multixpu.jl:63-75), as the single refined solver's final check.
static int instance_check_iteration(hipStream_t s, const sf_batch_instance& a, const BatchInstance& d, double* part,
                                    double* hscal, double* err_out)
{
    PTParams pp = d.pt[0];
    pp.flags = PT_WRITE_TAU | PT_WRITE_RX;
    pp.max_waves = -1;  // one wave per tile
    HIPCHK(launch_pt_fused(pp, s));
    if (a.use_fss) {
        int rc = enqueue_dt_raw(s, a.nx, a.ny, pp.vx_b, pp.vy_b, a.dx, a.dy, a.maxdisp, 0, part, d.scal);
        if (rc) return rc;
    } else {
        HIPCHK(launch_set_scalar(d.scal + SC_DT_CHECK, 0.0, s));
    }
    const SolveConsts k = solve_consts(a.nx, a.ny, a.dx, a.dy);
    ResYParams rq{};
    rq.nx = a.nx;
    rq.ny = a.ny;
    rq.tyy = pp.tyy;
    rq.txy = pp.txy;
    rq.P = pp.p_b;
    rq.rho = a.rho_vy;
    rq.vx = pp.vx_b;
    rq.vy = pp.vy_b;
    rq.dt = d.scal + SC_DT_CHECK;
    rq.ry = d.ry;
    rq.g = a.g_y;
    rq.rdx = k.rdx;
    rq.rdy = k.rdy;
    HIPCHK(launch_pt_resid_y(rq, s));
    const double n_rx = (double)(a.nx - 2) * (a.ny - 1), n_ry = (double)(a.nx - 1) * (a.ny - 2),
                 n_dv = (double)(a.nx - 1) * (a.ny - 1);
    RedParams rp{};
    rp.a[0] = pp.rx;
    rp.n[0] = (int64_t)n_rx;
    rp.a[1] = d.ry;
    rp.n[1] = (int64_t)n_ry;
    rp.a[2] = pp.divv;
    rp.n[2] = (int64_t)n_dv;
    rp.part = part;
    rp.mode = RED_SUM;
    const int nb = red_num_blocks((int64_t)n_dv);
    HIPCHK(launch_reduce(rp, 3, nb, s));
    HIPCHK(launch_finalize_err(part, nb, n_rx, n_ry, n_dv, d.scal, s));
    HIPCHK(hipMemcpyAsync(a.Vx_s, pp.vx_b, sizeof(double) * (size_t)a.nx * (a.ny + 1), hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(a.Vy_s, pp.vy_b, sizeof(double) * (size_t)(a.nx + 1) * a.ny, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(a.P, pp.p_b, sizeof(double) * (size_t)(a.nx - 1) * (a.ny - 1), hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(hscal, d.scal, SC_COUNT * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    *err_out = hscal[SC_ERR];
    return SF_OK;
}

// The refined wide mode: every instance runs the algorithm of sf_solve_stokes_refined
// (FP32 windows of the correction iteration driven by the FP64 residual, FP64 update,
// dt refresh and convergence test at the reference's iteration numbers, a final FP64 check
// iteration), and one launch runs one FP32 correction iteration of every active instance.
int batch_run_wide_refined(const RefineEnv& env, const sf_batch_instance* in, int n,
                                  const std::vector<BatchInstance>& desc, std::vector<CorrParams>& cdesc,
                                  const std::vector<int>& order, double budget, std::vector<BatchResult>& res)
{
    hipStream_t s = env.w.s;
    const std::vector<RefineBufs>& rb = *env.rb;
    struct Run {
        int id, cur, rebases;
        int64_t n32, n64, ncheck, ndtupdate, checked_at, window, left, wstart;
        double err_k, err_ref, dt_k, ws;
        bool converged;
    };
    res.assign((size_t)n, BatchResult{});
    auto resid_err = [&](Run& r, double* err) -> int {
        const sf_batch_instance& a = in[r.id];
        const RefineBufs& b = rb[(size_t)r.id];
        const SolveConsts k = solve_consts(a.nx, a.ny, a.dx, a.dy);
        const int64_t np = (int64_t)(a.nx - 1) * (a.ny - 1);
        const int nbr = red_num_blocks(np);
        ResidParams rp{};
        rp.nx = a.nx;
        rp.ny = a.ny;
        rp.vx = a.Vx_s;
        rp.vy = a.Vy_s;
        rp.P = a.P;
        rp.rho = a.rho_vy;
        rp.mu_b = a.mu_b;
        rp.mu_p = a.mu_p;
        rp.fx = b.fx;
        rp.fy = b.fy;
        rp.fc = b.fc;
        rp.part = b.part;
        rp.dt = r.dt_k;
        rp.g = a.g_y;
        rp.rdx = k.rdx;
        rp.rdy = k.rdy;
        HIPCHK(launch_resid_f64(rp, nbr, s));
        double* h = env.hpart + (size_t)3 * RED_MAX_BLOCKS * r.id;
        HIPCHK(hipMemcpyAsync(h, b.part, 3 * (size_t)nbr * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        double sx = 0.0, sy = 0.0, sc = 0.0;
        for (int q = 0; q < nbr; ++q) {
            sx += h[q];
            sy += h[nbr + q];
            sc += h[2 * nbr + q];
        }
        const double n_rx = (double)(a.nx - 2) * (a.ny - 1), n_ry = (double)(a.nx - 1) * (a.ny - 2);
        *err = std::max(std::max(sx / n_rx, sy / n_ry), sc / (double)np);
        return SF_OK;
    };
    // Runs checks / starts the next window; sets done when the instance has finished.
    auto begin_window = [&](Run& r, bool* done) -> int {
        const sf_batch_instance& a = in[r.id];
        const BatchInstance& d = desc[(size_t)r.id];
        const RefineBufs& b = rb[(size_t)r.id];
        const int64_t nmx = (int64_t)(a.nx - 2) * (a.ny - 1), nmy = (int64_t)(a.nx - 1) * (a.ny - 2);
        *done = false;
        for (;;) {
            if (r.n32 >= a.iter_max) {
                *done = true;
                return SF_OK;
            }
            int rc = resid_err(r, &r.err_k);
            if (rc) return rc;
            if (!std::isfinite(r.err_k)) {
                *done = true;
                return SF_OK;
            }
            if (r.n32 > 0 && r.n32 % r.ncheck == 0 && r.n32 != r.checked_at && r.err_k <= a.eps) {
                r.checked_at = r.n32;
                float* const mxc = r.cur ? b.mx2 : b.mx;
                float* const myc = r.cur ? b.my2 : b.my;
                HIPCHK(launch_cast_f64(a.dVxdtau, mxc, nmx, s));
                HIPCHK(launch_cast_f64(a.dVydtau, myc, nmy, s));
                HIPCHK(launch_set_scalar(d.scal + SC_DT, r.dt_k, s));
                rc = instance_check_iteration(s, a, d, b.part, env.w.hscal + (size_t)SC_COUNT * r.id, &r.err_ref);
                if (rc) return rc;
                r.n64 += 1;
                HIPCHK(launch_cast_f32(mxc, a.dVxdtau, nmx, s));
                HIPCHK(launch_cast_f32(myc, a.dVydtau, nmy, s));
                if (r.err_ref <= a.eps) {
                    r.converged = true;
                    *done = true;
                    return SF_OK;
                }
                continue;
            }
            const int64_t nvx = (int64_t)a.nx * (a.ny + 1), nvy = (int64_t)(a.nx + 1) * a.ny,
                          np = (int64_t)(a.nx - 1) * (a.ny - 1);
            for (int q = 0; q < 2; ++q) {
                HIPCHK(launch_fill_f32(b.vx[q], nvx, 0.f, s));
                HIPCHK(launch_fill_f32(b.vy[q], nvy, 0.f, s));
                HIPCHK(launch_fill_f32(b.p[q], np, 0.f, s));
            }
            const float dtf = (float)r.dt_k;
            if (cdesc[CDESC_PER_INST * (size_t)r.id].dt != dtf) {
                for (int q = 0; q < CDESC_PER_INST; ++q) cdesc[CDESC_PER_INST * (size_t)r.id + q].dt = dtf;
                env.hdt[r.id] = dtf;
                for (int q = 0; q < CDESC_PER_INST; ++q)
                    HIPCHK(hipMemcpyAsync((char*)(env.cdesc_dev + CDESC_PER_INST * (size_t)r.id + q) +
                                              offsetof(CorrParams, dt),
                                          env.hdt + r.id, sizeof(float), hipMemcpyHostToDevice, s));
            }
            int64_t steps = std::min(r.window, a.iter_max - r.n32);
            steps = std::min(steps, r.ncheck - r.n32 % r.ncheck);
            if (a.use_fss) steps = std::min(steps, r.ndtupdate - r.n32 % r.ndtupdate);
            r.left = steps;
            r.wstart = r.n32;
            // both correction buffers are zero, so the window starts at the momentum's parity
            return SF_OK;
        }
    };
    auto end_window = [&](Run& r) -> int {
        const sf_batch_instance& a = in[r.id];
        const BatchInstance& d = desc[(size_t)r.id];
        const RefineBufs& b = rb[(size_t)r.id];
        HIPCHK(launch_update_f64(a.Vx_s, b.vx[r.cur], (int64_t)a.nx * (a.ny + 1), s));
        HIPCHK(launch_update_f64(a.Vy_s, b.vy[r.cur], (int64_t)(a.nx + 1) * a.ny, s));
        HIPCHK(launch_update_f64(a.P, b.p[r.cur], (int64_t)(a.nx - 1) * (a.ny - 1), s));
        r.rebases += 1;
        // the reference refreshes the stabilisation dt every ndtupdate iterations (:60-62)
        if (a.use_fss && (r.n32 / r.ndtupdate) > (r.wstart / r.ndtupdate)) {
            int rc = enqueue_dt_raw(s, a.nx, a.ny, a.Vx_s, a.Vy_s, a.dx, a.dy, a.maxdisp, 0, b.part, d.scal);
            if (rc) return rc;
            double* hs = env.w.hscal + (size_t)SC_COUNT * r.id;
            HIPCHK(hipMemcpyAsync(hs, d.scal, SC_COUNT * sizeof(double), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            r.dt_k = hs[SC_DT_CHECK];
        }
        return SF_OK;
    };
    auto finish = [&](const Run& r) -> int {
        const sf_batch_instance& a = in[r.id];
        const BatchInstance& d = desc[(size_t)r.id];
        const RefineBufs& b = rb[(size_t)r.id];
        if (!r.converged) {
            HIPCHK(launch_cast_f64(a.dVxdtau, r.cur ? b.mx2 : b.mx, (int64_t)(a.nx - 2) * (a.ny - 1), s));
            HIPCHK(launch_cast_f64(a.dVydtau, r.cur ? b.my2 : b.my, (int64_t)(a.nx - 1) * (a.ny - 2), s));
        }
        HIPCHK(launch_pad_mirror(a.nx, a.ny, a.Vx_s, a.Vy_s, a.Vx_pad, a.Vy_pad, s));
        int rc = enqueue_dt_raw(s, a.nx, a.ny, a.Vx_s, a.Vy_s, a.dx, a.dy, a.maxdisp, 0, b.part, d.scal);
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(env.w.hscal + (size_t)SC_COUNT * r.id, d.scal, SC_COUNT * sizeof(double),
                              hipMemcpyDeviceToHost, s));
        BatchResult& o = res[(size_t)r.id];
        o.err = (!r.converged && std::isfinite(r.err_k)) ? r.err_k : r.err_ref;
        o.niter = r.n32 + r.n64;
        o.converged = r.converged;
        o.dt_last = r.dt_k;
        return SF_OK;
    };
    std::vector<Run> act;
    size_t q = 0;
    double ws_active = 0.0;
    WideCorrArgs wa, w2, w3;
    std::memset(&wa, 0, sizeof(wa));
    std::memset(&w2, 0, sizeof(w2));
    std::memset(&w3, 0, sizeof(w3));
    wa.desc = env.cdesc_dev;
    w2.desc = env.cdesc_dev;
    w3.desc = env.cdesc_dev;
    int rc = SF_OK;
    while (!act.empty() || q < order.size()) {
        while (q < order.size() && (int)act.size() < WIDE_MAX &&
               (act.empty() || ws_active + refine_working_set(in[order[q]].nx, in[order[q]].ny) <= budget)) {
            const int id = order[q++];
            const sf_batch_instance& a = in[id];
            const RefineBufs& b = rb[(size_t)id];
            const int mn = std::min(a.nx, a.ny);
            Run r{};
            r.id = id;
            r.ncheck = std::min<int64_t>(5 * (int64_t)mn, 2000);
            r.ndtupdate = 10 * (int64_t)mn;
            r.window = env.rebase > 0 ? env.rebase : r.ncheck;
            r.checked_at = -1;
            r.err_ref = 2.0 * a.eps;
            r.dt_k = a.use_fss ? a.dt : 0.0;
            r.ws = refine_working_set(a.nx, a.ny);
            HIPCHK(launch_cast_f32(b.rho, a.rho_vy, (int64_t)(a.nx + 1) * a.ny, s));
            HIPCHK(launch_cast_f32(b.mub, a.mu_b, (int64_t)a.nx * a.ny, s));
            HIPCHK(launch_cast_f32(b.mup, a.mu_p, (int64_t)(a.nx - 1) * (a.ny - 1), s));
            HIPCHK(launch_cast_f32(b.mx, a.dVxdtau, (int64_t)(a.nx - 2) * (a.ny - 1), s));
            HIPCHK(launch_cast_f32(b.my, a.dVydtau, (int64_t)(a.nx - 1) * (a.ny - 2), s));
            bool done = false;
            rc = begin_window(r, &done);
            if (rc) return rc;
            if (done) {
                rc = finish(r);
                if (rc) return rc;
            } else {
                act.push_back(r);
                ws_active += r.ws;
            }
        }
        if (act.empty()) continue;
        // three (or two) sweeps per pass for every instance with that many left in its window
        wa.n_active = 0;
        w2.n_active = 0;
        w3.n_active = 0;
        int tiles = 0, tiles2 = 0, tiles3 = 0;
        for (size_t k = 0; k < act.size(); ++k) {
            const size_t base = CDESC_PER_INST * (size_t)act[k].id;
            if (act[k].left >= 3) {
                w3.tile_start[w3.n_active] = tiles3;
                w3.slot_desc[w3.n_active] = (int)base + 4 + act[k].cur;
                tiles3 += cdesc[base + 4].ntiles;
                w3.n_active += 1;
            } else if (act[k].left >= 2) {
                w2.tile_start[w2.n_active] = tiles2;
                w2.slot_desc[w2.n_active] = (int)base + 2 + act[k].cur;
                tiles2 += cdesc[base + 2].ntiles;
                w2.n_active += 1;
            } else {
                wa.tile_start[wa.n_active] = tiles;
                wa.slot_desc[wa.n_active] = (int)base + act[k].cur;
                tiles += cdesc[base].ntiles;
                wa.n_active += 1;
            }
        }
        wa.tile_start[wa.n_active] = tiles;
        w2.tile_start[w2.n_active] = tiles2;
        w3.tile_start[w3.n_active] = tiles3;
        if (w3.n_active) HIPCHK(launch_wide_corr3(w3, s));
        if (w2.n_active) HIPCHK(launch_wide_corr2(w2, s));
        if (wa.n_active) HIPCHK(launch_wide_corr(wa, s));
        std::vector<Run> keep;
        keep.reserve(act.size());
        for (Run& r : act) {
            const int64_t adv = r.left >= 3 ? 3 : (r.left >= 2 ? 2 : 1);
            r.cur ^= 1;
            r.n32 += adv;
            r.left -= adv;
            bool done = false;
            if (r.left == 0) {
                rc = end_window(r);
                if (rc) return rc;
                rc = begin_window(r, &done);
                if (rc) return rc;
            }
            if (done) {
                rc = finish(r);
                if (rc) return rc;
                ws_active -= r.ws;
            } else {
                keep.push_back(r);
            }
        }
        act.swap(keep);
    }
    HIPCHK(hipStreamSynchronize(s));
    for (int i = 0; i < n; ++i) res[(size_t)i].dt_new = env.w.hscal[(size_t)SC_COUNT * i + SC_DT_CHECK];
    return SF_OK;
}

}  // namespace sf

extern "C" int sf_solve_stokes_batched(sf_ctx* c, const sf_batch_instance* in, int n, const sf_batch_params* prm,
                                       sf_solve_out* outs)
{
    const char* who = "sf_solve_stokes_batched";
    if (!c || (n > 0 && (!in || !outs))) return set_err(SF_ERR_ARG, "%s: NULL argument", who);
    if (n < 0) return set_err(SF_ERR_ARG, "%s: n < 0", who);
    if (n == 0) return SF_OK;
    const auto t_start = std::chrono::steady_clock::now();
    hipStream_t s = c->stream;
    int threads, rows;
    int rc = batch_threads(prm, &threads, &rows, who);
    if (rc) return rc;
    std::vector<size_t> off((size_t)n);
    size_t need = 0;
    for (int i = 0; i < n; ++i) {
        rc = batch_check(in[i], i, who);
        if (rc) return rc;
        off[(size_t)i] = need;
        need += batch_scratch_doubles(in[i].nx, in[i].ny);
    }
    if (need > c->b_pool_cap) {
        if (c->b_pool) HIPCHK(hipFree(c->b_pool));
        c->b_pool = nullptr;
        c->b_pool_cap = 0;
        HIPCHK(dalloc(&c->b_pool, need));
        c->b_pool_cap = need;
    }
    if (n > c->b_cap) {
        for (void* p : {(void*)c->b_inst, (void*)c->b_order, (void*)c->b_queue, (void*)c->b_res})
            if (p) HIPCHK(hipFree(p));
        c->b_inst = nullptr;
        c->b_order = nullptr;
        c->b_queue = nullptr;
        c->b_res = nullptr;
        c->b_cap = 0;
        HIPCHK(dalloc(&c->b_inst, (size_t)n));
        HIPCHK(dalloc(&c->b_order, (size_t)n));
        HIPCHK(dalloc(&c->b_queue, 1));
        HIPCHK(dalloc(&c->b_res, (size_t)n));
        c->b_cap = n;
    }
    const int mode = prm ? prm->mode : SF_BATCH_AUTO;
    const int precision = prm ? prm->precision : 0;
    std::vector<BatchInstance> host((size_t)n);
    for (int i = 0; i < n; ++i) {
        BatchShape sh = mode == SF_BATCH_GRID ? grid_shape(in[i].nx, in[i].ny) : batch_shape(c, in[i].nx, in[i].ny);
        if (mode != SF_BATCH_GRID) sh.copies = n;
        host[(size_t)i] = batch_descriptor(in[i], rows, c->b_pool + off[(size_t)i], nullptr, sh);
    }
    std::vector<std::pair<int, int>> dims((size_t)n);
    for (int i = 0; i < n; ++i) dims[(size_t)i] = {in[i].nx, in[i].ny};
    const std::vector<int> order = batch_order(dims);
    HIPCHK(hipMemcpyAsync(c->b_inst, host.data(), sizeof(BatchInstance) * (size_t)n, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->b_order, order.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, s));
    std::vector<BatchResult> res((size_t)n);
    if (c->b_hscal_cap < n) {
        if (c->b_hscal) HIPCHK(hipHostFree(c->b_hscal));
        c->b_hscal = nullptr;
        c->b_hscal_cap = 0;
        HIPCHK(hipHostMalloc((void**)&c->b_hscal, sizeof(double) * SC_COUNT * (size_t)n, hipHostMallocDefault));
        c->b_hscal_cap = n;
    }
    if (precision != 0 && precision != 1) return set_err(SF_ERR_ARG, "%s: precision %d not 0 or 1", who, precision);
    if (precision == 1 && (mode == SF_BATCH_PERSISTENT || mode == SF_BATCH_GRID))
        return set_err(SF_ERR_ARG, "%s: FP32 refinement runs in the wide mode only", who);
    if (precision == 1) {
        // the refined working sets are FP32; 256 mixed 64..512 instances: 11.10 s at 256 MiB, 11.07 s at
        // 512, 10.44-10.47 s at 1024 and 2048 MiB (one box, one process)
        const double budget = (prm && prm->budget_mb > 0.0) ? prm->budget_mb * 1048576.0 : 1024.0 * 1048576.0;
        size_t nf = 0;
        std::vector<size_t> foff((size_t)n);
        for (int i = 0; i < n; ++i) {
            foff[(size_t)i] = nf;
            nf += refine_floats(in[i].nx, in[i].ny);
        }
        rc = batch_refine_reserve(c, n, nf);
        if (rc) return rc;
        std::vector<RefineBufs> rb((size_t)n);
        std::vector<CorrParams> cdesc(CDESC_PER_INST * (size_t)n);
        for (int i = 0; i < n; ++i) {
            rb[(size_t)i] = refine_carve(in[i].nx, in[i].ny, c->b_pool32 + foff[(size_t)i],
                                         c->b_part + (size_t)3 * RED_MAX_BLOCKS * i);
            refine_descriptors(in[i], rb[(size_t)i], prm ? prm->rows : 0, &cdesc[CDESC_PER_INST * (size_t)i]);
        }
        HIPCHK(hipMemcpyAsync(c->b_cdesc, cdesc.data(), sizeof(CorrParams) * CDESC_PER_INST * (size_t)n,
                              hipMemcpyHostToDevice, s));
        const RefineEnv env{{s, c->b_inst, c->part, c->b_hscal}, &rb, c->b_cdesc, c->b_hpart, c->b_hdt, 0};
        rc = batch_run_wide_refined(env, in, n, host, cdesc, order, budget, res);
        if (rc) return rc;
    } else if (mode == SF_BATCH_PERSISTENT) {
        const int nblocks = batch_grid(c, n, threads, prm ? prm->max_blocks : 0);
        HIPCHK(launch_batch_solve(c->b_inst, c->b_order, n, c->b_queue, c->b_res, threads, nblocks, s));
        HIPCHK(hipMemcpyAsync(res.data(), c->b_res, sizeof(BatchResult) * (size_t)n, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    } else if (mode == SF_BATCH_GRID) {
        // one instance after another, each solved by one launch over the whole chip
        rc = grid_reserve(c, n);
        if (rc) return rc;
        const int k = (prm && prm->depth >= 1 && prm->depth <= 3) ? prm->depth : 3;
        HIPCHK(hipMemsetAsync(c->g_sync, 0, sizeof(unsigned long long) * GRID_SYNC_WORDS * (size_t)n, s));
        const double wait_s = grid_wait_seconds();
        const uint64_t wait_ticks = grid_wait_limit();
        for (int qi = 0; qi < n; ++qi) {
            const int i = order[(size_t)qi];
            const GridShape gs = grid_launch_shape(host[(size_t)i], k);
            const GridArgs ga{c->b_inst + i, c->b_res + i, c->g_sync + (size_t)GRID_SYNC_WORDS * i, c->g_part,
                              gs.nwg, gs.waves, wait_ticks};
            HIPCHK(launch_grid_solve(ga, k, s));
        }
        std::vector<unsigned long long> sync((size_t)GRID_SYNC_WORDS * n);
        HIPCHK(hipMemcpyAsync(res.data(), c->b_res, sizeof(BatchResult) * (size_t)n, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(sync.data(), c->g_sync, sizeof(unsigned long long) * GRID_SYNC_WORDS * (size_t)n,
                              hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        for (int i = 0; i < n; ++i)
            if (sync[(size_t)GRID_SYNC_WORDS * i + GRID_SYNC_STICKY])
                return set_err(SF_ERR_HIP, "%s: instance %d: a grid barrier wait exceeded %g s (workgroups not all resident?)",
                               who, i, wait_s);
    } else if (mode == SF_BATCH_WIDE || mode == SF_BATCH_AUTO) {
        // a working set past the Infinity Cache still pays with multi-iteration passes: 256 mixed
        // 64..512 instances, one box, one process: 17.69 s at 256 MiB, 16.93 at 384, 16.24 at 768, 15.35-
        // 15.38 at 1536 and 3072 MiB (64 slots full of 512^2 instances: 18,432 waves, nine full rounds of
        // the 2,048 resident ones); 128 slots were slower (15.5 s)
        const double budget = (prm && prm->budget_mb > 0.0) ? prm->budget_mb * 1048576.0 : 1536.0 * 1048576.0;
        const WideEnv env{s, c->b_inst, c->part, c->b_hscal};
        rc = batch_run_wide(env, in, n, host, order, budget, (prm && prm->depth > 1) ? 2 : 1, res);
        if (rc) return rc;
    } else {
        return set_err(SF_ERR_ARG, "%s: unknown mode %d", who, mode);
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    for (int i = 0; i < n; ++i) {
        const BatchResult& r = res[(size_t)i];
        sf_solve_out& o = outs[i];
        std::memset(&o, 0, sizeof(o));
        o.dt_new = r.dt_new;
        o.err = r.err;
        o.niter = r.niter;
        o.converged = r.converged;
        o.dt_last = r.dt_last;
        o.runtime = secs;  // the whole batch
    }
    return SF_OK;
}


// ---------------------------------------------------------------------------------------
// autotuner of the wide passes: per grid size, the multi-iteration pass (2, 3 or 4 PT iterations
// per launch) and its tile rows, timed over an active set of that size's instances (copies of
// the caller's problems in scratch allocated here), as the wide scheduler launches them

extern "C" int sf_batch_autotune(sf_ctx* c, const sf_batch_instance* in, int n, const sf_batch_params* prm, int iters,
                                 sf_batch_tuning* table, int table_cap, int* n_table)
{
    const char* who = "sf_batch_autotune";
    if (!c || (n > 0 && !in)) return set_err(SF_ERR_ARG, "%s: NULL argument", who);
    if (table_cap > 0 && !table) return set_err(SF_ERR_ARG, "%s: NULL table with table_cap > 0", who);
    if (n_table) *n_table = 0;
    if (iters < 1) iters = 30;
    for (int i = 0; i < n; ++i) {
        const int rc = batch_check(in[i], i, who);
        if (rc) return rc;
    }
    hipStream_t s = c->stream;
    const double budget = (prm && prm->budget_mb > 0.0) ? prm->budget_mb * 1048576.0 : 1536.0 * 1048576.0;
    std::vector<std::pair<int, int>> sizes;
    for (int i = 0; i < n; ++i)
        if (std::find(sizes.begin(), sizes.end(), std::make_pair(in[i].nx, in[i].ny)) == sizes.end())
            sizes.push_back({in[i].nx, in[i].ny});
    int nt = 0;
    std::vector<BatchTune> found;
    hipEvent_t ev[2] = {nullptr, nullptr};
    double* pool = nullptr;
    BatchInstance* dev_desc = nullptr;
    int rc = SF_OK;
    auto fail = [&](hipError_t e) { return set_err(SF_ERR_HIP, "%s: %s", who, hipGetErrorString(e)); };
    hipError_t e = hipEventCreate(&ev[0]);
    if (e == hipSuccess) e = hipEventCreate(&ev[1]);
    if (e == hipSuccess) e = dalloc(&dev_desc, (size_t)WIDE_MAX);
    if (e != hipSuccess) rc = fail(e);
    for (size_t z = 0; z < sizes.size() && rc == SF_OK; ++z) {
        const int nx = sizes[z].first, ny = sizes[z].second;
        std::vector<int> members;
        for (int i = 0; i < n; ++i)
            if (in[i].nx == nx && in[i].ny == ny) members.push_back(i);
        const double ws = batch_working_set(nx, ny);
        // an active set as the scheduler forms it: as many of this size as the budget holds
        const int R = std::max(1, std::min({(int)members.size(), (int)WIDE_MAX, (int)(budget / ws)}));
        const size_t gx = nx, gy = ny, np = (gx - 1) * (gy - 1), nvx = gx * (gy + 1), nvy = (gx + 1) * gy;
        const size_t ndx = (gx - 2) * (gy - 1), ndy = (gx - 1) * (gy - 2);
        const size_t lens[10] = {np, nvx, nvy, ndx, ndy, (gx + 2) * (gy + 1), (gx + 1) * (gy + 2), nvy, gx * gy, np};
        size_t per = batch_scratch_doubles(nx, ny);
        for (size_t l : lens) per += l;
        if (pool) (void)hipFree(pool);
        pool = nullptr;
        e = dalloc(&pool, per * (size_t)R);
        if (e != hipSuccess) {
            rc = fail(e);
            break;
        }
        // copies of the caller's instances of this size (state and materials)
        std::vector<sf_batch_instance> cp((size_t)R);
        std::vector<double*> scr((size_t)R);
        for (int j = 0; j < R && e == hipSuccess; ++j) {
            const sf_batch_instance& src = in[members[(size_t)j % members.size()]];
            double* q = pool + per * (size_t)j;
            sf_batch_instance a = src;
            double** dst[10] = {&a.P, &a.Vx_s, &a.Vy_s, &a.dVxdtau, &a.dVydtau, &a.Vx_pad, &a.Vy_pad, nullptr, nullptr,
                                nullptr};
            const double* from[10] = {src.P, src.Vx_s, src.Vy_s, src.dVxdtau, src.dVydtau, src.Vx_pad, src.Vy_pad,
                                      src.rho_vy, src.mu_b, src.mu_p};
            double* mat[3] = {nullptr, nullptr, nullptr};
            for (int t = 0; t < 10 && e == hipSuccess; ++t) {
                e = hipMemcpyAsync(q, from[t], lens[t] * sizeof(double), hipMemcpyDeviceToDevice, s);
                if (t < 7) *dst[t] = q;
                else mat[t - 7] = q;
                q += lens[t];
            }
            a.rho_vy = mat[0];
            a.mu_b = mat[1];
            a.mu_p = mat[2];
            cp[(size_t)j] = a;
            scr[(size_t)j] = q;
        }
        if (e != hipSuccess) {
            rc = fail(e);
            break;
        }
        BatchTune best{nx, ny, BatchShape()};
        double best_t[5] = {1e300, 1e300, 1e300, 1e300, 1e300};  // per k: ns per node and iteration
        int best_rows[5] = {0, 0, 0, 0, 0};
        const int rows_c[] = {4, 8, 12, 16, 24, 32, 48, 64};
        for (int k : {2, 3, 4}) {
            for (int rows : rows_c) {
                if (rows > ny && rows != rows_c[0]) continue;
                BatchShape sh;
                sh.kmax = k;
                sh.rows2 = rows;
                sh.rows3 = rows;
                sh.rows4 = rows;
                std::vector<BatchInstance> desc((size_t)R);
                for (int j = 0; j < R; ++j) {
                    cp[(size_t)j].dt = 0.0;
                    desc[(size_t)j] = batch_descriptor(cp[(size_t)j], 0, scr[(size_t)j], nullptr, sh);
                }
                e = hipMemcpyAsync(dev_desc, desc.data(), sizeof(BatchInstance) * (size_t)R, hipMemcpyHostToDevice, s);
                for (int j = 0; j < R && e == hipSuccess; ++j) e = launch_set_scalar(desc[(size_t)j].scal + SC_DT, 0.0, s);
                WideArgs w;
                std::memset(&w, 0, sizeof(w));
                w.desc = dev_desc;
                w.n_active = R;
                int tiles = 0;
                for (int j = 0; j < R; ++j) {
                    w.tile_start[j] = tiles;
                    w.slot_inst[j] = j;
                    tiles += desc[(size_t)j].wpt[2 * k - 2].ntiles;
                }
                w.tile_start[R] = tiles;
                const int reps = std::max(4, (iters + k - 1) / k) & ~1;  // even: the state ends where it began
                float ms = 0.f;
                for (int rep = -2; rep < reps && e == hipSuccess; ++rep) {  // two untimed launches first
                    if (rep == 0) e = hipEventRecord(ev[0], s);
                    for (int j = 0; j < R; ++j) w.slot_pf[j] = rep & 1;
                    if (e == hipSuccess)
                        e = (k == 4) ? launch_wide_pt4(w, s) : (k == 3) ? launch_wide_pt3(w, s) : launch_wide_pt2(w, s);
                }
                if (e == hipSuccess) e = hipEventRecord(ev[1], s);
                if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
                if (e != hipSuccess) break;
                const double t_ns = 1e6 * (double)ms / ((double)reps * k * R * (double)nx * ny);
                if (table && nt < table_cap) table[nt] = sf_batch_tuning{nx, ny, k, rows, R, 0, t_ns};
                ++nt;
                if (t_ns < best_t[k]) {
                    best_t[k] = t_ns;
                    best_rows[k] = rows;
                }
            }
            if (e != hipSuccess) break;
        }
        if (e != hipSuccess) {
            rc = fail(e);
            break;
        }
        best.sh.kmax = 2;
        for (int k = 3; k <= 4; ++k)
            if (best_t[k] <= best_t[best.sh.kmax]) best.sh.kmax = k;
        best.sh.rows2 = best_rows[2];
        best.sh.rows3 = best_rows[3];
        best.sh.rows4 = best_rows[4];
        found.push_back(best);
    }
    (void)hipStreamSynchronize(s);
    if (pool) (void)hipFree(pool);
    pool = nullptr;
    if (dev_desc) (void)hipFree(dev_desc);
    for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
    if (rc) return rc;

    // The choices are checked on the whole batch: every instance capped at a few hundred PT
    // iterations (copies of the caller's states; the materials are only read), once with the size
    // defaults and once per size with that size's choice added; a choice stays only where the
    // capped run gets faster, since one size's tiles share each launch with the others'.
    std::vector<BatchTune> saved = c->btune;
    for (const BatchTune& t : found)
        saved.erase(std::remove_if(saved.begin(), saved.end(),
                                   [&](const BatchTune& u) { return u.nx == t.nx && u.ny == t.ny; }),
                    saved.end());
    std::vector<sf_batch_instance> cp(in, in + n);
    std::vector<size_t> soff((size_t)n);
    size_t need = 0;
    for (int i = 0; i < n; ++i) {
        const size_t gx = in[i].nx, gy = in[i].ny;
        soff[(size_t)i] = need;
        need += (gx - 1) * (gy - 1) + gx * (gy + 1) + (gx + 1) * gy + (gx - 2) * (gy - 1) + (gx - 1) * (gy - 2) +
                (gx + 2) * (gy + 1) + (gx + 1) * (gy + 2);
    }
    hipError_t e2 = dalloc(&pool, need);
    if (e2 != hipSuccess) {
        c->btune = saved;
        return set_err(SF_ERR_HIP, "%s: %s", who, hipGetErrorString(e2));
    }
    const int64_t cap_iters = 240;
    auto capped_run = [&](double* secs) -> int {
        for (int i = 0; i < n && e2 == hipSuccess; ++i) {
            const size_t gx = in[i].nx, gy = in[i].ny;
            double* q = pool + soff[(size_t)i];
            const size_t len[7] = {(gx - 1) * (gy - 1), gx * (gy + 1), (gx + 1) * gy, (gx - 2) * (gy - 1),
                                   (gx - 1) * (gy - 2), (gx + 2) * (gy + 1), (gx + 1) * (gy + 2)};
            const double* from[7] = {in[i].P, in[i].Vx_s, in[i].Vy_s, in[i].dVxdtau, in[i].dVydtau, in[i].Vx_pad,
                                     in[i].Vy_pad};
            double** to[7] = {&cp[(size_t)i].P, &cp[(size_t)i].Vx_s, &cp[(size_t)i].Vy_s, &cp[(size_t)i].dVxdtau,
                              &cp[(size_t)i].dVydtau, &cp[(size_t)i].Vx_pad, &cp[(size_t)i].Vy_pad};
            for (int t = 0; t < 7 && e2 == hipSuccess; ++t) {
                e2 = hipMemcpyAsync(q, from[t], len[t] * sizeof(double), hipMemcpyDeviceToDevice, s);
                *to[t] = q;
                q += len[t];
            }
            cp[(size_t)i].iter_max = std::min<int64_t>(in[i].iter_max, cap_iters);
        }
        if (e2 == hipSuccess) e2 = hipStreamSynchronize(s);
        if (e2 != hipSuccess) return set_err(SF_ERR_HIP, "%s: %s", who, hipGetErrorString(e2));
        sf_batch_params wp = prm ? *prm : sf_batch_params{};
        wp.mode = SF_BATCH_WIDE;
        wp.precision = 0;
        std::vector<sf_solve_out> outs((size_t)n);
        const auto t0 = std::chrono::steady_clock::now();
        const int r = sf_solve_stokes_batched(c, cp.data(), n, &wp, outs.data());
        *secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return r;
    };
    std::vector<BatchTune> kept;
    double t_best = 0.0;
    c->btune = saved;
    rc = capped_run(&t_best);
    for (size_t z = 0; z < found.size() && rc == SF_OK; ++z) {
        c->btune = saved;
        c->btune.insert(c->btune.end(), kept.begin(), kept.end());
        c->btune.push_back(found[z]);
        double t = 0.0;
        rc = capped_run(&t);
        if (rc == SF_OK && t < 0.99 * t_best) {
            kept.push_back(found[z]);
            t_best = t;
        }
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(pool);
    c->btune = saved;
    if (rc) return rc;
    c->btune.insert(c->btune.end(), kept.begin(), kept.end());
    for (int q = 0; q < nt && q < table_cap; ++q)  // mark the candidates the check kept
        for (const BatchTune& t : kept)
            if (table[q].nx == t.nx && table[q].ny == t.ny && table[q].iters_per_launch == t.sh.kmax &&
                table[q].rows == (t.sh.kmax == 4 ? t.sh.rows4 : t.sh.kmax == 3 ? t.sh.rows3 : t.sh.rows2))
                table[q].kept = 1;
    if (n_table) *n_table = nt;
    return SF_OK;
}

extern "C" int sf_batch_tuning_clear(sf_ctx* c)
{
    if (!c) return set_err(SF_ERR_ARG, "sf_batch_tuning_clear: NULL ctx");
    c->btune.clear();
    return SF_OK;
}
