// sf_noalloc.hip -- sf_solve_stokes_noalloc: the FP64 solve of sf_solve_stokes with its scratch
// (PT ping-pong state, check-iteration scratch, reduction partials, device scalars) carved from
// one static device array instead of a context, for callers that cannot allocate device memory.
// It calls no hipMalloc / hipFree / hipHostMalloc and creates no events.
#include <mutex>

#include "sf_host.h"

using namespace sf;

// Scratch of the no-allocation solve: one copy per GPU, held until the process ends.
static __device__ __align__(256) unsigned char scratch[SF_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_dev[kMaxDevices] = {};  // device address of `scratch`, per device

// The current device's address of `scratch`.  In host code the array's name is only a stand-in
// for the symbol; its device address comes from hipGetSymbolAddress, looked up once per device
// with that device current.
int scratch_base(unsigned char** out)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return set_err(SF_ERR_ARG, "sf_solve_stokes_noalloc: device %d outside 0..%d", dev, kMaxDevices - 1);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_err(SF_ERR_HIP, "sf_solve_stokes_noalloc: hipGetSymbolAddress(scratch) failed: %s",
                           hipGetErrorString(e));
        g_scratch_dev[dev] = (unsigned char*)p;
    }
    *out = g_scratch_dev[dev];
    return SF_OK;
}

// Bump carving of 256-byte aligned float64 buffers; with base == nullptr it only counts bytes.
struct Carver {
    unsigned char* base;
    size_t used = 0;
    double* take(size_t n)
    {
        double* p = base ? (double*)(base + used) : nullptr;
        used += (std::max<size_t>(n, 1) * sizeof(double) + 255) & ~(size_t)255;
        return p;
    }
};

// Lays the solve's scratch out in c: ping-pong state, the second dV (multi-iteration passes),
// check-iteration arrays the caller did not pass, partials and scalars.
void carve(Carver& c, int nx, int ny, bool no_txx, bool no_tyy, bool no_txy, bool no_divv, bool no_rx, bool no_ry,
           SolveScratch* w)
{
    const size_t gx = nx, gy = ny;
    const size_t nvx = gx * (gy + 1), nvy = (gx + 1) * gy, np = (gx - 1) * (gy - 1);
    const size_t ndx = (gx - 2) * (gy - 1), ndy = (gx - 1) * (gy - 2);
    w->pt.vx_b = c.take(nvx);
    w->pt.vy_b = c.take(nvy);
    w->pt.p_b = c.take(np);
    w->pt.dvx_b = c.take(ndx);
    w->pt.dvy_b = c.take(ndy);
    w->txx = no_txx ? c.take(np) : nullptr;
    w->tyy = no_tyy ? c.take(np) : nullptr;
    w->txy = no_txy ? c.take(gx * gy) : nullptr;
    w->divv = no_divv ? c.take(np) : nullptr;
    w->rx = no_rx ? c.take(ndx) : nullptr;
    w->ry = no_ry ? c.take(ndy) : nullptr;
    w->part = c.take(3 * (size_t)RED_MAX_BLOCKS);
    w->scal = c.take((size_t)SC_COUNT);
}

}  // namespace

extern "C" int sf_solve_stokes_noalloc(void* hip_stream, double* P, double* Vx_pad, double* Vy_pad,
                                       const double* rho_vy, const double* mu_b, const double* mu_p, double* txx,
                                       double* tyy, double* txy, double* divV, double* dtauP, double* Rx, double* Ry,
                                       double* dVx, double* dVy, double* dtauVx, double* dtauVy, double* Vx_s,
                                       double* Vy_s, double g_y, double dx, double dy, int nx, int ny, double dt,
                                       double maxdisp, int use_fss, double eps, int64_t iter_max, int print_info,
                                       sf_solve_out* out)
{
    const char* who = "sf_solve_stokes_noalloc";
    if (nx < 3 || ny < 3) return set_err(SF_ERR_ARG, "%s: need Nx, Ny >= 3 (got %d, %d)", who, nx, ny);
    if ((int64_t)(nx + 2) * (ny + 2) >= (int64_t)1 << 31)
        return set_err(SF_ERR_SIZE, "%s: grid %d x %d exceeds the 32-bit cell index", who, nx, ny);
    SolveScratch w{};
    Carver need{nullptr};
    carve(need, nx, ny, !txx, !tyy, !txy, !divV, !Rx, !Ry, &w);
    if (need.used > SF_NOALLOC_SCRATCH_BYTES)
        return set_err(SF_ERR_SIZE,
                       "%s: a %d x %d solve needs %zu bytes of scratch, more than the static %zu (pass the txx, tyy, "
                       "txy, divV, Rx, Ry arrays to need less, or use sf_solve_stokes)",
                       who, nx, ny, need.used, (size_t)SF_NOALLOC_SCRATCH_BYTES);
    unsigned char* base = nullptr;
    int rc = scratch_base(&base);
    if (rc) return rc;
    Carver cut{base};
    carve(cut, nx, ny, !txx, !tyy, !txy, !divV, !Rx, !Ry, &w);
    double hscal[SC_COUNT] = {};
    w.s = (hipStream_t)hip_stream;
    w.hscal = hscal;
    w.ev0 = w.ev1 = nullptr;  // timing window on the host clock
    w.seg = 0;                // launch shapes: the size defaults of a new context
    w.depth = 0;
    w.max_waves = -1;
    w.fuse = 0;
    w.rows = 0;
    w.tm = nullptr;
    return solve_core(w, who, P, Vx_pad, Vy_pad, rho_vy, mu_b, mu_p, txx, tyy, txy, divV, dtauP, Rx, Ry, dVx, dVy,
                      dtauVx, dtauVy, Vx_s, Vy_s, g_y, dx, dy, nx, ny, dt, maxdisp, use_fss, eps, iter_max,
                      print_info, out);
}
