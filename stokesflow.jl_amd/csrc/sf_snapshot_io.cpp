// sf_snapshot_io.cpp -- snapshot files of the timestep loop: saveStats, saveMarkers and saveGrid
// (StokesFlow2D_multixpu.jl:212-299), the files viz.jl reads back (viz.jl:11-161).
//
// Format: MAT-file level 5 (MathWorks' published "MAT-File Format" description: a 128-byte
// header, then one miMATRIX data element per variable with array flags, dimensions, name and
// the real part, every subelement padded to 8 bytes).  The reference writes its files through
// MAT.jl, which writes the HDF5-based v7.3 layout and reads level 5 as well, so viz.jl reads
// these files unchanged; level 5 needs no HDF5 library.  Variable names, classes (Float32 for
// the fields, Int64 for dims and nt), shapes (vectors as n x 1, grids as nx x ny, column-major)
// and values (FP64 -> FP32 round to nearest, marker shift added in FP64 first) follow the
// reference.
//
// Large arrays stream from the device in chunks: the FP32 cast of chunk k+1 and its copy into
// pinned memory run on the context's stream while chunk k is written to the file.
#include "sf_host.h"

#include <cstdint>
#include <string>

using namespace sf;

namespace {

enum { miINT8 = 1, miINT32 = 5, miUINT32 = 6, miSINGLE = 7, miINT64 = 12, miMATRIX = 14 };
enum { mxSINGLE_CLASS = 7, mxINT64_CLASS = 14 };

struct MatFile {
    FILE* f = nullptr;
    bool ok = true;
    ~MatFile()
    {
        if (f) fclose(f);
    }
    void put(const void* p, size_t n)
    {
        if (ok && n && fwrite(p, 1, n, f) != n) ok = false;
    }
    void u32(uint32_t v) { put(&v, 4); }
    void pad(size_t n)
    {
        static const uint8_t z[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        put(z, (8 - n % 8) % 8);
    }
};

size_t pad8(size_t n) { return (n + 7) & ~(size_t)7; }

int mat_open(MatFile& m, const char* path, const char* who)
{
    if (!path) return set_err(SF_ERR_ARG, "%s: path is NULL", who);
    m.f = fopen(path, "wb");
    if (!m.f) return set_err(SF_ERR_ARG, "%s: cannot open %s for writing", who, path);
    char text[116];
    std::memset(text, ' ', sizeof(text));
    const char* desc = "MATLAB 5.0 MAT-file, written by libsfhip (StokesFlow snapshot)";
    std::memcpy(text, desc, std::strlen(desc));
    m.put(text, sizeof(text));
    const uint8_t subsys[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    m.put(subsys, 8);
    const uint16_t version = 0x0100;
    m.put(&version, 2);
    m.put("IM", 2);  // little-endian files read 'MI' as a 16-bit value
    return SF_OK;
}

// Writes the miMATRIX header of a real numeric variable (everything up to the real part's
// data); the caller then writes data_bytes bytes of data and the padding.
void mat_var_header(MatFile& m, const char* name, int cls, int mi_type, int32_t d0, int32_t d1, size_t data_bytes)
{
    const size_t nlen = std::strlen(name);
    const size_t name_bytes = nlen <= 4 ? 8 : 8 + pad8(nlen);
    const size_t body = 16 /* flags */ + 16 /* 2 dims */ + name_bytes + 8 + pad8(data_bytes);
    m.u32(miMATRIX);
    m.u32((uint32_t)body);
    m.u32(miUINT32);  // array flags
    m.u32(8);
    m.u32((uint32_t)cls);
    m.u32(0);
    m.u32(miINT32);  // dimensions
    m.u32(8);
    m.put(&d0, 4);
    m.put(&d1, 4);
    if (nlen <= 4) {  // small data element: size and type in one word
        m.u32((uint32_t)(nlen << 16) | miINT8);
        char b[4] = {0, 0, 0, 0};
        std::memcpy(b, name, nlen);
        m.put(b, 4);
    } else {
        m.u32(miINT8);
        m.u32((uint32_t)nlen);
        m.put(name, nlen);
        m.pad(nlen);
    }
    m.u32((uint32_t)mi_type);
    m.u32((uint32_t)data_bytes);
}

int check_size(int64_t d0, int64_t d1, size_t elem, const char* name, const char* who)
{
    if (d0 < 0 || d1 < 0 || d0 > INT32_MAX || d1 > INT32_MAX)
        return set_err(SF_ERR_SIZE, "%s: %s: dimensions %lld x %lld exceed the format's 32-bit limit", who, name,
                       (long long)d0, (long long)d1);
    const double bytes = (double)d0 * (double)d1 * (double)elem + 256.0;
    if (bytes >= 4294967296.0)
        return set_err(SF_ERR_SIZE, "%s: %s: %.0f bytes exceed the format's 2^32-byte element limit", who, name,
                       bytes);
    return SF_OK;
}

// Streams a device FP64 array as an FP32 variable: cast + D2H of chunk k+1 overlap the file
// write of chunk k (two device and two pinned staging buffers).
struct Stager {
    static constexpr int64_t CH = (int64_t)1 << 23;  // floats per chunk (32 MiB)
    float* dev[2] = {nullptr, nullptr};
    float* host[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    ~Stager()
    {
        for (int b = 0; b < 2; ++b) {
            if (dev[b]) (void)hipFree(dev[b]);
            if (host[b]) (void)hipHostFree(host[b]);
            if (done[b]) (void)hipEventDestroy(done[b]);
        }
    }
    int init(const char* who)
    {
        for (int b = 0; b < 2; ++b) {
            hipError_t e = dalloc(&dev[b], (size_t)CH);
            if (e == hipSuccess) e = hipHostMalloc((void**)&host[b], sizeof(float) * (size_t)CH, hipHostMallocDefault);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&done[b], hipEventDisableTiming);
            if (e != hipSuccess) return set_err(SF_ERR_HIP, "%s: staging buffers: %s", who, hipGetErrorString(e));
        }
        return SF_OK;
    }
    int write(MatFile& m, hipStream_t s, const double* src, int64_t n, double shift, const char* who)
    {
        const int64_t nch = (n + CH - 1) / CH;
        auto enqueue = [&](int64_t k) -> int {
            const int b = (int)(k & 1);
            const int64_t off = k * CH, len = std::min(CH, n - off);
            HIPCHK(launch_snap_cast(src + off, dev[b], len, shift, s));
            HIPCHK(hipMemcpyAsync(host[b], dev[b], sizeof(float) * (size_t)len, hipMemcpyDeviceToHost, s));
            HIPCHK(hipEventRecord(done[b], s));
            return SF_OK;
        };
        int rc = nch > 0 ? enqueue(0) : SF_OK;
        for (int64_t k = 0; k < nch && rc == SF_OK; ++k) {
            if (k + 1 < nch) rc = enqueue(k + 1);
            if (rc) break;
            const int b = (int)(k & 1);
            const hipError_t e = hipEventSynchronize(done[b]);
            if (e != hipSuccess) return set_err(SF_ERR_HIP, "%s: %s", who, hipGetErrorString(e));
            m.put(host[b], sizeof(float) * (size_t)std::min(CH, n - k * CH));
        }
        // the buffer of chunk k+1 is reused by chunk k+3 only after its event was waited on
        return rc;
    }
};

int put_device_var(MatFile& m, Stager& st, hipStream_t s, const char* name, const double* src, int64_t d0, int64_t d1,
                   double shift, const char* who)
{
    int rc = check_size(d0, d1, 4, name, who);
    if (rc) return rc;
    const size_t bytes = (size_t)d0 * (size_t)d1 * 4;
    mat_var_header(m, name, mxSINGLE_CLASS, miSINGLE, (int32_t)d0, (int32_t)d1, bytes);
    rc = st.write(m, s, src, d0 * d1, shift, who);
    if (rc) return rc;
    m.pad(bytes);
    return SF_OK;
}

// A host coordinate vector (n x 1, FP32) with x_k = a*k + b evaluated as the reference does.
template <class F>
void put_coords(MatFile& m, const char* name, int n, F f)
{
    std::vector<float> v((size_t)n);
    for (int k = 0; k < n; ++k) v[(size_t)k] = (float)f(k + 1);  // 1-based index, as in Julia
    mat_var_header(m, name, mxSINGLE_CLASS, miSINGLE, n, 1, sizeof(float) * (size_t)n);
    m.put(v.data(), sizeof(float) * (size_t)n);
    m.pad(sizeof(float) * (size_t)n);
}

int finish(MatFile& m, const char* path, const char* who)
{
    if (!m.ok) return set_err(SF_ERR_ARG, "%s: writing %s failed", who, path);
    FILE* f = m.f;
    m.f = nullptr;
    if (fclose(f) != 0) return set_err(SF_ERR_ARG, "%s: closing %s failed", who, path);
    return SF_OK;
}

}  // namespace

extern "C" int sf_save_stats(const char* path, const int64_t dims[3], int64_t nt)
{
    const char* who = "sf_save_stats";
    if (!dims) return set_err(SF_ERR_ARG, "%s: dims is NULL", who);
    MatFile m;
    int rc = mat_open(m, path, who);
    if (rc) return rc;
    mat_var_header(m, "dims", mxINT64_CLASS, miINT64, 3, 1, 24);
    m.put(dims, 24);
    mat_var_header(m, "nt", mxINT64_CLASS, miINT64, 1, 1, 8);
    m.put(&nt, 8);
    return finish(m, path, who);
}

extern "C" int sf_save_markers(sf_ctx* c, const char* path, const double* x_m, const double* y_m, const double* rho_m,
                               int64_t nm, double x_shift, double y_shift)
{
    const char* who = "sf_save_markers";
    if (!c) return set_err(SF_ERR_ARG, "%s: NULL ctx", who);
    if (nm < 0) return set_err(SF_ERR_ARG, "%s: nm < 0", who);
    if (nm > 0 && (!x_m || !y_m || !rho_m)) return set_err(SF_ERR_ARG, "%s: a marker array is NULL", who);
    int rc = check_size(nm, 1, 4, "x_m", who);
    if (rc) return rc;
    MatFile m;
    rc = mat_open(m, path, who);
    if (rc) return rc;
    Stager st;
    rc = st.init(who);
    if (rc) return rc;
    hipStream_t s = c->stream;
    rc = put_device_var(m, st, s, "x_m", x_m, nm, 1, x_shift, who);
    if (!rc) rc = put_device_var(m, st, s, "y_m", y_m, nm, 1, y_shift, who);
    if (!rc) rc = put_device_var(m, st, s, "rho_m", rho_m, nm, 1, 0.0, who);
    if (rc) return rc;
    return finish(m, path, who);
}

extern "C" int sf_save_grid(sf_ctx* c, const char* path, int nx, int ny, double dx, double dy, const double* mu_b,
                            const double* P, const double* Vx_pad, const double* Vy_pad)
{
    const char* who = "sf_save_grid";
    if (!c) return set_err(SF_ERR_ARG, "%s: NULL ctx", who);
    if (nx < 3 || ny < 3) return set_err(SF_ERR_ARG, "%s: need Nx, Ny >= 3 (got %d, %d)", who, nx, ny);
    if (!mu_b || !P || !Vx_pad || !Vy_pad) return set_err(SF_ERR_ARG, "%s: a grid array is NULL", who);
    int rc = check_size(nx + 2, ny + 2, 4, "Vx", who);
    if (rc) return rc;
    MatFile m;
    rc = mat_open(m, path, who);
    if (rc) return rc;
    Stager st;
    rc = st.init(who);
    if (rc) return rc;
    hipStream_t s = c->stream;
    // node coordinates of StokesFlow2D_multixpu.jl:117-124 (local, 1-based index i)
    put_coords(m, "x", nx, [&](int i) { return (i - 1) * dx; });
    put_coords(m, "y", ny, [&](int i) { return (i - 1) * dy; });
    rc = put_device_var(m, st, s, "mu_b", mu_b, nx, ny, 0.0, who);
    if (rc) return rc;
    put_coords(m, "x_p", nx - 1, [&](int i) { return (i - 1) * dx + 0.5 * dx; });
    put_coords(m, "y_p", ny - 1, [&](int i) { return (i - 1) * dy + 0.5 * dy; });
    rc = put_device_var(m, st, s, "P", P, nx - 1, ny - 1, 0.0, who);
    if (rc) return rc;
    put_coords(m, "x_vx", nx + 2, [&](int i) { return (i - 2) * dx; });
    put_coords(m, "y_vx", ny + 1, [&](int i) { return (i - 1) * dy - 0.5 * dy; });
    rc = put_device_var(m, st, s, "Vx", Vx_pad, nx + 2, ny + 1, 0.0, who);
    if (rc) return rc;
    put_coords(m, "x_vy", nx + 1, [&](int i) { return (i - 1) * dx - 0.5 * dx; });
    put_coords(m, "y_vy", ny + 2, [&](int i) { return (i - 2) * dy; });
    rc = put_device_var(m, st, s, "Vy", Vy_pad, nx + 1, ny + 2, 0.0, who);
    if (rc) return rc;
    return finish(m, path, who);
}
