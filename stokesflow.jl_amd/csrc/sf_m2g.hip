// sf_m2g.hip -- deterministic marker -> grid interpolation for gfx950.
//
// Reference: bilinearMarkerToGrid! (MarkerToGrid.jl:18-38).  The reference's GPU path adds
// 8 float64 atomics per marker (:58-100) and is order-nondeterministic; its CPU path
// (:110-155) adds the markers sequentially in index order m = 1..Nm.  This
// implementation reproduces the CPU path bit for bit:
//
//   1. key:     per marker the top-left node (ix, iy) of common.jl:22-33 with the same
//               IEEE operations; markers outside [0,Nx]x[0,Ny] are counted ("erroneous
//               marker position") and dropped.  cell = ix*(Ny+1) + iy (column-major: see m2g_keys).
//   2. count + exclusive scan of markers per cell (integer atomics: exact).
//   3. scatter marker indices into their cell segment, then sort each segment by marker
//      index, and store per slot the marker's (dxij, dyij, val).
//   4. node:    node (A, B) (1-based) receives contributions from cells (A,B) as top-left,
//               (A-1,B) as top-right, (A,B-1) as bottom-left and (A-1,B-1) as bottom-right
//               (:124-151).  Merging the four index-sorted segments by marker index
//               replays exactly the CPU accumulation order, so wt_sum and val_wt_sum are
//               the CPU sums; safeDivision (:164-172) keeps the old value where wt == 0.
#include "sf_device.h"
#include "sf_internal.h"

namespace sf {

enum { CNT_BAD = 0, CNT_EMPTY = 1, CNT_NONFINITE = 2 };
#define SCAN_ITEMS 8
#define SCAN_BLOCK 256
#define SCAN_TILE (SCAN_ITEMS * SCAN_BLOCK)

// floor(Int, v) in the reference throws unless v fits an Int64.
__device__ inline bool key_ok(double v) { return __builtin_isfinite(v) && v > -9.2e18 && v < 9.2e18; }

// topleftIndexRelDist (common.jl:22-33), 1-based ix/iy.
__device__ inline void topleft(double xmin, double ymin, double x, double y, double dx, double dy, int64_t& ix,
                               int64_t& iy, double& dxij, double& dyij)
{
    ix = (int64_t)floor((x - xmin) / dx) + 1;
    iy = (int64_t)floor((y - ymin) / dy) + 1;
    const double x_ix = xmin + (double)(ix - 1) * dx;
    const double y_iy = ymin + (double)(iy - 1) * dy;
    dxij = (x - x_ix) / dx;
    dyij = (y - y_iy) / dy;
}

__global__ __launch_bounds__(256) void fill_int(int* __restrict__ a, int64_t n, int v)
{
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) a[k] = v;
}

__global__ __launch_bounds__(256) void fill_dbl(double* __restrict__ a, int64_t n, double v)
{
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) a[k] = v;
}

__global__ void zero_counters(unsigned long long* c)
{
    if (threadIdx.x < 4) c[threadIdx.x] = 0ull;
}

// Runs of equal keys in consecutive lanes (markers in index order share cells: the x-outer /
// y-inner set-up puts md consecutive markers in one cell): per lane, whether it heads a run, the
// run length (heads only) and the lane of its run's head.  Every lane of the wave calls this.
struct KeyRun {
    bool head;
    int len, head_lane;
};
__device__ inline KeyRun key_run(int k)
{
    const int lane = (int)__lane_id();
    const int kp = dpp_i32<0x138>(k);  // the key of lane - 1 (wave_shr:1; lane 0 reads 0)
    const bool valid = k >= 0;
    const bool head = valid && (lane == 0 || kp != k);
    const unsigned long long heads = __ballot(head);
    const unsigned long long brk = __ballot(head || !valid);  // a run ends where the next one starts
    KeyRun r;
    r.head = head;
    const unsigned long long after = lane == 63 ? 0ull : (brk >> (lane + 1));
    r.len = head ? (after ? __builtin_ctzll(after) + 1 : 64 - lane) : 0;
    const unsigned long long upto = heads & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
    r.head_lane = upto ? 63 - __builtin_clzll(upto) : lane;
    return r;
}

__global__ __launch_bounds__(256) void m2g_keys(const double* __restrict__ x_m, const double* __restrict__ y_m, int64_t nm,
                                                double xmin, double ymin, double dx, double dy, int gnx, int gny,
                                                int colmajor, int* __restrict__ key, int* __restrict__ count,
                                                unsigned long long* __restrict__ counters)
{
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int k = -1;
    if (m < nm) {
        const double x = x_m[m], y = y_m[m];
        if (!key_ok((x - xmin) / dx) || !key_ok((y - ymin) / dy)) {
            atomicAdd(&counters[CNT_NONFINITE], 1ull);
        } else {
            int64_t ix, iy;
            double dxij, dyij;
            topleft(xmin, ymin, x, y, dx, dy, ix, iy, dxij, dyij);
            if (!(ix >= 0 && ix <= gnx && iy >= 0 && iy <= gny)) {
                atomicAdd(&counters[CNT_BAD], 1ull);
            } else {
                // the sort numbers cells column-major (ix slowest), the order in which the x-outer /
                // y-inner marker set-up visits them, so the per-cell gathers below read nearly in
                // order; the exported keys are row-major (sf_marker_cell_keys)
                k = colmajor ? (int)(ix * (int64_t)(gny + 1) + iy) : (int)(iy * (int64_t)(gnx + 1) + ix);
            }
        }
        key[m] = k;
    }
    // one counter atomic per run of equal keys instead of one per marker
    const KeyRun run = key_run(k);
    if (run.head) atomicAdd(&count[k], run.len);
}

__global__ __launch_bounds__(SCAN_BLOCK) void scan_block_sums(const int* __restrict__ cnt, int64_t n,
                                                              int* __restrict__ bsum)
{
    __shared__ int sh[SCAN_BLOCK];
    const int t = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)t * SCAN_ITEMS;
    int s = 0;
    for (int i = 0; i < SCAN_ITEMS; ++i)
        if (base + i < n) s += cnt[base + i];
    sh[t] = s;
    __syncthreads();
    for (int w = SCAN_BLOCK / 2; w > 0; w >>= 1) {
        if (t < w) sh[t] += sh[t + w];
        __syncthreads();
    }
    if (t == 0) bsum[blockIdx.x] = sh[0];
}

// Exclusive scan of the (few thousand) block sums by one thread.
__global__ void scan_top(int* __restrict__ bsum, int nblk)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int acc = 0;
    for (int b = 0; b < nblk; ++b) {
        int v = bsum[b];
        bsum[b] = acc;
        acc += v;
    }
}

__global__ __launch_bounds__(SCAN_BLOCK) void scan_apply(const int* __restrict__ cnt, int64_t n,
                                                         const int* __restrict__ bsum, int* __restrict__ off)
{
    __shared__ int sh[SCAN_BLOCK];
    const int t = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)t * SCAN_ITEMS;
    int v[SCAN_ITEMS];
    int s = 0;
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        v[i] = (base + i < n) ? cnt[base + i] : 0;
        s += v[i];
    }
    sh[t] = s;
    __syncthreads();
    // Hillis-Steele inclusive scan of thread sums
    for (int w = 1; w < SCAN_BLOCK; w <<= 1) {
        int add = (t >= w) ? sh[t - w] : 0;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    int run = bsum[blockIdx.x] + sh[t] - s;
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        if (base + i < n) {
            off[base + i] = run;
            if (base + i == n - 1) off[n] = run + v[i];
        }
        run += v[i];
    }
}

__global__ __launch_bounds__(256) void m2g_scatter(const int* __restrict__ key, int64_t nm, const int* __restrict__ off,
                                                   int* __restrict__ fill, int* __restrict__ rec_idx)
{
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = m < nm ? key[m] : -1;
    // a run of equal keys takes its slots with one atomic; its markers fill them in index order
    const KeyRun run = key_run(k);
    int base = 0;
    if (run.head) base = atomicAdd(&fill[k], run.len);
    base = __shfl(base, run.head_lane);
    if (k < 0) return;
    rec_idx[off[k] + base + ((int)__lane_id() - run.head_lane)] = (int)m;
}

// Sort each cell segment by marker index (insertion sort: segments hold ~md^2 markers in a few
// index-ordered runs).
__global__ __launch_bounds__(256) void m2g_sort_cells(const int* __restrict__ off, int64_t ncells, int* __restrict__ rec_idx)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ncells) return;
    const int b = off[k], e = off[k + 1];
    for (int s = b + 1; s < e; ++s) {
        const int v = rec_idx[s];
        int q = s - 1;
        while (q >= b && rec_idx[q] > v) {
            rec_idx[q + 1] = rec_idx[q];
            --q;
        }
        rec_idx[q + 1] = v;
    }
}

// The per-slot records used by the node merge, one thread per slot (coalesced in slot order).
__global__ __launch_bounds__(256) void m2g_records(const int* __restrict__ rec_idx, const int* __restrict__ nrec,
                                                   const double* __restrict__ x_m, const double* __restrict__ y_m,
                                                   const double* __restrict__ val_m, double xmin, double ymin,
                                                   double dx, double dy, M2GRec* __restrict__ rec)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= *nrec) return;  // the valid markers' slots (the scan's total)
    const int m = rec_idx[s];
    int64_t ix, iy;
    double dxij, dyij;
    topleft(xmin, ymin, x_m[m], y_m[m], dx, dy, ix, iy, dxij, dyij);
    M2GRec r;
    r.dx = dxij;
    r.dy = dyij;
    r.val = val_m[m];
    r.idx = m;
    r.pad = 0;
    rec[s] = r;
}


__device__ inline double corner_weight(int corner, double dxij, double dyij)
{
    // MarkerToGrid.jl:127,133,141,147 -- 0 TL, 1 TR, 2 BL, 3 BR
    switch (corner) {
        case 0: return (1.0 - dxij) * (1.0 - dyij);
        case 1: return dxij * (1.0 - dyij);
        case 2: return (1.0 - dxij) * dyij;
        default: return dxij * dyij;
    }
}

__global__ __launch_bounds__(256) void m2g_nodes(const int* __restrict__ off, const M2GRec* __restrict__ rec, int gnx,
                                                 int gny, double* __restrict__ val_grid, double* __restrict__ val_wt_sum,
                                                 double* __restrict__ wt_sum, int ws_nx,
                                                 unsigned long long* __restrict__ counters)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;  // y fastest: cells are numbered column-major
    const int a = blockIdx.y;
    if (a >= gnx || b >= gny) return;
    const int A = a + 1, B = b + 1;  // 1-based node
    const int64_t ld = gny + 1;
    // cells: TL (A,B), TR (A-1,B), BL (A,B-1), BR (A-1,B-1)
    const int64_t kc[4] = {(int64_t)A * ld + B, (int64_t)(A - 1) * ld + B, (int64_t)A * ld + (B - 1),
                           (int64_t)(A - 1) * ld + (B - 1)};
    int h[4], e[4], head[4];
    for (int q = 0; q < 4; ++q) {
        h[q] = off[kc[q]];
        e[q] = off[kc[q] + 1];
        head[q] = (h[q] < e[q]) ? rec[h[q]].idx : 0x7fffffff;
    }
    double wt = 0.0, vw = 0.0;
    while (true) {
        int q = 0;
        int best = head[0];
        for (int t = 1; t < 4; ++t)
            if (head[t] < best) {
                best = head[t];
                q = t;
            }
        if (best == 0x7fffffff) break;
        const int s = h[q];
        const M2GRec r = rec[s];
        const double w = corner_weight(q, r.dx, r.dy);
        wt = wt + w;
        vw = vw + w * r.val;
        h[q] = s + 1;
        head[q] = (h[q] < e[q]) ? rec[h[q]].idx : 0x7fffffff;
    }
    if (wt_sum) {
        wt_sum[(int64_t)a + (int64_t)b * ws_nx] = wt;
        val_wt_sum[(int64_t)a + (int64_t)b * ws_nx] = vw;
    }
    if (wt != 0.0)
        val_grid[(int64_t)a + (int64_t)b * gnx] = vw / wt;
    else
        atomicAdd(&counters[CNT_EMPTY], 1ull);
}

static int grid1d(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

hipError_t m2g_cell_keys(M2GScratch& S, double xmin, double ymin, int gnx, int gny, const double* x_m,
                         const double* y_m, int64_t nm, double dx, double dy, int* keys, int* counts, hipStream_t st)
{
    const int64_t ncells = (int64_t)(gnx + 1) * (gny + 1);
    hipLaunchKernelGGL(zero_counters, dim3(1), dim3(64), 0, st, S.counters);
    hipLaunchKernelGGL(fill_int, dim3(1024), dim3(256), 0, st, counts, ncells, 0);
    if (nm > 0)
        hipLaunchKernelGGL(m2g_keys, dim3(grid1d(nm, 256)), dim3(256), 0, st, x_m, y_m, nm, xmin, ymin, dx, dy, gnx,
                           gny, 0, keys, counts, S.counters);
    return hipGetLastError();
}

hipError_t m2g_run(M2GScratch& S, double xmin, double ymin, double* val_grid, int gnx, int gny, const double* x_m,
                   const double* y_m, const double* val_m, int64_t nm, double dx, double dy, double* val_wt_sum,
                   double* wt_sum, int ws_nx, int ws_ny, hipStream_t st)
{
    const int64_t ncells = (int64_t)(gnx + 1) * (gny + 1);
    const int nblk = (int)((ncells + SCAN_TILE - 1) / SCAN_TILE);
    const int fillgrid = 1024;
    hipLaunchKernelGGL(zero_counters, dim3(1), dim3(64), 0, st, S.counters);
    hipLaunchKernelGGL(fill_int, dim3(fillgrid), dim3(256), 0, st, S.count, ncells, 0);
    hipLaunchKernelGGL(fill_int, dim3(fillgrid), dim3(256), 0, st, S.fill, ncells, 0);
    if (wt_sum) {
        hipLaunchKernelGGL(fill_dbl, dim3(fillgrid), dim3(256), 0, st, wt_sum, (int64_t)ws_nx * ws_ny, 0.0);
        hipLaunchKernelGGL(fill_dbl, dim3(fillgrid), dim3(256), 0, st, val_wt_sum, (int64_t)ws_nx * ws_ny, 0.0);
    }
    if (nm > 0)
        hipLaunchKernelGGL(m2g_keys, dim3(grid1d(nm, 256)), dim3(256), 0, st, x_m, y_m, nm, xmin, ymin, dx, dy, gnx,
                           gny, 1, S.key, S.count, S.counters);
    hipLaunchKernelGGL(scan_block_sums, dim3(nblk), dim3(SCAN_BLOCK), 0, st, S.count, ncells, S.blocksums);
    hipLaunchKernelGGL(scan_top, dim3(1), dim3(64), 0, st, S.blocksums, nblk);
    hipLaunchKernelGGL(scan_apply, dim3(nblk), dim3(SCAN_BLOCK), 0, st, S.count, ncells, S.blocksums, S.offset);
    if (nm > 0)
        hipLaunchKernelGGL(m2g_scatter, dim3(grid1d(nm, 256)), dim3(256), 0, st, S.key, nm, S.offset, S.fill,
                           S.rec_idx);
    hipLaunchKernelGGL(m2g_sort_cells, dim3(grid1d(ncells, 256)), dim3(256), 0, st, S.offset, ncells, S.rec_idx);
    if (nm > 0)  // the valid markers fill slots 0 .. offset[ncells] - 1 (at most nm)
        hipLaunchKernelGGL(m2g_records, dim3(grid1d(nm, 256)), dim3(256), 0, st, S.rec_idx, S.offset + ncells, x_m,
                           y_m, val_m, xmin, ymin, dx, dy, S.rec);
    hipLaunchKernelGGL(m2g_nodes, dim3(grid1d(gny, 256), gnx), dim3(256), 0, st, S.offset, S.rec, gnx, gny,
                       val_grid, val_wt_sum, wt_sum, ws_nx, S.counters);
    return hipGetLastError();
}

}  // namespace sf
