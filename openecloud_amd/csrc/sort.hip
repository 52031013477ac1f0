// Stable LSD radix sort of 64-bit keys with 32-bit payloads, and the particle permutation.
//
// Each pass handles 8 key bits: a per-tile digit histogram (digit-major layout, so one
// exclusive scan gives every tile's output offset per digit), then a stable scatter whose
// in-tile ranks come from a wave64 multisplit: the lanes holding the same digit are found
// with eight ballots, their order within the wave from a popcount of the lower-lane mask,
// and the order across waves and sub-tiles from LDS counters.
#include "common.h"
#include "particles_dev.h"
#include "scan.h"

namespace {

constexpr int kThreads = 256;
constexpr int kItems = 16;
constexpr int kTile = kThreads * kItems;
constexpr int kRadix = 256;

__global__ void __launch_bounds__(kThreads)
    k_digit_hist(const uint64_t *__restrict__ keys, uint32_t n, int shift, uint32_t nTiles,
                 uint32_t *__restrict__ hist)
{
    __shared__ uint32_t cnt[kRadix];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    uint64_t base = (uint64_t)blockIdx.x * kTile;
    for (int k = 0; k < kItems; ++k) {
        uint64_t i = base + (uint64_t)k * kThreads + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 0xFF], 1u);
    }
    __syncthreads();
    hist[(uint64_t)threadIdx.x * nTiles + blockIdx.x] = cnt[threadIdx.x];
}

__global__ void __launch_bounds__(kThreads)
    k_digit_scatter(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ vals, uint32_t n,
                    int shift, uint32_t nTiles, const uint32_t *__restrict__ off,
                    uint64_t *__restrict__ keysOut, uint32_t *__restrict__ valsOut)
{
    __shared__ uint32_t run[kRadix];
    __shared__ uint32_t wcnt[kThreads / 64][kRadix];
    __shared__ uint32_t base_off[kRadix];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    run[t] = 0;
    base_off[t] = off[(uint64_t)t * nTiles + blockIdx.x];
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint64_t base = (uint64_t)blockIdx.x * kTile;
    __syncthreads();
    for (int k = 0; k < kItems; ++k) {
        uint64_t i = base + (uint64_t)k * kThreads + t;
        bool valid = i < n;
        uint64_t key = valid ? keys[i] : 0ull;
        uint32_t d = (uint32_t)((key >> shift) & 0xFF);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            bool bit = (d >> b) & 1u;
            uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        uint32_t lrank = (uint32_t)__popcll(peers & lt);
#pragma unroll
        for (int w = 0; w < kThreads / 64; ++w) wcnt[w][t] = 0;
        __syncthreads();
        if (valid && lrank == 0) wcnt[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t r = run[d] + lrank;
            for (int w = 0; w < wave; ++w) r += wcnt[w][d];
            uint32_t dst = base_off[d] + r;
            keysOut[dst] = key;
            valsOut[dst] = vals[i];
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int w = 0; w < kThreads / 64; ++w) add += wcnt[w][t];
        run[t] += add;
        __syncthreads();
    }
}

__global__ void k_iota(uint32_t *v, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = i;
}

// Order-preserving map of a double (or float) column to uint64 keys.
template <typename T>
__global__ void k_column_keys(Soa<T> p, int col, uint32_t n, uint64_t *keys)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = (double)p(col, i);
    uint64_t u = (uint64_t)__double_as_longlong(v);
    keys[i] = (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

template <typename T>
__global__ void k_gather_rows(Soa<T> src, Soa<T> dst, const uint32_t *__restrict__ perm, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t s = perm[i];
#pragma unroll
    for (int c = 0; c < 6; ++c) dst(c, i) = src(c, s);
}

}  // namespace

static uint32_t sort_tiles(uint32_t n) { return (n + kTile - 1) / kTile; }

OECL_EXPORT uint64_t oecl_sort_work_bytes(uint32_t n)
{
    uint64_t nt = sort_tiles(n) ? sort_tiles(n) : 1;
    uint64_t h = (uint64_t)kRadix * nt;
    // hist, offsets (h + 1), scan work, alternate keys and values
    return 4 * (h + (h + 1) + 2 * (oecl_scan_tiles((uint32_t)h) + 2) + 16) + 12ull * n + 1024;
}

// Sorts keys[n] ascending (stable), permuting vals along; bits: number of low key bits that
// can be non-zero (rounded up to whole bytes).  Results end in keys/vals.
OECL_EXPORT int oecl_radix_sort_u64(uint64_t *keys, uint32_t *vals, uint32_t n, int bits, void *work,
                                    void *stream)
{
    if (!keys || !vals || !work || bits < 1 || bits > 64) return OECL_ERR_ARG;
    if (n < 2) return OECL_OK;
    hipStream_t s = (hipStream_t)stream;
    uint32_t nt = sort_tiles(n);
    uint64_t h = (uint64_t)kRadix * nt;
    char *w = (char *)work;
    uint32_t *hist = (uint32_t *)w;
    uint32_t *off = hist + h;
    uint32_t *scanw = off + h + 1;
    uint64_t scanWords = 2 * (oecl_scan_tiles((uint32_t)h) + 2) + 16;
    char *alt = (char *)(scanw + scanWords);
    alt = (char *)(((uintptr_t)alt + 15) & ~(uintptr_t)15);
    uint64_t *k2 = (uint64_t *)alt;
    uint32_t *v2 = (uint32_t *)(k2 + n);
    uint64_t *kin = keys, *kout = k2;
    uint32_t *vin = vals, *vout = v2;
    int passes = (bits + 7) / 8;
    for (int ps = 0; ps < passes; ++ps) {
        int shift = 8 * ps;
        hipLaunchKernelGGL(k_digit_hist, dim3(nt), dim3(kThreads), 0, s, kin, n, shift, nt, hist);
        int rc = oecl_launch_scan_u32(hist, (uint32_t)h, off, scanw, s);
        if (rc) return rc;
        hipLaunchKernelGGL(k_digit_scatter, dim3(nt), dim3(kThreads), 0, s, kin, vin, n, shift, nt, off,
                           kout, vout);
        uint64_t *tk = kin; kin = kout; kout = tk;
        uint32_t *tv = vin; vin = vout; vout = tv;
    }
    if (kin != keys) {
        OECL_HIP(hipMemcpyAsync(keys, kin, 8ull * n, hipMemcpyDeviceToDevice, s), "sort copy");
        OECL_HIP(hipMemcpyAsync(vals, vin, 4ull * n, hipMemcpyDeviceToDevice, s), "sort copy");
    }
    OECL_CHECK_LAUNCH("oecl_radix_sort_u64");
    return OECL_OK;
}

// Particles.sortByColumn (particles.pyx:49-50 -> _quicksortByColumn :431-459): ascending by
// one column, stable; the sorted block is written to `spare` (same stride).
OECL_EXPORT int oecl_sort_by_column(int dtype, const void *data, void *spare, uint64_t stride,
                                    uint32_t n, int column, uint64_t *keys, uint32_t *perm,
                                    void *work, void *stream)
{
    if (!data || !spare || !keys || !perm || !work || column < 0 || column > 5) return OECL_ERR_ARG;
    if (n == 0) return OECL_OK;
    hipStream_t s = (hipStream_t)stream;
    dim3 gr(oecl_blocks(n, 256)), bl(256);
    if (dtype == OECL_F64)
        hipLaunchKernelGGL(k_column_keys<double>, gr, bl, 0, s, Soa<double>{(double *)data, stride}, column,
                           n, keys);
    else if (dtype == OECL_F32)
        hipLaunchKernelGGL(k_column_keys<float>, gr, bl, 0, s, Soa<float>{(float *)data, stride}, column, n,
                           keys);
    else
        return OECL_ERR_ARG;
    hipLaunchKernelGGL(k_iota, gr, bl, 0, s, perm, n);
    int rc = oecl_radix_sort_u64(keys, perm, n, 64, work, stream);
    if (rc) return rc;
    if (dtype == OECL_F64)
        hipLaunchKernelGGL(k_gather_rows<double>, gr, bl, 0, s, Soa<double>{(double *)data, stride},
                           Soa<double>{(double *)spare, stride}, perm, n);
    else
        hipLaunchKernelGGL(k_gather_rows<float>, gr, bl, 0, s, Soa<float>{(float *)data, stride},
                           Soa<float>{(float *)spare, stride}, perm, n);
    OECL_CHECK_LAUNCH("oecl_sort_by_column");
    return OECL_OK;
}

template <typename T>
__global__ void k_cell_keys(oecl_grid_t g, Soa<T> p, uint32_t n, uint64_t *keys, uint32_t *vals)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Weights<double> w = grid_weights<double>((double)p(0, i), (double)p(1, i), 1. / g.dx, 1. / g.dy,
                                             g.lxHalfExt, g.lyHalfExt, g.nx, g.ny);
    keys[i] = w.cell;
    vals[i] = i;
}

// Orders the particles by grid cell (stable) into `spare`: consecutive particles then share
// cells, which keeps the field gather in cache and lets the deposit sum in LDS.
OECL_EXPORT int oecl_sort_by_cell(const oecl_grid_t *g, int dtype, const void *data, void *spare,
                                  uint64_t stride, uint32_t n, uint64_t *keys, uint32_t *perm, void *work,
                                  void *stream)
{
    if (!g || !data || !spare || !keys || !perm || !work) return OECL_ERR_ARG;
    if (n == 0) return OECL_OK;
    hipStream_t s = (hipStream_t)stream;
    dim3 gr(oecl_blocks(n, 256)), bl(256);
    if (dtype == OECL_F64)
        hipLaunchKernelGGL(k_cell_keys<double>, gr, bl, 0, s, *g, Soa<double>{(double *)data, stride}, n, keys,
                           perm);
    else if (dtype == OECL_F32)
        hipLaunchKernelGGL(k_cell_keys<float>, gr, bl, 0, s, *g, Soa<float>{(float *)data, stride}, n, keys,
                           perm);
    else
        return OECL_ERR_ARG;
    int bits = 1;
    while ((1ull << bits) < g->np) ++bits;
    int rc = oecl_radix_sort_u64(keys, perm, n, bits, work, stream);
    if (rc) return rc;
    return oecl_permute_particles(dtype, data, spare, stride, perm, n, stream);
}

namespace {

// 16 low bits of v spread to the even bit positions (Morton interleave).
__device__ __forceinline__ uint32_t spread_bits(uint32_t v)
{
    v &= 0xFFFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

// Sort key of slot i of a batched particle block: (run << cellBits) | bin, where the bin of a
// live particle is its cell (ix, iy) >> binShift in row-major (order 0) or Morton (order 1)
// order, and a slot past the run's live count takes the largest bin, so that it sorts to the
// end of its own run's segment.
template <typename T>
__global__ void k_run_cell_keys(oecl_grid_t g, Soa<T> p, uint32_t cap, const uint32_t *__restrict__ segOff,
                                const uint32_t *__restrict__ blockRun, const uint32_t *__restrict__ nDev, int order,
                                int binShift, int cellBits, uint64_t *__restrict__ keys, uint32_t *__restrict__ vals)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cap) return;
    const uint32_t r = blockRun[i / kPushBlock];
    uint64_t bin = (1ull << cellBits) - 1;
    if (i - segOff[r] < nDev[r]) {
        const Weights<double> w = grid_weights<double>((double)p(0, i), (double)p(1, i), 1. / g.dx, 1. / g.dy,
                                                       g.lxHalfExt, g.lyHalfExt, g.nx, g.ny);
        const uint32_t bx = w.ix >> binShift, by = w.iy >> binShift;
        bin = order ? (uint64_t)(spread_bits(bx) | (spread_bits(by) << 1))
                    : (uint64_t)bx + (uint64_t)by * (((g.nx - 1) >> binShift) + 1);
    }
    keys[i] = ((uint64_t)r << cellBits) | bin;
    vals[i] = i;
}

}  // namespace

OECL_EXPORT int oecl_sort_runs_by_cell(const oecl_grid_t *g, int dtype, const void *data, void *spare, uint64_t stride,
                                       uint32_t R, const uint32_t *segOff, const uint32_t *blockRun,
                                       const uint32_t *nDev, int order, int binShift, uint64_t *keys, uint32_t *perm,
                                       void *work, void *stream)
{
    if (!g || !data || !spare || !segOff || !blockRun || !nDev || !keys || !perm || !work || R == 0) return OECL_ERR_ARG;
    if (order < 0 || order > 1 || binShift < 0 || binShift > 8 || stride % kPushBlock || stride > 0xFFFFFFFFull)
        return OECL_ERR_ARG;
    if (g->nx > 65536 || g->ny > 65536) return OECL_ERR_ARG;
    const uint32_t cap = (uint32_t)stride;
    if (cap == 0) return OECL_OK;
    // bits of the largest live bin + 1 (the dead key), and of the run index
    const uint64_t bx = ((g->nx - 1) >> binShift) + 1, by = ((g->ny - 1) >> binShift) + 1;
    uint64_t maxBin;
    if (order) {
        int b = 0;
        while ((1ull << b) < (bx > by ? bx : by)) ++b;
        maxBin = (1ull << (2 * b)) - 1;
    } else {
        maxBin = bx * by - 1;
    }
    int cellBits = 1, runBits = 0;
    while ((1ull << cellBits) <= maxBin + 1) ++cellBits;
    while ((1ull << runBits) < R) ++runBits;
    hipStream_t s = (hipStream_t)stream;
    const dim3 gr(oecl_blocks(cap, 256)), bl(256);
    if (dtype == OECL_F64)
        hipLaunchKernelGGL(k_run_cell_keys<double>, gr, bl, 0, s, *g, Soa<double>{(double *)data, stride}, cap, segOff,
                           blockRun, nDev, order, binShift, cellBits, keys, perm);
    else if (dtype == OECL_F32)
        hipLaunchKernelGGL(k_run_cell_keys<float>, gr, bl, 0, s, *g, Soa<float>{(float *)data, stride}, cap, segOff,
                           blockRun, nDev, order, binShift, cellBits, keys, perm);
    else
        return OECL_ERR_ARG;
    int rc = oecl_radix_sort_u64(keys, perm, cap, cellBits + runBits, work, stream);
    if (rc) return rc;
    return oecl_permute_particles(dtype, data, spare, stride, perm, cap, stream);
}

// dst[:, i] = src[:, perm[i]] for i < n.
OECL_EXPORT int oecl_permute_particles(int dtype, const void *src, void *dst, uint64_t stride,
                                       const uint32_t *perm, uint32_t n, void *stream)
{
    if (!src || !dst || !perm) return OECL_ERR_ARG;
    if (n == 0) return OECL_OK;
    hipStream_t s = (hipStream_t)stream;
    dim3 gr(oecl_blocks(n, 256)), bl(256);
    if (dtype == OECL_F64)
        hipLaunchKernelGGL(k_gather_rows<double>, gr, bl, 0, s, Soa<double>{(double *)src, stride},
                           Soa<double>{(double *)dst, stride}, perm, n);
    else if (dtype == OECL_F32)
        hipLaunchKernelGGL(k_gather_rows<float>, gr, bl, 0, s, Soa<float>{(float *)src, stride},
                           Soa<float>{(float *)dst, stride}, perm, n);
    else
        return OECL_ERR_ARG;
    OECL_CHECK_LAUNCH("oecl_permute_particles");
    return OECL_OK;
}

// ---------------------------------------------------------------------------------------------
// The reference's deposit order (particles.pyx:336-351): rho is zeroed, then each particle in
// index order adds (w_corner * weight) * q to its four corners.  Every node therefore receives
// its contributions in increasing particle index, summed from 0.  The ordered deposit
// reproduces that sum exactly: a stable sort of the particle indices by cell gives, for every
// cell, its particles in index order; node j's contributions come from four cells (j as the
// lower-left corner, j - 1 as lower-right, j - nx as upper-left, j - nx - 1 as upper-right),
// and one thread per node merges those four lists by particle index.
namespace {

__global__ void k_zero_u32(uint32_t *a, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}

// inCell from calcGridWeights is at most np - nx - 2 (the cell index is clamped to the grid), so
// every list below stays inside the node arrays; an index past np is clamped to the last node.
__global__ void k_ordered_keys(const uint32_t *__restrict__ inCell, uint32_t n, uint32_t np, uint64_t *keys,
                               uint32_t *vals, uint32_t *counts)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t c = inCell[i];
    if (c >= np) c = np - 1;
    keys[i] = c;
    vals[i] = i;
    atomicAdd(counts + c, 1u);
}

template <typename T>
__global__ void k_ordered_nodes(Soa<T> p, const T *__restrict__ w4, uint64_t ws, const uint32_t *__restrict__ perm,
                                const uint32_t *__restrict__ start, uint32_t nx, uint32_t np, double q,
                                double *__restrict__ rho)
{
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= np) return;
    // list r: the particles of cell j - off[r], whose corner r is node j
    uint32_t a[4], e[4];
    const uint32_t off[4] = {0u, 1u, nx, nx + 1u};
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        if (j >= off[r]) {
            const uint32_t c = j - off[r];
            a[r] = start[c];
            e[r] = start[c + 1];
        } else {
            a[r] = e[r] = 0;
        }
    }
    double acc = 0.;
    for (;;) {
        int h = -1;
        uint32_t best = 0xFFFFFFFFu;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            if (a[r] < e[r]) {
                const uint32_t i = perm[a[r]];
                if (i < best) {
                    best = i;
                    h = r;
                }
            }
        }
        if (h < 0) break;
        acc += (double)w4[(uint64_t)h * ws + best] * (double)p(5, best) * q;
        ++a[h];
    }
    rho[j] = acc;
}

uint64_t align16(uint64_t b) { return (b + 15) & ~15ull; }

}  // namespace

OECL_EXPORT uint64_t oecl_charge_ordered_work_bytes(uint32_t n, uint32_t np)
{
    return align16(8ull * n) + align16(4ull * n) + align16(4ull * (np + 2)) + align16(4ull * (np + 2)) +
           align16(oecl_scan_work_bytes(np + 1)) + align16(oecl_sort_work_bytes(n)) + 64;
}

OECL_EXPORT int oecl_charge_to_grid_ordered(const oecl_grid_t *g, int dtype, const void *data, uint64_t stride,
                                            uint32_t n, const uint32_t *inCell, const void *w4, uint64_t wstride,
                                            double q, double *rho, void *work, void *stream)
{
    if (!g || !rho || !work || g->np < 1 || g->nx < 2) return OECL_ERR_ARG;
    if (n > 0 && (!data || !inCell || !w4)) return OECL_ERR_ARG;
    if (dtype != OECL_F64 && dtype != OECL_F32) return OECL_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    const uint32_t np = g->np;
    char *w = (char *)work;
    uint64_t *keys = (uint64_t *)w;
    w += align16(8ull * n);
    uint32_t *perm = (uint32_t *)w;
    w += align16(4ull * n);
    uint32_t *counts = (uint32_t *)w;
    w += align16(4ull * (np + 2));
    uint32_t *start = (uint32_t *)w;
    w += align16(4ull * (np + 2));
    uint32_t *scanw = (uint32_t *)w;
    w += align16(oecl_scan_work_bytes(np + 1));
    void *sortw = w;
    hipLaunchKernelGGL(k_zero_u32, dim3(oecl_blocks(np + 2, 256)), dim3(256), 0, s, counts, np + 2);
    if (n > 0)
        hipLaunchKernelGGL(k_ordered_keys, dim3(oecl_blocks(n, 256)), dim3(256), 0, s, inCell, n, np, keys, perm,
                           counts);
    int bits = 1;
    while ((1ull << bits) < np) ++bits;
    if (n > 1) {
        int rc = oecl_radix_sort_u64(keys, perm, n, bits, sortw, stream);
        if (rc) return rc;
    }
    // start[c] = particles in cells before c; start[np] = n
    int rc = oecl_launch_scan_u32(counts, np, start, scanw, s);
    if (rc) return rc;
    if (dtype == OECL_F64)
        hipLaunchKernelGGL(k_ordered_nodes<double>, dim3(oecl_blocks(np, 256)), dim3(256), 0, s,
                           Soa<double>{(double *)data, stride}, (const double *)w4, wstride, perm, start, g->nx, np,
                           q, rho);
    else
        hipLaunchKernelGGL(k_ordered_nodes<float>, dim3(oecl_blocks(np, 256)), dim3(256), 0, s,
                           Soa<float>{(float *)data, stride}, (const float *)w4, wstride, perm, start, g->nx, np, q,
                           rho);
    OECL_CHECK_LAUNCH("oecl_charge_to_grid_ordered");
    return OECL_OK;
}
