// Wall kernels (particleBoundary.pyx) and the add/remove compaction (particles.pyx:94-269).
//
// Compaction never scans serially: one pass counts holes per 4096-particle tile, a single
// block scans the tile counts, and the consuming kernels recompute each hole's rank inside
// its tile with a block scan.  Only particles that move are read or written.
#include <mutex>

#include "common.h"
#include "scan.h"
#include "wall.h"

static char g_err[512] = "";

void oecl_set_error(const char *where, hipError_t e)
{
    snprintf(g_err, sizeof(g_err), "%s: %s", where, hipGetErrorString(e));
}
void oecl_set_error_msg(const char *msg) { snprintf(g_err, sizeof(g_err), "%s", msg); }

OECL_EXPORT const char *oecl_last_error(void) { return g_err; }

OECL_EXPORT int oecl_device_sync(void)
{
    OECL_HIP(hipDeviceSynchronize(), "oecl_device_sync");
    return OECL_OK;
}

namespace {
// Plain streaming copy, 16 bytes per lane per iteration (grid-stride).
__global__ void __launch_bounds__(256)
    k_copy16(uint4 *__restrict__ dst, const uint4 *__restrict__ src, uint64_t n16)
{
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * 256)
        dst[i] = src[i];
}
}  // namespace

OECL_EXPORT int oecl_copy_bytes(void *dst, const void *src, uint64_t bytes, void *stream)
{
    if ((!dst || !src) && bytes) return OECL_ERR_ARG;
    if (((uintptr_t)dst | (uintptr_t)src | bytes) & 15) return OECL_ERR_ARG;
    const uint64_t n16 = bytes / 16;
    if (n16 == 0) return OECL_OK;
    const uint64_t blocks = (n16 + 255) / 256;
    hipLaunchKernelGGL(k_copy16, dim3((unsigned)(blocks < 8192 ? blocks : 8192)), dim3(256), 0, (hipStream_t)stream,
                       (uint4 *)dst, (const uint4 *)src, n16);
    OECL_CHECK_LAUNCH("oecl_copy_bytes");
    return OECL_OK;
}

namespace {

// ---------------------------------------------------------------------------- scans
__global__ void __launch_bounds__(OECL_SCAN_THREADS)
    k_tile_count_zero(const uint8_t *__restrict__ flags, uint32_t n, uint32_t *__restrict__ counts)
{
    __shared__ uint32_t lds[4];
    uint64_t base = (uint64_t)blockIdx.x * OECL_SCAN_TILE + threadIdx.x * OECL_SCAN_ITEMS;
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < OECL_SCAN_ITEMS; ++k) {
        uint64_t i = base + k;
        if (i < n && flags[i] == 0) ++c;
    }
    uint32_t tot;
    block_excl_scan(c, lds, &tot);
    if (threadIdx.x == 0) counts[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(OECL_SCAN_THREADS)
    k_tile_sum_u32(const uint32_t *__restrict__ in, uint32_t n, uint32_t *__restrict__ counts)
{
    __shared__ uint32_t lds[4];
    uint64_t base = (uint64_t)blockIdx.x * OECL_SCAN_TILE + threadIdx.x * OECL_SCAN_ITEMS;
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < OECL_SCAN_ITEMS; ++k) {
        uint64_t i = base + k;
        if (i < n) c += in[i];
    }
    uint32_t tot;
    block_excl_scan(c, lds, &tot);
    if (threadIdx.x == 0) counts[blockIdx.x] = tot;
}

// Single block: offsets[t] = sum(counts[0..t)), offsets[nTiles] = total.
__global__ void __launch_bounds__(OECL_SCAN_THREADS)
    k_scan_counts(const uint32_t *__restrict__ counts, uint32_t nTiles, uint32_t *__restrict__ offsets)
{
    __shared__ uint32_t lds[4];
    uint32_t carry = 0;
    for (uint32_t b = 0; b < nTiles; b += OECL_SCAN_THREADS) {
        uint32_t i = b + threadIdx.x;
        uint32_t v = i < nTiles ? counts[i] : 0;
        uint32_t tot;
        uint32_t ex = block_excl_scan(v, lds, &tot);
        if (i < nTiles) offsets[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) offsets[nTiles] = carry;
}

__global__ void __launch_bounds__(OECL_SCAN_THREADS)
    k_scan_apply_u32(const uint32_t *__restrict__ in, uint32_t n, const uint32_t *__restrict__ tileOff,
                     uint32_t *__restrict__ out)
{
    __shared__ uint32_t lds[4];
    uint64_t base = (uint64_t)blockIdx.x * OECL_SCAN_TILE + threadIdx.x * OECL_SCAN_ITEMS;
    uint32_t v[OECL_SCAN_ITEMS];
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < OECL_SCAN_ITEMS; ++k) {
        uint64_t i = base + k;
        v[k] = i < n ? in[i] : 0;
        c += v[k];
    }
    uint32_t tot;
    uint32_t run = tileOff[blockIdx.x] + block_excl_scan(c, lds, &tot);
#pragma unroll
    for (int k = 0; k < OECL_SCAN_ITEMS; ++k) {
        uint64_t i = base + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == OECL_SCAN_THREADS - 1) out[n] = run;
}

// ---------------------------------------------------------------------------- inside
template <typename T>
__global__ void k_index_inside(WallDev w, Soa<T> p, uint32_t n, uint8_t *isInside, int32_t *clipped)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T x = p(0, i), y = p(1, i);
    T x0 = x, y0 = y;
    isInside[i] = wall_inside<T>(w, x, y);
    if (w.kind == OECL_WALL_CUTCELL && (x != x0 || y != y0)) {
        p(0, i) = x;
        p(1, i) = y;
        if (clipped) *clipped = 1;
    }
}

// saveAbsorbed copy: hole of rank r -> absorbed row r (AoS, 6 values).
template <typename T>
__global__ void __launch_bounds__(OECL_SCAN_THREADS)
    k_save_absorbed(Soa<T> p, uint32_t n, const uint8_t *__restrict__ flags,
                    const uint32_t *__restrict__ tileOff, T *__restrict__ absorbed, uint32_t absCap)
{
    for_each_ranked(flags, n, tileOff, [&](uint32_t i, bool hole, uint32_t r) {
        if (hole && r < absCap) {
#pragma unroll
            for (int c = 0; c < 6; ++c) absorbed[6 * (uint64_t)r + c] = p(c, i);
        }
    });
}

__global__ void k_copy_u32(const uint32_t *src, uint32_t *dst) { *dst = *src; }

// ---------------------------------------------------------------------------- wall hits
template <typename T>
__global__ void k_wall_interaction(WallDev w, T *ap, uint32_t na, double dt, T *rts, T *nv,
                                   uint32_t *notFound)
{
    uint32_t ii = blockIdx.x * blockDim.x + threadIdx.x;
    if (ii >= na) return;
    T *p = ap + 6 * (uint64_t)ii;
    double x = p[0], y = p[1], vx = p[2], vy = p[3];
    double r = 0., n0 = 0., n1 = 0.;
    if (w.kind == OECL_WALL_RECT) {
        wall_rect_hit(w, x, y, vx, vy, r, n0, n1);
    } else if (w.kind == OECL_WALL_ELLIP) {
        // particleBoundary.pyx:252-274 and :277-297.
        const double eps = 1.e-6;
        double ax = 4. / ((w.lx - eps * w.dx) * (w.lx - eps * w.dx));
        double ay = 4. / ((w.ly - eps * w.dy) * (w.ly - eps * w.dy));
        double c = x * x * ax + y * y * ay - 1.;
        double b = x * vx * ax + y * vy * ay;
        double a = vx * vx * ax + vy * vy * ay;
        r = (b - sqrt(b * b - a * c)) / a;
        x -= r * vx;
        y -= r * vy;
        double ax2 = w.ax * w.ax, ay2 = w.ay * w.ay;
        double ninv = 1. / sqrt(x * x * ax2 + y * y * ay2);
        n0 = -x * w.ax * ninv;
        n1 = -y * w.ay * ninv;
    } else {
        // particleBoundary.pyx:373-457 and :460-481.
        const double se = 1.e-6;
        uint32_t inCell = wall_cell(w, x, y);
        double rx = -dt * vx, ry = -dt * vy;
        int xDir[2] = {0, rx < 0. ? -1 : 1};
        int yDir[2] = {0, ry < 0. ? -1 : 1};
        bool stop = false;
        for (int a = 0; a < 2 && !stop; ++a) {
            for (int b = 0; b < 2; ++b) {
                uint32_t c = inCell + xDir[a] + yDir[b] * w.nx;
                uint16_t f = w.insideFaces[c];
                if (f != 0 && f != 8) {
                    int32_t i0 = w.cpi[2 * c], i1 = w.cpi[2 * c + 1];
                    double sx = w.bp[2 * i1] - w.bp[2 * i0];
                    double sy = w.bp[2 * i1 + 1] - w.bp[2 * i0 + 1];
                    double rs = rx * sy - ry * sx;
                    if (rs != 0.) {
                        double qx = w.bp[2 * i0] - x;
                        double qy = w.bp[2 * i0 + 1] - y;
                        double t = (qx * sy - qy * sx) / rs;
                        double u = (qx * ry - qy * rx) / rs;
                        if (-se <= t && t <= 1. + se && -se <= u && u <= 1. + se) {
                            x += t * rx * (1. + se);
                            y += t * ry * (1. + se);
                            r = t * (1. + se) * dt;
                            stop = true;
                            break;
                        }
                    }
                }
            }
        }
        if (!stop && notFound) atomicAdd(notFound, 1u);
        uint32_t c2 = wall_cell(w, x, y);
        n0 = w.ccn[2 * c2];
        n1 = w.ccn[2 * c2 + 1];
    }
    p[0] = wall_store<T>(x);
    p[1] = wall_store<T>(y);
    rts[ii] = (T)r;
    nv[2 * (uint64_t)ii] = (T)n0;
    nv[2 * (uint64_t)ii + 1] = (T)n1;
}

// ---------------------------------------------------------------------------- compaction
struct CompactState {
    uint32_t n;         // old count
    uint32_t nAdd;
    const uint32_t *holes;   // device: total hole count (offsets[nTiles])
    uint32_t *holePos;       // [holes]
};

// Pass 1: holePos[r] = position of the hole of rank r.
__global__ void __launch_bounds__(OECL_SCAN_THREADS)
    k_hole_positions(const uint8_t *__restrict__ flags, uint32_t n,
                     const uint32_t *__restrict__ tileOff, uint32_t *__restrict__ holePos)
{
    for_each_ranked(flags, n, tileOff, [&](uint32_t i, bool hole, uint32_t r) {
        if (hole) holePos[r] = i;
    });
}

// Additions: add[s] -> hole s (s < holes) or -> n + s - holes.  Used by modes 1 and 3.
template <typename T>
__global__ void k_place_additions(Soa<T> p, const T *__restrict__ add, uint32_t nAdd, uint32_t n,
                                  const uint32_t *__restrict__ holes,
                                  const uint32_t *__restrict__ holePos, uint64_t capacity,
                                  uint32_t *status)
{
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nAdd) return;
    uint32_t h = *holes;
    uint64_t dst = s < h ? holePos[s] : (uint64_t)n + (s - h);
    if (dst >= capacity) {
        atomicOr(status, 1u);
        return;
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) p(c, dst) = add[6 * (uint64_t)s + c];
}

// Tail fill (mode 0 / mode 3 and mode 1 when nAdd >= holes): the kept particle at p >= newN
// whose rank counted from the end is j moves to hole nAdd + j.
template <typename T>
__global__ void __launch_bounds__(OECL_SCAN_THREADS)
    k_tail_fill(Soa<T> p, const uint8_t *__restrict__ flags, uint32_t n, uint32_t nAdd,
                const uint32_t *__restrict__ tileOff, const uint32_t *__restrict__ holes,
                const uint32_t *__restrict__ holePos)
{
    uint32_t h = *holes;
    if (nAdd >= h) return;
    uint32_t newN = n - h + nAdd;
    for_each_ranked(flags, n, tileOff, [&](uint32_t i, bool hole, uint32_t hb) {
        if (!hole && i >= newN) {
            // kept particles in (i, n): (n - 1 - i) - (h - hb)
            uint32_t j = (n - 1 - i) - (h - hb);
            uint32_t dst = holePos[nAdd + j];
#pragma unroll
            for (int c = 0; c < 6; ++c) p(c, dst) = p(c, i);
        }
    });
}

// Mode 1 with nAdd < holes: kept particles after the last filled hole are shifted down
// stably (particles.pyx:238-245), through the spare block.
template <typename T>
__global__ void __launch_bounds__(OECL_SCAN_THREADS)
    k_shift_to_spare(Soa<T> p, Soa<T> sp, const uint8_t *__restrict__ flags, uint32_t n,
                     uint32_t nAdd, const uint32_t *__restrict__ tileOff,
                     const uint32_t *__restrict__ holes, const uint32_t *__restrict__ holePos)
{
    uint32_t h = *holes;
    if (nAdd == 0 || nAdd >= h) return;
    uint32_t p0 = holePos[nAdd - 1] + 1;
    for_each_ranked(flags, n, tileOff, [&](uint32_t i, bool hole, uint32_t hb) {
        if (!hole && i >= p0) {
            uint32_t dst = i - (hb - nAdd);
#pragma unroll
            for (int c = 0; c < 6; ++c) sp(c, dst) = p(c, i);
        }
    });
}

template <typename T>
__global__ void k_copy_back(Soa<T> p, Soa<T> sp, uint32_t n, uint32_t nAdd,
                            const uint32_t *__restrict__ holes, const uint32_t *__restrict__ holePos)
{
    uint32_t h = *holes;
    if (nAdd == 0 || nAdd >= h) return;
    uint32_t p0 = holePos[nAdd - 1] + 1;
    uint32_t newN = n - h + nAdd;
    for (uint64_t i = p0 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < newN;
         i += (uint64_t)gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < 6; ++c) p(c, i) = sp(c, i);
    }
}

// Mode 2: stable keep then append, into the spare block.
template <typename T>
__global__ void __launch_bounds__(OECL_SCAN_THREADS)
    k_stable_to_spare(Soa<T> p, Soa<T> sp, const uint8_t *__restrict__ flags, uint32_t n,
                      const uint32_t *__restrict__ tileOff)
{
    for_each_ranked(flags, n, tileOff, [&](uint32_t i, bool hole, uint32_t hb) {
        if (!hole) {
            uint32_t dst = i - hb;
#pragma unroll
            for (int c = 0; c < 6; ++c) sp(c, dst) = p(c, i);
        }
    });
}

template <typename T>
__global__ void k_append_to_spare(Soa<T> sp, const T *__restrict__ add, uint32_t nAdd, uint32_t n,
                                  const uint32_t *__restrict__ holes, uint64_t capacity,
                                  uint32_t *status)
{
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nAdd) return;
    uint64_t dst = (uint64_t)(n - *holes) + s;
    if (dst >= capacity) {
        atomicOr(status, 1u);
        return;
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) sp(c, dst) = add[6 * (uint64_t)s + c];
}

__global__ void k_new_count(uint32_t n, uint32_t nAdd, const uint32_t *holes, uint32_t *newCount)
{
    *newCount = n - *holes + nAdd;
}

}  // namespace

int oecl_launch_hole_scan(const uint8_t *flags, uint32_t n, uint32_t *counts, uint32_t *offsets,
                          hipStream_t s)
{
    uint32_t nt = oecl_scan_tiles(n);
    if (nt == 0) nt = 1;
    hipLaunchKernelGGL(k_tile_count_zero, dim3(nt), dim3(OECL_SCAN_THREADS), 0, s, flags, n, counts);
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(OECL_SCAN_THREADS), 0, s, counts, nt, offsets);
    OECL_CHECK_LAUNCH("hole scan");
    return OECL_OK;
}

int oecl_launch_scan_u32(const uint32_t *in, uint32_t n, uint32_t *offsets, uint32_t *work,
                         hipStream_t s)
{
    uint32_t nt = oecl_scan_tiles(n);
    if (nt == 0) nt = 1;
    uint32_t *counts = work, *tileOff = work + nt;
    hipLaunchKernelGGL(k_tile_sum_u32, dim3(nt), dim3(OECL_SCAN_THREADS), 0, s, in, n, counts);
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(OECL_SCAN_THREADS), 0, s, counts, nt, tileOff);
    hipLaunchKernelGGL(k_scan_apply_u32, dim3(nt), dim3(OECL_SCAN_THREADS), 0, s, in, n, tileOff,
                       offsets);
    OECL_CHECK_LAUNCH("u32 scan");
    return OECL_OK;
}

OECL_EXPORT uint64_t oecl_scan_work_bytes(uint32_t n)
{
    uint64_t nt = oecl_scan_tiles(n) + 1;
    return 4 * (2 * nt + 8);
}

OECL_EXPORT int oecl_scan_u32(const uint32_t *in, uint32_t n, uint32_t *offsetsOut, void *work,
                              void *stream)
{
    if (!offsetsOut || !work || (n > 0 && !in)) return OECL_ERR_ARG;
    return oecl_launch_scan_u32(in, n, offsetsOut, (uint32_t *)work, (hipStream_t)stream);
}

OECL_EXPORT int oecl_index_inside(const oecl_wall_t *wall, int dtype, void *data, uint64_t stride,
                                  uint32_t n, uint8_t *isInside, int32_t *clipped, void *stream)
{
    if (!data || !isInside) return OECL_ERR_ARG;
    if (n == 0) return OECL_OK;
    WallDev wd;
    int rc = wall_to_device(wall, &wd);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    if (dtype == OECL_F64)
        hipLaunchKernelGGL(k_index_inside<double>, dim3(oecl_blocks(n, 256)), dim3(256), 0, s, wd,
                           Soa<double>{(double *)data, stride}, n, isInside, clipped);
    else if (dtype == OECL_F32)
        hipLaunchKernelGGL(k_index_inside<float>, dim3(oecl_blocks(n, 256)), dim3(256), 0, s, wd,
                           Soa<float>{(float *)data, stride}, n, isInside, clipped);
    else
        return OECL_ERR_ARG;
    OECL_CHECK_LAUNCH("oecl_index_inside");
    return OECL_OK;
}

OECL_EXPORT int oecl_save_absorbed(int dtype, const void *data, uint64_t stride, uint32_t n,
                                   const uint8_t *isInside, void *absorbed, uint32_t absCap,
                                   uint32_t *nAbsorbed, void *work, void *stream)
{
    if (!isInside || !nAbsorbed || !work || (n > 0 && !data)) return OECL_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    uint32_t nt = oecl_scan_tiles(n);
    if (nt == 0) nt = 1;
    uint32_t *counts = (uint32_t *)work, *off = counts + nt;
    int rc = oecl_launch_hole_scan(isInside, n, counts, off, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_copy_u32, dim3(1), dim3(1), 0, s, off + nt, nAbsorbed);
    if (n > 0 && absorbed && absCap > 0) {
        if (dtype == OECL_F64)
            hipLaunchKernelGGL(k_save_absorbed<double>, dim3(nt), dim3(OECL_SCAN_THREADS), 0, s,
                               Soa<double>{(double *)data, stride}, n, isInside, off,
                               (double *)absorbed, absCap);
        else if (dtype == OECL_F32)
            hipLaunchKernelGGL(k_save_absorbed<float>, dim3(nt), dim3(OECL_SCAN_THREADS), 0, s,
                               Soa<float>{(float *)data, stride}, n, isInside, off,
                               (float *)absorbed, absCap);
        else
            return OECL_ERR_ARG;
    }
    OECL_CHECK_LAUNCH("oecl_save_absorbed");
    return OECL_OK;
}

// ---------------------------------------------------------------- saveAbsorbed without allocation
// The hole scan's tile counts and offsets (oecl_scan_work_bytes(n) bytes) live in this array
// instead of a workspace argument.  Every kernel writes the words it later reads (the tile
// counts, then the offsets), so nothing has to start at zero.  Each GPU holds its own copy at
// its own address; host code uses that address, never the array's name.
namespace {
constexpr uint64_t kSaveAbsorbedScratchBytes = 128u << 10;
__device__ __align__(256) uint32_t scratch[kSaveAbsorbedScratchBytes / 4];
constexpr int kMaxDevices = 64;
void *g_scratchDev[kMaxDevices];
std::mutex g_scratchLock;

int save_absorbed_scratch(void **out)
{
    int dev = 0;
    OECL_HIP(hipGetDevice(&dev), "save_absorbed_noalloc: current device");
    if (dev < 0 || dev >= kMaxDevices) {
        oecl_set_error_msg("oecl_save_absorbed_noalloc: device index beyond 63");
        return OECL_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_scratchLock);
    if (!g_scratchDev[dev]) {
        void *p = nullptr;
        OECL_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "save_absorbed_noalloc: scratch address");
        g_scratchDev[dev] = p;
    }
    *out = g_scratchDev[dev];
    return OECL_OK;
}
}  // namespace

OECL_EXPORT uint32_t oecl_save_absorbed_noalloc_max(void)
{
    // largest n whose scan workspace fits: 4 (2 (tiles + 1) + 8) bytes
    const uint64_t tiles = kSaveAbsorbedScratchBytes / 8 - 5;
    const uint64_t n = tiles * OECL_SCAN_TILE;
    return n > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)n;
}

OECL_EXPORT int oecl_save_absorbed_noalloc(int dtype, const void *data, uint64_t stride, uint32_t n,
                                           const uint8_t *isInside, void *absorbed, uint32_t absCap,
                                           uint32_t *nAbsorbed, void *stream)
{
    if (!isInside || !nAbsorbed || (n > 0 && !data)) return OECL_ERR_ARG;
    if (dtype != OECL_F64 && dtype != OECL_F32) return OECL_ERR_ARG;
    if (oecl_scan_work_bytes(n) > kSaveAbsorbedScratchBytes) {
        char m[160];
        snprintf(m, sizeof(m), "oecl_save_absorbed_noalloc: n = %u exceeds the scratch limit of %u particles", n,
                 oecl_save_absorbed_noalloc_max());
        oecl_set_error_msg(m);
        return OECL_ERR_CAPACITY;
    }
    void *work = nullptr;
    const int rc = save_absorbed_scratch(&work);
    if (rc != OECL_OK) return rc;
    return oecl_save_absorbed(dtype, data, stride, n, isInside, absorbed, absCap, nAbsorbed, work, stream);
}

OECL_EXPORT int oecl_wall_interaction(const oecl_wall_t *wall, int dtype, void *absorbed,
                                      uint32_t nAbs, double dt, void *remaining, void *normals,
                                      uint32_t *notFound, void *stream)
{
    if (nAbs == 0) return OECL_OK;
    if (!absorbed || !remaining || !normals) return OECL_ERR_ARG;
    WallDev wd;
    int rc = wall_to_device(wall, &wd);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    if (dtype == OECL_F64)
        hipLaunchKernelGGL(k_wall_interaction<double>, dim3(oecl_blocks(nAbs, 256)), dim3(256), 0, s,
                           wd, (double *)absorbed, nAbs, dt, (double *)remaining,
                           (double *)normals, notFound);
    else if (dtype == OECL_F32)
        hipLaunchKernelGGL(k_wall_interaction<float>, dim3(oecl_blocks(nAbs, 256)), dim3(256), 0, s,
                           wd, (float *)absorbed, nAbs, dt, (float *)remaining, (float *)normals,
                           notFound);
    else
        return OECL_ERR_ARG;
    OECL_CHECK_LAUNCH("oecl_wall_interaction");
    return OECL_OK;
}

OECL_EXPORT uint64_t oecl_compact_work_bytes(uint32_t n, uint32_t nAdd)
{
    (void)nAdd;
    uint64_t nt = oecl_scan_tiles(n) + 1;
    return 4 * (2 * nt + 16) + 4 * (uint64_t)n + 256;
}

template <typename T>
static int compact_typed(T *data, T *spare, uint64_t stride, uint32_t n, const uint8_t *keep,
                         const T *add, uint32_t nAdd, int mode, uint32_t *newCount, void *work,
                         hipStream_t s)
{
    uint32_t nt = oecl_scan_tiles(n);
    if (nt == 0) nt = 1;
    uint32_t *counts = (uint32_t *)work;
    uint32_t *off = counts + nt;
    uint32_t *status = off + nt + 1;
    uint32_t *holePos = status + 8;
    uint32_t *holes = off + nt;
    Soa<T> p{data, stride}, sp{spare, stride};
    int rc = oecl_launch_hole_scan(keep, n, counts, off, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_hole_positions, dim3(nt), dim3(OECL_SCAN_THREADS), 0, s, keep, n, off,
                       holePos);
    if (mode == 2) {
        if (!spare) return OECL_ERR_ARG;
        hipLaunchKernelGGL(k_stable_to_spare<T>, dim3(nt), dim3(OECL_SCAN_THREADS), 0, s, p, sp,
                           keep, n, off);
        if (nAdd)
            hipLaunchKernelGGL(k_append_to_spare<T>, dim3(oecl_blocks(nAdd, 256)), dim3(256), 0, s,
                               sp, add, nAdd, n, holes, stride, status);
    } else {
        if (mode == 1) {
            if (!spare) return OECL_ERR_ARG;
            // Shift first: it reads kept particles after the last filled hole, which the
            // additions never overwrite.
            hipLaunchKernelGGL(k_shift_to_spare<T>, dim3(nt), dim3(OECL_SCAN_THREADS), 0, s, p, sp,
                               keep, n, nAdd, off, holes, holePos);
            hipLaunchKernelGGL(k_copy_back<T>, dim3(1024), dim3(256), 0, s, p, sp, n, nAdd, holes,
                               holePos);
        }
        if (nAdd)
            hipLaunchKernelGGL(k_place_additions<T>, dim3(oecl_blocks(nAdd, 256)), dim3(256), 0, s, p,
                               add, nAdd, n, holes, holePos, stride, status);
        // Modes 0 and 3 fill the holes left below the new count from the tail; mode 1 has
        // no such holes after its shift (or none at all when nAdd >= holes).
        if (mode == 0 || mode == 3)
            hipLaunchKernelGGL(k_tail_fill<T>, dim3(nt), dim3(OECL_SCAN_THREADS), 0, s, p, keep, n,
                               nAdd, off, holes, holePos);
    }
    if (newCount)
        hipLaunchKernelGGL(k_new_count, dim3(1), dim3(1), 0, s, n, nAdd, holes, newCount);
    OECL_CHECK_LAUNCH("oecl_add_and_remove");
    return OECL_OK;
}

OECL_EXPORT int oecl_add_and_remove(int dtype, void *data, void *spare, uint64_t stride, uint32_t n,
                                    const uint8_t *keep, const void *add, uint32_t nAdd, int mode,
                                    uint32_t *newCount, void *work, void *stream)
{
    if (!data || !keep || !work || mode < 0 || mode > 3) return OECL_ERR_ARG;
    if (nAdd > 0 && !add) return OECL_ERR_ARG;
    if (mode == 0 && nAdd != 0) return OECL_ERR_ARG;
    if (mode == 1 && nAdd == 0) mode = 0;  // particles.pyx:104-107: only removals
    hipStream_t s = (hipStream_t)stream;
    if (dtype == OECL_F64)
        return compact_typed<double>((double *)data, (double *)spare, stride, n, keep,
                                     (const double *)add, nAdd, mode, newCount, work, s);
    if (dtype == OECL_F32)
        return compact_typed<float>((float *)data, (float *)spare, stride, n, keep,
                                    (const float *)add, nAdd, mode, newCount, work, s);
    return OECL_ERR_ARG;
}
