// Macro-particle count control (particleManagement.pyx:35-238) on the device.
//
// globalRandom: weight rescale, then a uniformly random kill set (the k smallest of per-
// particle Philox keys, ties broken by index) or uniformly random clones.
// localRandom: the reference's weight thresholds and split rule; the merge replaces the
// serial kd-tree nearest-neighbour search by a Morton (Z-order) sort of the quantised 5-D
// weighted phase-space coordinates, after which particles that are neighbours in Z order and
// within the reference's distance limit are merged in groups of up to G, lightest groups
// first.  Every merge gives the survivor the sum of the group's weights.
#include "common.h"
#include "wall.h"

#define PURPOSE_MANAGE 3u

namespace {

template <typename T>
__global__ void k_scale_weights(Soa<T> p, uint32_t n, double f)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p(5, i) = (T)((double)p(5, i) * f);
}

__global__ void k_fill_u8(uint8_t *a, uint32_t n, uint8_t v)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_random_keys(uint32_t n, uint32_t stream, uint32_t s0, uint32_t s1, uint64_t *keys,
                              uint32_t *vals)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Philox4 r = philox4x32_10(i, 0u, stream, PURPOSE_MANAGE, s0, s1);
    keys[i] = ((uint64_t)r.v[0] << 32) | i;
    vals[i] = i;
}

__global__ void k_mark_first(const uint32_t *perm, uint32_t k, uint8_t *keep)
{
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < k) keep[perm[r]] = 0;
}

template <typename T>
__global__ void k_clone_rows(Soa<T> p, uint32_t n, uint32_t nAdd, uint32_t stream, uint32_t s0,
                             uint32_t s1, T *add)
{
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nAdd) return;
    double u = philox_uniform(s, 2u, stream, PURPOSE_MANAGE, s0, s1);
    uint32_t idx = (uint32_t)(u * n);
    if (idx >= n) idx = n - 1;
#pragma unroll
    for (int c = 0; c < 6; ++c) add[6 * (uint64_t)s + c] = p(c, idx);
}

// Upper bound (first index whose value exceeds v) of three values in the sorted weights,
// as particleManagement.pyx:244-261 (_searchSorted) returns it.
template <typename T>
__global__ void k_search3(Soa<T> p, uint32_t n, double v0, double v1, double v2, uint32_t *out)
{
    int t = threadIdx.x;
    if (t >= 3) return;
    double v = t == 0 ? v0 : (t == 1 ? v1 : v2);
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if ((double)p(5, mid) > v) hi = mid;
        else lo = mid + 1;
    }
    out[t] = lo;
}

// Split of the heaviest particles (particleManagement.pyx:130-163): both halves get half
// the weight; the proposed displaced positions are built from integer cell indices exactly
// as the reference builds them and are used only if both pass the wall's isInside.
template <typename T>
__global__ void k_split(Soa<T> p, uint32_t first, uint32_t cnt, WallDev w, double dxi, double dyi,
                        double lxHalf, double lyHalf, uint32_t stream, uint32_t s0, uint32_t s1, T *add)
{
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= cnt) return;
    uint64_t i = first + k;
    double wt = (double)p(5, i) * 0.5;
    p(5, i) = (T)wt;
    T *a = add + 6 * (uint64_t)k;
    a[2] = p(2, i);
    a[3] = p(3, i);
    a[4] = p(4, i);
    a[5] = (T)wt;
    double x = p(0, i), y = p(1, i);
    uint32_t cx = (uint32_t)((x + lxHalf) * dxi);
    uint32_t cy = (uint32_t)((y + lyHalf) * dyi);
    double X = cx, Y = cy, r;
    double u0 = philox_uniform(k, 0u, stream, PURPOSE_MANAGE + 1u, s0, s1);
    double u1 = philox_uniform(k, 1u, stream, PURPOSE_MANAGE + 1u, s0, s1);
    r = X < 0.5 ? u0 * fmin(X, 0.5 - X) : u0 * fmin(X - 0.5, 1. - X);
    double n0x = X + r, n1x = X - r;
    r = Y < 0.5 ? u1 * fmin(Y, 0.5 - Y) : u1 * fmin(Y - 0.5, 1. - Y);
    double n0y = Y + r, n1y = Y - r;
    if (wall_is_inside_point(w, n0x, n0y) == 0 || wall_is_inside_point(w, n1x, n1y) == 0) {
        a[0] = (T)x;
        a[1] = (T)y;
    } else {
        p(0, i) = (T)n0x;
        p(1, i) = (T)n0y;
        a[0] = (T)n1x;
        a[1] = (T)n1y;
    }
}

// Per-block min / max of the 5 weighted coordinates over [0, n).
template <typename T>
__global__ void __launch_bounds__(256) k_bbox(Soa<T> p, uint32_t n, const double *__restrict__ pw,
                                              double *__restrict__ part)
{
    __shared__ double lo[256][5], hi[256][5];
    double l[5], h[5];
    for (int c = 0; c < 5; ++c) {
        l[c] = INFINITY;
        h[c] = -INFINITY;
    }
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
        for (int c = 0; c < 5; ++c) {
            double v = (double)p(c, i) * pw[c];
            l[c] = fmin(l[c], v);
            h[c] = fmax(h[c], v);
        }
    for (int c = 0; c < 5; ++c) {
        lo[threadIdx.x][c] = l[c];
        hi[threadIdx.x][c] = h[c];
    }
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st)
            for (int c = 0; c < 5; ++c) {
                lo[threadIdx.x][c] = fmin(lo[threadIdx.x][c], lo[threadIdx.x + st][c]);
                hi[threadIdx.x][c] = fmax(hi[threadIdx.x][c], hi[threadIdx.x + st][c]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 5) {
        part[blockIdx.x * 10 + threadIdx.x] = lo[0][threadIdx.x];
        part[blockIdx.x * 10 + 5 + threadIdx.x] = hi[0][threadIdx.x];
    }
}

__global__ void k_bbox_final(const double *part, uint32_t nb, double *box)
{
    int c = threadIdx.x;
    if (c >= 5) return;
    double l = INFINITY, h = -INFINITY;
    for (uint32_t b = 0; b < nb; ++b) {
        l = fmin(l, part[b * 10 + c]);
        h = fmax(h, part[b * 10 + 5 + c]);
    }
    box[c] = l;
    box[5 + c] = h;
}

__device__ __forceinline__ uint64_t spread5(uint32_t v)
{
    // places bit b of a 12-bit value at bit 5*b
    uint64_t r = 0;
#pragma unroll
    for (int b = 0; b < 12; ++b) r |= (uint64_t)((v >> b) & 1u) << (5 * b);
    return r;
}

template <typename T>
__global__ void k_morton(Soa<T> p, uint32_t n, const double *__restrict__ pw, const double *__restrict__ box,
                         uint64_t *keys, uint32_t *vals)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t key = 0;
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        double lo = box[c], hi = box[5 + c];
        double v = (double)p(c, i) * pw[c];
        double f = hi > lo ? (v - lo) / (hi - lo) : 0.;
        uint32_t q = (uint32_t)fmin(4095., fmax(0., f * 4096.));
        key |= spread5(q) << c;
    }
    keys[i] = key;
    vals[i] = i;
}

// Group g = Z-order positions [G g, G g + G).  Members within sqrt(disLimit) (squared
// distance compared with disLimit, as particleManagement.pyx:185 compares) of the group's
// first member are eligible; the group's sort key is the weight of its lightest member,
// or all ones when fewer than two members are eligible.  elig[g] is a member bit mask.
template <typename T>
__global__ void k_group_keys(Soa<T> p, uint32_t nc, const uint32_t *__restrict__ zperm,
                             const double *__restrict__ pw, int G, double disLimit, uint64_t *gkeys,
                             uint32_t *gvals, uint8_t *elig)
{
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t ng = (nc + G - 1) / G;
    if (g >= ng) return;
    uint32_t a = zperm[(uint64_t)g * G];
    double pa[5];
    for (int c = 0; c < 5; ++c) pa[c] = (double)p(c, a) * pw[c];
    double wmin = (double)p(5, a);
    uint8_t mask = 1;
    int cnt = 1;
    for (int k = 1; k < G; ++k) {
        uint64_t pos = (uint64_t)g * G + k;
        if (pos >= nc) break;
        uint32_t b = zperm[pos];
        double d2 = 0.;
        for (int c = 0; c < 5; ++c) {
            double d = (double)p(c, b) * pw[c] - pa[c];
            d2 += d * d;
        }
        if (d2 < disLimit) {
            mask |= (uint8_t)(1u << k);
            ++cnt;
            wmin = fmin(wmin, (double)p(5, b));
        }
    }
    elig[g] = mask;
    gvals[g] = g;
    gkeys[g] = cnt >= 2 ? (uint64_t)__double_as_longlong(wmin) : ~0ull;
}

// Merge the first `take` groups of the sorted group order.  Survivor chosen with
// probability proportional to weight; it receives the group's total weight.  Scheme 1
// (pairs) rescales the survivor's velocity as particleManagement.pyx:209-232 does.
template <typename T>
__global__ void k_group_merge(Soa<T> p, uint32_t nc, const uint32_t *__restrict__ zperm,
                              const uint32_t *__restrict__ gorder, const uint8_t *__restrict__ elig,
                              uint32_t take, int G, int scheme, uint32_t stream, uint32_t s0, uint32_t s1,
                              uint8_t *keep, uint32_t *merged)
{
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= take) return;
    uint32_t g = gorder[r];
    uint8_t mask = elig[g];
    uint32_t idx[8];
    double wsum = 0.;
    int cnt = 0;
    for (int k = 0; k < G; ++k)
        if (mask & (1u << k)) {
            uint64_t pos = (uint64_t)g * G + k;
            if (pos >= nc) break;
            idx[cnt] = zperm[pos];
            wsum += (double)p(5, idx[cnt]);
            ++cnt;
        }
    if (cnt < 2) return;
    double u = philox_uniform(g, 4u, stream, PURPOSE_MANAGE + 2u, s0, s1) * wsum;
    int sv = cnt - 1;
    double acc = 0.;
    for (int k = 0; k < cnt; ++k) {
        acc += (double)p(5, idx[k]);
        if (u < acc) {
            sv = k;
            break;
        }
    }
    uint32_t s = idx[sv];
    if (scheme == 1 && cnt == 2) {
        uint32_t o = idx[1 - sv];
        double ws = p(5, s), wo = p(5, o), W = ws + wo;
        double vs2 = (double)p(2, s) * p(2, s) + (double)p(3, s) * p(3, s) + (double)p(4, s) * p(4, s);
        double vo2 = (double)p(2, o) * p(2, o) + (double)p(3, o) * p(3, o) + (double)p(4, o) * p(4, o);
        double f = ws / W + vo2 / vs2 * wo / W;
        p(2, s) = (T)(p(2, s) * f);
        p(3, s) = (T)(p(3, s) * f);
        p(4, s) = (T)(p(4, s) * f);
    }
    p(5, s) = (T)wsum;
    for (int k = 0; k < cnt; ++k)
        if (k != sv) keep[idx[k]] = 0;
    atomicAdd(merged, (uint32_t)(cnt - 1));
}

}  // namespace

OECL_EXPORT int oecl_scale_weights(int dtype, void *data, uint64_t stride, uint32_t n, double f,
                                   void *stream)
{
    if (!data) return OECL_ERR_ARG;
    if (n == 0) return OECL_OK;
    hipStream_t s = (hipStream_t)stream;
    if (dtype == OECL_F64)
        hipLaunchKernelGGL(k_scale_weights<double>, dim3(oecl_blocks(n, 256)), dim3(256), 0, s,
                           Soa<double>{(double *)data, stride}, n, f);
    else if (dtype == OECL_F32)
        hipLaunchKernelGGL(k_scale_weights<float>, dim3(oecl_blocks(n, 256)), dim3(256), 0, s,
                           Soa<float>{(float *)data, stride}, n, f);
    else
        return OECL_ERR_ARG;
    OECL_CHECK_LAUNCH("oecl_scale_weights");
    return OECL_OK;
}

OECL_EXPORT int oecl_fill_u8(uint8_t *a, uint32_t n, uint8_t v, void *stream)
{
    if (!a) return OECL_ERR_ARG;
    if (n) hipLaunchKernelGGL(k_fill_u8, dim3(oecl_blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, a, n, v);
    OECL_CHECK_LAUNCH("oecl_fill_u8");
    return OECL_OK;
}

// Uniform random subset of k of the n particles: keep[i] = 0 for the chosen ones.
OECL_EXPORT int oecl_random_kill(uint32_t n, uint32_t k, uint32_t stream, uint32_t s0, uint32_t s1,
                                 uint8_t *keep, uint64_t *keys, uint32_t *perm, void *work, void *st)
{
    if (!keep || !keys || !perm || !work || k > n) return OECL_ERR_ARG;
    hipStream_t s = (hipStream_t)st;
    hipLaunchKernelGGL(k_fill_u8, dim3(oecl_blocks(n, 256)), dim3(256), 0, s, keep, n, (uint8_t)1);
    hipLaunchKernelGGL(k_random_keys, dim3(oecl_blocks(n, 256)), dim3(256), 0, s, n, stream, s0, s1, keys,
                       perm);
    int rc = oecl_radix_sort_u64(keys, perm, n, 64, work, st);
    if (rc) return rc;
    if (k) hipLaunchKernelGGL(k_mark_first, dim3(oecl_blocks(k, 256)), dim3(256), 0, s, perm, k, keep);
    OECL_CHECK_LAUNCH("oecl_random_kill");
    return OECL_OK;
}

OECL_EXPORT int oecl_random_clones(int dtype, const void *data, uint64_t stride, uint32_t n,
                                   uint32_t nAdd, uint32_t stream, uint32_t s0, uint32_t s1, void *add,
                                   void *st)
{
    if (!data || !add || n == 0) return OECL_ERR_ARG;
    if (nAdd == 0) return OECL_OK;
    hipStream_t s = (hipStream_t)st;
    if (dtype == OECL_F64)
        hipLaunchKernelGGL(k_clone_rows<double>, dim3(oecl_blocks(nAdd, 256)), dim3(256), 0, s,
                           Soa<double>{(double *)data, stride}, n, nAdd, stream, s0, s1, (double *)add);
    else if (dtype == OECL_F32)
        hipLaunchKernelGGL(k_clone_rows<float>, dim3(oecl_blocks(nAdd, 256)), dim3(256), 0, s,
                           Soa<float>{(float *)data, stride}, n, nAdd, stream, s0, s1, (float *)add);
    else
        return OECL_ERR_ARG;
    OECL_CHECK_LAUNCH("oecl_random_clones");
    return OECL_OK;
}

OECL_EXPORT int oecl_search_sorted_weights(int dtype, const void *data, uint64_t stride, uint32_t n,
                                           double v0, double v1, double v2, uint32_t *out, void *st)
{
    if (!data || !out || n == 0) return OECL_ERR_ARG;
    hipStream_t s = (hipStream_t)st;
    if (dtype == OECL_F64)
        hipLaunchKernelGGL(k_search3<double>, dim3(1), dim3(64), 0, s, Soa<double>{(double *)data, stride},
                           n, v0, v1, v2, out);
    else if (dtype == OECL_F32)
        hipLaunchKernelGGL(k_search3<float>, dim3(1), dim3(64), 0, s, Soa<float>{(float *)data, stride}, n,
                           v0, v1, v2, out);
    else
        return OECL_ERR_ARG;
    OECL_CHECK_LAUNCH("oecl_search_sorted_weights");
    return OECL_OK;
}

OECL_EXPORT int oecl_split_heaviest(const oecl_wall_t *wall, int dtype, void *data, uint64_t stride,
                                    uint32_t first, uint32_t cnt, double dx, double dy, double lx,
                                    double ly, uint32_t stream, uint32_t s0, uint32_t s1, void *add,
                                    void *st)
{
    if (!data || !add) return OECL_ERR_ARG;
    if (cnt == 0) return OECL_OK;
    WallDev wd;
    int rc = wall_to_device(wall, &wd);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)st;
    if (dtype == OECL_F64)
        hipLaunchKernelGGL(k_split<double>, dim3(oecl_blocks(cnt, 256)), dim3(256), 0, s,
                           Soa<double>{(double *)data, stride}, first, cnt, wd, 1. / dx, 1. / dy, lx * 0.5,
                           ly * 0.5, stream, s0, s1, (double *)add);
    else if (dtype == OECL_F32)
        hipLaunchKernelGGL(k_split<float>, dim3(oecl_blocks(cnt, 256)), dim3(256), 0, s,
                           Soa<float>{(float *)data, stride}, first, cnt, wd, 1. / dx, 1. / dy, lx * 0.5,
                           ly * 0.5, stream, s0, s1, (float *)add);
    else
        return OECL_ERR_ARG;
    OECL_CHECK_LAUNCH("oecl_split_heaviest");
    return OECL_OK;
}

// Morton merge over particles [0, nc).  pw[5] phase-space weights (device), G group size
// (2..8), disLimit the squared-distance limit, target merges toMerge.  keep[nc] is set;
// *merged (device) receives the number of particles removed.
OECL_EXPORT uint64_t oecl_merge_work_bytes(uint32_t nc)
{
    return oecl_sort_work_bytes(nc) + 24ull * nc + 8192 + 512 * 80;
}

OECL_EXPORT int oecl_morton_merge(int dtype, void *data, uint64_t stride, uint32_t nc, const double *pw,
                                  int G, double disLimit, uint32_t toMerge, int scheme, uint32_t stream,
                                  uint32_t s0, uint32_t s1, uint8_t *keep, uint32_t *merged,
                                  uint32_t *hostGroupsUsed, void *work, void *st)
{
    if (!data || !pw || !keep || !merged || !work || G < 2 || G > 8) return OECL_ERR_ARG;
    hipStream_t s = (hipStream_t)st;
    hipLaunchKernelGGL(k_fill_u8, dim3(oecl_blocks(nc ? nc : 1, 256)), dim3(256), 0, s, keep, nc, (uint8_t)1);
    hipLaunchKernelGGL(k_fill_u8, dim3(1), dim3(64), 0, s, (uint8_t *)merged, 4u, (uint8_t)0);
    if (nc < 2 || toMerge == 0) {
        if (hostGroupsUsed) *hostGroupsUsed = 0;
        return OECL_OK;
    }
    char *w = (char *)work;
    uint64_t *keys = (uint64_t *)w;
    uint32_t *zperm = (uint32_t *)(keys + nc);
    uint8_t *elig = (uint8_t *)(zperm + nc);
    double *part = (double *)(((uintptr_t)(elig + nc) + 15) & ~(uintptr_t)15);
    double *box = part + 512 * 10;
    char *sw = (char *)(box + 16);
    uint32_t ng = (nc + G - 1) / G;
    // the group keys reuse the tail of the work after the Z sort is done
    Soa<double> pd{(double *)data, stride};
    Soa<float> pf{(float *)data, stride};
    if (dtype == OECL_F64) hipLaunchKernelGGL(k_bbox<double>, dim3(512), dim3(256), 0, s, pd, nc, pw, part);
    else hipLaunchKernelGGL(k_bbox<float>, dim3(512), dim3(256), 0, s, pf, nc, pw, part);
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(64), 0, s, part, 512u, box);
    if (dtype == OECL_F64)
        hipLaunchKernelGGL(k_morton<double>, dim3(oecl_blocks(nc, 256)), dim3(256), 0, s, pd, nc, pw, box, keys,
                           zperm);
    else
        hipLaunchKernelGGL(k_morton<float>, dim3(oecl_blocks(nc, 256)), dim3(256), 0, s, pf, nc, pw, box, keys,
                           zperm);
    int rc = oecl_radix_sort_u64(keys, zperm, nc, 64, sw, st);
    if (rc) return rc;
    uint64_t *gkeys = keys;   // Morton keys no longer needed
    uint32_t *gvals = (uint32_t *)(sw + oecl_sort_work_bytes(nc));
    if (dtype == OECL_F64)
        hipLaunchKernelGGL(k_group_keys<double>, dim3(oecl_blocks(ng, 256)), dim3(256), 0, s, pd, nc, zperm, pw,
                           G, disLimit, gkeys, gvals, elig);
    else
        hipLaunchKernelGGL(k_group_keys<float>, dim3(oecl_blocks(ng, 256)), dim3(256), 0, s, pf, nc, zperm, pw,
                           G, disLimit, gkeys, gvals, elig);
    rc = oecl_radix_sort_u64(gkeys, gvals, ng, 64, sw, st);
    if (rc) return rc;
    // Each eligible group removes at least one particle: merging the first toMerge sorted
    // groups (all eligible when their key is not all ones) reaches at least toMerge
    // removals if enough eligible groups exist; a host pass trims to exactly what is needed.
    uint64_t *hk = (uint64_t *)malloc(8ull * ng);
    uint8_t *hel = (uint8_t *)malloc(ng);
    uint32_t *hv = (uint32_t *)malloc(4ull * ng);
    OECL_HIP(hipMemcpyAsync(hk, gkeys, 8ull * ng, hipMemcpyDeviceToHost, s), "merge copy");
    OECL_HIP(hipMemcpyAsync(hv, gvals, 4ull * ng, hipMemcpyDeviceToHost, s), "merge copy");
    OECL_HIP(hipMemcpyAsync(hel, elig, ng, hipMemcpyDeviceToHost, s), "merge copy");
    OECL_HIP(hipStreamSynchronize(s), "merge sync");
    uint32_t take = 0, removed = 0;
    while (take < ng && removed < toMerge && hk[take] != ~0ull) {
        removed += (uint32_t)__builtin_popcount(hel[hv[take]]) - 1;
        ++take;
    }
    free(hk);
    free(hel);
    free(hv);
    if (hostGroupsUsed) *hostGroupsUsed = take;
    if (take) {
        if (dtype == OECL_F64)
            hipLaunchKernelGGL(k_group_merge<double>, dim3(oecl_blocks(take, 128)), dim3(128), 0, s, pd, nc,
                               zperm, gvals, elig, take, G, scheme, stream, s0, s1, keep, merged);
        else
            hipLaunchKernelGGL(k_group_merge<float>, dim3(oecl_blocks(take, 128)), dim3(128), 0, s, pf, nc,
                               zperm, gvals, elig, take, G, scheme, stream, s0, s1, keep, merged);
    }
    OECL_CHECK_LAUNCH("oecl_morton_merge");
    return OECL_OK;
}
