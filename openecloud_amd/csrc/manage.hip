// Macro-particle count control (particleManagement.pyx:35-238) on the device.
//
// globalRandom: weight rescale, then a uniformly random kill set (the k smallest of per-
// particle Philox keys, ties broken by index) or uniformly random clones.
// localRandom: the reference's weight thresholds and split rule; the merge keeps the reference's
// rule -- each particle, lightest first, is merged with its nearest neighbour in the weighted
// 5-D phase space (x, y, lambda vx, lambda vy, lambda vz) when the squared distance is below
// the limit, the survivor chosen with probability proportional to weight and given the pair's
// total weight -- and replaces the serial kd-tree loop by rounds of parallel matching: every
// particle still in the pool finds its exact nearest neighbour in the pool through a
// Morton-ordered tree (the weighted points sorted by their 5-D Z-order code, leaves of 8
// consecutive points, an implicit binary tree of bounding boxes over the leaves, branch and
// bound queries), a pair holds both of its particles only when it is the lightest-ranked claim
// on each (atomic minimum locks), and the accepted pairs are taken in rank order until the
// requested number of merges is reached.
#include "common.h"
#include "particles_dev.h"
#include "scan.h"
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
constexpr uint32_t NN_NONE = 0xFFFFFFFFu;    // no proposal (out of the pool)
constexpr uint32_t NN_RETIRE = 0xFFFFFFFEu;  // no neighbour within the limit: leaves the pool
constexpr uint32_t NN_MERGED = 0xFFFFFFFEu;  // lock value of a particle merged in this round
#ifndef OECL_NN_LEAF
#define OECL_NN_LEAF 8
#endif
constexpr uint32_t NN_LEAF = OECL_NN_LEAF;   // points per leaf of the tree

__global__ void k_zero_u32(uint32_t *a, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}

// Bounding box of the weighted 5-D points: per-block partial minima and maxima, then a final pass.
template <typename T>
__global__ void __launch_bounds__(256) k_bbox(Soa<T> p, uint32_t n, const double *__restrict__ pw, double *part)
{
    __shared__ double sh[10][256];
    double lo[5], hi[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        lo[c] = INFINITY;
        hi[c] = -INFINITY;
    }
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            const double v = (double)p(c, i) * pw[c];
            lo[c] = fmin(lo[c], v);
            hi[c] = fmax(hi[c], v);
        }
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        sh[c][threadIdx.x] = lo[c];
        sh[5 + c][threadIdx.x] = hi[c];
    }
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
#pragma unroll
            for (int c = 0; c < 5; ++c) {
                sh[c][threadIdx.x] = fmin(sh[c][threadIdx.x], sh[c][threadIdx.x + st]);
                sh[5 + c][threadIdx.x] = fmax(sh[5 + c][threadIdx.x], sh[5 + c][threadIdx.x + st]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 10) part[blockIdx.x * 10 + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ void k_bbox_final(const double *part, uint32_t nb, double *box)
{
    const int c = threadIdx.x;
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
    // bit b of a 12-bit value goes to bit 5 b
    uint64_t r = 0;
#pragma unroll
    for (int b = 0; b < 12; ++b) r |= (uint64_t)((v >> b) & 1u) << (5 * b);
    return r;
}

// 5-D Z-order code of each weighted point: 12 bits per coordinate, one cell size for all five
// (the largest extent of the bounding box over 4096), so that the cells of every level are cubes
// in the metric the search uses.  (Scaled per coordinate instead, the cells were some ten times
// longer along the velocities than along the positions for a thermal load, and every search ball
// met many more of those slab-shaped leaf boxes.)
template <typename T>
__global__ void k_morton(Soa<T> p, uint32_t n, const double *__restrict__ pw, const double *__restrict__ box,
                         uint64_t *keys, uint32_t *vals, uint8_t *gsize)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double span = 0.;
#pragma unroll
    for (int c = 0; c < 5; ++c) span = fmax(span, box[5 + c] - box[c]);
    uint64_t key = 0;
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        const double lo = box[c];
        const double v = (double)p(c, i) * pw[c];
        const double f = span > 0. ? (v - lo) / span : 0.;
        const uint32_t q = (uint32_t)fmin(4095., fmax(0., f * 4096.));
        key |= spread5(q) << c;
    }
    keys[i] = key;
    vals[i] = i;
    gsize[i] = 1;
}

// The weighted points in Z-order: pz[c * ldz + pos] = p(c, zperm[pos]) * pw[c] (a leaf's points
// are then 8 consecutive values per coordinate), zpos[zperm[pos]] = pos, and the pool sizes in
// Z-order, gz[pos] = gsize[zperm[pos]] (= 1 here).
template <typename T>
__global__ void k_zorder_points(Soa<T> p, uint32_t nc, const uint32_t *__restrict__ zperm,
                                const double *__restrict__ pw, double *pz, uint64_t ldz, uint32_t *zpos, uint8_t *gz)
{
    const uint32_t pos = blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= nc) return;
    const uint32_t b = zperm[pos];
#pragma unroll
    for (int c = 0; c < 5; ++c) pz[c * ldz + pos] = (double)p(c, b) * pw[c];
    zpos[b] = pos;
    gz[pos] = 1;
}

// Leaf boxes: leaf l holds sorted positions [8 l, 8 l + 8) (those below nc); node L + l of the
// heap-ordered tree (root 1, children 2k and 2k + 1).  Empty leaves get an empty box.
__global__ void k_leaf_boxes(uint32_t nc, const double *__restrict__ pz, uint64_t ldz, uint32_t L, double *blo,
                             double *bhi)
{
    const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= L) return;
    double lo[5], hi[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        lo[c] = INFINITY;
        hi[c] = -INFINITY;
    }
    for (uint32_t k = 0; k < NN_LEAF; ++k) {
        const uint64_t pos = (uint64_t)l * NN_LEAF + k;
        if (pos >= nc) break;
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            const double v = pz[c * ldz + pos];
            lo[c] = fmin(lo[c], v);
            hi[c] = fmax(hi[c], v);
        }
    }
    const uint64_t node = (uint64_t)L + l;
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        blo[node * 5 + c] = lo[c];
        bhi[node * 5 + c] = hi[c];
    }
}

// Inner boxes of one level: nodes [first, first + cnt), each the union of its two children.
__global__ void k_inner_boxes(uint32_t first, uint32_t cnt, double *blo, double *bhi)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const uint64_t k = (uint64_t)first + t, a = 2 * k, b = 2 * k + 1;
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        blo[k * 5 + c] = fmin(blo[a * 5 + c], blo[b * 5 + c]);
        bhi[k * 5 + c] = fmax(bhi[a * 5 + c], bhi[b * 5 + c]);
    }
}

__device__ __forceinline__ double box_dist2(const double *__restrict__ blo, const double *__restrict__ bhi,
                                            uint64_t k, const double *q)
{
    double d2 = 0.;
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        const double lo = blo[k * 5 + c], hi = bhi[k * 5 + c];
        const double d = q[c] < lo ? lo - q[c] : (q[c] > hi ? q[c] - hi : 0.);
        d2 += d * d;
    }
    return d2;   // +inf for an empty box
}

// Proposal of candidate a (rank = index, lightest first): its exact nearest neighbour in the pool
// (gsize != 0, groups of at most G particles together), ties to the lower index, found by branch
// and bound over the tree (boxes farther than the best distance pruned); a box containing
// particles that left the pool still bounds the pooled ones, and the boxes are refitted after
// every round in which a survivor that stays in the pool has moved (groups with scheme 1), so
// the search stays exact.  NN_RETIRE when none lies within the squared-distance limit
// (particleManagement.pyx:185 compares the squared distance with factorDisLimit * ds).  From the
// second round on, a proposal a -> b of the previous round stands without a new search when
// neither a nor b took part in a merge since (mrk): the pool has only lost particles, b is still
// in it and still the nearest.  Resets a's lock for this round.
__global__ void k_nn_propose(uint32_t nc, const uint32_t *__restrict__ zperm, uint32_t L,
                             const double *__restrict__ blo, const double *__restrict__ bhi,
                             const double *__restrict__ pz, uint64_t ldz, const uint8_t *__restrict__ gz,
                             const uint8_t *__restrict__ gsize, const uint8_t *__restrict__ mrk, double disLimit,
                             int G, uint32_t round, uint32_t *__restrict__ prop, uint32_t *__restrict__ lock,
                             uint32_t *__restrict__ cnt)
{
    // thread t searches for the t-th point in Z-order, so the lanes of a wave are neighbours in
    // space and walk nearly the same part of the tree
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nc) return;
    const uint32_t a = zperm[t];
    lock[a] = NN_NONE;
    const uint32_t sa = gz[t];
    if (sa == 0 || (int)sa >= G) {
        prop[a] = NN_NONE;
        return;
    }
    if (round > 0) {
        const uint32_t prev = prop[a];
        if (prev < nc && !mrk[a] && !mrk[prev] && gsize[prev] != 0 && (int)(sa + gsize[prev]) <= G) {
            atomicAdd(cnt, 1u);
            return;   // prop[a] stands
        }
    }
    double q[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) q[c] = pz[c * ldz + t];
    // a neighbour at or beyond the limit is rejected anyway (the reference's test is strict), so
    // the search starts with the limit as the best distance: only the ball inside it is searched,
    // and the result is the same
    double best = disLimit;
    uint32_t bi = NN_NONE;
    auto scan_leaf = [&](uint32_t leaf) {
        const uint64_t base = (uint64_t)leaf * NN_LEAF;
        for (uint32_t j = 0; j < NN_LEAF; ++j) {
            const uint64_t pos = base + j;
            if (pos >= nc) break;
            const uint32_t sb = gz[pos];
            if (sb == 0 || (int)(sa + sb) > G || pos == t) continue;
            double d2 = 0.;
#pragma unroll
            for (int c = 0; c < 5; ++c) {
                const double d = pz[c * ldz + pos] - q[c];
                d2 += d * d;
            }
            if (d2 <= best) {
                const uint32_t b = zperm[pos];
                if (d2 < best || (bi != NN_NONE && b < bi)) {
                    best = d2;
                    bi = b;
                }
            }
        }
    };
    // the candidate's own leaf and its neighbours in Z-order first: they hold the points nearest in
    // the ordering, which bounds the walk below tightly from the start
    const uint32_t own = t / NN_LEAF;
    scan_leaf(own);
    if (own > 0) scan_leaf(own - 1);
    if (own + 1 < L) scan_leaf(own + 1);
    // stackless depth-first walk of the heap-ordered tree (root 1, children 2k and 2k + 1, leaves
    // L..2L-1): descend into the left child of a box within reach; after a leaf or a box out of
    // reach, go to the right sibling of the nearest left child on the way up (done at the root).
    // The result, the least (distance, index) within the limit, does not depend on the order in
    // which points are met (a point met twice changes nothing).
    uint32_t k = 1;
    for (;;) {
        if (box_dist2(blo, bhi, k, q) <= best) {
            if (k < L) {
                k = 2 * k;
                continue;
            }
            const uint32_t leaf = k - L;
            if (leaf + 1 < own || leaf > own + 1) scan_leaf(leaf);
        }
        while (k & 1u) k >>= 1;
        if (k == 0) break;
        ++k;
    }
    if (bi == NN_NONE) {
        prop[a] = NN_RETIRE;
        return;
    }
    prop[a] = bi;
    atomicAdd(cnt, 1u);
}

// Locks: each particle keeps the lowest-ranked proposal that involves it.
__global__ void k_nn_lock(uint32_t nc, const uint32_t *__restrict__ prop, uint32_t *lock, uint8_t *mrk)
{
    const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nc) return;
    mrk[a] = 0;   // the proposals have read the previous round's marks
    const uint32_t b = prop[a];
    if (b >= nc) return;
    atomicMin(lock + a, a);
    atomicMin(lock + b, a);
}

__global__ void k_nn_accept(uint32_t nc, const uint32_t *__restrict__ prop, const uint32_t *__restrict__ lock,
                            uint32_t *acc)
{
    const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nc) return;
    const uint32_t b = prop[a];
    acc[a] = (b < nc && lock[a] == a && lock[b] == a) ? 1u : 0u;
}

// Accepted pair a -> prop[a] of rank order off[a] < need: the survivor is a with probability
// w_a / (w_a + w_b) (particleManagement.pyx:186-191) and takes the pair's weight; scheme 1 rescales
// its velocity to conserve the pair's kinetic energy (:209-232).  pairs (optional) records
// (survivor, removed) at slot base + off[a].  Groups: the survivor stays in the pool while its
// group has fewer than G particles (G = 2: the reference's pairs, the survivor leaves it).
template <typename T>
__global__ void k_nn_merge(Soa<T> p, uint32_t nc, const uint32_t *__restrict__ prop, const uint32_t *__restrict__ acc,
                           const uint32_t *__restrict__ off, uint32_t need, int G, int scheme, uint32_t round,
                           uint32_t stream, uint32_t s0, uint32_t s1, uint8_t *gsize, uint32_t *lock, uint8_t *keep,
                           uint32_t *merged, uint32_t *pairs, uint32_t base, const uint32_t *__restrict__ zpos,
                           uint8_t *gz, uint8_t *mrk, double *pz, uint64_t ldz, const double *__restrict__ pw)
{
    const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nc || !acc[a] || off[a] >= need) return;
    const uint32_t b = prop[a];
    const double wa = (double)p(5, a), wb = (double)p(5, b), W = wa + wb;
    const double u = philox_uniform(a, round, stream, PURPOSE_MANAGE + 2u, s0, s1);
    const bool keepA = u < wa / W;
    const uint32_t sv = keepA ? a : b, rm = keepA ? b : a;
    if (scheme == 1) {
        const double ws = keepA ? wa : wb, wo = keepA ? wb : wa;
        const double vs2 = (double)p(2, sv) * p(2, sv) + (double)p(3, sv) * p(3, sv) + (double)p(4, sv) * p(4, sv);
        const double vo2 = (double)p(2, rm) * p(2, rm) + (double)p(3, rm) * p(3, rm) + (double)p(4, rm) * p(4, rm);
        const double f = ws / W + vo2 / vs2 * wo / W;
        p(2, sv) = (T)(p(2, sv) * f);
        p(3, sv) = (T)(p(3, sv) * f);
        p(4, sv) = (T)(p(4, sv) * f);
        // the survivor's point in the Z-ordered copy follows
        const uint32_t zs = zpos[sv];
        for (int c = 2; c < 5; ++c) pz[c * ldz + zs] = (double)p(c, sv) * pw[c];
    }
    p(5, sv) = (T)W;
    keep[rm] = 0;
    const int ss = (int)gsize[a] + (int)gsize[b];
    gsize[rm] = 0;
    gsize[sv] = (uint8_t)(ss >= G ? 0 : ss);
    gz[zpos[rm]] = 0;
    gz[zpos[sv]] = gsize[sv];
    mrk[a] = mrk[b] = 1;
    lock[a] = lock[b] = NN_MERGED;
    if (pairs) {
        pairs[2 * (uint64_t)(base + off[a])] = sv;
        pairs[2 * (uint64_t)(base + off[a]) + 1] = rm;
    }
    atomicAdd(merged, 1u);
}

// Candidates whose neighbour search failed leave the pool, unless a pair took them this round.
__global__ void k_nn_retire(uint32_t nc, const uint32_t *__restrict__ prop, const uint32_t *__restrict__ lock,
                            uint8_t *gsize, const uint32_t *__restrict__ zpos, uint8_t *gz)
{
    const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nc) return;
    if (prop[a] == NN_RETIRE && lock[a] != NN_MERGED) {
        gsize[a] = 0;
        gz[zpos[a]] = 0;
    }
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

// Nearest-neighbour merge over particles [0, nc) (sorted by weight, lightest first).
static uint64_t al16(uint64_t b) { return (b + 15) & ~15ull; }

static uint32_t nn_leaves(uint32_t nc)
{
    const uint64_t buckets = ((uint64_t)nc + NN_LEAF - 1) / NN_LEAF;
    uint32_t L = 1;
    while (L < buckets) L <<= 1;
    return L;
}

OECL_EXPORT uint64_t oecl_merge_work_bytes(uint32_t nc)
{
    const uint64_t L = nn_leaves(nc);
    return al16(8ull * nc) + 5 * al16(4ull * nc + 4) + al16(nc) + al16(8ull * 512 * 10) + al16(8ull * 16) +
           2 * al16(8ull * 5 * 2 * L) + al16(oecl_scan_work_bytes(nc + 1)) + al16(oecl_sort_work_bytes(nc)) +
           al16(40ull * nc) + al16(4ull * nc) + 2 * al16(nc) + 256;
}

OECL_EXPORT int oecl_nn_merge(int dtype, void *data, uint64_t stride, uint32_t nc, const double *pw, int G,
                              double disLimit, uint32_t toMerge, int scheme, uint32_t stream, uint32_t s0, uint32_t s1,
                              uint8_t *keep, uint32_t *merged, uint32_t *pairs, uint32_t *hostMerged,
                              uint32_t *hostRounds, void *work, void *st)
{
    if (!data || !pw || !keep || !merged || !work || G < 2 || G > 8) return OECL_ERR_ARG;
    if (dtype != OECL_F64 && dtype != OECL_F32) return OECL_ERR_ARG;
    hipStream_t s = (hipStream_t)st;
    if (hostMerged) *hostMerged = 0;
    if (hostRounds) *hostRounds = 0;
    hipLaunchKernelGGL(k_fill_u8, dim3(oecl_blocks(nc ? nc : 1, 256)), dim3(256), 0, s, keep, nc, (uint8_t)1);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, merged, 1u);
    if (nc < 2 || toMerge == 0) {
        OECL_CHECK_LAUNCH("oecl_nn_merge");
        return OECL_OK;
    }
    const uint32_t L = nn_leaves(nc);
    char *w = (char *)work;
    uint64_t *keys = (uint64_t *)w;
    w += al16(8ull * nc);
    uint32_t *zperm = (uint32_t *)w;
    w += al16(4ull * nc + 4);
    uint32_t *prop = (uint32_t *)w;
    w += al16(4ull * nc + 4);
    uint32_t *lock = (uint32_t *)w;
    w += al16(4ull * nc + 4);
    uint32_t *acc = (uint32_t *)w;
    w += al16(4ull * nc + 4);
    uint32_t *off = (uint32_t *)w;
    w += al16(4ull * nc + 4);
    uint8_t *gsize = (uint8_t *)w;
    w += al16(nc);
    double *part = (double *)w;
    w += al16(8ull * 512 * 10);
    double *box = (double *)w;   // box[10], then the round's proposal count as a word
    uint32_t *cnt = (uint32_t *)(box + 10);
    w += al16(8ull * 16);
    double *blo = (double *)w;
    w += al16(8ull * 5 * 2 * L);
    double *bhi = (double *)w;
    w += al16(8ull * 5 * 2 * L);
    uint32_t *scanw = (uint32_t *)w;
    w += al16(oecl_scan_work_bytes(nc + 1));
    double *pz = (double *)w;     // the weighted points in Z-order, 5 rows of ldz
    const uint64_t ldz = nc;
    w += al16(40ull * nc);
    uint32_t *zpos = (uint32_t *)w;
    w += al16(4ull * nc);
    uint8_t *gz = (uint8_t *)w;
    w += al16(nc);
    uint8_t *mrk = (uint8_t *)w;
    w += al16(nc);
    void *sortw = w;
    const dim3 gr(oecl_blocks(nc, 256)), bl(256);
    Soa<double> pd{(double *)data, stride};
    Soa<float> pf{(float *)data, stride};
    // the tree: Z-order sort of the weighted points, leaf boxes, inner boxes level by level
    if (dtype == OECL_F64) hipLaunchKernelGGL(k_bbox<double>, dim3(512), bl, 0, s, pd, nc, pw, part);
    else hipLaunchKernelGGL(k_bbox<float>, dim3(512), bl, 0, s, pf, nc, pw, part);
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(64), 0, s, part, 512u, box);
    if (dtype == OECL_F64) hipLaunchKernelGGL(k_morton<double>, gr, bl, 0, s, pd, nc, pw, box, keys, zperm, gsize);
    else hipLaunchKernelGGL(k_morton<float>, gr, bl, 0, s, pf, nc, pw, box, keys, zperm, gsize);
    int rc = oecl_radix_sort_u64(keys, zperm, nc, 64, sortw, st);
    if (rc) return rc;
    if (dtype == OECL_F64)
        hipLaunchKernelGGL(k_zorder_points<double>, gr, bl, 0, s, pd, nc, zperm, pw, pz, ldz, zpos, gz);
    else
        hipLaunchKernelGGL(k_zorder_points<float>, gr, bl, 0, s, pf, nc, zperm, pw, pz, ldz, zpos, gz);
    hipLaunchKernelGGL(k_fill_u8, gr, bl, 0, s, mrk, nc, (uint8_t)0);
    hipLaunchKernelGGL(k_leaf_boxes, dim3(oecl_blocks(L, 256)), bl, 0, s, nc, pz, ldz, L, blo, bhi);
    for (uint32_t first = L >> 1; first >= 1; first >>= 1)
        hipLaunchKernelGGL(k_inner_boxes, dim3(oecl_blocks(first, 256)), bl, 0, s, first, first, blo, bhi);
    uint32_t done = 0, rounds = 0, h[2] = {0, 0};
    for (;;) {
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, cnt, 1u);
        hipLaunchKernelGGL(k_nn_propose, gr, bl, 0, s, nc, zperm, L, blo, bhi, pz, ldz, gz, gsize, mrk, disLimit, G,
                           rounds, prop, lock, cnt);
        hipLaunchKernelGGL(k_nn_lock, gr, bl, 0, s, nc, prop, lock, mrk);
        hipLaunchKernelGGL(k_nn_accept, gr, bl, 0, s, nc, prop, lock, acc);
        rc = oecl_launch_scan_u32(acc, nc, off, scanw, s);
        if (rc) return rc;
        if (dtype == OECL_F64)
            hipLaunchKernelGGL(k_nn_merge<double>, gr, bl, 0, s, pd, nc, prop, acc, off, toMerge - done, G, scheme,
                               rounds, stream, s0, s1, gsize, lock, keep, merged, pairs, done, zpos, gz, mrk, pz, ldz,
                               pw);
        else
            hipLaunchKernelGGL(k_nn_merge<float>, gr, bl, 0, s, pf, nc, prop, acc, off, toMerge - done, G, scheme,
                               rounds, stream, s0, s1, gsize, lock, keep, merged, pairs, done, zpos, gz, mrk, pz, ldz,
                               pw);
        hipLaunchKernelGGL(k_nn_retire, gr, bl, 0, s, nc, prop, lock, gsize, zpos, gz);
        if (G > 2 && scheme == 1) {
            // survivors that stay in the pool had their velocity rescaled: refit the boxes
            hipLaunchKernelGGL(k_leaf_boxes, dim3(oecl_blocks(L, 256)), bl, 0, s, nc, pz, ldz, L, blo, bhi);
            for (uint32_t first = L >> 1; first >= 1; first >>= 1)
                hipLaunchKernelGGL(k_inner_boxes, dim3(oecl_blocks(first, 256)), bl, 0, s, first, first, blo, bhi);
        }
        OECL_CHECK_LAUNCH("oecl_nn_merge round");
        OECL_HIP(hipMemcpyAsync(&h[0], merged, 4, hipMemcpyDeviceToHost, s), "merge count");
        OECL_HIP(hipMemcpyAsync(&h[1], cnt, 4, hipMemcpyDeviceToHost, s), "merge count");
        OECL_HIP(hipStreamSynchronize(s), "merge round");
        ++rounds;
        const bool progress = h[0] > done;
        done = h[0];
        // Every round with proposals merges at least the pair of the lowest-ranked proposer
        // (no lower rank can hold either of its locks), so the loop ends when enough merges
        // are made or nobody proposes any more; the round cap is a guard.
        if (done >= toMerge || h[1] == 0 || !progress || rounds >= 4096) break;
    }
    if (hostMerged) *hostMerged = done;
    if (hostRounds) *hostRounds = rounds;
    return OECL_OK;
}
