// One whole plain time step of the buildup loop (example01_basic.py:52-97 order) launched
// from one host call, with every count kept on the device: no host synchronisation inside
// a step, so steps queue back to back and can be captured into a HIP graph.
//
//   q = rho + s(step) * beam profile                 (example01_basic.py:81)
//   phi = A^-1 q (capacitance direct solve), E = phiToE phi        (poissonSolver.pyx:83-88)
//   gather + Boris push + wall test + deposit of survivors -> rho   (particles.pyx:52-92,
//                                                                     particleBoundary indexInside)
//   absorbed rows in index order                    (particleBoundary.pyx:38-65)
//   interaction point + normal, Furman emission     (particleBoundary.pyx, particleEmitter.pyx:154-308)
//   deposit of the secondaries -> rho
//   compaction: secondaries fill the first holes, the rest append; remaining holes below the
//   new count are filled from the tail (particles.pyx:222-269 rules)
//
// Kernels are launched over the capacity (or the hit / secondary capacities) and read the
// live counts from device memory.  Overflows set status bits instead of writing past a
// buffer; the host checks the status word.
#include <rocblas/rocblas.h>

#include "common.h"
#include "furman_dev.h"
#include "particles_dev.h"
#include "scan.h"
#include "wall.h"

enum {
    ST_ABS_OVERFLOW = 1u,   // more wall hits than absCap
    ST_SEC_OVERFLOW = 2u,   // more secondaries than secCap
    ST_CAP_OVERFLOW = 4u    // particle count would exceed the capacity
};

namespace {

__global__ void k_step_begin(const uint32_t *nDev, const uint32_t *stepDev, uint32_t *hist,
                             uint32_t histCap)
{
    uint32_t s = *stepDev;
    if (s < histCap) hist[s] = *nDev;
}

__global__ void k_rhs_dev(uint32_t np, const double *__restrict__ rho, const double *__restrict__ table,
                          uint32_t nScale, const uint32_t *__restrict__ stepDev,
                          const double *__restrict__ prof, double *__restrict__ q)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    uint32_t s = *stepDev;
    double sc = (prof && table && s < nScale) ? table[s] : 0.;
    q[i] = sc != 0. ? rho[i] + sc * prof[i] : rho[i];
}

__global__ void k_zero_d(double *a, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.;
}

template <typename T, int TYPEB>
__global__ void __launch_bounds__(256)
    k_push_dev(oecl_grid_t g, WallDev wall, Soa<T> p, const uint32_t *__restrict__ nDev,
               const double *__restrict__ E, const double *__restrict__ B, double b0, double b1,
               double b2, double dt, double k, double q, uint8_t *__restrict__ isInside,
               double *__restrict__ rho)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < *nDev;
    // FP32 storage is widened on load: weights, gather and the Boris rotation run in FP64
    // registers (the kernel is bound by memory traffic, not arithmetic).
    using C = double;
    const C dxi = 1. / g.dx, dyi = 1. / g.dy, lxh = g.lxHalfExt, lyh = g.lyHalfExt;
    bool dep = false;
    Weights<C> w2{};
    double wq = 0.;
    if (live) {
        C x = p(0, i), y = p(1, i), vx = p(2, i), vy = p(3, i), vz = p(4, i);
        Weights<C> w = grid_weights<C>(x, y, dxi, dyi, lxh, lyh, g.nx, g.ny);
        gather_push<C, TYPEB>(w, g.nx, g.np, E, B, b0, b1, b2, dt, k, x, y, vx, vy, vz);
        uint8_t in = wall_inside<C>(wall, x, y);
        p(0, i) = (T)x;
        p(1, i) = (T)y;
        p(2, i) = (T)vx;
        p(3, i) = (T)vy;
        if (TYPEB != 0) p(4, i) = (T)vz;
        isInside[i] = in;
        if (in) {
            dep = true;
            w2 = grid_weights<C>((C)(T)x, (C)(T)y, dxi, dyi, lxh, lyh, g.nx, g.ny);
            wq = (double)p(5, i);
        }
    }
    block_deposit<512>(rho, g.nx, dep, w2.cell, w2.wLL * wq * q, w2.wUL * wq * q, w2.wLU * wq * q,
                       w2.wUU * wq * q);
}

__global__ void __launch_bounds__(OECL_SCAN_THREADS)
    k_tile_holes_dev(const uint8_t *__restrict__ flags, const uint32_t *__restrict__ nDev,
                     uint32_t *__restrict__ counts)
{
    __shared__ uint32_t lds[4];
    uint32_t n = *nDev;
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
    k_scan_tiles_dev(const uint32_t *__restrict__ counts, uint32_t nTiles, uint32_t *__restrict__ offsets)
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

// Absorbed rows (hole of rank r -> row r) and the hole positions.
template <typename T>
__global__ void __launch_bounds__(OECL_SCAN_THREADS)
    k_collect_dev(Soa<T> p, const uint32_t *__restrict__ nDev, const uint8_t *__restrict__ flags,
                  const uint32_t *__restrict__ tileOff, T *__restrict__ absorbed, uint32_t absCap,
                  uint32_t *__restrict__ holePos)
{
    uint32_t n = *nDev;
    for_each_ranked(flags, n, tileOff, [&](uint32_t i, bool hole, uint32_t r) {
        if (hole) {
            holePos[r] = i;
            if (r < absCap) {
#pragma unroll
                for (int c = 0; c < 6; ++c) absorbed[6 * (uint64_t)r + c] = p(c, i);
            }
        }
    });
}

// Wall interaction, normals and the first Furman loop for hit j (< min(holes, absCap)).
template <typename T>
__global__ void k_hits_dev(WallDev w, oecl_furman_t P, T *ap, T *nv, T *rts, double dt,
                           const uint32_t *__restrict__ holes, uint32_t absCap, int emit, uint32_t s0,
                           uint32_t s1, const uint32_t *__restrict__ stepDev, uint32_t *__restrict__ nSec,
                           double *__restrict__ impact, uint32_t *stats)
{
    uint32_t jj = blockIdx.x * blockDim.x + threadIdx.x;
    if (jj >= absCap) return;
    uint32_t h = *holes;
    uint32_t na = h < absCap ? h : absCap;
    if (jj >= na) {
        nSec[jj] = 0;
        return;
    }
    T *a = ap + 6 * (uint64_t)jj;
    double x = a[0], y = a[1], vx = a[2], vy = a[3];
    double r = 0., n0 = 0., n1 = 0.;
    if (w.kind == OECL_WALL_RECT) {
        const double eps = 1e-6;
        double lxh = w.lx * 0.5 - w.dx * eps, lyh = w.ly * 0.5 - w.dy * eps;
        int xi = 1, yi = 1, lw = 0;
        double tx, ty;
        if (x > lxh) xi = 2;
        else if (x < -lxh) xi = 0;
        if (y > lyh) yi = 2;
        else if (y < -lyh) yi = 0;
        if (xi == 0 && yi == 1) { r = (x + lxh) / vx; lw = 0; }
        else if (xi == 2 && yi == 1) { r = (x - lxh) / vx; lw = 1; }
        else if (xi == 1 && yi == 0) { r = (y + lyh) / vy; lw = 2; }
        else if (xi == 1 && yi == 2) { r = (y - lyh) / vy; lw = 3; }
        else if (xi == 0 && yi == 0) {
            tx = (x + lxh) / vx; ty = (y + lyh) / vy;
            if (tx > ty) { r = tx; lw = 0; } else { r = ty; lw = 1; }
        } else if (xi == 2 && yi == 0) {
            tx = (x - lxh) / vx; ty = (y + lyh) / vy;
            if (tx > ty) { r = tx; lw = 2; } else { r = ty; lw = 1; }
        } else if (xi == 0 && yi == 2) {
            tx = (x + lxh) / vx; ty = (y - lyh) / vy;
            if (tx > ty) { r = tx; lw = 0; } else { r = ty; lw = 3; }
        } else {
            tx = (x - lxh) / vx; ty = (y - lyh) / vy;
            if (tx > ty) { r = tx; lw = 2; } else { r = ty; lw = 3; }
        }
        x -= r * vx;
        y -= r * vy;
        n0 = lw == 0 ? 1. : (lw == 1 ? -1. : 0.);
        n1 = lw == 2 ? 1. : (lw == 3 ? -1. : 0.);
    } else if (w.kind == OECL_WALL_ELLIP) {
        const double eps = 1.e-6;
        double ax = 4. / ((w.lx - eps * w.dx) * (w.lx - eps * w.dx));
        double ay = 4. / ((w.ly - eps * w.dy) * (w.ly - eps * w.dy));
        double c = x * x * ax + y * y * ay - 1.;
        double b = x * vx * ax + y * vy * ay;
        double aa = vx * vx * ax + vy * vy * ay;
        r = (b - sqrt(b * b - aa * c)) / aa;
        x -= r * vx;
        y -= r * vy;
        double ax2 = w.ax * w.ax, ay2 = w.ay * w.ay;
        double ninv = 1. / sqrt(x * x * ax2 + y * y * ay2);
        n0 = -x * w.ax * ninv;
        n1 = -y * w.ay * ninv;
    } else {
        const double se = 1.e-6;
        uint32_t inCell = wall_cell(w, x, y);
        double rx = -dt * vx, ry = -dt * vy;
        int xDir[2] = {0, rx < 0. ? -1 : 1};
        int yDir[2] = {0, ry < 0. ? -1 : 1};
        bool stop = false;
        for (int ia = 0; ia < 2 && !stop; ++ia) {
            for (int ib = 0; ib < 2; ++ib) {
                uint32_t c = inCell + xDir[ia] + yDir[ib] * w.nx;
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
        if (!stop) atomicAdd(stats + 3, 1u);
        uint32_t c2 = wall_cell(w, x, y);
        n0 = w.ccn[2 * c2];
        n1 = w.ccn[2 * c2 + 1];
    }
    a[0] = (T)x;
    a[1] = (T)y;
    rts[jj] = (T)r;
    T *nn = nv + 2 * (uint64_t)jj;
    nn[0] = (T)n0;
    nn[1] = (T)n1;
    if (!emit) {
        nSec[jj] = 0;
        return;
    }
    furman::Rng rng{nullptr, jj, *stepDev, s0, s1};
    nSec[jj] = furman::pass1<T>(P, a, nn, rng, impact + 4 * (uint64_t)jj);
}

template <typename T>
__global__ void k_emit_dev(oecl_furman_t P, const T *__restrict__ ap, const T *__restrict__ nv,
                           const uint32_t *__restrict__ holes, uint32_t absCap,
                           const uint32_t *__restrict__ nSec, const uint32_t *__restrict__ off,
                           const double *__restrict__ impact, uint32_t s0, uint32_t s1,
                           const uint32_t *__restrict__ stepDev, T *__restrict__ out, uint32_t secCap,
                           uint32_t *stats)
{
    uint32_t jj = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t h = *holes;
    uint32_t na = h < absCap ? h : absCap;
    if (jj >= na) return;
    uint32_t ns = nSec[jj];
    if (ns == 0) return;
    furman::Rng rng{nullptr, jj, *stepDev, s0, s1};
    uint32_t base = off[jj];
    furman::pass2<T>(P, ap + 6 * (uint64_t)jj, nv + 2 * (uint64_t)jj, ns, impact[4 * (uint64_t)jj],
                     impact + 4 * (uint64_t)(na - 1), rng, [&](uint32_t ii) -> T * {
                         uint32_t k = base + ii;
                         if (k >= secCap) {
                             atomicOr(stats + 2, (uint32_t)ST_SEC_OVERFLOW);
                             return nullptr;
                         }
                         return out + 6 * (uint64_t)k;
                     });
}

__device__ __forceinline__ uint32_t n_added(const uint32_t *secOff, uint32_t absCap, uint32_t secCap)
{
    uint32_t t = secOff[absCap];
    return t < secCap ? t : secCap;
}

template <typename T>
__global__ void k_deposit_sec_dev(oecl_grid_t g, const T *__restrict__ a, const uint32_t *__restrict__ secOff,
                                  uint32_t absCap, uint32_t secCap, double q, double *rho)
{
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_added(secOff, absCap, secCap)) return;
    const T *r = a + 6 * (uint64_t)k;
    Weights<T> w = grid_weights<T>(r[0], r[1], (T)(1. / g.dx), (T)(1. / g.dy), (T)g.lxHalfExt,
                                   (T)g.lyHalfExt, g.nx, g.ny);
    deposit_corners<T>(rho, w, g.nx, (double)r[5], q);
}

template <typename T>
__global__ void k_place_dev(Soa<T> p, const T *__restrict__ add, const uint32_t *__restrict__ secOff,
                            uint32_t absCap, uint32_t secCap, const uint32_t *__restrict__ nDev,
                            const uint32_t *__restrict__ holes, const uint32_t *__restrict__ holePos,
                            uint64_t capacity, uint32_t *stats)
{
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_added(secOff, absCap, secCap)) return;
    uint32_t h = *holes, n = *nDev;
    uint64_t dst = s < h ? holePos[s] : (uint64_t)n + (s - h);
    if (dst >= capacity) {
        atomicOr(stats + 2, (uint32_t)ST_CAP_OVERFLOW);
        return;
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) p(c, dst) = add[6 * (uint64_t)s + c];
}

template <typename T>
__global__ void __launch_bounds__(OECL_SCAN_THREADS)
    k_tail_dev(Soa<T> p, const uint8_t *__restrict__ flags, const uint32_t *__restrict__ nDev,
               const uint32_t *__restrict__ secOff, uint32_t absCap, uint32_t secCap,
               const uint32_t *__restrict__ tileOff, const uint32_t *__restrict__ holes,
               const uint32_t *__restrict__ holePos)
{
    uint32_t n = *nDev, h = *holes, nAdd = n_added(secOff, absCap, secCap);
    if (nAdd >= h) return;
    uint32_t newN = n - h + nAdd;
    for_each_ranked(flags, n, tileOff, [&](uint32_t i, bool hole, uint32_t hb) {
        if (!hole && i >= newN) {
            uint32_t j = (n - 1 - i) - (h - hb);
            uint32_t dst = holePos[nAdd + j];
#pragma unroll
            for (int c = 0; c < 6; ++c) p(c, dst) = p(c, i);
        }
    });
}

__global__ void k_step_end(uint32_t *nDev, const uint32_t *holes, const uint32_t *secOff, uint32_t absCap,
                           uint32_t secCap, uint64_t capacity, uint32_t *stepDev, uint32_t *stats)
{
    uint32_t n = *nDev, h = *holes;
    uint32_t tot = secOff[absCap];
    uint32_t nAdd = tot < secCap ? tot : secCap;
    uint64_t newN = (uint64_t)n - h + nAdd;
    if (h > absCap) stats[2] |= ST_ABS_OVERFLOW;
    if (tot > secCap) stats[2] |= ST_SEC_OVERFLOW;
    if (newN > capacity) {
        stats[2] |= ST_CAP_OVERFLOW;
        newN = capacity;
    }
    *nDev = (uint32_t)newN;
    stats[0] += h;
    stats[1] += nAdd;
    stats[4] = h;
    stats[5] = nAdd;
    *stepDev += 1;
}

}  // namespace

OECL_EXPORT uint64_t oecl_step_work_bytes(uint64_t capacity, uint32_t absCap)
{
    uint64_t nt = (capacity + OECL_SCAN_TILE - 1) / OECL_SCAN_TILE + 1;
    uint64_t ntA = oecl_scan_tiles(absCap) + 1;
    return 4 * (2 * nt + 8 + capacity + 2 * ntA + 8) + 1024;
}

template <typename T>
static int step_typed(const oecl_grid_t *g, const WallDev &wd, const oecl_furman_t *P, oecl_fastpoisson *fp,
                      const oecl_step_args *a, hipStream_t s)
{
    const uint64_t cap = a->stride;
    const uint32_t np = g->np;
    Soa<T> p{(T *)a->data, a->stride};
    uint32_t nT = (uint32_t)((cap + OECL_SCAN_TILE - 1) / OECL_SCAN_TILE);
    uint32_t *counts = (uint32_t *)a->work;
    uint32_t *off = counts + nT + 1;
    uint32_t *holes = off + nT;
    uint32_t *holePos = off + nT + 8;
    uint32_t *scanw = holePos + cap;
    hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(1), 0, s, a->nDev, a->stepDev, a->countHist, a->histCap);
    hipLaunchKernelGGL(k_rhs_dev, dim3(oecl_blocks(np, 256)), dim3(256), 0, s, np, a->rho, a->beamScale,
                       a->nScale, a->stepDev, a->profile, a->q);
    int rc = oecl_fastpoisson_solve(fp, a->q, a->phi, (void *)s);
    if (rc) return rc;
    rc = oecl_csr_spmv(2 * np, a->e2nRowPtr, a->e2nCol, a->e2nVal, a->phi, a->E, (void *)s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_zero_d, dim3(oecl_blocks(np, 256)), dim3(256), 0, s, a->rho, np);
    double kk = a->typeB == 0 ? a->qm * a->dt : 0.5 * a->qm * a->dt;
    dim3 gp(oecl_blocks(cap, 256)), bp(256);
    if (a->evPushBegin) OECL_HIP(hipEventRecord((hipEvent_t)a->evPushBegin, s), "event");
    if (a->typeB == 0)
        hipLaunchKernelGGL((k_push_dev<T, 0>), gp, bp, 0, s, *g, wd, p, a->nDev, a->E, a->bField, 0., 0., 0.,
                           a->dt, kk, a->qe, a->isInside, a->rho);
    else if (a->typeB == 1)
        hipLaunchKernelGGL((k_push_dev<T, 1>), gp, bp, 0, s, *g, wd, p, a->nDev, a->E, a->bField, a->b0, a->b1,
                           a->b2, a->dt, kk, a->qe, a->isInside, a->rho);
    else
        hipLaunchKernelGGL((k_push_dev<T, 2>), gp, bp, 0, s, *g, wd, p, a->nDev, a->E, a->bField, 0., 0., 0.,
                           a->dt, kk, a->qe, a->isInside, a->rho);
    if (a->evPushEnd) OECL_HIP(hipEventRecord((hipEvent_t)a->evPushEnd, s), "event");
    hipLaunchKernelGGL(k_tile_holes_dev, dim3(nT), dim3(OECL_SCAN_THREADS), 0, s, a->isInside, a->nDev, counts);
    hipLaunchKernelGGL(k_scan_tiles_dev, dim3(1), dim3(OECL_SCAN_THREADS), 0, s, counts, nT, off);
    hipLaunchKernelGGL(k_collect_dev<T>, dim3(nT), dim3(OECL_SCAN_THREADS), 0, s, p, a->nDev, a->isInside, off,
                       (T *)a->absorbed, a->absCap, holePos);
    hipLaunchKernelGGL(k_hits_dev<T>, dim3(oecl_blocks(a->absCap, 128)), dim3(128), 0, s, wd, *P,
                       (T *)a->absorbed, (T *)a->normals, (T *)a->remaining, a->dt, holes, a->absCap, a->emit,
                       a->seed0, a->seed1, a->stepDev, a->nSec, a->impact, a->stats);
    rc = oecl_launch_scan_u32(a->nSec, a->absCap, a->secOff, scanw, s);
    if (rc) return rc;
    if (a->emit)
        hipLaunchKernelGGL(k_emit_dev<T>, dim3(oecl_blocks(a->absCap, 64)), dim3(64), 0, s, *P,
                           (const T *)a->absorbed, (const T *)a->normals, holes, a->absCap, a->nSec, a->secOff,
                           a->impact, a->seed0, a->seed1, a->stepDev, (T *)a->secondaries, a->secCap, a->stats);
    hipLaunchKernelGGL(k_deposit_sec_dev<T>, dim3(oecl_blocks(a->secCap, 256)), dim3(256), 0, s, *g,
                       (const T *)a->secondaries, a->secOff, a->absCap, a->secCap, a->qe, a->rho);
    hipLaunchKernelGGL(k_place_dev<T>, dim3(oecl_blocks(a->secCap, 256)), dim3(256), 0, s, p,
                       (const T *)a->secondaries, a->secOff, a->absCap, a->secCap, a->nDev, holes, holePos, cap,
                       a->stats);
    hipLaunchKernelGGL(k_tail_dev<T>, dim3(nT), dim3(OECL_SCAN_THREADS), 0, s, p, a->isInside, a->nDev, a->secOff,
                       a->absCap, a->secCap, off, holes, holePos);
    hipLaunchKernelGGL(k_step_end, dim3(1), dim3(1), 0, s, a->nDev, holes, a->secOff, a->absCap, a->secCap, cap,
                       a->stepDev, a->stats);
    OECL_CHECK_LAUNCH("oecl_step_run");
    return OECL_OK;
}

OECL_EXPORT int oecl_step_run(const oecl_grid_t *g, const oecl_wall_t *wall, const oecl_furman_t *P,
                              oecl_fastpoisson *fp, const oecl_step_args *a, void *stream)
{
    if (!g || !wall || !P || !fp || !a || !a->data || !a->nDev || !a->isInside || !a->rho || !a->q ||
        !a->phi || !a->E || !a->e2nRowPtr || !a->e2nCol || !a->e2nVal || !a->absorbed || !a->normals ||
        !a->remaining || !a->nSec || !a->secOff || !a->impact || !a->secondaries || !a->stepDev ||
        !a->stats || !a->work || a->absCap == 0 || a->secCap == 0)
        return OECL_ERR_ARG;
    if (a->typeB == 2 && !a->bField) return OECL_ERR_ARG;
    WallDev wd;
    int rc = wall_to_device(wall, &wd);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    if (a->dtype == OECL_F64) return step_typed<double>(g, wd, P, fp, a, s);
    if (a->dtype == OECL_F32) return step_typed<float>(g, wd, P, fp, a, s);
    return OECL_ERR_ARG;
}

OECL_EXPORT int oecl_event_create(void **ev)
{
    if (!ev) return OECL_ERR_ARG;
    hipEvent_t e;
    OECL_HIP(hipEventCreate(&e), "hipEventCreate");
    *ev = (void *)e;
    return OECL_OK;
}

OECL_EXPORT int oecl_event_destroy(void *ev)
{
    if (ev) OECL_HIP(hipEventDestroy((hipEvent_t)ev), "hipEventDestroy");
    return OECL_OK;
}

OECL_EXPORT int oecl_event_elapsed_ms(void *begin, void *end, float *ms)
{
    if (!begin || !end || !ms) return OECL_ERR_ARG;
    OECL_HIP(hipEventSynchronize((hipEvent_t)end), "hipEventSynchronize");
    OECL_HIP(hipEventElapsedTime(ms, (hipEvent_t)begin, (hipEvent_t)end), "hipEventElapsedTime");
    return OECL_OK;
}
