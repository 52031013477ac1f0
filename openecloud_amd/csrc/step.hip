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
// Kernels read the live counts from device memory: the push grid covers the capacity and
// whole blocks past the count return at once; the per-hit and per-secondary kernels loop
// over the live counts with small fixed grids.  Overflows set status bits instead of writing
// past a buffer; the host checks the status word.
#include "common.h"
#include "fastpoisson_internal.h"
#include "furman_dev.h"
#include "secem_dev.h"
#include "photo_dev.h"
#include "particles_dev.h"
#include "scan.h"
#include "wall.h"

enum {
    ST_ABS_OVERFLOW = 1u,   // more wall hits than absCap
    ST_SEC_OVERFLOW = 2u,   // more secondaries than secCap
    ST_CAP_OVERFLOW = 4u,   // particle count would exceed the capacity
    ST_PHOTO_OVERFLOW = 8u  // more photoelectrons in one step than photoCap
};

namespace {

constexpr uint32_t PB = kPushBlock;  // particles per push block (one deposit tile per block)
constexpr uint32_t CT = 1024;     // threads of the single-block kernels
constexpr uint32_t NZMAX = 16384; // push waves with holes sorted in LDS (more: full scan)
constexpr uint32_t MAXSEC = 10;   // secondaries per impact (Furman M)
// step scratch counters
enum { M_HOLES = 0, M_NZ = 1, M_PHOTO = 2, M_COUNT = 16 };

// The runs of a step: R independent buildups sharing the grid, the wall and the Poisson
// operator (R = 1 for the plain step).  Run r's particles live in [seg0(r), seg0(r) + nDev[r])
// of one SoA block, inside its segment [seg0(r), seg1(r)), a multiple of PB, so that every
// push block belongs to one run; its hits use rows [abs0(r), abs0(r) + absN(r)) of the
// per-hit arrays.  Grid arrays (rho, q, phi: np nodes, E: 2 np) and per-run state (nDev,
// stats[8], misc[M_COUNT], countHist[histCap], beam scale[nScale]) are stored run after run.
// A single run uses no tables: segment [0, cap), hit rows [0, absCap), the emitter parameters
// and Philox key passed by value.
struct Runs {
    uint32_t R;
    uint32_t absCap;
    uint64_t cap;
    const uint32_t *segOff;       // [R + 1] or nullptr
    const uint32_t *blockRun;     // [cap / PB] or nullptr
    const uint32_t *absOff;       // [R + 1] or nullptr
    const oecl_furman_t *furman;  // [R] or nullptr
    const uint32_t *keys;         // [2 R] or nullptr
    __device__ __forceinline__ uint64_t seg0(uint32_t r) const { return segOff ? segOff[r] : 0; }
    __device__ __forceinline__ uint64_t seg1(uint32_t r) const { return segOff ? segOff[r + 1] : cap; }
    __device__ __forceinline__ uint32_t abs0(uint32_t r) const { return absOff ? absOff[r] : 0; }
    __device__ __forceinline__ uint32_t absN(uint32_t r) const
    {
        return absOff ? absOff[r + 1] - absOff[r] : absCap;
    }
    __device__ __forceinline__ uint32_t block_run(uint32_t b) const { return blockRun ? blockRun[b] : 0; }
    __device__ __forceinline__ uint32_t key0(uint32_t r, uint32_t k) const { return keys ? keys[2 * r] : k; }
    __device__ __forceinline__ uint32_t key1(uint32_t r, uint32_t k) const { return keys ? keys[2 * r + 1] : k; }
};

// Start of a step, one thread per node, blockIdx.y = run:
//   the Poisson right-hand side q = rho + s(step) * beam profile (example01_basic.py:81),
//   masked and written straight into the direct solver's box vector (inside nodes keep q,
//   other box nodes 0; the potential outside the box is set to 0);
//   rho = 0 for this step's deposit; thread 0 records the particle count of the step;
//   the step's photoelectrons (a per-step table, single run only) are generated into their
//   slots.
// The per-run bookkeeping at the start of a step, thread i of run `run` (gridDim.x * blockDim.x
// threads per run): the step's particle count into the history, the counters, and the step's
// photoelectrons (a per-step table, single run only) generated into their slots.
template <typename T>
__device__ __forceinline__ void begin_step_run(uint32_t i, uint32_t run, uint32_t s, const uint32_t *__restrict__ nDev,
                                               uint32_t *__restrict__ hist, uint32_t histCap,
                                               uint32_t *__restrict__ misc, WallDev w,
                                               const uint32_t *__restrict__ photoCount,
                                               const double *__restrict__ photoWeight, uint32_t nPhoto,
                                               uint32_t photoCap, uint32_t s0, uint32_t s1, double e0, double sigmaE,
                                               double mass, T *__restrict__ photoSlots, uint32_t *stats)
{
    if (hist) hist += (size_t)run * histCap;
    misc += run * M_COUNT;
    nDev += run;
    stats += 8 * run;
    if (run) photoCount = nullptr;
    uint32_t nph = 0;
    if (photoCount && s < nPhoto) {
        nph = photoCount[s];
        if (nph > photoCap) nph = photoCap;
    }
    if (i == 0) {
        if (hist && s < histCap) hist[s] = *nDev;
        misc[M_NZ] = 0;
        misc[M_PHOTO] = nph;
        if (photoCount && s < nPhoto && photoCount[s] > photoCap) atomicOr(stats + 2, (uint32_t)ST_PHOTO_OVERFLOW);
    }
    for (uint32_t k = i; k < nph; k += gridDim.x * blockDim.x)
        photo::emit_one<T>(w, k, s, s0, s1, e0, sigmaE, photoWeight[s], mass, photoSlots + 6 * (uint64_t)k);
}

// Start of a step, one thread per node, blockIdx.y = run:
//   the Poisson right-hand side q = rho + s(step) * beam profile (example01_basic.py:81),
//   masked and written straight into the direct solver's box vector (inside nodes keep q,
//   other box nodes 0; the potential outside the box is set to 0);
//   rho = 0 for this step's deposit; the run's bookkeeping (begin_step_run).
// Used when the caller solves (externalPhi); the step's own direct solve gathers q in its first
// pass instead (k_begin_edges).
template <typename T>
__global__ void k_begin_dev(uint32_t np, double *__restrict__ rho, const double *__restrict__ table,
                            uint32_t nScale, const uint32_t *__restrict__ stepDev,
                            const double *__restrict__ prof, FpBox box, double *__restrict__ phi,
                            const uint32_t *__restrict__ nDev, uint32_t *__restrict__ hist, uint32_t histCap,
                            uint32_t *__restrict__ misc, WallDev w, const uint32_t *__restrict__ photoCount,
                            const double *__restrict__ photoWeight, uint32_t nPhoto, uint32_t photoCap,
                            uint32_t s0, uint32_t s1, double e0, double sigmaE, double mass,
                            T *__restrict__ photoSlots, uint32_t *stats)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t s = *stepDev, run = blockIdx.y;
    rho += (size_t)run * np;
    phi += (size_t)run * np;
    box.b += (size_t)run * box.bx * (box.ny - 4);
    if (table) table += (size_t)run * nScale;
    begin_step_run<T>(i, run, s, nDev, hist, histCap, misc, w, photoCount, photoWeight, nPhoto, photoCap, s0, s1, e0,
                      sigmaE, mass, photoSlots, stats);
    if (i >= np) return;
    const double sc = (prof && table && s < nScale) ? table[s] : 0.;
    const double r = rho[i];
    const double qv = sc != 0. ? r + sc * prof[i] : r;
    rho[i] = 0.;
    const uint32_t ix = i % box.nx, iy = i / box.nx;
    if (ix < 2 || iy < 2 || ix >= box.nx - 2 || iy >= box.ny - 2) phi[i] = 0.;
    else box.b[(ix - 2) + (iy - 2) * box.bx] = box.mask[i] ? qv : 0.;
}

// Start of a step whose direct solve gathers q = rho + s(step) * profile from the grid in its
// first pass (fastpoisson_solve_grid, which also zeroes rho on grid rows 2 .. ny-3): the run's
// bookkeeping (begin_step_run), rho = 0 on grid rows 0, 1, ny-2 and ny-1, and phi = 0 outside the
// box (those rows and grid columns 0, 1, nx-2, nx-1), which the solve never writes.  blockIdx.y
// = run.
template <typename T>
__global__ void k_begin_edges(uint32_t nx, uint32_t ny, double *__restrict__ rho, const uint32_t *__restrict__ stepDev,
                              double *__restrict__ phi, const uint32_t *__restrict__ nDev,
                              uint32_t *__restrict__ hist, uint32_t histCap, uint32_t *__restrict__ misc, WallDev w,
                              const uint32_t *__restrict__ photoCount, const double *__restrict__ photoWeight,
                              uint32_t nPhoto, uint32_t photoCap, uint32_t s0, uint32_t s1, double e0,
                              double sigmaE, double mass, T *__restrict__ photoSlots, uint32_t *stats)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t s = *stepDev, run = blockIdx.y;
    const size_t np = (size_t)nx * ny;
    rho += (size_t)run * np;
    phi += (size_t)run * np;
    begin_step_run<T>(i, run, s, nDev, hist, histCap, misc, w, photoCount, photoWeight, nPhoto, photoCap, s0, s1, e0,
                      sigmaE, mass, photoSlots, stats);
    const uint32_t nRows = 4 * nx, nEdge = nRows + 4 * (ny - 4);
    for (uint32_t k = i; k < nEdge; k += gridDim.x * blockDim.x) {
        size_t node;
        if (k < nRows) {
            const uint32_t rr = k / nx, c = k - rr * nx;
            node = (size_t)(rr < 2 ? rr : ny - 4 + rr) * nx + c;
            rho[node] = 0.;
        } else {
            const uint32_t kk = k - nRows, rr = 2 + kk / 4, c = kk % 4;
            node = (size_t)rr * nx + (c < 2 ? c : nx - 4 + c);
        }
        phi[node] = 0.;
    }
}

// The push maps hardware block b to logical block xcd_block_chunk(b, n, PUSH_XCD_CHUNK)
// (common.h): every XCD works through whole chunks of consecutive blocks, whose particles share a
// few rows of cells after a cell sort, so that XCD's L2 holds that part of E and rho.
#ifndef OECL_PUSH_XCD_CHUNK
#define OECL_PUSH_XCD_CHUNK 32
#endif
constexpr uint32_t PUSH_XCD_CHUNK = OECL_PUSH_XCD_CHUNK;

// Holes (absorbed particles) of one wave in index order: the wave's count goes to
// counts[w], the indices to list[w * 64 + rank] (w = global index of the 64-particle group the
// wave handles), and a wave with holes appends w to nzList (in no particular order).  All 64
// lanes of the wave call it.
__device__ __forceinline__ void wave_holes(bool dead, uint32_t i, uint32_t w, uint32_t *__restrict__ counts,
                                           uint32_t *__restrict__ list, uint32_t *__restrict__ misc,
                                           uint32_t *__restrict__ nzList)
{
    const int lane = threadIdx.x & 63;
    const uint64_t bal = __ballot(dead);
    if (dead) list[(size_t)w * 64 + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull))] = i;
    if (lane == 0) {
        const uint32_t tot = (uint32_t)__popcll(bal);
        counts[w] = tot;
        if (tot) nzList[atomicAdd(misc + M_NZ, 1u)] = w;
    }
}

#ifndef OECL_PUSH_PER_THREAD
#define OECL_PUSH_PER_THREAD 1
#endif
// Particles per push thread: the block's PB particles are handled by PB / KP threads, each
// taking particles tid, tid + PB / KP, ... (two independent chains of dependent loads and
// arithmetic per thread when KP = 2).
constexpr int KP = OECL_PUSH_PER_THREAD;
constexpr uint32_t PT = PB / KP;   // threads per push block

// Gather + Boris push + wall test + hole list + deposit of the survivors, KP particle slots per
// thread; the block's run gives its E, rho, counters and live count.
// The pushes are latency-bound: the compiler is held to the registers of 8 waves per SIMD (its
// own choice for the mixed push was 7, by 100 scalar registers; 8 made the push of 2e7 particles
// 12 % faster).
#ifndef OECL_PUSH_WAVES_PER_EU
#define OECL_PUSH_WAVES_PER_EU 8
#endif

template <typename T, int TYPEB>
__global__ void __launch_bounds__(PT)
    __attribute__((amdgpu_waves_per_eu(TYPEB == 2 ? 7 : OECL_PUSH_WAVES_PER_EU)))  // 3-D B: no spills
    k_push_dev(Runs runs, oecl_grid_t g, WallDev wall, Soa<T> p, const uint32_t *__restrict__ nDev,
               const double *__restrict__ E, const double *__restrict__ Eil, const double *__restrict__ B, double b0, double b1,
               double b2, double dt, double k, double q, double dxi, double dyi, uint8_t *__restrict__ isInside,
               double *__restrict__ rho, uint32_t *__restrict__ counts, uint32_t *__restrict__ list,
               uint32_t *__restrict__ misc, uint32_t *__restrict__ nzList, DepCfg dc)
{
    __shared__ double tiles[PB / 64 * 512];
    __shared__ uint32_t dbox[4];
    const uint32_t blk = xcd_block_chunk(blockIdx.x, gridDim.x, PUSH_XCD_CHUNK);
    const uint32_t run = runs.block_run(blk);
    const uint64_t sBeg = runs.seg0(run);
    const uint32_t n = nDev[run];
    const uint64_t first = (uint64_t)blk * PB;
    if (first - sBeg >= n) return;  // whole block past the run's live particles
    E += (size_t)run * 2 * g.np;
    if (Eil) Eil += (size_t)run * 2 * g.np;
    rho += (size_t)run * g.np;
    misc += run * M_COUNT;
    nzList += sBeg / 64;
    // FP32 storage is widened on load: weights, gather and the Boris rotation run in FP64
    // registers.
    using C = double;
    // dxi = 1 / dx and dyi = 1 / dy come from the host (the same correctly rounded quotients)
    const C lxh = g.lxHalfExt, lyh = g.lyHalfExt;
    bool dep[KP], dead[KP];
    Weights<C> w2[KP];
    double wq[KP];
#pragma unroll
    for (int kk = 0; kk < KP; ++kk) {
        const uint32_t i = (uint32_t)first + kk * PT + threadIdx.x;
        dep[kk] = dead[kk] = false;
        w2[kk] = Weights<C>{};
        wq[kk] = 0.;
        if (i - sBeg < n) {
            C x = p(0, i), y = p(1, i), vx = p(2, i), vy = p(3, i), vz = p(4, i);
            Weights<C> w = grid_weights<C>(x, y, dxi, dyi, lxh, lyh, g.nx, g.ny);
            gather_push<C, TYPEB>(w, g.nx, g.np, E, B, b0, b1, b2, dt, k, x, y, vx, vy, vz, Eil);
            // the wall test sees the position as stored (FP32 storage: rounded)
            x = (C)(T)x;
            y = (C)(T)y;
            uint8_t in = wall_inside<C>(wall, x, y);
            p(0, i) = (T)x;
            p(1, i) = (T)y;
            p(2, i) = (T)vx;
            p(3, i) = (T)vy;
            if (TYPEB != 0) p(4, i) = (T)vz;
            if (isInside) isInside[i] = in;
            if (in) {
                dep[kk] = true;
                w2[kk] = grid_weights<C>((C)(T)x, (C)(T)y, dxi, dyi, lxh, lyh, g.nx, g.ny);
                wq[kk] = (double)p(5, i);
            } else {
                dead[kk] = true;
            }
        }
    }
#pragma unroll
    for (int kk = 0; kk < KP; ++kk) {
        const uint32_t i = (uint32_t)first + kk * PT + threadIdx.x;
        wave_holes(dead[kk], i, blk * (PB / 64) + kk * (PT / 64) + (threadIdx.x >> 6), counts, list, misc, nzList);
        if (kk) __syncthreads();   // the previous deposit's readers of the tile and box are done
        block_deposit<PT / 64, 512 * KP>(rho, g.nx, dep[kk], w2[kk].ix, w2[kk].iy, w2[kk].wLL * wq[kk] * q,
                                         w2[kk].wUL * wq[kk] * q, w2[kk].wLU * wq[kk] * q, w2[kk].wUU * wq[kk] * q,
                                         tiles, dbox, dc);
    }
}

// The wall margin of the mixed push, in FP32: a position with
//   rectangle  |x| <= xlim and |y| <= ylim,
//   ellipse    x^2 ax + y^2 ay <= gmax
// lies at least the margin inside the wall.  (For an ellipse of semi-axes a, b, every point of
// the scaled ellipse rho E lies at least (1 - rho) min(a, b) inside E: the distance to the
// boundary is the minimum over directions n of h_E(n) - P.n >= (1 - rho) h_E(n), and the support
// function h_E of an ellipse centred on the origin is at least min(a, b).)  The limits are
// rounded toward the chamber's centre.  The cut-cell wall uses the grid's cut cells instead.
struct NearWall {
    float xlim, ylim;
    float ax, ay, gmax;
};

// FP32 constants of the mixed push: 1/dx, 1/dy, lxHalfExt/dx, lyHalfExt/dy, dt, k (q/m dt for
// typeB 0, q/m dt / 2 otherwise) and k b for a constant field (typeB 1).
struct Push32 {
    float dxi, dyi, ox, oy, dt, k, t0, t1, t2;
};

// The mixed push (pushArith 1, FP32 particles): gather and Boris push in FP32 from the FP32 field
// copies E32 ((Ex, Ey) per node, per run) and B32 ((Bx, By) per node, typeB 3); the deposit as
// k_push_dev's (FP64 weights of the stored position, FP64 charges, run sums and rho).  A
// particle within the margin of the wall before or after the FP32 push takes the FP64 path of
// k_push_dev<float, TYPEB> instead, from its stored state, with the FP64 fields E ([Ex | Ey]) and
// B (typeB 3: (Bx, By) per node; typeB 2: [Bx | By | Bz]): its result, wall test included, is
// that of k_push_dev bit for bit.  A particle outside the margin is inside the chamber as stored,
// with the margin far above FP32 rounding.
#ifndef OECL_MIXED_PER_THREAD
#define OECL_MIXED_PER_THREAD 1
#endif
// Particles per thread of the mixed push (MKP particle slots per thread, MPT threads per block of
// PB particles), as KP for k_push_dev.
constexpr int MKP = OECL_MIXED_PER_THREAD;
constexpr uint32_t MPT = PB / MKP;

// One particle's result in the mixed push: absorbed (dead) or deposited (dep) at cell (cx, cy)
// with the four corner charges.
struct MixedOut {
    bool dep, dead;
    uint32_t cx, cy;
    double q00, q10, q01, q11;
};

// CP: the cache-policy bits of the particle stream's loads and stores.  Non-temporal (2) when
// the particle state is larger than the last-level cache can usefully hold (kStreamBytes): each
// component is then read and written once a step and only evicts the field copies and rho that
// the gather and deposit reuse (2e7 particles: 274 -> 265 us a push).  A state that fits stays
// cached from one step to the next instead (0; 2e6 particles: non-temporal is 4 % slower).
constexpr uint64_t kStreamBytes = 128ull << 20;

template <int TYPEB, int CP>
__global__ void __launch_bounds__(MPT) __attribute__((amdgpu_waves_per_eu(OECL_PUSH_WAVES_PER_EU)))
    k_push_mixed(Runs runs, oecl_grid_t g, WallDev wall, NearWall nw, Push32 P, Soa<float> p,
                 const uint32_t *__restrict__ nDev, const double *__restrict__ E, const float *__restrict__ E32,
                 const double *__restrict__ B, const float *__restrict__ B32, double b0, double b1, double b2,
                 double dt, double k, double q, double dxi, double dyi, uint8_t *__restrict__ isInside,
                 double *__restrict__ rho, uint32_t *__restrict__ counts, uint32_t *__restrict__ list,
                 uint32_t *__restrict__ misc, uint32_t *__restrict__ nzList, DepCfg dc)
{
    __shared__ double tiles[PB / 64 * 512];
    __shared__ uint32_t dbox[4];
    const uint32_t blk = xcd_block_chunk(blockIdx.x, gridDim.x, PUSH_XCD_CHUNK);
    const uint32_t run = runs.block_run(blk);
    const uint64_t sBeg = runs.seg0(run);
    const uint32_t n = nDev[run];
    const uint64_t first = (uint64_t)blk * PB;
    if (first - sBeg >= n) return;  // whole block past the run's live particles
    E += (size_t)run * 2 * g.np;
    rho += (size_t)run * g.np;
    misc += run * M_COUNT;
    nzList += sBeg / 64;
    // buffer resources (the host checks that every size fits 32 bits): one per particle component,
    // the run's FP32 E copy and the FP32 B copy, 8 bytes per node
    const uint32_t sb = (uint32_t)(p.stride * 4u), nb8 = g.np * 8u, nx8 = g.nx * 8u;
    const BufRes rX = buf_res(p.base, sb), rY = buf_res(p.base + p.stride, sb);
    const BufRes rVx = buf_res(p.base + 2 * p.stride, sb), rVy = buf_res(p.base + 3 * p.stride, sb);
    const BufRes rVz = buf_res(p.base + 4 * p.stride, sb), rW = buf_res(p.base + 5 * p.stride, sb);
    const BufRes rE = buf_res(E32 + (size_t)run * 2 * g.np, nb8);
    auto one = [&](uint32_t i, MixedOut &o) {
        const uint32_t oi = i * 4u;
        constexpr int LC = CP, SC = CP;
        const float x0 = buf_ld_f32<LC>(rX, oi), y0 = buf_ld_f32<LC>(rY, oi), vx0 = buf_ld_f32<LC>(rVx, oi);
        const float vy0 = buf_ld_f32<LC>(rVy, oi), vz0 = buf_ld_f32<LC>(rVz, oi);
        // ---- FP32 gather and push: the four corners of a field at byte offsets oc, oc + 8 and
        // the same one row up (scalar offset nx8)
        const Loc32 l = locate32(x0, y0, P.dxi, P.dyi, P.ox, P.oy, g.nx, g.ny);
        const float gx = 1.f - l.fx, gy = 1.f - l.fy;
        const float wLL = gx * gy, wUL = l.fx * gy, wLU = gx * l.fy, wUU = l.fx * l.fy;
        const uint32_t row = l.iy * g.nx;
        const uint32_t oc = (row + l.ix) * 8u;
        const float2 ev = bilinear2_buf(rE, oc, nx8, wLL, wUL, wLU, wUU);
        float vx = vx0, vy = vy0, vz = vz0;
        const float kex = P.k * ev.x, key = P.k * ev.y;
        vx += kex;
        vy += key;
        if (TYPEB == 1) {
            boris_rotate32(vx, vy, vz, P.t0, P.t1, P.t2, kex, key);
        } else if (TYPEB == 3) {
            const float2 bv = bilinear2_buf(buf_res(B32, nb8), oc, nx8, wLL, wUL, wLU, wUU);
            boris_rotate32_bz0(vx, vy, vz, P.k * bv.x, P.k * bv.y, kex, key);
        } else if (TYPEB == 2) {
            const uint32_t c = row + l.ix;
            float bx = wLL * (float)B[c], by = wLL * (float)B[c + g.np], bz = wLL * (float)B[c + 2 * g.np];
            bx = fmaf(wUL, (float)B[c + 1], bx);
            by = fmaf(wUL, (float)B[c + 1 + g.np], by);
            bz = fmaf(wUL, (float)B[c + 1 + 2 * g.np], bz);
            bx = fmaf(wLU, (float)B[c + g.nx], bx);
            by = fmaf(wLU, (float)B[c + g.nx + g.np], by);
            bz = fmaf(wLU, (float)B[c + g.nx + 2 * g.np], bz);
            bx = fmaf(wUU, (float)B[c + g.nx + 1], bx);
            by = fmaf(wUU, (float)B[c + g.nx + 1 + g.np], by);
            bz = fmaf(wUU, (float)B[c + g.nx + 1 + 2 * g.np], bz);
            boris_rotate32(vx, vy, vz, P.k * bx, P.k * by, P.k * bz, kex, key);
        }
        const float x1 = fmaf(P.dt, vx, x0), y1 = fmaf(P.dt, vy, y0);
        // ---- the wall margin, before and after the push (NaN positions count as near)
        bool near;
        if (wall.kind == OECL_WALL_RECT) {
            near = !(fabsf(x0) <= nw.xlim && fabsf(y0) <= nw.ylim && fabsf(x1) <= nw.xlim && fabsf(y1) <= nw.ylim);
        } else if (wall.kind == OECL_WALL_ELLIP) {
            near = !(fmaf(x0 * x0, nw.ax, y0 * y0 * nw.ay) <= nw.gmax && fmaf(x1 * x1, nw.ax, y1 * y1 * nw.ay) <= nw.gmax);
        } else {
            // a cut or outside cell under the old or the new position; the new one located as the
            // FP64 wall test does (clipped to the grid: a clipped point lies in an edge cell,
            // which is outside the chamber)
            double xd = (double)x1, yd = (double)y1;
            if (!(xd <= wall.clipX)) xd = wall.clipX;
            if (!(xd >= -wall.clipX)) xd = -wall.clipX;
            if (!(yd <= wall.clipY)) yd = wall.clipY;
            if (!(yd >= -wall.clipY)) yd = -wall.clipY;
            near = wall.insideFaces[row + l.ix] != 8 || wall.insideFaces[wall_cell(wall, xd, yd)] != 8;
        }
        float xs = x1, ys = y1;
        uint8_t in = 1;
        if (near) {
            // ---- FP64 path: k_push_dev<float, TYPEB> for this particle
            double x = x0, y = y0, vxd = vx0, vyd = vy0, vzd = vz0;
            const Weights<double> w = grid_weights<double>(x, y, dxi, dyi, g.lxHalfExt, g.lyHalfExt, g.nx, g.ny);
            gather_push<double, TYPEB>(w, g.nx, g.np, E, B, b0, b1, b2, dt, k, x, y, vxd, vyd, vzd, nullptr);
            x = (double)(float)x;
            y = (double)(float)y;
            in = wall_inside<double>(wall, x, y);
            xs = (float)x;
            ys = (float)y;
            vx = (float)vxd;
            vy = (float)vyd;
            vz = (float)vzd;
        }
        buf_st_f32<SC>(xs, rX, oi);
        buf_st_f32<SC>(ys, rY, oi);
        buf_st_f32<SC>(vx, rVx, oi);
        buf_st_f32<SC>(vy, rVy, oi);
        if (TYPEB != 0) buf_st_f32<SC>(vz, rVz, oi);
        if (isInside) buf_st_u8(in, buf_res(isInside, (uint32_t)p.stride), i);
        if (in) {
            // the deposit of k_push_dev: FP64 weights of the stored position, FP64 charges
            o.dep = true;
            const Weights<double> w2 = grid_weights<double>((double)xs, (double)ys, dxi, dyi, g.lxHalfExt,
                                                            g.lyHalfExt, g.nx, g.ny);
            const double wq = (double)buf_ld_f32<LC>(rW, oi);
            o.cx = w2.ix;
            o.cy = w2.iy;
            o.q00 = w2.wLL * wq * q;
            o.q10 = w2.wUL * wq * q;
            o.q01 = w2.wLU * wq * q;
            o.q11 = w2.wUU * wq * q;
        } else {
            o.dead = true;
        }
    };
    MixedOut o[MKP];
#pragma unroll
    for (int kk = 0; kk < MKP; ++kk) {
        const uint32_t i = (uint32_t)first + kk * MPT + threadIdx.x;
        o[kk] = MixedOut{false, false, 0u, 0u, 0., 0., 0., 0.};
        if (i - sBeg < n) one(i, o[kk]);
    }
#pragma unroll
    for (int kk = 0; kk < MKP; ++kk) {
        const uint32_t i = (uint32_t)first + kk * MPT + threadIdx.x;
        wave_holes(o[kk].dead, i, blk * (PB / 64) + kk * (MPT / 64) + (threadIdx.x >> 6), counts, list, misc, nzList);
        if (kk) __syncthreads();   // the previous deposit's readers of the tile and box are done
        block_deposit<MPT / 64, 512 * MKP>(rho, g.nx, o[kk].dep, o[kk].cx, o[kk].cy, o[kk].q00, o[kk].q10, o[kk].q01,
                                           o[kk].q11, tiles, dbox, dc);
    }
}

// Exclusive prefix of v over the CT threads of the block (wsum: CT/64 LDS words); *tot the sum.
__device__ __forceinline__ uint32_t block_scan_ct(uint32_t v, uint32_t *wsum, uint32_t *tot)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    uint32_t off = 0, t = 0;
#pragma unroll
    for (int w = 0; w < (int)(CT / 64); ++w) {
        const uint32_t c = wsum[w];
        if (w < wave) off += c;
        t += c;
    }
    __syncthreads();
    *tot = t;
    return off + x - v;
}

// One block per run: all holes of the run in ascending index order (holePos, the first absCap
// of them) and their count misc[M_HOLES] (k_hits_dev copies the absorbed rows, hole r -> row
// r, over the whole grid).  The push waves that have holes (nzList, usually a few dozen) are
// sorted in LDS and their lists concatenated in wave order; beyond NZMAX such waves every
// wave's count is scanned instead.
__global__ void __launch_bounds__(CT)
    k_holes_collect(Runs runs, const uint32_t *__restrict__ nDev, const uint32_t *__restrict__ counts,
                    const uint32_t *__restrict__ list, const uint32_t *__restrict__ nzList,
                    uint32_t *__restrict__ holePos, uint32_t *__restrict__ misc)
{
    __shared__ uint32_t keys[NZMAX];
    __shared__ uint32_t wsum[CT / 64];
    const uint32_t run = blockIdx.x;
    const uint32_t w0 = (uint32_t)(runs.seg0(run) / 64), a0 = runs.abs0(run), absCap = runs.absN(run);
    misc += run * M_COUNT;
    nzList += w0;
    holePos += a0;
    const uint32_t n = nDev[run], nb = (n + 63) / 64, nz = misc[M_NZ];
    const uint32_t tid = threadIdx.x;
    const bool sorted = nz <= NZMAX;
    uint32_t N = 1;
    if (nz <= CT) {
        // one wave index per thread: its place is the number of smaller ones (the indices are
        // distinct), read from LDS by all threads at once
        const uint32_t mine = tid < nz ? nzList[tid] : 0u;
        if (tid < nz) keys[tid] = mine;
        __syncthreads();
        uint32_t rank = 0;
        for (uint32_t k = 0; k < nz; ++k) rank += keys[k] < mine ? 1u : 0u;
        __syncthreads();
        if (tid < nz) keys[rank] = mine;
        __syncthreads();
    } else if (sorted) {
        while (N < nz) N <<= 1;
        for (uint32_t k = tid; k < N; k += CT) keys[k] = k < nz ? nzList[k] : 0xFFFFFFFFu;
        __syncthreads();
        for (uint32_t k = 2; k <= N; k <<= 1) {
            for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                for (uint32_t i = tid; i < N; i += CT) {
                    const uint32_t ixj = i ^ j;
                    if (ixj > i) {
                        const uint32_t a = keys[i], b = keys[ixj];
                        if ((a > b) == ((i & k) == 0)) {
                            keys[i] = b;
                            keys[ixj] = a;
                        }
                    }
                }
                __syncthreads();
            }
        }
    }
    const uint32_t items = sorted ? nz : nb;
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < items; b0 += CT) {
        const uint32_t k = b0 + tid;
        const uint32_t blk = k < items ? (sorted ? keys[k] : w0 + k) : 0;
        const uint32_t c = k < items ? counts[blk] : 0;
        uint32_t tot;
        const uint32_t start = carry + block_scan_ct(c, wsum, &tot);
        for (uint32_t r = 0; r < c; ++r) {
            const uint32_t dst = start + r;
            if (dst < absCap) holePos[dst] = list[(size_t)blk * 64 + r];
        }
        carry += tot;
    }
    if (tid == 0) misc[M_HOLES] = carry;
}

__device__ __forceinline__ uint32_t n_hits(const uint32_t *misc, uint32_t absCap)
{
    const uint32_t h = misc[M_HOLES];
    return h < absCap ? h : absCap;
}

// The run of flattened hit t, hitOff[0 .. R] being the exclusive prefix of the runs' hit counts
// (the largest r with hitOff[r] <= t; runs without hits are passed over).
__device__ __forceinline__ uint32_t hit_run(const uint32_t *__restrict__ hitOff, uint32_t R, uint32_t t)
{
    uint32_t lo = 0, hi = R;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (hitOff[mid] <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Interaction point, remaining time and wall normal of an absorbed particle at (x, y) with
// velocity (vx, vy) (particleBoundary.pyx: AbsorbRectangular/Elliptical/ArbitraryCutCell
// calculateInteractionPoint and calculateNormalVectors).  found = false when the cut-cell
// search misses (counted by the caller).
__device__ __forceinline__ void hit_geom(const WallDev &w, double dt, double &x, double &y, double vx, double vy,
                                         double &r, double &n0, double &n1, bool &found)
{
    r = 0.;
    n0 = 0.;
    n1 = 0.;
    found = true;
    if (w.kind == OECL_WALL_RECT) {
        wall_rect_hit(w, x, y, vx, vy, r, n0, n1);
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
        found = stop;
        uint32_t c2 = wall_cell(w, x, y);
        n0 = w.ccn[2 * c2];
        n1 = w.ccn[2 * c2 + 1];
    }
}

// Wall interaction, normals (written in place of the absorbed row's position) and the first
// Furman loop for hit jj: nSec[jj] and impact[jj].
template <typename T>
__device__ __forceinline__ uint32_t hit_one(uint32_t jj, const WallDev &w, const oecl_furman_t &P,
                                            const oecl_secem_t &S, int useSecem, T *ap, T *nv, T *rts, double dt,
                                            int emit, uint32_t s0, uint32_t s1, uint32_t step,
                                            uint32_t *__restrict__ nSec, double *__restrict__ impact,
                                            uint32_t *stats)
{
    T *a = ap + 6 * (uint64_t)jj;
    double x = a[0], y = a[1], r, n0, n1;
    bool found;
    hit_geom(w, dt, x, y, a[2], a[3], r, n0, n1, found);
    if (!found) atomicAdd(stats + 3, 1u);
    a[0] = wall_store<T>(x);
    a[1] = wall_store<T>(y);
    rts[jj] = (T)r;
    T *nn = nv + 2 * (uint64_t)jj;
    nn[0] = (T)n0;
    nn[1] = (T)n1;
    uint32_t ns = 0;
    if (emit && useSecem) {
        const secem::Rng rng{nullptr, jj, step, s0, s1};
        ns = secem::pass1<T>(S, a, nn, rng, impact + 4 * (uint64_t)jj);
    } else if (emit) {
        furman::Rng rng{nullptr, jj, step, s0, s1};
        ns = furman::pass1<T>(P, a, nn, rng, impact + 4 * (uint64_t)jj);
    }
    nSec[jj] = ns;
    return ns;
}

// Wall hits of every run, grid-stride over the flattened hits (hitOff: the runs' hit offsets,
// or null for a single run).
template <typename T>
__global__ void __launch_bounds__(128)
    k_hits_dev(Runs runs, const uint32_t *__restrict__ hitOff, WallDev w, oecl_furman_t P, oecl_secem_t S,
               int useSecem, Soa<T> p, const uint32_t *__restrict__ holePos, T *ap, T *nv, T *rts, double dt,
               int emit, uint32_t s0, uint32_t s1,
               const uint32_t *__restrict__ stepDev, uint32_t *__restrict__ nSec, double *__restrict__ impact,
               const uint32_t *misc, uint32_t *stats)
{
    const uint32_t total = hitOff ? hitOff[runs.R] : n_hits(misc, runs.absN(0)), step = *stepDev;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        const uint32_t run = hitOff ? hit_run(hitOff, runs.R, t) : 0;
        const uint32_t jj = hitOff ? t - hitOff[run] : t, a0 = runs.abs0(run);
        const oecl_furman_t &Pr = runs.furman ? runs.furman[run] : P;
        {
            // the absorbed row of hole jj: the particle as the push left it
            const uint32_t i = holePos[a0 + jj];
            T *row = ap + 6 * (uint64_t)(a0 + jj);
#pragma unroll
            for (int c = 0; c < 6; ++c) row[c] = p(c, i);
        }
        hit_one<T>(jj, w, Pr, S, useSecem, ap + 6 * (uint64_t)a0, nv + 2 * (uint64_t)a0, rts + a0, dt, emit,
                   runs.key0(run, s0), runs.key1(run, s1), step, nSec + a0, impact + 4 * (uint64_t)a0,
                   stats + 8 * run);
    }
}

// hitOff[0 .. R]: exclusive prefix of the runs' hit counts (one block).
__global__ void __launch_bounds__(CT) k_hit_offsets(Runs runs, const uint32_t *__restrict__ misc, uint32_t *hitOff)
{
    __shared__ uint32_t wsum[CT / 64];
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < runs.R; b0 += CT) {
        const uint32_t r = b0 + threadIdx.x;
        const uint32_t c = r < runs.R ? n_hits(misc + r * M_COUNT, runs.absN(r)) : 0;
        uint32_t tot;
        const uint32_t o = block_scan_ct(c, wsum, &tot);
        if (r < runs.R) hitOff[r] = carry + o;
        carry += tot;
    }
    if (threadIdx.x == 0) hitOff[runs.R] = carry;
}

// Second loop of the simpler emitters (secem::pass2) into the hit's MAXSEC slots, one thread
// per hit, grid-stride.
template <typename T>
__global__ void __launch_bounds__(128)
    k_secem_emit_dev(Runs runs, const uint32_t *__restrict__ hitOff, oecl_secem_t S, const T *__restrict__ ap,
                     const T *__restrict__ nv, const uint32_t *__restrict__ nSec, const double *__restrict__ impact,
                     uint32_t s0, uint32_t s1, const uint32_t *__restrict__ stepDev, T *__restrict__ slots,
                     const uint32_t *misc)
{
    const uint32_t total = hitOff ? hitOff[runs.R] : n_hits(misc, runs.absN(0)), step = *stepDev;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        const uint32_t run = hitOff ? hit_run(hitOff, runs.R, t) : 0;
        const uint32_t jj = hitOff ? t - hitOff[run] : t, r = runs.abs0(run) + jj;
        const uint32_t ns = nSec[r];
        if (ns == 0) continue;
        const secem::Rng rng{nullptr, jj, step, runs.key0(run, s0), runs.key1(run, s1)};
        T *base = slots + 6 * (uint64_t)MAXSEC * r;
        secem::pass2<T>(S, ap + 6 * (uint64_t)r, nv + 2 * (uint64_t)r, impact + 4 * (uint64_t)r, rng,
                        [&](uint32_t k) -> T * { return k < ns && k < MAXSEC ? base + 6 * (uint64_t)k : nullptr; });
    }
}

// Second Furman loop into the hit's MAXSEC slots of `slots` (impact-major; k_place_dev makes
// the compact order).  Each wave takes 6 impacts at a time, one group of MAXSEC lanes per
// impact (furman::pass2_lane), grid-stride over the hits.  The single-secondary branch takes
// the LAST hit's first-loop values, as the reference does.
template <typename T>
__global__ void __launch_bounds__(64)
    k_emit_dev(Runs runs, const uint32_t *__restrict__ hitOff, oecl_furman_t P, const T *__restrict__ ap,
               const T *__restrict__ nv, const uint32_t *__restrict__ nSec, const double *__restrict__ impact,
               uint32_t s0, uint32_t s1, const uint32_t *__restrict__ stepDev, T *__restrict__ slots,
               const uint32_t *misc)
{
    constexpr uint32_t GPW = 64 / MAXSEC;  // impact groups per wave
    const uint32_t total = hitOff ? hitOff[runs.R] : n_hits(misc, runs.absN(0)), step = *stepDev;
    const uint32_t lane = threadIdx.x & 63, grp = lane / MAXSEC, ii = lane % MAXSEC;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nWaves = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t j0 = wave * GPW; j0 < total; j0 += nWaves * GPW) {  // wave-uniform bound
        const uint32_t t = j0 + grp;
        const bool ok = grp < GPW && t < total;
        const uint32_t run = ok && hitOff ? hit_run(hitOff, runs.R, t) : 0;
        const uint32_t jj = ok ? (hitOff ? t - hitOff[run] : t) : 0;
        const uint32_t a0 = runs.abs0(run), na = n_hits(misc + run * M_COUNT, runs.absN(run));
        const uint32_t r = a0 + jj;
        const uint32_t ns = ok ? nSec[r] : 0;
        const oecl_furman_t &Pr = runs.furman ? runs.furman[run] : P;
        furman::Rng rng{nullptr, jj, step, runs.key0(run, s0), runs.key1(run, s1)};
        T *o = ok && ii < ns ? slots + 6 * ((uint64_t)MAXSEC * r + ii) : nullptr;
        furman::pass2_lane<T>(Pr, ap + 6 * (uint64_t)r, nv + 2 * (uint64_t)r, ns, ok ? impact[4 * (uint64_t)r] : 0.,
                              impact + 4 * (uint64_t)(a0 + (na ? na - 1 : 0)), rng, ii, (int)(grp * MAXSEC), o);
    }
}

// One block per run: secOff = exclusive prefix of nSec over the hits (total at
// secOff[absCap]); secondary s (hit order, then emission order) is deposited, copied to
// secOut[s] (single run; secOut may be null) and placed in hole s, or appended after the live
// particles once the holes are used up.  With secOut, at most secCap secondaries are taken
// (more set ST_SEC_OVERFLOW); nothing is placed when the hits overflowed absCap
// (ST_ABS_OVERFLOW, set at the end of the step).  Run 0 advances the step counter.
template <typename T>
__global__ void __launch_bounds__(CT)
    k_place_dev(Runs runs, oecl_grid_t g, Soa<T> p, const T *__restrict__ slots, const T *__restrict__ photoSlots,
                T *__restrict__ secOut, uint32_t secCap, const uint32_t *__restrict__ nSec,
                uint32_t *__restrict__ secOff, uint32_t *__restrict__ nDev, const uint32_t *__restrict__ holePos,
                double q, double *rho, uint32_t *misc, uint32_t *stepDev, uint32_t *stats)
{
    __shared__ uint32_t wsum[CT / 64];
    const uint32_t run = blockIdx.x, a0 = runs.abs0(run), absCap = runs.absN(run);
    const uint64_t sBeg = runs.seg0(run), capacity = runs.seg1(run) - sBeg;
    misc += run * M_COUNT;
    stats += 8 * run;
    nDev += run;
    nSec += a0;
    secOff += a0 + run;
    holePos += a0;
    slots += 6 * (uint64_t)MAXSEC * a0;
    rho += (size_t)run * g.np;
    if (!secOut) secCap = 0xFFFFFFFFu;
    const uint32_t na = n_hits(misc, absCap), h = misc[M_HOLES], n = *nDev;
    const uint32_t tid = threadIdx.x;
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < na; b0 += CT) {
        const uint32_t r = b0 + tid;
        const uint32_t c = r < na ? nSec[r] : 0;
        uint32_t tot;
        const uint32_t o = block_scan_ct(c, wsum, &tot);
        if (r < na) secOff[r] = carry + o;
        carry += tot;
    }
    const uint32_t nPh = misc[M_PHOTO];
    if (tid == 0) secOff[absCap] = carry + nPh;
    __syncthreads();
    // secondary s (or photoelectron s - carry): deposit, copy out, place
    auto place = [&](uint32_t s, const T *r) {
        // FP64 weights of the stored position, as the push deposits the survivors
        Weights<double> w = grid_weights<double>((double)r[0], (double)r[1], 1. / g.dx, 1. / g.dy, g.lxHalfExt,
                                                 g.lyHalfExt, g.nx, g.ny);
        deposit_corners<double>(rho, w, g.nx, (double)r[5], q);
        if (secOut) {
#pragma unroll
            for (int c = 0; c < 6; ++c) secOut[6 * (uint64_t)s + c] = r[c];
        }
        if (h > absCap) return;
        if (s >= h && (uint64_t)n + (s - h) >= capacity) {
            atomicOr(stats + 2, (uint32_t)ST_CAP_OVERFLOW);
            return;
        }
        const uint64_t dst = s < h ? (uint64_t)holePos[s] : sBeg + n + (s - h);
#pragma unroll
        for (int c = 0; c < 6; ++c) p(c, dst) = r[c];
    };
    for (uint32_t jj = tid; jj < na; jj += CT) {
        const uint32_t ns = nSec[jj], base = secOff[jj];
        for (uint32_t ii = 0; ii < ns; ++ii) {
            const uint32_t s = base + ii;
            if (s >= secCap) {
                atomicOr(stats + 2, (uint32_t)ST_SEC_OVERFLOW);
                break;
            }
            place(s, slots + 6 * ((uint64_t)MAXSEC * jj + ii));
        }
    }
    for (uint32_t k = tid; k < nPh; k += CT) {
        const uint32_t s = carry + k;
        if (s >= secCap) {
            atomicOr(stats + 2, (uint32_t)ST_SEC_OVERFLOW);
            break;
        }
        place(s, photoSlots + 6 * (uint64_t)k);
    }
    const uint32_t tot = carry + nPh;
    const uint32_t nAdd = tot < secCap ? tot : secCap;
    __syncthreads();
    if (h <= absCap && nAdd < h) {
        const uint32_t cnt = h - nAdd;
        for (uint32_t t = tid; t < cnt; t += CT) {
            const uint32_t i = (uint32_t)sBeg + n - 1 - t;
            uint32_t lo = 0, hi = h;  // first index with holePos > i
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (holePos[mid] <= i) lo = mid + 1;
                else hi = mid;
            }
            if (lo > 0 && holePos[lo - 1] == i) continue;
            const uint32_t j = t - (h - lo);
            const uint32_t dst = holePos[nAdd + j];
#pragma unroll
            for (int c = 0; c < 6; ++c) p(c, dst) = p(c, i);
        }
    }
    if (tid != 0) return;
    if (tot > secCap) stats[2] |= ST_SEC_OVERFLOW;
    if (h > absCap) {
        stats[2] |= ST_ABS_OVERFLOW;  // compaction was skipped; the count stays
    } else {
        uint64_t newN = (uint64_t)n - h + nAdd;
        if (newN > capacity) {
            stats[2] |= ST_CAP_OVERFLOW;
            newN = capacity;
        }
        *nDev = (uint32_t)newN;
    }
    const uint32_t nPhPlaced = carry >= secCap ? 0u : min(nPh, secCap - carry);
    stats[0] += h;
    stats[1] += nAdd - nPhPlaced;
    stats[4] = h;
    stats[5] = nAdd;
    stats[6] += nPhPlaced;
    if (run == 0) *stepDev += 1;
}

unsigned grid_cap(uint64_t items, unsigned block, unsigned maxBlocks)
{
    const uint64_t b = (items + block - 1) / block;
    return (unsigned)(b < 1 ? 1 : (b > maxBlocks ? maxBlocks : b));
}

}  // namespace

// Work buffer of a step: per push wave hole counts and lists, the list of waves with holes,
// the ascending hole lists (absTotal), counters (M_COUNT per run), and MAXSEC secondary slots
// per hit.
static uint64_t slots_offset(uint64_t capacity, uint32_t absTotal, uint32_t R)
{
    const uint64_t nW = (capacity + PB - 1) / PB * (PB / 64);
    return ((4 * (nW * (64 + 2) + absTotal + (uint64_t)M_COUNT * R + R + 1)) + 255) / 256 * 256;
}

OECL_EXPORT uint32_t oecl_push_block_size(void) { return kPushBlock; }

OECL_EXPORT uint64_t oecl_step_work_bytes(uint64_t capacity, uint32_t absCap)
{
    return slots_offset(capacity, absCap, 1) + 8ull * 6 * MAXSEC * absCap + 1024;
}

OECL_EXPORT uint64_t oecl_step_batch_work_bytes(uint64_t capacity, uint32_t absTotal, uint32_t R)
{
    return slots_offset(capacity, absTotal, R) + 8ull * 6 * MAXSEC * absTotal + 1024;
}

namespace {

// E_r = phiToE phi_r for every run r = blockIdx.y (rows in the CSR's own column order, the
// arithmetic of oecl_csr_spmv); with yil, also the interleaved copy yil[2 (r % xStride) + r / xStride].
// E = phiToE phi per run (rows [Ex | Ey], nRows = 2 np); also into the interleaved copy yil when
// given: ilw doubles per node (2: (Ex, Ey); 4: (Ex, Ey, Bx, By), B left as it is).
__global__ void k_spmv_runs(uint32_t nRows, const int32_t *__restrict__ rowPtr, const int32_t *__restrict__ col,
                            const double *__restrict__ val, const double *__restrict__ x, uint32_t xStride,
                            double *__restrict__ y, double *__restrict__ yil, uint32_t ilw,
                            float *__restrict__ yil32)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nRows) return;
    x += (size_t)blockIdx.y * xStride;
    y += (size_t)blockIdx.y * nRows;
    double sum = 0.;
    for (int32_t k = rowPtr[r]; k < rowPtr[r + 1]; ++k) sum += val[k] * x[col[k]];
    y[r] = sum;
    const size_t o = (size_t)blockIdx.y * ilw * xStride + ilw * (r % xStride) + r / xStride;
    if (yil) yil[o] = sum;
    if (yil32) yil32[o] = (float)sum;
}

// E from phi by row class (e2nKind): the interior pattern (c phi[node - o], -c phi[node + o]),
// summed as the CSR rows are (0 + first product, then the second); empty rows give 0; listed
// rows are left to k_e_listed.  One thread per node, blockIdx.y the run; also into the
// interleaved copy yil ((Ex, Ey) per node) when given, and into its FP32 form yil32 when given.
__device__ __forceinline__ void e_struct_node(uint32_t r, uint32_t np, uint32_t nx, const uint8_t *__restrict__ kind,
                                              double cx, double cy, const double *__restrict__ phi,
                                              double *__restrict__ E, double *__restrict__ yil,
                                              float *__restrict__ yil32)
{
    if (r >= np) return;
    phi += (size_t)blockIdx.y * np;
    E += (size_t)blockIdx.y * 2 * np;
    if (yil) yil += (size_t)blockIdx.y * 2 * np + (size_t)2 * r;
    if (yil32) yil32 += (size_t)blockIdx.y * 2 * np + (size_t)2 * r;
    const uint32_t k = kind[r], kx = k & 3u, ky = k >> 2;
    double ex = 0., ey = 0.;
    if (kx == 1u) {
        ex += cx * phi[r - 1];
        ex += (-cx) * phi[r + 1];
    }
    if (ky == 1u) {
        ey += cy * phi[r - nx];
        ey += (-cy) * phi[r + nx];
    }
    if (kx != 2u) E[r] = ex;
    if (ky != 2u) E[np + r] = ey;
    if (yil) {
        if (kx != 2u && ky != 2u) {
            typedef double d2 __attribute__((ext_vector_type(2)));
            *reinterpret_cast<d2 *>(yil) = d2{ex, ey};   // one 16-byte store
        } else {
            if (kx != 2u) yil[0] = ex;
            if (ky != 2u) yil[1] = ey;
        }
    }
    if (yil32) {
        if (kx != 2u && ky != 2u) {
            *reinterpret_cast<float2 *>(yil32) = make_float2((float)ex, (float)ey);   // one 8-byte store
        } else {
            if (kx != 2u) yil32[0] = (float)ex;
            if (ky != 2u) yil32[1] = (float)ey;
        }
    }
}

// The listed rows of phiToE from the CSR (one thread per listed row, blockIdx.y the run).
__device__ __forceinline__ void e_listed_row(uint32_t t, uint32_t nList, const int32_t *__restrict__ rows,
                                             const int32_t *__restrict__ rowPtr, const int32_t *__restrict__ col,
                                             const double *__restrict__ val, const double *__restrict__ phi,
                                             uint32_t np, double *__restrict__ E, double *__restrict__ yil,
                                             uint32_t ilw, float *__restrict__ yil32)
{
    if (t >= nList) return;
    const uint32_t r = (uint32_t)rows[t];
    phi += (size_t)blockIdx.y * np;
    double sum = 0.;
    for (int32_t k = rowPtr[r]; k < rowPtr[r + 1]; ++k) sum += val[k] * phi[col[k]];
    E[(size_t)blockIdx.y * 2 * np + r] = sum;
    const size_t o = (size_t)blockIdx.y * ilw * np + ilw * (r % np) + r / np;
    if (yil) yil[o] = sum;
    if (yil32) yil32[o] = (float)sum;
}

// E = phiToE phi in one launch: blocks below sBlocks take the nodes by row class, the others
// the listed rows (the two write disjoint components, so their order does not matter).
__global__ void __launch_bounds__(256)
    k_e_rows(uint32_t sBlocks, uint32_t np, uint32_t nx, const uint8_t *__restrict__ kind, double cx, double cy,
             uint32_t nList, const int32_t *__restrict__ rows, const int32_t *__restrict__ rowPtr,
             const int32_t *__restrict__ col, const double *__restrict__ val, const double *__restrict__ phi,
             double *__restrict__ E, double *__restrict__ yil, uint32_t ilw, float *__restrict__ yil32)
{
    if (blockIdx.x < sBlocks)
        e_struct_node(blockIdx.x * 256 + threadIdx.x, np, nx, kind, cx, cy, phi, E, yil, yil32);
    else
        e_listed_row((blockIdx.x - sBlocks) * 256 + threadIdx.x, nList, rows, rowPtr, col, val, phi, np, E, yil, ilw,
                     yil32);
}

}  // namespace

// A float no larger than v (v >= 0 here: the limits of the wall margin).
static float float_down(double v)
{
    float f = (float)v;
    if ((double)f > v) f = nextafterf(f, -INFINITY);
    return f;
}

// The mixed push's wall margin of m cells (NearWall); a margin that leaves no room gives limits
// that no position meets, so every particle takes the FP64 path.
static NearWall near_wall(const WallDev &w, double m)
{
    NearWall nw{};
    nw.xlim = nw.ylim = -1.f;
    nw.gmax = -1.f;
    if (w.kind == OECL_WALL_RECT) {
        const double xl = w.lxHalf - m * w.dx, yl = w.lyHalf - m * w.dy;
        if (xl > 0. && yl > 0.) {
            nw.xlim = float_down(xl);
            nw.ylim = float_down(yl);
        }
    } else if (w.kind == OECL_WALL_ELLIP) {
        const double r = 1. - m * fmax(w.dx, w.dy) / fmin(0.5 * w.lx, 0.5 * w.ly);
        nw.ax = (float)w.ax;
        nw.ay = (float)w.ay;
        // FP32 evaluation of x^2 ax + y^2 ay is within a few 1e-7 of its value: keep 1e-5 clear
        if (r > 0.) nw.gmax = float_down(r * r * (1. - 1e-5));
    }
    return nw;
}

// Queues one step of every run.  R = runs.R; absTotal = rows of the per-hit arrays.
template <typename T>
static int step_typed(const oecl_grid_t *g, const WallDev &wd, const oecl_furman_t *P, oecl_fastpoisson *fp,
                      const oecl_step_args *a, const Runs &runs, hipStream_t s)
{
    const uint64_t cap = a->stride;
    const uint32_t np = g->np, R = runs.R, absTotal = a->absCap;
    Soa<T> p{(T *)a->data, a->stride};
    const uint64_t nBlk = (cap + PB - 1) / PB, nW = nBlk * (PB / 64);
    uint32_t *counts = (uint32_t *)a->work;
    uint32_t *list = counts + nW;
    uint32_t *nzList = list + nW * 64;
    uint32_t *holePos = nzList + nW;
    uint32_t *misc = holePos + absTotal;
    uint32_t *hitOff = R > 1 ? misc + (size_t)M_COUNT * R : nullptr;
    T *slots = (T *)((char *)a->work + slots_offset(cap, absTotal, R));
    int rc = fastpoisson_reserve(fp, R);
    if (rc) return rc;
    const FpBox box = fastpoisson_box(fp);
    if (a->externalPhi) {
        hipLaunchKernelGGL(k_begin_dev<T>, dim3(oecl_blocks(np, 256), R), dim3(256), 0, s, np, a->rho, a->beamScale,
                           a->nScale, a->stepDev, a->profile, box, a->phi, a->nDev, a->countHist, a->histCap, misc,
                           wd, a->photoCount, a->photoWeight, a->nPhoto, a->photoCap, a->seed0, a->seed1, a->photoE0,
                           a->photoSigmaE, P->particleMass, (T *)a->photoSlots, a->stats);
    } else {
        const uint64_t nEdge = 4ull * g->nx + 4ull * (g->ny - 4);
        const uint64_t nThr = nEdge > a->photoCap ? nEdge : a->photoCap;
        hipLaunchKernelGGL(k_begin_edges<T>, dim3(oecl_blocks(nThr, 256), R), dim3(256), 0, s, g->nx, g->ny, a->rho,
                           a->stepDev, a->phi, a->nDev, a->countHist, a->histCap, misc, wd, a->photoCount,
                           a->photoWeight, a->nPhoto, a->photoCap, a->seed0, a->seed1, a->photoE0, a->photoSigmaE,
                           P->particleMass, (T *)a->photoSlots, a->stats);
        rc = fastpoisson_solve_grid(fp, a->rho, a->profile, a->beamScale, a->beamScale ? a->nScale : 0, a->stepDev,
                                    a->phi, R, s);
        if (rc) return rc;
    }
    // mixed push (FP32 particles): E into its FP32 interleaved copy, the FP64 interleaved copy
    // is not used
    const bool mixed = sizeof(T) == 4 && a->pushArith == 1;
    // grid field with Bz = 0: B gathered from the (Bx, By) layout, E from Eil (or Eil32)
    const bool eb = a->typeB == 2 && a->Bil && (mixed ? a->Bil32 != nullptr : a->Eil != nullptr);
    double *il = mixed ? nullptr : a->Eil;
    float *il32 = mixed ? a->Eil32 : nullptr;
    const uint32_t ilw = 2u;
    if (a->e2nKind) {
        const uint32_t sBlocks = oecl_blocks(np, 256), lBlocks = a->nIrr ? oecl_blocks(a->nIrr, 256) : 0u;
        hipLaunchKernelGGL(k_e_rows, dim3(sBlocks + lBlocks, R), dim3(256), 0, s, sBlocks, np, g->nx, a->e2nKind,
                           a->e2nCx, a->e2nCy, a->nIrr, a->e2nIrr, a->e2nRowPtr, a->e2nCol, a->e2nVal, a->phi, a->E,
                           il, ilw, il32);
    } else {
        hipLaunchKernelGGL(k_spmv_runs, dim3(oecl_blocks(2 * np, 256), R), dim3(256), 0, s, 2 * np, a->e2nRowPtr,
                           a->e2nCol, a->e2nVal, a->phi, np, a->E, il, ilw, il32);
    }
    double kk = a->typeB == 0 ? a->qm * a->dt : 0.5 * a->qm * a->dt;
    const double dxi = 1. / g->dx, dyi = 1. / g->dy;
    const dim3 gp((unsigned)nBlk), bp(PT);
    const DepCfg dc{a->depSpan ? a->depSpan : (uint32_t)OECL_RUNSUM_SPAN, a->depTile ? 1u : 0u};
    if (a->evPushBegin) OECL_HIP(hipEventRecord((hipEvent_t)a->evPushBegin, s), "event");
    if (mixed) {
        const NearWall nw = near_wall(wd, a->nearCells);
        const Push32 P32{(float)dxi, (float)dyi, (float)(g->lxHalfExt * dxi), (float)(g->lyHalfExt * dyi),
                         (float)a->dt, (float)kk, (float)(kk * a->b0), (float)(kk * a->b1), (float)(kk * a->b2)};
        Soa<float> pf{(float *)a->data, a->stride};
        const dim3 gm((unsigned)nBlk), bm(MPT);
        // 24 bytes of FP32 state per particle slot
        const bool nt = 24ull * a->stride > kStreamBytes;
#define OECL_MIXED(TB, ...)                                                       \
    do {                                                                          \
        if (nt) hipLaunchKernelGGL((k_push_mixed<TB, 2>), gm, bm, 0, s, __VA_ARGS__); \
        else hipLaunchKernelGGL((k_push_mixed<TB, 0>), gm, bm, 0, s, __VA_ARGS__);    \
    } while (0)
        if (a->typeB == 0)
            OECL_MIXED(0, runs, *g, wd, nw, P32, pf, a->nDev, a->E, a->Eil32,
                               a->bField, nullptr, 0., 0., 0., a->dt, kk, a->qe, dxi, dyi, a->isInside, a->rho, counts,
                               list, misc, nzList, dc);
        else if (a->typeB == 1)
            OECL_MIXED(1, runs, *g, wd, nw, P32, pf, a->nDev, a->E, a->Eil32,
                               a->bField, nullptr, a->b0, a->b1, a->b2, a->dt, kk, a->qe, dxi, dyi, a->isInside,
                               a->rho, counts, list, misc, nzList, dc);
        else if (eb)
            OECL_MIXED(3, runs, *g, wd, nw, P32, pf, a->nDev, a->E, a->Eil32,
                               a->Bil, a->Bil32, 0., 0., 0., a->dt, kk, a->qe, dxi, dyi, a->isInside, a->rho, counts,
                               list, misc, nzList, dc);
        else
            OECL_MIXED(2, runs, *g, wd, nw, P32, pf, a->nDev, a->E, a->Eil32,
                               a->bField, nullptr, 0., 0., 0., a->dt, kk, a->qe, dxi, dyi, a->isInside, a->rho, counts,
                               list, misc, nzList, dc);
#undef OECL_MIXED
    } else if (a->typeB == 0)
        hipLaunchKernelGGL((k_push_dev<T, 0>), gp, bp, 0, s, runs, *g, wd, p, a->nDev, a->E, a->Eil, a->bField, 0., 0., 0.,
                           a->dt, kk, a->qe, dxi, dyi, a->isInside, a->rho, counts, list, misc, nzList, dc);
    else if (a->typeB == 1)
        hipLaunchKernelGGL((k_push_dev<T, 1>), gp, bp, 0, s, runs, *g, wd, p, a->nDev, a->E, a->Eil, a->bField, a->b0,
                           a->b1, a->b2, a->dt, kk, a->qe, dxi, dyi, a->isInside, a->rho, counts, list, misc, nzList, dc);
    else if (eb)
        hipLaunchKernelGGL((k_push_dev<T, 3>), gp, bp, 0, s, runs, *g, wd, p, a->nDev, a->E, a->Eil, a->Bil, 0., 0.,
                           0., a->dt, kk, a->qe, dxi, dyi, a->isInside, a->rho, counts, list, misc, nzList, dc);
    else
        hipLaunchKernelGGL((k_push_dev<T, 2>), gp, bp, 0, s, runs, *g, wd, p, a->nDev, a->E, a->Eil, a->bField, 0., 0., 0.,
                           a->dt, kk, a->qe, dxi, dyi, a->isInside, a->rho, counts, list, misc, nzList, dc);
    if (a->evPushEnd) OECL_HIP(hipEventRecord((hipEvent_t)a->evPushEnd, s), "event");
    hipLaunchKernelGGL(k_holes_collect, dim3(R), dim3(CT), 0, s, runs, a->nDev, counts, list, nzList, holePos, misc);
    const oecl_secem_t S = a->secem ? *a->secem : oecl_secem_t{};
    const int useSecem = a->secem ? 1 : 0;
    if (hitOff) hipLaunchKernelGGL(k_hit_offsets, dim3(1), dim3(CT), 0, s, runs, misc, hitOff);
    const unsigned hitBlocks = grid_cap(absTotal, 128, R > 1 ? 2048 : 512);
    hipLaunchKernelGGL(k_hits_dev<T>, dim3(hitBlocks), dim3(128), 0, s, runs, hitOff, wd, *P, S, useSecem, p,
                       holePos, (T *)a->absorbed, (T *)a->normals, (T *)a->remaining, a->dt, a->emit, a->seed0, a->seed1,
                       a->stepDev, a->nSec, a->impact, misc, a->stats);
    if (a->emit && useSecem)
        hipLaunchKernelGGL(k_secem_emit_dev<T>, dim3(hitBlocks), dim3(128), 0, s, runs, hitOff, S,
                           (const T *)a->absorbed, (const T *)a->normals, a->nSec, a->impact, a->seed0, a->seed1,
                           a->stepDev, slots, misc);
    else if (a->emit)
        hipLaunchKernelGGL(k_emit_dev<T>, dim3(grid_cap(absTotal, 6, R > 1 ? 4096 : 1024)), dim3(64), 0, s, runs,
                           hitOff, *P, (const T *)a->absorbed, (const T *)a->normals, a->nSec, a->impact, a->seed0,
                           a->seed1, a->stepDev, slots, misc);
    hipLaunchKernelGGL(k_place_dev<T>, dim3(R), dim3(CT), 0, s, runs, *g, p, (const T *)slots,
                       (const T *)a->photoSlots, (T *)a->secondaries, a->secCap, a->nSec, a->secOff, a->nDev, holePos,
                       a->qe, a->rho, misc, a->stepDev, a->stats);
    OECL_CHECK_LAUNCH("oecl_step_run");
    return OECL_OK;
}

static int step_check(const oecl_grid_t *g, const oecl_wall_t *wall, oecl_fastpoisson *fp, const oecl_step_args *a)
{
    if (!g || !wall || !fp || !a || !a->data || !a->nDev || !a->rho || !a->q || !a->phi || !a->E ||
        !a->e2nRowPtr || !a->e2nCol || !a->e2nVal || !a->absorbed || !a->normals || !a->remaining || !a->nSec ||
        !a->secOff || !a->impact || !a->stepDev || !a->stats || !a->work || a->absCap == 0)
        return OECL_ERR_ARG;
    if (a->typeB == 2 && !a->bField) return OECL_ERR_ARG;
    if (a->e2nKind && a->nIrr && !a->e2nIrr) return OECL_ERR_ARG;
    if (a->secem && a->secem->model != OECL_SECEM_STAINLESS && a->secem->model != OECL_SECEM_SIMPLE)
        return OECL_ERR_ARG;
    if (a->photoCount && (!a->photoWeight || !a->photoSlots || a->photoCap == 0)) return OECL_ERR_ARG;
    if (a->photoCount && wall->kind != OECL_WALL_RECT && wall->kind != OECL_WALL_ELLIP) {
        oecl_set_error_msg("photoelectrons: rectangular or elliptical wall only");
        return OECL_ERR_ARG;
    }
    if (a->dtype != OECL_F64 && a->dtype != OECL_F32) return OECL_ERR_ARG;
    if (a->pushArith != 0 && a->pushArith != 1) return OECL_ERR_ARG;
    if ((a->depSpan != 0 && (a->depSpan < 2 || a->depSpan > 64 || (a->depSpan & (a->depSpan - 1)))) ||
        (a->depTile != 0 && a->depTile != 1)) {
        oecl_set_error_msg("step: depSpan must be 0 or a power of two in 2..64, depTile 0 or 1");
        return OECL_ERR_ARG;
    }
    if (a->dtype == OECL_F32 && a->pushArith == 1) {
        if (!a->Eil32 || !(a->nearCells >= 1.)) {
            oecl_set_error_msg("step: the mixed push needs Eil32 and nearCells >= 1");
            return OECL_ERR_ARG;
        }
        if (a->typeB == 2 && a->Bil && !a->Bil32) {
            oecl_set_error_msg("step: the mixed push needs Bil32 with Bil");
            return OECL_ERR_ARG;
        }
        // the mixed push addresses a particle component and a grid field copy by 32-bit byte offsets
        if (a->stride * 4u > 0xFFFFFFFFull || (uint64_t)g->np * 8u > 0xFFFFFFFFull) {
            oecl_set_error_msg("step: the mixed push needs capacity * 4 and npExt * 8 below 2^32 bytes");
            return OECL_ERR_ARG;
        }
    }
    if (g->nx >= 0xFFFFu || g->ny >= 0xFFFFu) {   // the push's deposit packs cell indices in 16 bits
        oecl_set_error_msg("step: nxExt and nyExt must be below 65535");
        return OECL_ERR_ARG;
    }
    return OECL_OK;
}

OECL_EXPORT int oecl_step_run(const oecl_grid_t *g, const oecl_wall_t *wall, const oecl_furman_t *P,
                              oecl_fastpoisson *fp, const oecl_step_args *a, void *stream)
{
    int rc = step_check(g, wall, fp, a);
    if (rc) return rc;
    if (!P || !a->secondaries || a->secCap == 0) return OECL_ERR_ARG;
    WallDev wd;
    rc = wall_to_device(wall, &wd);
    if (rc) return rc;
    const Runs runs{1u, a->absCap, a->stride, nullptr, nullptr, nullptr, nullptr, nullptr};
    hipStream_t s = (hipStream_t)stream;
    if (a->dtype == OECL_F64) return step_typed<double>(g, wd, P, fp, a, runs, s);
    return step_typed<float>(g, wd, P, fp, a, runs, s);
}

OECL_EXPORT int oecl_step_run_batch(const oecl_grid_t *g, const oecl_wall_t *wall, const oecl_runs_t *rs,
                                    oecl_fastpoisson *fp, const oecl_step_args *a, void *stream)
{
    int rc = step_check(g, wall, fp, a);
    if (rc) return rc;
    if (!rs || rs->R == 0 || rs->R > 65535 || !rs->segOff || !rs->blockRun || !rs->absOff || !rs->keys)
        return OECL_ERR_ARG;
    if (a->emit && !a->secem && !rs->furman) return OECL_ERR_ARG;
    if (a->photoCount) {
        oecl_set_error_msg("step batch: no photoelectron source (use oecl_step_run)");
        return OECL_ERR_ARG;
    }
    if (a->stride % PB) {
        oecl_set_error_msg("step batch: the capacity must be a multiple of the push block (256)");
        return OECL_ERR_ARG;
    }
    WallDev wd;
    rc = wall_to_device(wall, &wd);
    if (rc) return rc;
    const oecl_furman_t P{};  // the runs' emitter parameters come from rs->furman
    const Runs runs{rs->R, a->absCap, a->stride, rs->segOff, rs->blockRun, rs->absOff, rs->furman, rs->keys};
    hipStream_t s = (hipStream_t)stream;
    if (a->dtype == OECL_F64) return step_typed<double>(g, wd, &P, fp, a, runs, s);
    return step_typed<float>(g, wd, &P, fp, a, runs, s);
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
