// Per-particle device helpers shared by the particle kernels and the fused step.
#pragma once
#include "common.h"

// particles.pyx:400-418 for one particle, with the cell index clamped into the grid so an
// escaped particle can never address memory outside the node arrays (the reference has no
// bounds check there; inside the grid the result is unchanged).
template <typename T>
struct Weights {
    uint32_t cell, ix, iy;
    T wLL, wUL, wLU, wUU;
};

template <typename T>
__device__ __forceinline__ Weights<T> grid_weights(T x, T y, T dxi, T dyi, T lxh, T lyh,
                                                   uint32_t nx, uint32_t ny)
{
    T fx = (x + lxh) * dxi;
    T fy = (y + lyh) * dyi;
    T fxc = fx < T(0) ? T(0) : fx;
    T fyc = fy < T(0) ? T(0) : fy;
    uint32_t ix = (uint32_t)fxc;
    uint32_t iy = (uint32_t)fyc;
    if (ix > nx - 2) ix = nx - 2;
    if (iy > ny - 2) iy = ny - 2;
    fx = fx - (T)ix;
    fy = fy - (T)iy;
    Weights<T> w;
    w.cell = ix + iy * nx;
    w.ix = ix;
    w.iy = iy;
    w.wLL = (1 - fx) * (1 - fy);
    w.wUL = fx * (1 - fy);
    w.wLU = (1 - fx) * fy;
    w.wUU = fx * fy;
    return w;
}

// Four FP64 atomic adds: (w_k * weight) * q at the cell's corners (particles.pyx:344-351).
template <typename T>
__device__ __forceinline__ void deposit_corners(double *rho, const Weights<T> &w, uint32_t nx,
                                                double wt, double q)
{
    unsafeAtomicAdd(rho + w.cell, (double)w.wLL * wt * q);
    unsafeAtomicAdd(rho + w.cell + 1, (double)w.wUL * wt * q);
    unsafeAtomicAdd(rho + w.cell + nx, (double)w.wLU * wt * q);
    unsafeAtomicAdd(rho + w.cell + nx + 1, (double)w.wUU * wt * q);
}

#ifndef OECL_DEPOSIT_RUN_SUMS
#define OECL_DEPOSIT_RUN_SUMS 1
#endif

// ---- wave-level reductions through DPP (data-parallel primitives: VALU lane moves within
// rows of 16 lanes, row_bcast across rows), which cost no LDS bandwidth.  All 64 lanes of
// the wave must be active when these are called.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v, uint32_t ident)
{
    // lanes outside ROWMASK, or whose source lane is out of the row, receive `ident`
    return (uint32_t)__builtin_amdgcn_update_dpp((int)ident, (int)v, CTRL, ROWMASK, 0xF, false);
}

// Lane move with every row written: lanes whose source lies outside their row read 0
// (bound_ctrl), so no identity has to be placed in the destination first.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp0_u32(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, true);
}

template <int CTRL, int ROWMASK>
__device__ __forceinline__ uint32_t dpp_any_u32(uint32_t v)
{
    if (ROWMASK == 0xF) return dpp0_u32<CTRL>(v);
    return dpp_u32<CTRL, ROWMASK>(v, 0u);
}

template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_f64(double v)
{
    const uint32_t lo = (uint32_t)__double2loint(v), hi = (uint32_t)__double2hiint(v);
    return __hiloint2double((int)dpp_any_u32<CTRL, ROWMASK>(hi), (int)dpp_any_u32<CTRL, ROWMASK>(lo));
}

// One step of a segmented inclusive scan: a lane that does not start a segment inside its
// current range adds its predecessor's partial sums and inherits its start flag.  Lanes
// without a predecessor in range receive 0 for both.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ void seg_scan_step(uint32_t &f, double &q0, double &q1, double &q2, double &q3)
{
    const uint32_t fp = dpp_any_u32<CTRL, ROWMASK>(f);
    const double p0 = dpp_f64<CTRL, ROWMASK>(q0), p1 = dpp_f64<CTRL, ROWMASK>(q1);
    const double p2 = dpp_f64<CTRL, ROWMASK>(q2), p3 = dpp_f64<CTRL, ROWMASK>(q3);
    if (!f) {
        q0 += p0;
        q1 += p1;
        q2 += p2;
        q3 += p3;
        f = fp;
    }
}

// Sums each run of equal cells among consecutive lanes of the wave: a segmented prefix scan
// (row_shr 1, 2, 4, 8 within rows, then row_bcast 15 and 31 across rows, as far as the span
// needs).  The LAST lane of a run (within its span group) receives the run's totals and returns
// true; lanes with c == 0xFFFFFFFF never do.
// Atomics to one address from many lanes of a wave are serialised by the hardware, so summing
// the runs first is what makes the deposit of cell-sorted particles cheap.
// Span of the run sums: runs are summed within aligned groups of this many lanes (2 scan steps
// for 4 instead of 6 for the whole wave); a run crossing a group boundary gives one head per
// group, i.e. a few more LDS atomics.  4 measured fastest for the cell-sorted 2e7-particle push
// on the 2048^2 grid (about 5 particles per cell), 2-3 % ahead of 64.
#ifndef OECL_RUNSUM_SPAN
#define OECL_RUNSUM_SPAN 4
#endif
// S (a power of two, 2-64, the same in all lanes of the wave) may be a run-time value: the scan
// steps it needs are then taken by wave-uniform branches.
__device__ __forceinline__ bool wave_run_sum(uint32_t c, double &q0, double &q1, double &q2, double &q3,
                                             uint32_t S = OECL_RUNSUM_SPAN)
{
    // runs are cut at aligned groups of S lanes (S = 64: the whole wave)
    const uint32_t lane = threadIdx.x & 63;
    // the neighbours' cells: within a row of 16 lanes by DPP row shifts when the groups do not
    // cross rows (a row's first lane starts a group, its last lane ends one, so the 0 that
    // bound_ctrl gives them is never used), else by cross-lane shuffles
    uint32_t cp, cn;
    if (S <= 16) {
        cp = dpp0_u32<0x111>(c);   // row_shr:1
        cn = dpp0_u32<0x101>(c);   // row_shl:1
    } else {
        cp = __shfl_up(c, 1, 64);
        cn = __shfl_down(c, 1, 64);
    }
    uint32_t f = ((lane & (S - 1)) == 0 || cp != c) ? 1u : 0u;
    const bool last = (lane & (S - 1)) == S - 1 || cn != c;
    seg_scan_step<0x111, 0xF>(f, q0, q1, q2, q3);                  // row_shr:1
    if (S > 2) seg_scan_step<0x112, 0xF>(f, q0, q1, q2, q3);       // row_shr:2
    if (S > 4) seg_scan_step<0x114, 0xF>(f, q0, q1, q2, q3);       // row_shr:4
    if (S > 8) seg_scan_step<0x118, 0xF>(f, q0, q1, q2, q3);       // row_shr:8
    if (S > 16) {
        seg_scan_step<0x142, 0xA>(f, q0, q1, q2, q3);  // row_bcast:15 into rows 1 and 3
        seg_scan_step<0x143, 0xC>(f, q0, q1, q2, q3);  // row_bcast:31 into rows 2 and 3
    }
    return last && c != 0xFFFFFFFFu;
}

// Minimum (MIN) or maximum over the wave, uniform result.
// Rows of 16 lanes are reduced by rotations (row_ror 1, 2, 4, 8: after them every lane holds its
// row's result), which read no lane outside the row and so need no identity value.
template <bool MIN>
__device__ __forceinline__ uint32_t wave_minmax(uint32_t v)
{
    uint32_t o;
    o = dpp0_u32<0x121>(v);
    v = MIN ? min(v, o) : max(v, o);
    o = dpp0_u32<0x122>(v);
    v = MIN ? min(v, o) : max(v, o);
    o = dpp0_u32<0x124>(v);
    v = MIN ? min(v, o) : max(v, o);
    o = dpp0_u32<0x128>(v);
    v = MIN ? min(v, o) : max(v, o);
    const uint32_t a = __builtin_amdgcn_readlane(v, 15), b = __builtin_amdgcn_readlane(v, 31);
    const uint32_t c = __builtin_amdgcn_readlane(v, 47), d = __builtin_amdgcn_readlane(v, 63);
    return MIN ? min(min(a, b), min(c, d)) : max(max(a, b), max(c, d));
}

// Minimum (MIN) or maximum over the wave of two 16-bit fields at once (v = lo | hi << 16;
// packed 16-bit integer min/max per field), uniform result.
template <bool MIN>
__device__ __forceinline__ uint32_t wave_minmax2x16(uint32_t v)
{
    typedef unsigned short u2 __attribute__((ext_vector_type(2)));
    auto op = [](uint32_t a, uint32_t b) {
        const u2 x = __builtin_bit_cast(u2, a), y = __builtin_bit_cast(u2, b);
        return __builtin_bit_cast(uint32_t, MIN ? __builtin_elementwise_min(x, y) : __builtin_elementwise_max(x, y));
    };
    v = op(v, dpp0_u32<0x121>(v));   // row_ror:1
    v = op(v, dpp0_u32<0x122>(v));   // row_ror:2
    v = op(v, dpp0_u32<0x124>(v));   // row_ror:4
    v = op(v, dpp0_u32<0x128>(v));   // row_ror:8
    const uint32_t a = __builtin_amdgcn_readlane(v, 15), b = __builtin_amdgcn_readlane(v, 31);
    const uint32_t c = __builtin_amdgcn_readlane(v, 47), d = __builtin_amdgcn_readlane(v, 63);
    return op(op(a, b), op(c, d));
}

// Adds the non-zero nodes of an LDS tile of nt <= 2048 nodes (tw nodes per row, rows of nodes
// from (X0, Y0)) to rho, one global FP64 atomic each; thread t of `step` takes nodes t, t + step,
// ...  The row of tile node k is floor((k + 0.5) / tw) in FP32 with the hardware reciprocal:
// (k + 0.5) / tw lies at least 0.5 / tw from an integer, the computed quotient (below
// 2048 / tw) within 1.5 * 2^-23 of its value, so the floor is exact for k < 2048.
__device__ __forceinline__ void tile_to_rho(double *__restrict__ rho, uint32_t nx, const double *tile, uint32_t nt,
                                            uint32_t tw, uint32_t X0, uint32_t Y0, uint32_t t, uint32_t step)
{
    const float rtw = __builtin_amdgcn_rcpf((float)tw);
    for (uint32_t k = t; k < nt; k += step) {
        const double v = tile[k];
        if (v != 0.) {
            const uint32_t r = (uint32_t)(((float)k + 0.5f) * rtw);
            unsafeAtomicAdd(rho + ((X0 + (k - r * tw)) + (Y0 + r) * nx), v);
        }
    }
}

// Wave-cooperative deposit (all 64 lanes call it; valid = false when the lane has no
// particle), with the particle's cell (ix, iy) and its four corner charges.  Runs of equal
// cells are summed first (wave_run_sum).  The run totals are then summed in the wave's own
// LDS tile `win` covering the bounding box of the wave's cells (nodes [X0, X1+1] x
// [Y0, Y1+1]), and each touched node is added to rho with one global FP64 atomic.  Particles
// sorted by cell, or only locally in order (a few cells of drift since the last sort), keep
// the box small; a wave whose box exceeds WN nodes adds its run totals to rho directly.  No
// block barrier is involved: LDS accesses of one wave are ordered.
template <int WN>
__device__ __forceinline__ void wave_deposit(double *__restrict__ rho, uint32_t nx, bool valid, uint32_t ix,
                                             uint32_t iy, double q00, double q10, double q01, double q11,
                                             double *win)
{
    static_assert(WN <= 2048, "tile_to_rho handles tiles of at most 2048 nodes");
    const int lane = threadIdx.x & 63;
    const uint32_t c = valid ? ix + iy * nx : 0xFFFFFFFFu;
    if (!valid) q00 = q10 = q01 = q11 = 0.;
    const bool head = wave_run_sum(c, q00, q10, q01, q11);
    const uint32_t X0 = wave_minmax<true>(head ? ix : 0xFFFFFFFFu), X1 = wave_minmax<false>(head ? ix : 0u);
    const uint32_t Y0 = wave_minmax<true>(head ? iy : 0xFFFFFFFFu), Y1 = wave_minmax<false>(head ? iy : 0u);
    if (X0 == 0xFFFFFFFFu) return;  // nothing to deposit in this wave (wave-uniform)
    const uint32_t wx = X1 - X0 + 2, wy = Y1 - Y0 + 2;
    if ((uint64_t)wx * wy > WN) {  // wave-uniform
        if (head) {
            const uint32_t cc = ix + iy * nx;
            unsafeAtomicAdd(rho + cc, q00);
            unsafeAtomicAdd(rho + cc + 1, q10);
            unsafeAtomicAdd(rho + cc + nx, q01);
            unsafeAtomicAdd(rho + cc + nx + 1, q11);
        }
        return;
    }
    const uint32_t nw = wx * wy;
    for (uint32_t k = lane; k < nw; k += 64) win[k] = 0.;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (head) {
        const uint32_t o = (ix - X0) + (iy - Y0) * wx;
        atomicAdd(&win[o], q00);
        atomicAdd(&win[o + 1], q10);
        atomicAdd(&win[o + wx], q01);
        atomicAdd(&win[o + wx + 1], q11);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    tile_to_rho(rho, nx, win, nw, wx, X0, Y0, lane, 64);
}

// Block-cooperative deposit: every thread of the block calls it, once.  As wave_deposit, but the
// run totals of all NW waves go to one LDS tile covering the bounding box of the block's cells
// when that box fits NW * WN nodes, so each touched node costs one global FP64 atomic per block
// instead of one per wave: those atomics, on nodes that consecutive sorted blocks share, are what
// bounds the push (adding the run totals to rho directly made it 3.4x slower).  A block whose box
// is larger uses per-wave tiles in the same LDS (WN nodes each), as wave_deposit does.  tile:
// NW * WN doubles of LDS; box: 4 words of LDS.
// The deposit's tile choices, the same for the whole launch: span, the run-sum span of
// wave_run_sum; waveTiles, per-wave tiles only (no block tile, no block barrier).
struct DepCfg {
    uint32_t span;
    uint32_t waveTiles;
};

template <int NW, int WN>
__device__ __forceinline__ void block_deposit(double *__restrict__ rho, uint32_t nx, bool valid, uint32_t ix,
                                              uint32_t iy, double q00, double q10, double q01, double q11,
                                              double *tile, uint32_t *box,
                                              DepCfg dc = DepCfg{OECL_RUNSUM_SPAN, 0u})
{
    static_assert(NW * WN <= 2048, "tile_to_rho handles tiles of at most 2048 nodes");
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t c = valid ? ix + iy * nx : 0xFFFFFFFFu;
    if (!valid) q00 = q10 = q01 = q11 = 0.;
#if OECL_DEPOSIT_RUN_SUMS
    const bool head = wave_run_sum(c, q00, q10, q01, q11, dc.span);
#else
    const bool head = valid;
#endif
    // the heads' bounding box in cells (ix, iy < 0xFFFF: both fit 16 bits; no head: X0 = 0xFFFFFFFF)
    const uint32_t xy = ix | iy << 16;
    const uint32_t mn = wave_minmax2x16<true>(head ? xy : 0xFFFFFFFFu), mx = wave_minmax2x16<false>(head ? xy : 0u);
    const uint32_t X0 = mn == 0xFFFFFFFFu ? 0xFFFFFFFFu : mn & 0xFFFFu, X1 = mx & 0xFFFFu;
    const uint32_t Y0 = mn >> 16, Y1 = mx >> 16;
    if (!dc.waveTiles) {   // launch-uniform
        if (threadIdx.x == 0) {
            box[0] = 0xFFFFFFFFu;
            box[1] = 0u;
            box[2] = 0xFFFFFFFFu;
            box[3] = 0u;
        }
        __syncthreads();
        if (lane == 0 && X0 != 0xFFFFFFFFu) {
            atomicMin(&box[0], X0);
            atomicMax(&box[1], X1);
            atomicMin(&box[2], Y0);
            atomicMax(&box[3], Y1);
        }
        __syncthreads();
        const uint32_t B0 = box[0], B1 = box[1], C0 = box[2], C1 = box[3];
        if (B0 == 0xFFFFFFFFu) return;  // nothing to deposit in this block (block-uniform)
        const uint32_t bw = B1 - B0 + 2, bh = C1 - C0 + 2;
        if ((uint64_t)bw * bh <= (uint64_t)NW * WN) {  // block-uniform
            const uint32_t nb = bw * bh;
            for (uint32_t k = threadIdx.x; k < nb; k += NW * 64) tile[k] = 0.;
            __syncthreads();
            if (head) {
                const uint32_t o = (ix - B0) + (iy - C0) * bw;
                atomicAdd(&tile[o], q00);
                atomicAdd(&tile[o + 1], q10);
                atomicAdd(&tile[o + bw], q01);
                atomicAdd(&tile[o + bw + 1], q11);
            }
            __syncthreads();
            tile_to_rho(rho, nx, tile, nb, bw, B0, C0, threadIdx.x, NW * 64);
            return;
        }
    }
    if (X0 == 0xFFFFFFFFu) return;  // wave-uniform
    double *win = tile + wv * WN;
    const uint32_t wx = X1 - X0 + 2, wy = Y1 - Y0 + 2;
    if ((uint64_t)wx * wy > WN) {  // wave-uniform
        if (head) {
            unsafeAtomicAdd(rho + c, q00);
            unsafeAtomicAdd(rho + c + 1, q10);
            unsafeAtomicAdd(rho + c + nx, q01);
            unsafeAtomicAdd(rho + c + nx + 1, q11);
        }
        return;
    }
    const uint32_t nw = wx * wy;
    for (uint32_t k = lane; k < nw; k += 64) win[k] = 0.;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (head) {
        const uint32_t o = (ix - X0) + (iy - Y0) * wx;
        atomicAdd(&win[o], q00);
        atomicAdd(&win[o + 1], q10);
        atomicAdd(&win[o + wx], q01);
        atomicAdd(&win[o + wx + 1], q11);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    tile_to_rho(rho, nx, win, nw, wx, X0, Y0, lane, 64);
}

// Boris rotation + second half kick of particles.pyx:272-321, from the velocity after the
// first half kick.
template <typename T>
__device__ __forceinline__ void boris_rotate(T &vx, T &vy, T &vz, T t0, T t1, T t2, T kex,
                                             T key)
{
    T vs0 = vx + vy * t2 - vz * t1;
    T vs1 = vy + vz * t0 - vx * t2;
    T vs2 = vz + vx * t1 - vy * t0;
    T f = T(2) / (T(1) + t0 * t0 + t1 * t1 + t2 * t2);
    T s0 = t0 * f, s1 = t1 * f, s2 = t2 * f;
    vx += vs1 * s2 - vs2 * s1 + kex;
    vy += vs2 * s0 - vs0 * s2 + key;
    vz += vs0 * s1 - vs1 * s0;
}

// E at the particle from the reference's layout [Ex | Ey] (particles.pyx:354-372 order).
template <typename T>
__device__ __forceinline__ void gather_e(const Weights<T> &w, uint32_t nx, uint32_t np, const double *__restrict__ E,
                                         T &ex, T &ey)
{
    const uint32_t c = w.cell;
    ex = w.wLL * (T)E[c];
    ey = w.wLL * (T)E[c + np];
    ex += w.wUL * (T)E[c + 1];
    ey += w.wUL * (T)E[c + 1 + np];
    ex += w.wLU * (T)E[c + nx];
    ey += w.wLU * (T)E[c + nx + np];
    ex += w.wUU * (T)E[c + nx + 1];
    ey += w.wUU * (T)E[c + nx + 1 + np];
}

// The same from an interleaved copy Eil[2 node + {0, 1}] = (Ex, Ey): one 16-byte load per node,
// the same products in the same order.
template <typename T>
__device__ __forceinline__ void gather_e_il(const Weights<T> &w, uint32_t nx, const double *__restrict__ Eil, T &ex,
                                            T &ey)
{
    typedef double d2 __attribute__((ext_vector_type(2)));
    const d2 *E2 = reinterpret_cast<const d2 *>(Eil);
    const uint32_t c = w.cell;
    const d2 a = E2[c], b = E2[c + 1], cc = E2[c + nx], d = E2[c + nx + 1];
    ex = w.wLL * (T)a.x;
    ey = w.wLL * (T)a.y;
    ex += w.wUL * (T)b.x;
    ey += w.wUL * (T)b.y;
    ex += w.wLU * (T)cc.x;
    ey += w.wLU * (T)cc.y;
    ex += w.wUU * (T)d.x;
    ey += w.wUU * (T)d.y;
}

// Gather E (and B for TYPEB 2) at the particle and apply the Boris push (particles.pyx:
// 272-333 with _fieldToParticles's accumulation order).  Positions advance in x and y only.
// With Eil (not null) E is read from the interleaved copy.  TYPEB 3 is TYPEB 2 for a field with
// Bz = 0: E from Eil, B from the (Bx, By) layout B; the same products and sums (the Bz sum of
// TYPEB 2 is +0 there).
template <typename T, int TYPEB>
__device__ __forceinline__ void gather_push(const Weights<T> &w, uint32_t nx, uint32_t np,
                                            const double *__restrict__ E,
                                            const double *__restrict__ B, T b0, T b1, T b2, T dt,
                                            T k, T &x, T &y, T &vx, T &vy, T &vz,
                                            const double *__restrict__ Eil = nullptr)
{
    const uint32_t c = w.cell;
    T ex, ey;
    if (TYPEB == 3) {
        // E from Eil (or E when Eil is null), B from B in the (Bx, By) layout; Bz is zero
        typedef double d2 __attribute__((ext_vector_type(2)));
        if (Eil) gather_e_il<T>(w, nx, Eil, ex, ey);
        else gather_e<T>(w, nx, np, E, ex, ey);
        vx += k * ex;
        vy += k * ey;
        const d2 *B2 = reinterpret_cast<const d2 *>(B);
        const d2 ba = B2[c], bb = B2[c + 1], bc = B2[c + nx], bd = B2[c + nx + 1];
        T bx = w.wLL * (T)ba.x, by = w.wLL * (T)ba.y;
        bx += w.wUL * (T)bb.x;
        by += w.wUL * (T)bb.y;
        bx += w.wLU * (T)bc.x;
        by += w.wLU * (T)bc.y;
        bx += w.wUU * (T)bd.x;
        by += w.wUU * (T)bd.y;
        const T bz = T(0);
        boris_rotate<T>(vx, vy, vz, k * bx, k * by, k * bz, k * ex, k * ey);
        x += dt * vx;
        y += dt * vy;
        return;
    }
    if (Eil) gather_e_il<T>(w, nx, Eil, ex, ey);
    else gather_e<T>(w, nx, np, E, ex, ey);
    vx += k * ex;
    vy += k * ey;
    if (TYPEB != 0) {
        T t0, t1, t2;
        if (TYPEB == 1) {
            t0 = k * b0;
            t1 = k * b1;
            t2 = k * b2;
        } else {
            T bx = w.wLL * (T)B[c], by = w.wLL * (T)B[c + np], bz = w.wLL * (T)B[c + 2 * np];
            bx += w.wUL * (T)B[c + 1];
            by += w.wUL * (T)B[c + 1 + np];
            bz += w.wUL * (T)B[c + 1 + 2 * np];
            bx += w.wLU * (T)B[c + nx];
            by += w.wLU * (T)B[c + nx + np];
            bz += w.wLU * (T)B[c + nx + 2 * np];
            bx += w.wUU * (T)B[c + nx + 1];
            by += w.wUU * (T)B[c + nx + 1 + np];
            bz += w.wUU * (T)B[c + nx + 1 + 2 * np];
            t0 = k * bx;
            t1 = k * by;
            t2 = k * bz;
        }
        boris_rotate<T>(vx, vy, vz, t0, t1, t2, k * ex, k * ey);
    }
    x += dt * vx;
    y += dt * vy;
}

// ---- FP32 arithmetic for FP32 particles (the step's mixed push) ------------------------------
// Bilinear location of (x, y) in FP32: fx = x / dx + ox with ox = lxHalfExt / dx (one fused
// multiply-add), the cell clamped into the grid as grid_weights does, the fractions left in
// (fx, fy).
struct Loc32 {
    uint32_t ix, iy;
    float fx, fy;
};

__device__ __forceinline__ Loc32 locate32(float x, float y, float dxi, float dyi, float ox, float oy, uint32_t nx,
                                          uint32_t ny)
{
    float fx = fmaf(x, dxi, ox), fy = fmaf(y, dyi, oy);
    uint32_t ix = (uint32_t)fmaxf(fx, 0.f), iy = (uint32_t)fmaxf(fy, 0.f);
    ix = min(ix, nx - 2);
    iy = min(iy, ny - 2);
    Loc32 l;
    l.ix = ix;
    l.iy = iy;
    l.fx = fx - (float)ix;
    l.fy = fy - (float)iy;
    return l;
}

// Bilinear interpolation of a two-component node field stored as float2 per node, at cell
// (row pointers r0 = field + row iy, r1 = the next row) with corner weights wLL..wUU.
__device__ __forceinline__ float2 bilinear2(const float2 *__restrict__ r0, const float2 *__restrict__ r1, uint32_t ix,
                                           float wLL, float wUL, float wLU, float wUU)
{
    const float2 a = r0[ix], b = r0[ix + 1], c = r1[ix], d = r1[ix + 1];
    float2 v;
    v.x = fmaf(wUU, d.x, fmaf(wLU, c.x, fmaf(wUL, b.x, wLL * a.x)));
    v.y = fmaf(wUU, d.y, fmaf(wLU, c.y, fmaf(wUL, b.y, wLL * a.y)));
    return v;
}

// The same from a buffer resource: the corners at byte offsets oc, oc + 8 (row iy) and oc + nx8,
// oc + nx8 + 8 (row iy + 1).
__device__ __forceinline__ float2 bilinear2_buf(BufRes r, uint32_t oc, uint32_t nx8, float wLL, float wUL, float wLU,
                                               float wUU)
{
    const float2 a = buf_ld_f32x2(r, oc), b = buf_ld_f32x2(r, oc + 8u);
    const float2 c = buf_ld_f32x2(r, oc, nx8), d = buf_ld_f32x2(r, oc + 8u, nx8);
    float2 v;
    v.x = fmaf(wUU, d.x, fmaf(wLU, c.x, fmaf(wUL, b.x, wLL * a.x)));
    v.y = fmaf(wUU, d.y, fmaf(wLU, c.y, fmaf(wUL, b.y, wLL * a.y)));
    return v;
}

// Boris rotation + second half kick (particles.pyx:272-321) in FP32 with fused multiply-adds,
// t = (t0, t1, t2), from the velocity after the first half kick; 1 / (1 + |t|^2) from the
// hardware reciprocal (1 ulp).
__device__ __forceinline__ void boris_rotate32(float &vx, float &vy, float &vz, float t0, float t1, float t2,
                                               float kex, float key)
{
    const float vs0 = fmaf(-vz, t1, fmaf(vy, t2, vx));
    const float vs1 = fmaf(-vx, t2, fmaf(vz, t0, vy));
    const float vs2 = fmaf(-vy, t0, fmaf(vx, t1, vz));
    const float f = 2.f * __builtin_amdgcn_rcpf(fmaf(t2, t2, fmaf(t1, t1, fmaf(t0, t0, 1.f))));
    const float s0 = t0 * f, s1 = t1 * f, s2 = t2 * f;
    vx = fmaf(-vs2, s1, fmaf(vs1, s2, vx)) + kex;
    vy = fmaf(-vs0, s2, fmaf(vs2, s0, vy)) + key;
    vz = fmaf(-vs1, s0, fmaf(vs0, s1, vz));
}

// Boris rotation with Bz = 0 (t2 = 0): the same as boris_rotate32 without the t2 terms.
__device__ __forceinline__ void boris_rotate32_bz0(float &vx, float &vy, float &vz, float t0, float t1, float kex,
                                                   float key)
{
    const float vs0 = fmaf(-vz, t1, vx);
    const float vs1 = fmaf(vz, t0, vy);
    const float vs2 = fmaf(-vy, t0, fmaf(vx, t1, vz));
    const float f = 2.f * __builtin_amdgcn_rcpf(fmaf(t1, t1, fmaf(t0, t0, 1.f)));
    const float s0 = t0 * f, s1 = t1 * f;
    vx = fmaf(-vs2, s1, vx) + kex;
    vy = fmaf(vs2, s0, vy) + key;
    vz = fmaf(-vs1, s0, fmaf(vs0, s1, vz));
}
