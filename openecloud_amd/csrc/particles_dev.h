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

// Sums each run of equal cells among consecutive lanes of the wave (an inclusive suffix
// scan that stops at run ends, log2(64) shuffle steps); the first lane of a run receives
// the run's totals and returns true.  Lanes with c == 0xFFFFFFFF never return true.  Atomics
// to one address from many lanes of a wave are serialised by the hardware, so summing the
// runs first is what makes the deposit of cell-sorted particles cheap.
__device__ __forceinline__ bool wave_run_sum(uint32_t c, double &q0, double &q1, double &q2, double &q3)
{
    const int lane = threadIdx.x & 63;
    const uint32_t cn = __shfl_down(c, 1, 64), cp = __shfl_up(c, 1, 64);
    bool tail = lane == 63 || cn != c;
    const bool head = lane == 0 || cp != c;
    // After the step with offset d, a lane holds the sum over [lane, min(lane + 2d - 1, run end)]
    // and tail says whether that range reaches the run end; a lane that has not reached it
    // lies at least d below the run end, so lane + d is in the same run and in range.
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const double o0 = __shfl_down(q0, d, 64), o1 = __shfl_down(q1, d, 64);
        const double o2 = __shfl_down(q2, d, 64), o3 = __shfl_down(q3, d, 64);
        const bool ot = __shfl_down((int)tail, d, 64) != 0;
        if (!tail) {
            q0 += o0;
            q1 += o1;
            q2 += o2;
            q3 += o3;
            tail = ot;
        }
    }
    return head && c != 0xFFFFFFFFu;
}

// Block-cooperative deposit for 256-thread blocks; every thread of the block must call it
// (valid = false when the thread has no particle), with the particle's cell (ix, iy) and
// its four corner charges.  Runs of equal cells within a wave are summed first
// (wave_run_sum).  The run totals are then summed in an LDS tile covering the bounding box
// of the block's cells (nodes [X0, X1+1] x [Y0, Y1+1]), and each touched node of the tile
// is added to rho with one global atomic.  Particles sorted by cell, or only locally in
// order (a few cells of drift since the last sort), keep the box small; a block whose box
// exceeds WN nodes adds its run totals to rho directly.
template <int WN>
__device__ __forceinline__ void block_deposit(double *__restrict__ rho, uint32_t nx, bool valid, uint32_t ix,
                                              uint32_t iy, double q00, double q10, double q01, double q11)
{
    __shared__ double win[WN];
    __shared__ uint32_t bb[4][4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t c = valid ? ix + iy * nx : 0xFFFFFFFFu;
    if (!valid) q00 = q10 = q01 = q11 = 0.;
    const bool head = wave_run_sum(c, q00, q10, q01, q11);
    uint32_t x0 = head ? ix : 0xFFFFFFFFu, x1 = head ? ix : 0u;
    uint32_t y0 = head ? iy : 0xFFFFFFFFu, y1 = head ? iy : 0u;
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        x0 = min(x0, (uint32_t)__shfl_xor(x0, d, 64));
        x1 = max(x1, (uint32_t)__shfl_xor(x1, d, 64));
        y0 = min(y0, (uint32_t)__shfl_xor(y0, d, 64));
        y1 = max(y1, (uint32_t)__shfl_xor(y1, d, 64));
    }
    if (lane == 0) {
        bb[wave][0] = x0;
        bb[wave][1] = x1;
        bb[wave][2] = y0;
        bb[wave][3] = y1;
    }
    __syncthreads();
    uint32_t X0 = bb[0][0], X1 = bb[0][1], Y0 = bb[0][2], Y1 = bb[0][3];
#pragma unroll
    for (int w = 1; w < 4; ++w) {
        X0 = min(X0, bb[w][0]);
        X1 = max(X1, bb[w][1]);
        Y0 = min(Y0, bb[w][2]);
        Y1 = max(Y1, bb[w][3]);
    }
    if (X0 == 0xFFFFFFFFu) return;  // no particle to deposit in this block (block-uniform)
    const uint32_t wx = X1 - X0 + 2, wy = Y1 - Y0 + 2;
    if ((uint64_t)wx * wy > WN) {   // block-uniform
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
    for (uint32_t k = tid; k < nw; k += 256) win[k] = 0.;
    __syncthreads();
    if (head) {
        const uint32_t o = (ix - X0) + (iy - Y0) * wx;
        atomicAdd(&win[o], q00);
        atomicAdd(&win[o + 1], q10);
        atomicAdd(&win[o + wx], q01);
        atomicAdd(&win[o + wx + 1], q11);
    }
    __syncthreads();
    for (uint32_t k = tid; k < nw; k += 256) {
        const double v = win[k];
        if (v != 0.) {
            const uint32_t r = k / wx;
            unsafeAtomicAdd(rho + (X0 + (k - r * wx)) + (size_t)(Y0 + r) * nx, v);
        }
    }
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

// Gather E (and B for TYPEB 2) at the particle and apply the Boris push (particles.pyx:
// 272-333 with _fieldToParticles's accumulation order).  Positions advance in x and y only.
template <typename T, int TYPEB>
__device__ __forceinline__ void gather_push(const Weights<T> &w, uint32_t nx, uint32_t np,
                                            const double *__restrict__ E,
                                            const double *__restrict__ B, T b0, T b1, T b2, T dt,
                                            T k, T &x, T &y, T &vx, T &vy, T &vz)
{
    const uint32_t c = w.cell;
    T ex = w.wLL * (T)E[c];
    T ey = w.wLL * (T)E[c + np];
    ex += w.wUL * (T)E[c + 1];
    ey += w.wUL * (T)E[c + 1 + np];
    ex += w.wLU * (T)E[c + nx];
    ey += w.wLU * (T)E[c + nx + np];
    ex += w.wUU * (T)E[c + nx + 1];
    ey += w.wUU * (T)E[c + nx + 1 + np];
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
