// Particle-side kernels: bilinear weights, charge deposition, field gather, Boris push,
// and the fused gather + push + wall test + deposit kernel of the plain step.
//
// Arithmetic follows the reference loops term for term (particles.pyx:272-418) and the
// library is built with -ffp-contract=off, so the FP64 path is bit-identical to the
// reference's own FP64 arithmetic for the same inputs.
#include "common.h"
#include "wall.h"

namespace {

// particles.pyx:400-418 for one particle, with the cell index clamped into the grid so an
// escaped particle can never address memory outside the node arrays (the reference has no
// bounds check there; inside the grid the result is unchanged).
template <typename T>
struct Weights {
    uint32_t cell;
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
    w.wLL = (1 - fx) * (1 - fy);
    w.wUL = fx * (1 - fy);
    w.wLU = (1 - fx) * fy;
    w.wUU = fx * fy;
    return w;
}

template <typename T>
__global__ void k_calc_weights(oecl_grid_t g, Soa<T> p, uint32_t n, uint32_t *inCell, T *w4,
                               uint64_t ws)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Weights<T> w = grid_weights<T>(p(0, i), p(1, i), (T)(1. / g.dx), (T)(1. / g.dy),
                                   (T)g.lxHalfExt, (T)g.lyHalfExt, g.nx, g.ny);
    inCell[i] = w.cell;
    w4[i] = w.wLL;
    w4[ws + i] = w.wUL;
    w4[2 * ws + i] = w.wLU;
    w4[3 * ws + i] = w.wUU;
}

__global__ void k_zero(double *a, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.;
}

// Four FP64 atomic adds per particle: wk * weight * q as in particles.pyx:344-351.
template <typename T>
__device__ __forceinline__ void deposit4(double *rho, uint32_t c, uint32_t nx, double wLL,
                                         double wUL, double wLU, double wUU, double wq)
{
    unsafeAtomicAdd(rho + c, wLL * wq);
    unsafeAtomicAdd(rho + c + 1, wUL * wq);
    unsafeAtomicAdd(rho + c + nx, wLU * wq);
    unsafeAtomicAdd(rho + c + nx + 1, wUU * wq);
}

template <typename T>
__global__ void k_charge_to_grid(Soa<T> p, uint32_t n, const uint32_t *inCell, const T *w4,
                                 uint64_t ws, double q, uint32_t nx, double *rho)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double wt = (double)p(5, i);
    uint32_t c = inCell[i];
    // Reference order: (w_k * weight) * q.
    unsafeAtomicAdd(rho + c, (double)w4[i] * wt * q);
    unsafeAtomicAdd(rho + c + 1, (double)w4[ws + i] * wt * q);
    unsafeAtomicAdd(rho + c + nx, (double)w4[2 * ws + i] * wt * q);
    unsafeAtomicAdd(rho + c + nx + 1, (double)w4[3 * ws + i] * wt * q);
}

template <typename T>
__global__ void k_deposit_add(oecl_grid_t g, Soa<T> p, uint32_t n, double q, double *rho)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Weights<T> w = grid_weights<T>(p(0, i), p(1, i), (T)(1. / g.dx), (T)(1. / g.dy),
                                   (T)g.lxHalfExt, (T)g.lyHalfExt, g.nx, g.ny);
    double wt = (double)p(5, i);
    unsafeAtomicAdd(rho + w.cell, (double)w.wLL * wt * q);
    unsafeAtomicAdd(rho + w.cell + 1, (double)w.wUL * wt * q);
    unsafeAtomicAdd(rho + w.cell + g.nx, (double)w.wLU * wt * q);
    unsafeAtomicAdd(rho + w.cell + g.nx + 1, (double)w.wUU * wt * q);
}

// Same deposit for an AoS block [n][6] (the emitter's secondaries).
template <typename T>
__global__ void k_deposit_add_aos(oecl_grid_t g, const T *__restrict__ a, uint32_t n, double q,
                                  double *rho)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T *r = a + 6 * (uint64_t)i;
    Weights<T> w = grid_weights<T>(r[0], r[1], (T)(1. / g.dx), (T)(1. / g.dy), (T)g.lxHalfExt,
                                   (T)g.lyHalfExt, g.nx, g.ny);
    double wt = (double)r[5];
    unsafeAtomicAdd(rho + w.cell, (double)w.wLL * wt * q);
    unsafeAtomicAdd(rho + w.cell + 1, (double)w.wUL * wt * q);
    unsafeAtomicAdd(rho + w.cell + g.nx, (double)w.wLU * wt * q);
    unsafeAtomicAdd(rho + w.cell + g.nx + 1, (double)w.wUU * wt * q);
}

// particles.pyx:354-397: the first corner assigns, the other three accumulate, in order.
template <typename T, int DIMS>
__global__ void k_field_to_particles(const double *f, const uint32_t *inCell, const T *w4,
                                     uint64_t ws, uint32_t n, uint32_t nx, uint32_t np, T *out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t c = inCell[i];
    T a = w4[i], b = w4[ws + i], cc = w4[2 * ws + i], d = w4[3 * ws + i];
#pragma unroll
    for (int k = 0; k < DIMS; ++k) {
        const double *fk = f + (uint64_t)k * np;
        T v = a * (T)fk[c];
        v += b * (T)fk[c + 1];
        v += cc * (T)fk[c + nx];
        v += d * (T)fk[c + nx + 1];
        out[(uint64_t)DIMS * i + k] = v;
    }
}

// Boris rotation + second half kick + drift of particles.pyx:272-321, from the velocity
// after the first half kick.
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

template <typename T>
__global__ void k_boris_push(Soa<T> p, uint32_t n, const T *eAt, const T *bAt, T b0, T b1,
                             T b2, int typeB, T dt, T k)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T vx = p(2, i), vy = p(3, i), vz = p(4, i);
    T ex = eAt[2 * (uint64_t)i], ey = eAt[2 * (uint64_t)i + 1];
    if (typeB == 0) {
        vx += k * ex;
        vy += k * ey;
    } else {
        vx += k * ex;
        vy += k * ey;
        T t0, t1, t2;
        if (typeB == 1) {
            t0 = k * b0; t1 = k * b1; t2 = k * b2;
        } else {
            t0 = k * bAt[3 * (uint64_t)i];
            t1 = k * bAt[3 * (uint64_t)i + 1];
            t2 = k * bAt[3 * (uint64_t)i + 2];
        }
        boris_rotate<T>(vx, vy, vz, t0, t1, t2, k * ex, k * ey);
        p(4, i) = vz;
    }
    p(2, i) = vx;
    p(3, i) = vy;
    p(0, i) += dt * vx;
    p(1, i) += dt * vy;
}

// One fused particle step.  The gather reproduces _fieldToParticles's accumulation order,
// so the pushed state is identical to the unfused kernels.
template <typename T, int TYPEB>
__global__ void __launch_bounds__(256) k_push_fused(oecl_grid_t g, WallDev wall, Soa<T> p,
                                                    uint32_t n, const double *__restrict__ E,
                                                    const double *__restrict__ B, T b0, T b1,
                                                    T b2, T dt, T k, double q,
                                                    uint8_t *__restrict__ isInside,
                                                    double *__restrict__ rhoNext)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T dxi = (T)(1. / g.dx), dyi = (T)(1. / g.dy);
    const T lxh = (T)g.lxHalfExt, lyh = (T)g.lyHalfExt;
    T x = p(0, i), y = p(1, i), vx = p(2, i), vy = p(3, i), vz = p(4, i);
    Weights<T> w = grid_weights<T>(x, y, dxi, dyi, lxh, lyh, g.nx, g.ny);
    const uint32_t c = w.cell, nx = g.nx, np = g.np;
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
            t0 = k * b0; t1 = k * b1; t2 = k * b2;
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
            t0 = k * bx; t1 = k * by; t2 = k * bz;
        }
        boris_rotate<T>(vx, vy, vz, t0, t1, t2, k * ex, k * ey);
        p(4, i) = vz;
    }
    x += dt * vx;
    y += dt * vy;
    uint8_t in = wall_inside<T>(wall, x, y);  // may clip x, y (cut-cell wall)
    p(0, i) = x;
    p(1, i) = y;
    p(2, i) = vx;
    p(3, i) = vy;
    isInside[i] = in;
    if (rhoNext != nullptr && in) {
        Weights<T> w2 = grid_weights<T>(x, y, dxi, dyi, lxh, lyh, g.nx, g.ny);
        double wt = (double)p(5, i);
        unsafeAtomicAdd(rhoNext + w2.cell, (double)w2.wLL * wt * q);
        unsafeAtomicAdd(rhoNext + w2.cell + 1, (double)w2.wUL * wt * q);
        unsafeAtomicAdd(rhoNext + w2.cell + nx, (double)w2.wLU * wt * q);
        unsafeAtomicAdd(rhoNext + w2.cell + nx + 1, (double)w2.wUU * wt * q);
    }
}

}  // namespace

OECL_EXPORT int oecl_calc_grid_weights(const oecl_grid_t *g, int dtype, const void *data,
                                       uint64_t stride, uint32_t n, uint32_t *inCell, void *w4,
                                       uint64_t wstride, void *stream)
{
    if (!g || !data || !inCell || !w4 || g->nx < 2 || g->ny < 2) return OECL_ERR_ARG;
    if (n == 0) return OECL_OK;
    hipStream_t s = (hipStream_t)stream;
    if (dtype == OECL_F64)
        hipLaunchKernelGGL(k_calc_weights<double>, dim3(oecl_blocks(n, 256)), dim3(256), 0, s, *g,
                           Soa<double>{(double *)data, stride}, n, inCell, (double *)w4, wstride);
    else if (dtype == OECL_F32)
        hipLaunchKernelGGL(k_calc_weights<float>, dim3(oecl_blocks(n, 256)), dim3(256), 0, s, *g,
                           Soa<float>{(float *)data, stride}, n, inCell, (float *)w4, wstride);
    else
        return OECL_ERR_ARG;
    OECL_CHECK_LAUNCH("oecl_calc_grid_weights");
    return OECL_OK;
}

OECL_EXPORT int oecl_charge_to_grid(const oecl_grid_t *g, int dtype, const void *data,
                                    uint64_t stride, uint32_t n, const uint32_t *inCell,
                                    const void *w4, uint64_t wstride, double q, double *rho,
                                    void *stream)
{
    if (!g || !rho) return OECL_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_zero, dim3(oecl_blocks(g->np, 256)), dim3(256), 0, s, rho, g->np);
    if (n > 0) {
        if (!data || !inCell || !w4) return OECL_ERR_ARG;
        if (dtype == OECL_F64)
            hipLaunchKernelGGL(k_charge_to_grid<double>, dim3(oecl_blocks(n, 256)), dim3(256), 0,
                               s, Soa<double>{(double *)data, stride}, n, inCell,
                               (const double *)w4, wstride, q, g->nx, rho);
        else if (dtype == OECL_F32)
            hipLaunchKernelGGL(k_charge_to_grid<float>, dim3(oecl_blocks(n, 256)), dim3(256), 0, s,
                               Soa<float>{(float *)data, stride}, n, inCell, (const float *)w4,
                               wstride, q, g->nx, rho);
        else
            return OECL_ERR_ARG;
    }
    OECL_CHECK_LAUNCH("oecl_charge_to_grid");
    return OECL_OK;
}

OECL_EXPORT int oecl_deposit_add(const oecl_grid_t *g, int dtype, const void *data,
                                 uint64_t stride, uint32_t n, double q, double *rho, void *stream)
{
    if (!g || !rho) return OECL_ERR_ARG;
    if (n == 0) return OECL_OK;
    if (!data) return OECL_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    if (dtype == OECL_F64)
        hipLaunchKernelGGL(k_deposit_add<double>, dim3(oecl_blocks(n, 256)), dim3(256), 0, s, *g,
                           Soa<double>{(double *)data, stride}, n, q, rho);
    else if (dtype == OECL_F32)
        hipLaunchKernelGGL(k_deposit_add<float>, dim3(oecl_blocks(n, 256)), dim3(256), 0, s, *g,
                           Soa<float>{(float *)data, stride}, n, q, rho);
    else
        return OECL_ERR_ARG;
    OECL_CHECK_LAUNCH("oecl_deposit_add");
    return OECL_OK;
}

OECL_EXPORT int oecl_deposit_add_aos(const oecl_grid_t *g, int dtype, const void *aos, uint32_t n,
                                     double q, double *rho, void *stream)
{
    if (!g || !rho) return OECL_ERR_ARG;
    if (n == 0) return OECL_OK;
    if (!aos) return OECL_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    if (dtype == OECL_F64)
        hipLaunchKernelGGL(k_deposit_add_aos<double>, dim3(oecl_blocks(n, 256)), dim3(256), 0, s, *g,
                           (const double *)aos, n, q, rho);
    else if (dtype == OECL_F32)
        hipLaunchKernelGGL(k_deposit_add_aos<float>, dim3(oecl_blocks(n, 256)), dim3(256), 0, s, *g,
                           (const float *)aos, n, q, rho);
    else
        return OECL_ERR_ARG;
    OECL_CHECK_LAUNCH("oecl_deposit_add_aos");
    return OECL_OK;
}

OECL_EXPORT int oecl_field_to_particles(const oecl_grid_t *g, int dtype, const double *field,
                                        const uint32_t *inCell, const void *w4, uint64_t wstride,
                                        uint32_t n, int dims, void *out, void *stream)
{
    if (!g || !field || (dims != 2 && dims != 3)) return OECL_ERR_ARG;
    if (n == 0) return OECL_OK;
    if (!inCell || !w4 || !out) return OECL_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    dim3 gr(oecl_blocks(n, 256)), bl(256);
    if (dtype == OECL_F64) {
        if (dims == 2)
            hipLaunchKernelGGL((k_field_to_particles<double, 2>), gr, bl, 0, s, field, inCell,
                               (const double *)w4, wstride, n, g->nx, g->np, (double *)out);
        else
            hipLaunchKernelGGL((k_field_to_particles<double, 3>), gr, bl, 0, s, field, inCell,
                               (const double *)w4, wstride, n, g->nx, g->np, (double *)out);
    } else if (dtype == OECL_F32) {
        if (dims == 2)
            hipLaunchKernelGGL((k_field_to_particles<float, 2>), gr, bl, 0, s, field, inCell,
                               (const float *)w4, wstride, n, g->nx, g->np, (float *)out);
        else
            hipLaunchKernelGGL((k_field_to_particles<float, 3>), gr, bl, 0, s, field, inCell,
                               (const float *)w4, wstride, n, g->nx, g->np, (float *)out);
    } else {
        return OECL_ERR_ARG;
    }
    OECL_CHECK_LAUNCH("oecl_field_to_particles");
    return OECL_OK;
}

OECL_EXPORT int oecl_boris_push(int dtype, void *data, uint64_t stride, uint32_t n,
                                const void *eAt, const void *bAt, const double *bConst,
                                int typeB, double dt, double qm, void *stream)
{
    if (!data || !eAt || typeB < 0 || typeB > 2) return OECL_ERR_ARG;
    if (typeB == 1 && !bConst) return OECL_ERR_ARG;
    if (typeB == 2 && !bAt) return OECL_ERR_ARG;
    if (n == 0) return OECL_OK;
    hipStream_t s = (hipStream_t)stream;
    // particles.pyx:84 (q/m*dt, full kick without B) and :87/:91 (0.5*q/m*dt).
    double k = typeB == 0 ? qm * dt : 0.5 * qm * dt;
    double b0 = typeB == 1 ? bConst[0] : 0., b1 = typeB == 1 ? bConst[1] : 0.,
           b2 = typeB == 1 ? bConst[2] : 0.;
    dim3 gr(oecl_blocks(n, 256)), bl(256);
    if (dtype == OECL_F64)
        hipLaunchKernelGGL(k_boris_push<double>, gr, bl, 0, s, Soa<double>{(double *)data, stride},
                           n, (const double *)eAt, (const double *)bAt, b0, b1, b2, typeB, dt, k);
    else if (dtype == OECL_F32)
        hipLaunchKernelGGL(k_boris_push<float>, gr, bl, 0, s, Soa<float>{(float *)data, stride}, n,
                           (const float *)eAt, (const float *)bAt, (float)b0, (float)b1,
                           (float)b2, typeB, (float)dt, (float)k);
    else
        return OECL_ERR_ARG;
    OECL_CHECK_LAUNCH("oecl_boris_push");
    return OECL_OK;
}

template <typename T>
static int launch_fused(const oecl_grid_t *g, const WallDev &wd, void *data, uint64_t stride,
                        uint32_t n, const double *E, const double *B, const double *bConst,
                        int typeB, double dt, double qm, double q, uint8_t *isInside,
                        double *rhoNext, hipStream_t s)
{
    double k = typeB == 0 ? qm * dt : 0.5 * qm * dt;
    T b0 = typeB == 1 ? (T)bConst[0] : T(0), b1 = typeB == 1 ? (T)bConst[1] : T(0),
      b2 = typeB == 1 ? (T)bConst[2] : T(0);
    Soa<T> p{(T *)data, stride};
    dim3 gr(oecl_blocks(n, 256)), bl(256);
    if (typeB == 0)
        hipLaunchKernelGGL((k_push_fused<T, 0>), gr, bl, 0, s, *g, wd, p, n, E, B, b0, b1, b2,
                           (T)dt, (T)k, q, isInside, rhoNext);
    else if (typeB == 1)
        hipLaunchKernelGGL((k_push_fused<T, 1>), gr, bl, 0, s, *g, wd, p, n, E, B, b0, b1, b2,
                           (T)dt, (T)k, q, isInside, rhoNext);
    else
        hipLaunchKernelGGL((k_push_fused<T, 2>), gr, bl, 0, s, *g, wd, p, n, E, B, b0, b1, b2,
                           (T)dt, (T)k, q, isInside, rhoNext);
    OECL_CHECK_LAUNCH("oecl_push_fused");
    return OECL_OK;
}

OECL_EXPORT int oecl_push_fused(const oecl_grid_t *g, const oecl_wall_t *wall, int dtype,
                                void *data, uint64_t stride, uint32_t n, const double *E,
                                const double *bField, const double *bConst, int typeB, double dt,
                                double qm, double q, uint8_t *isInside, double *rhoNext,
                                void *stream)
{
    if (!g || !wall || !data || !E || !isInside || typeB < 0 || typeB > 2) return OECL_ERR_ARG;
    if (typeB == 1 && !bConst) return OECL_ERR_ARG;
    if (typeB == 2 && !bField) return OECL_ERR_ARG;
    if (n == 0) return OECL_OK;
    WallDev wd;
    int rc = wall_to_device(wall, &wd);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    if (dtype == OECL_F64)
        return launch_fused<double>(g, wd, data, stride, n, E, bField, bConst, typeB, dt, qm, q,
                                    isInside, rhoNext, s);
    if (dtype == OECL_F32)
        return launch_fused<float>(g, wd, data, stride, n, E, bField, bConst, typeB, dt, qm, q,
                                   isInside, rhoNext, s);
    return OECL_ERR_ARG;
}
