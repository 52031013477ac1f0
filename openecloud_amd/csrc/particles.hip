// Particle-side kernels: bilinear weights, charge deposition, field gather, Boris push,
// and the fused gather + push + wall test + deposit kernel of the plain step.
//
// Arithmetic follows the reference loops term for term (particles.pyx:272-418) and the
// library is built with -ffp-contract=off, so the FP64 path is bit-identical to the
// reference's own FP64 arithmetic for the same inputs.
#include "common.h"
#include "particles_dev.h"
#include "wall.h"

namespace {

template <typename T>
__global__ void k_calc_weights(oecl_grid_t g, Soa<T> p, uint32_t n, uint32_t *inCell, T *w4,
                               uint64_t ws)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // FP64 arithmetic (particles.pyx:400-418), weights stored in the particle dtype
    Weights<double> w = grid_weights<double>((double)p(0, i), (double)p(1, i), 1. / g.dx, 1. / g.dy,
                                             g.lxHalfExt, g.lyHalfExt, g.nx, g.ny);
    inCell[i] = w.cell;
    w4[i] = (T)w.wLL;
    w4[ws + i] = (T)w.wUL;
    w4[2 * ws + i] = (T)w.wLU;
    w4[3 * ws + i] = (T)w.wUU;
}

__global__ void k_zero(double *a, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.;
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
    // FP64 weights of the stored position, as the push deposits (FP32 storage: in-cell
    // coordinates of a 2048-cell grid keep only ~6e-5 of a cell in FP32)
    Weights<double> w = grid_weights<double>((double)p(0, i), (double)p(1, i), 1. / g.dx, 1. / g.dy,
                                             g.lxHalfExt, g.lyHalfExt, g.nx, g.ny);
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
    Weights<double> w = grid_weights<double>((double)r[0], (double)r[1], 1. / g.dx, 1. / g.dy, g.lxHalfExt,
                                             g.lyHalfExt, g.nx, g.ny);
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
                                                    const double *__restrict__ B, double b0,
                                                    double b1, double b2, double dt, double k,
                                                    double q,
                                                    uint8_t *__restrict__ isInside,
                                                    double *__restrict__ rhoNext)
{
    __shared__ double tiles[4][512];
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < n;
    // FP32 storage is widened on load; the arithmetic runs in FP64 registers.
    using C = double;
    const C dxi = 1. / g.dx, dyi = 1. / g.dy, lxh = g.lxHalfExt, lyh = g.lyHalfExt;
    bool dep = false;
    Weights<C> w2{};
    double wq = 0.;
    if (live) {
        C x = p(0, i), y = p(1, i), vx = p(2, i), vy = p(3, i), vz = p(4, i);
        Weights<C> w = grid_weights<C>(x, y, dxi, dyi, lxh, lyh, g.nx, g.ny);
        gather_push<C, TYPEB>(w, g.nx, g.np, E, B, b0, b1, b2, dt, k, x, y, vx, vy, vz);
        x = (C)(T)x;  // the wall test sees the position as stored (FP32 storage: rounded)
        y = (C)(T)y;
        uint8_t in = wall_inside<C>(wall, x, y);  // may clip x, y (cut-cell wall)
        p(0, i) = (T)x;
        p(1, i) = (T)y;
        p(2, i) = (T)vx;
        p(3, i) = (T)vy;
        if (TYPEB != 0) p(4, i) = (T)vz;
        isInside[i] = in;
        if (rhoNext != nullptr && in) {
            dep = true;
            w2 = grid_weights<C>((C)(T)x, (C)(T)y, dxi, dyi, lxh, lyh, g.nx, g.ny);
            wq = (double)p(5, i);
        }
    }
    if (rhoNext == nullptr) return;  // uniform: every lane of a wave reaches the deposit
    wave_deposit<512>(rhoNext, g.nx, dep, w2.ix, w2.iy, w2.wLL * wq * q, w2.wUL * wq * q, w2.wLU * wq * q,
                      w2.wUU * wq * q, tiles[threadIdx.x >> 6]);
}

// Moments of the particle set, FP64, fixed reduction order (bit-reproducible):
//   m[0] = sum w, m[1] = sum w |v|^2, m[2] = sum w^2 |v|^2, m[3] = sum w x, m[4] = sum w y,
//   m[5] = count.  Block partials first, then one block sums them.
constexpr int kMomThreads = 256;
constexpr int kMom = 6;

template <typename T>
__global__ void __launch_bounds__(kMomThreads)
    k_moments_partial(Soa<T> p, uint32_t n, double *__restrict__ partials)
{
    __shared__ double lds[kMomThreads / 64][kMom];
    double v[kMom] = {0., 0., 0., 0., 0., 0.};
    for (uint32_t i = blockIdx.x * kMomThreads + threadIdx.x; i < n; i += gridDim.x * kMomThreads) {
        double w = p(5, i), vx = p(2, i), vy = p(3, i), vz = p(4, i);
        double v2 = vx * vx + vy * vy + vz * vz;
        v[0] += w;
        v[1] += w * v2;
        v[2] += w * w * v2;
        v[3] += w * (double)p(0, i);
        v[4] += w * (double)p(1, i);
        v[5] += 1.;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < kMom; ++k) {
        double x = v[k];
        for (int d = 32; d > 0; d >>= 1) x += __shfl_down(x, d, 64);
        if (lane == 0) lds[wave][k] = x;
    }
    __syncthreads();
    if (threadIdx.x < kMom) {
        double s = 0.;
        for (int w = 0; w < kMomThreads / 64; ++w) s += lds[w][threadIdx.x];
        partials[blockIdx.x * kMom + threadIdx.x] = s;
    }
}

__global__ void k_moments_final(const double *__restrict__ partials, uint32_t nb, double *__restrict__ out)
{
    if (threadIdx.x < kMom) {
        double s = 0.;
        for (uint32_t b = 0; b < nb; ++b) s += partials[b * kMom + threadIdx.x];
        out[threadIdx.x] = s;
    }
}

}  // namespace

OECL_EXPORT int oecl_particle_moments(int dtype, const void *data, uint64_t stride, uint32_t n,
                                      double *out, double *work, void *stream)
{
    if (!out || !work || (n && !data)) return OECL_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    const uint32_t nb = 512;
    if (dtype == OECL_F64)
        hipLaunchKernelGGL(k_moments_partial<double>, dim3(nb), dim3(kMomThreads), 0, s,
                           Soa<double>{(double *)data, stride}, n, work);
    else if (dtype == OECL_F32)
        hipLaunchKernelGGL(k_moments_partial<float>, dim3(nb), dim3(kMomThreads), 0, s,
                           Soa<float>{(float *)data, stride}, n, work);
    else
        return OECL_ERR_ARG;
    hipLaunchKernelGGL(k_moments_final, dim3(1), dim3(64), 0, s, work, nb, out);
    OECL_CHECK_LAUNCH("oecl_particle_moments");
    return OECL_OK;
}

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
    double b0 = typeB == 1 ? bConst[0] : 0., b1 = typeB == 1 ? bConst[1] : 0.,
           b2 = typeB == 1 ? bConst[2] : 0.;
    Soa<T> p{(T *)data, stride};
    dim3 gr(oecl_blocks(n, 256)), bl(256);
    if (typeB == 0)
        hipLaunchKernelGGL((k_push_fused<T, 0>), gr, bl, 0, s, *g, wd, p, n, E, B, b0, b1, b2,
                           dt, k, q, isInside, rhoNext);
    else if (typeB == 1)
        hipLaunchKernelGGL((k_push_fused<T, 1>), gr, bl, 0, s, *g, wd, p, n, E, B, b0, b1, b2,
                           dt, k, q, isInside, rhoNext);
    else
        hipLaunchKernelGGL((k_push_fused<T, 2>), gr, bl, 0, s, *g, wd, p, n, E, B, b0, b1, b2,
                           dt, k, q, isInside, rhoNext);
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
