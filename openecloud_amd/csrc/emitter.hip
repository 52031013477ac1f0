// Furman-Pivi secondary emission (particleEmitter.pyx:154-308), the two simpler emitters
// (SecElecEmitter, SimpleSecEmitter: particleEmitter.pyx:316-596, drawn as in secem_dev.h), the
// uniform loader (particleEmitter.pyx:41-72) and the photoelectron source.
//
// Furman random numbers: either an injected FP64 table [nAbs][35] or Philox4x32-10 with counter
// (impact, slot/2, step, purpose).  The slot of every draw is fixed per impact:
//   0 nSec pick, 1 branch pick, 2-3 Box-Muller pair, 4 the further nSec == 1 draw,
//   5..13 thetaK, 14 y, 15+2i / 16+2i theta / phi of secondary i.
#include "common.h"
#include "furman_dev.h"
#include "photo_dev.h"
#include "secem_dev.h"
#include "scan.h"
#include "wall.h"

#define PURPOSE_LOADER 2u

namespace {

constexpr double kPi = 3.141592653589793;  // constants.pyx:9

template <typename T>
__global__ void k_furman_pass1(oecl_furman_t P, const T *__restrict__ ap, const T *__restrict__ nv,
                               uint32_t n0, const double *__restrict__ table, uint32_t s0,
                               uint32_t s1, uint32_t step, uint32_t *__restrict__ nSec,
                               double *__restrict__ impact)
{
    uint32_t jj = blockIdx.x * blockDim.x + threadIdx.x;
    if (jj >= n0) return;
    furman::Rng rng{table, jj, step, s0, s1};
    nSec[jj] = furman::pass1<T>(P, ap + 6 * (uint64_t)jj, nv + 2 * (uint64_t)jj, rng,
                                impact + 4 * (uint64_t)jj);
}

template <typename T>
__global__ void k_furman_pass2(oecl_furman_t P, const T *__restrict__ ap, const T *__restrict__ nv,
                               uint32_t n0, const uint32_t *__restrict__ nSec,
                               const uint32_t *__restrict__ offs, const double *__restrict__ impact,
                               const double *__restrict__ table, uint32_t s0, uint32_t s1,
                               uint32_t step, T *__restrict__ out, uint32_t secCap)
{
    uint32_t jj = blockIdx.x * blockDim.x + threadIdx.x;
    if (jj >= n0) return;
    uint32_t ns = nSec[jj];
    if (ns == 0) return;
    furman::Rng rng{table, jj, step, s0, s1};
    uint32_t base = offs[jj];
    furman::pass2<T>(P, ap + 6 * (uint64_t)jj, nv + 2 * (uint64_t)jj, ns, impact[4 * (uint64_t)jj],
                     impact + 4 * (uint64_t)(n0 - 1), rng, [&](uint32_t ii) -> T * {
                         uint32_t k = base + ii;
                         return k < secCap ? out + 6 * (uint64_t)k : nullptr;
                     });
}

// Uniform loader: particle i tries block t of its Philox stream for (x, y), t < 64.
template <typename T>
__global__ void k_homo_loader(WallDev w, Soa<T> p, uint32_t n, double weightInit, double vsigma,
                              int randomizeWeight, uint32_t s0, uint32_t s1, uint32_t *failed)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x = 0., y = 0.;
    bool ok = false;
    for (uint32_t t = 0; t < 64 && !ok; ++t) {
        Philox4 r = philox4x32_10(i, t, 0u, PURPOSE_LOADER, s0, s1);
        x = (u01_from_words(r.v[0], r.v[1]) - 0.5) * w.lx;
        y = (u01_from_words(r.v[2], r.v[3]) - 0.5) * w.ly;
        ok = wall_is_inside_point(w, x, y);
    }
    if (!ok) {
        atomicAdd(failed, 1u);
        x = 0.;
        y = 0.;
    }
    Philox4 r1 = philox4x32_10(i, 64u, 0u, PURPOSE_LOADER, s0, s1);
    Philox4 r2 = philox4x32_10(i, 65u, 0u, PURPOSE_LOADER, s0, s1);
    double a0 = u01_from_words(r1.v[0], r1.v[1]), a1 = u01_from_words(r1.v[2], r1.v[3]);
    double a2 = u01_from_words(r2.v[0], r2.v[1]), a3 = u01_from_words(r2.v[2], r2.v[3]);
    double rad0 = sqrt(-2. * log(1. - a0)), rad1 = sqrt(-2. * log(1. - a2));
    double z0 = rad0 * cos(2. * kPi * a1), z1 = rad0 * sin(2. * kPi * a1);
    double z2 = rad1 * cos(2. * kPi * a3);
    double wt = weightInit;
    if (randomizeWeight) {
        Philox4 r3 = philox4x32_10(i, 66u, 0u, PURPOSE_LOADER, s0, s1);
        wt = weightInit * (0.5 + u01_from_words(r3.v[0], r3.v[1]));
    }
    p(0, i) = (T)x;
    p(1, i) = (T)y;
    p(2, i) = (T)(z0 * vsigma);
    p(3, i) = (T)(z1 * vsigma);
    p(4, i) = (T)(z2 * vsigma);
    p(5, i) = (T)wt;
}

template <typename T>
__global__ void k_secem_pass1(oecl_secem_t P, const T *__restrict__ ap, const T *__restrict__ nv, uint32_t n0,
                              const double *__restrict__ table, uint32_t s0, uint32_t s1, uint32_t step,
                              uint32_t *__restrict__ nSec, double *__restrict__ impact)
{
    const uint32_t jj = blockIdx.x * blockDim.x + threadIdx.x;
    if (jj >= n0) return;
    const secem::Rng rng{table, jj, step, s0, s1};
    nSec[jj] = secem::pass1<T>(P, ap + 6 * (uint64_t)jj, nv + 2 * (uint64_t)jj, rng, impact + 4 * (uint64_t)jj);
}

template <typename T>
__global__ void k_secem_pass2(oecl_secem_t P, const T *__restrict__ ap, const T *__restrict__ nv, uint32_t n0,
                              const uint32_t *__restrict__ offs, const double *__restrict__ impact,
                              const double *__restrict__ table, uint32_t s0, uint32_t s1, uint32_t step,
                              T *__restrict__ out, uint32_t secCap)
{
    const uint32_t jj = blockIdx.x * blockDim.x + threadIdx.x;
    if (jj >= n0) return;
    const uint32_t base = offs[jj];
    if (offs[jj + 1] == base) return;
    const secem::Rng rng{table, jj, step, s0, s1};
    secem::pass2<T>(P, ap + 6 * (uint64_t)jj, nv + 2 * (uint64_t)jj, impact + 4 * (uint64_t)jj, rng,
                    [&](uint32_t k) -> T * { return base + k < secCap ? out + 6 * (uint64_t)(base + k) : nullptr; });
}

template <typename T>
__global__ void k_photo_emit(WallDev w, T *out, uint32_t n, uint32_t step, uint32_t s0, uint32_t s1, double e0,
                             double sigmaE, double weight, double mass)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) photo::emit_one<T>(w, k, step, s0, s1, e0, sigmaE, weight, mass, out + 6 * (uint64_t)k);
}

}  // namespace

OECL_EXPORT int oecl_photo_emit(const oecl_wall_t *wall, int dtype, void *out, uint32_t n, uint32_t step,
                                uint32_t seed0, uint32_t seed1, double e0, double sigmaE, double weight,
                                double particleMass, void *stream)
{
    if (!wall || (n && !out)) return OECL_ERR_ARG;
    if (wall->kind != OECL_WALL_RECT && wall->kind != OECL_WALL_ELLIP) {
        oecl_set_error_msg("photoelectrons: rectangular or elliptical wall only");
        return OECL_ERR_ARG;
    }
    if (n == 0) return OECL_OK;
    WallDev wd;
    int rc = wall_to_device(wall, &wd);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    if (dtype == OECL_F64)
        hipLaunchKernelGGL(k_photo_emit<double>, dim3(oecl_blocks(n, 128)), dim3(128), 0, s, wd, (double *)out, n,
                           step, seed0, seed1, e0, sigmaE, weight, particleMass);
    else if (dtype == OECL_F32)
        hipLaunchKernelGGL(k_photo_emit<float>, dim3(oecl_blocks(n, 128)), dim3(128), 0, s, wd, (float *)out, n,
                           step, seed0, seed1, e0, sigmaE, weight, particleMass);
    else
        return OECL_ERR_ARG;
    OECL_CHECK_LAUNCH("oecl_photo_emit");
    return OECL_OK;
}

OECL_EXPORT int oecl_furman_pass1(const oecl_furman_t *prm, int dtype, const void *absorbed,
                                  const void *normals, uint32_t nAbs, const double *uniforms,
                                  uint32_t seed0, uint32_t seed1, uint32_t step, uint32_t *nSec,
                                  double *impact, void *stream)
{
    if (!prm || !nSec || !impact) return OECL_ERR_ARG;
    if (nAbs == 0) return OECL_OK;
    if (!absorbed || !normals) return OECL_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    dim3 gr(oecl_blocks(nAbs, 128)), bl(128);
    if (dtype == OECL_F64)
        hipLaunchKernelGGL(k_furman_pass1<double>, gr, bl, 0, s, *prm, (const double *)absorbed,
                           (const double *)normals, nAbs, uniforms, seed0, seed1, step, nSec,
                           impact);
    else if (dtype == OECL_F32)
        hipLaunchKernelGGL(k_furman_pass1<float>, gr, bl, 0, s, *prm, (const float *)absorbed,
                           (const float *)normals, nAbs, uniforms, seed0, seed1, step, nSec,
                           impact);
    else
        return OECL_ERR_ARG;
    OECL_CHECK_LAUNCH("oecl_furman_pass1");
    return OECL_OK;
}

OECL_EXPORT int oecl_furman_pass2(const oecl_furman_t *prm, int dtype, const void *absorbed,
                                  const void *normals, uint32_t nAbs, const uint32_t *nSec,
                                  const uint32_t *offsets, const double *impact,
                                  const double *uniforms, uint32_t seed0, uint32_t seed1,
                                  uint32_t step, void *secondaries, uint32_t secCap, void *stream)
{
    if (!prm || !nSec || !offsets || !impact) return OECL_ERR_ARG;
    if (nAbs == 0 || secCap == 0) return OECL_OK;
    if (!absorbed || !normals || !secondaries) return OECL_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    dim3 gr(oecl_blocks(nAbs, 64)), bl(64);
    if (dtype == OECL_F64)
        hipLaunchKernelGGL(k_furman_pass2<double>, gr, bl, 0, s, *prm, (const double *)absorbed,
                           (const double *)normals, nAbs, nSec, offsets, impact, uniforms, seed0,
                           seed1, step, (double *)secondaries, secCap);
    else if (dtype == OECL_F32)
        hipLaunchKernelGGL(k_furman_pass2<float>, gr, bl, 0, s, *prm, (const float *)absorbed,
                           (const float *)normals, nAbs, nSec, offsets, impact, uniforms, seed0,
                           seed1, step, (float *)secondaries, secCap);
    else
        return OECL_ERR_ARG;
    OECL_CHECK_LAUNCH("oecl_furman_pass2");
    return OECL_OK;
}

OECL_EXPORT int oecl_secem_pass1(const oecl_secem_t *prm, int dtype, const void *absorbed, const void *normals,
                                 uint32_t nAbs, const double *uniforms, uint32_t seed0, uint32_t seed1,
                                 uint32_t step, uint32_t *nSec, double *impact, void *stream)
{
    if (!prm || !nSec || !impact) return OECL_ERR_ARG;
    if (prm->model != OECL_SECEM_STAINLESS && prm->model != OECL_SECEM_SIMPLE) return OECL_ERR_ARG;
    if (nAbs == 0) return OECL_OK;
    if (!absorbed || !normals) return OECL_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    const dim3 gr(oecl_blocks(nAbs, 128)), bl(128);
    if (dtype == OECL_F64)
        hipLaunchKernelGGL(k_secem_pass1<double>, gr, bl, 0, s, *prm, (const double *)absorbed,
                           (const double *)normals, nAbs, uniforms, seed0, seed1, step, nSec, impact);
    else if (dtype == OECL_F32)
        hipLaunchKernelGGL(k_secem_pass1<float>, gr, bl, 0, s, *prm, (const float *)absorbed,
                           (const float *)normals, nAbs, uniforms, seed0, seed1, step, nSec, impact);
    else
        return OECL_ERR_ARG;
    OECL_CHECK_LAUNCH("oecl_secem_pass1");
    return OECL_OK;
}

OECL_EXPORT int oecl_secem_pass2(const oecl_secem_t *prm, int dtype, const void *absorbed, const void *normals,
                                 uint32_t nAbs, const uint32_t *offsets, const double *impact,
                                 const double *uniforms, uint32_t seed0, uint32_t seed1, uint32_t step,
                                 void *secondaries, uint32_t secCap, void *stream)
{
    if (!prm || !offsets || !impact) return OECL_ERR_ARG;
    if (prm->model != OECL_SECEM_STAINLESS && prm->model != OECL_SECEM_SIMPLE) return OECL_ERR_ARG;
    if (nAbs == 0 || secCap == 0) return OECL_OK;
    if (!absorbed || !normals || !secondaries) return OECL_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    const dim3 gr(oecl_blocks(nAbs, 64)), bl(64);
    if (dtype == OECL_F64)
        hipLaunchKernelGGL(k_secem_pass2<double>, gr, bl, 0, s, *prm, (const double *)absorbed,
                           (const double *)normals, nAbs, offsets, impact, uniforms, seed0, seed1, step,
                           (double *)secondaries, secCap);
    else if (dtype == OECL_F32)
        hipLaunchKernelGGL(k_secem_pass2<float>, gr, bl, 0, s, *prm, (const float *)absorbed,
                           (const float *)normals, nAbs, offsets, impact, uniforms, seed0, seed1, step,
                           (float *)secondaries, secCap);
    else
        return OECL_ERR_ARG;
    OECL_CHECK_LAUNCH("oecl_secem_pass2");
    return OECL_OK;
}

OECL_EXPORT int oecl_homo_loader(const oecl_wall_t *wall, int dtype, void *data, uint64_t stride,
                                 uint32_t n, double weightInit, double vsigma, int randomizeWeight,
                                 uint32_t seed0, uint32_t seed1, uint32_t *failed, void *stream)
{
    if (!data || !failed) return OECL_ERR_ARG;
    if (n == 0) return OECL_OK;
    WallDev wd;
    int rc = wall_to_device(wall, &wd);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    if (dtype == OECL_F64)
        hipLaunchKernelGGL(k_homo_loader<double>, dim3(oecl_blocks(n, 256)), dim3(256), 0, s, wd,
                           Soa<double>{(double *)data, stride}, n, weightInit, vsigma,
                           randomizeWeight, seed0, seed1, failed);
    else if (dtype == OECL_F32)
        hipLaunchKernelGGL(k_homo_loader<float>, dim3(oecl_blocks(n, 256)), dim3(256), 0, s, wd,
                           Soa<float>{(float *)data, stride}, n, weightInit, vsigma,
                           randomizeWeight, seed0, seed1, failed);
    else
        return OECL_ERR_ARG;
    OECL_CHECK_LAUNCH("oecl_homo_loader");
    return OECL_OK;
}
