// Furman-Pivi secondary emission (particleEmitter.pyx:154-308) and the uniform loader
// (particleEmitter.pyx:41-72).
//
// Random numbers: either an injected FP64 table [nAbs][35] or Philox4x32-10 with counter
// (impact, slot/2, step, purpose).  The slot of every draw is fixed per impact:
//   0 nSec pick, 1 branch pick, 2-3 Box-Muller pair, 4 the further nSec == 1 draw,
//   5..13 thetaK, 14 y, 15+2i / 16+2i theta / phi of secondary i.
#include "common.h"
#include "scan.h"
#include "specfun.h"
#include "wall.h"

#define FURMAN_SLOTS 35
#define PURPOSE_FURMAN 1u
#define PURPOSE_LOADER 2u

namespace {

constexpr double kPi = 3.141592653589793;            // constants.pyx:9
constexpr double kQe = 1.602176565e-19;              // constants.pyx:5

struct Rng {
    const double *table;
    uint32_t item, step, s0, s1;
    __device__ __forceinline__ double u(int k) const
    {
        if (table) return table[(uint64_t)item * FURMAN_SLOTS + k];
        return philox_uniform(item, (uint32_t)k, step, PURPOSE_FURMAN, s0, s1);
    }
};

__device__ __forceinline__ double clipd(double x, double lo, double hi)
{
    if (x <= hi && x >= lo) return x;
    if (x < lo) return lo;
    return hi;
}

template <typename T>
__global__ void k_furman_pass1(oecl_furman_t P, const T *__restrict__ ap, const T *__restrict__ nv,
                               uint32_t n0, const double *__restrict__ table, uint32_t s0,
                               uint32_t s1, uint32_t step, uint32_t *__restrict__ nSec,
                               double *__restrict__ impact)
{
    uint32_t jj = blockIdx.x * blockDim.x + threadIdx.x;
    if (jj >= n0) return;
    const int m = 10;
    double vx = ap[6 * (uint64_t)jj + 2], vy = ap[6 * (uint64_t)jj + 3], vz = ap[6 * (uint64_t)jj + 4];
    double nx0 = nv[2 * (uint64_t)jj], ny0 = nv[2 * (uint64_t)jj + 1];
    double eFactor = 0.5 * P.particleMass / kQe;
    double v0 = sqrt(vx * vx + vy * vy + vz * vz);
    double e0 = eFactor * v0 * v0;
    double theta0 = acos((vx * nx0 + vy * ny0) / v0);
    if (theta0 > 0.5 * kPi) theta0 = kPi - theta0;
    theta0 = clipd(theta0, 0., P.theta0max);
    double ct = cos(theta0);
    double deltaE0 = P.p1EInf + (P.p1Ehat - P.p1EInf) * exp(-pow(fabs(e0 - P.eEHat) / P.w, P.p) / P.p);
    double deltaE = deltaE0 * (1 + P.e1 * (1 - pow(ct, P.e2)));
    double deltaR0 = P.p1RInf * (1 - exp(-pow(e0 / P.eR, P.r)));
    double deltaR = deltaR0 * (1 + P.r1 * (1 - pow(ct, P.r2)));
    double eHat = P.eHat0 * (1 + P.t3 * (1 - pow(ct, P.t4)));
    double deltaTS0 = P.deltaTSHat * P.s * e0 / eHat / (P.s - 1 + pow(e0 / eHat, P.s));
    double deltaTS = deltaTS0 * (1 + P.t1 * (1 - pow(ct, P.t2)));
    double temp = deltaE + deltaR;
    if (temp >= P.deltaERmax) {
        deltaE = deltaE / temp * P.deltaERmax;
        deltaR = deltaR / temp * P.deltaERmax;
    }
    double dtsp = deltaTS / (1 - deltaE - deltaR);
    if (dtsp >= m) dtsp = m;
    else if (dtsp < 0) dtsp = 0;
    double pr = dtsp / m;
    Rng rng{table, jj, step, s0, s1};
    double u = rng.u(0);
    double cum = 0., pN1 = 0.;
    uint32_t ns = m;
    for (int ii = 0; ii < m; ++ii) {
        double pn = sf::binom_pmf(ii, pr, m) * (1 - deltaE - deltaR);
        if (ii == 1) {
            pn += deltaE + deltaR;
            pN1 = pn;
        }
        cum += pn;
        if (u <= cum) {
            ns = ii;
            break;
        }
    }
    if (ns == 0) pN1 = sf::binom_pmf(1, pr, m) * (1 - deltaE - deltaR) + deltaE + deltaR;
    nSec[jj] = ns;
    impact[4 * (uint64_t)jj] = e0;
    impact[4 * (uint64_t)jj + 1] = deltaE;
    impact[4 * (uint64_t)jj + 2] = deltaR;
    impact[4 * (uint64_t)jj + 3] = pN1;
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
    Rng rng{table, jj, step, s0, s1};
    double e0 = impact[4 * (uint64_t)jj];
    double eSec[10];
    if (ns == 1) {
        // particleEmitter.pyx:245-246: deltaE, deltaR and pN[1] are those left over from
        // the LAST impact of the first loop.
        const double *last = impact + 4 * (uint64_t)(n0 - 1);
        double aE = last[1] / last[3];
        double aR = last[2] / last[3];
        double u = rng.u(1);
        if (u < aE) {
            double z = sqrt(-2. * log(1. - rng.u(2))) * cos(2. * kPi * rng.u(3));
            if (P.scaleSigmaE == 0) eSec[0] = e0 - fabs(P.sigmaE * z);
            else eSec[0] = e0 - fabs(P.sigmaE * sqrt(e0 / 300) * z);
            if (eSec[0] < 0.) eSec[0] = rng.u(4) * e0;
        } else if (u < aE + aR) {
            eSec[0] = e0 * pow(rng.u(4), 1. / (1. + P.q));
        } else {
            double a = P.pSmallN[0];
            eSec[0] = P.epsN[0] * sf::gammaincinv(a, rng.u(4) * sf::gammainc(a, e0 / P.epsN[0]));
        }
    } else {
        double pn = P.pSmallN[ns - 1], en = P.epsN[ns - 1];
        double p0 = sf::gammainc(ns * pn, e0 / en);
        double thetaK[9];
        for (uint32_t ii = 0; ii < ns - 1; ++ii)
            thetaK[ii] = asin(sqrt(sf::betaincinv(pn * (ns - ii - 1), pn, rng.u(5 + ii))));
        double y = sqrt(sf::gammaincinv(ns * pn, rng.u(14) * p0));
        double yK[10];
        yK[0] = y * cos(thetaK[0]);
        yK[ns - 1] = y;
        for (uint32_t ii = 0; ii < ns - 1; ++ii) yK[ns - 1] *= sin(thetaK[ii]);
        for (uint32_t ii = 1; ii < ns - 1; ++ii) {
            yK[ii] = y * cos(thetaK[ii]);
            for (uint32_t kk = 0; kk < ii; ++kk) yK[ii] *= sin(thetaK[kk]);
        }
        for (uint32_t ii = 0; ii < ns; ++ii) eSec[ii] = en * (yK[ii] * yK[ii]);
    }
    double eFactorInv = 1. / (0.5 * P.particleMass / kQe);
    const T *a = ap + 6 * (uint64_t)jj;
    double n0x = nv[2 * (uint64_t)jj], n0y = nv[2 * (uint64_t)jj + 1];
    uint32_t base = offs[jj];
    for (uint32_t ii = 0; ii < ns; ++ii) {
        uint32_t k = base + ii;
        if (k >= secCap) return;
        double vSec = sqrt(eFactorInv * eSec[ii]);
        double th = asin(rng.u(15 + 2 * ii));
        double ct = cos(th), st = sin(th);
        double ph = rng.u(16 + 2 * ii) * kPi * 2.;
        double cp = cos(ph), sp = sin(ph);
        double t1 = ct * n0x - st * n0y;
        double t2 = st * n0x + ct * n0y;
        T *o = out + 6 * (uint64_t)k;
        o[0] = a[0];
        o[1] = a[1];
        o[5] = a[5];
        o[2] = (T)(vSec * ((cp + n0x * n0x * (1 - cp)) * t1 + n0x * n0y * (1 - cp) * t2));
        o[3] = (T)(vSec * ((cp + n0y * n0y * (1 - cp)) * t2 + n0x * n0y * (1 - cp) * t1));
        o[4] = (T)(vSec * (-sp * n0y * t1 + sp * n0x * t2));
    }
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

}  // namespace

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
