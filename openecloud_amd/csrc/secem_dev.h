// The two simpler secondary emitters of particleEmitter.pyx, for one impact, shared by the
// standalone emitter entry points and the one-call step:
//   OECL_SECEM_STAINLESS  SecElecEmitter (particleEmitter.pyx:316-482): the Furman-Pivi yields of
//       stainless steel, but at most one elastic and one rediffused secondary (a Bernoulli draw
//       each) and floor(deltaTS) or floor(deltaTS) + 1 true secondaries with log-normal energies;
//   OECL_SECEM_SIMPLE     SimpleSecEmitter (particleEmitter.pyx:489-596): fixed yields 0.05 and
//       0.5, the true-secondary yield curve with peak 1.4, and every secondary reflected.
// Secondaries of an impact come in the reference's order: elastic, rediffused, true ones.
//
// RNG is any functor with double u(int slot); the slot of every draw is fixed per impact:
//   0 elastic pick, 1 rediffused pick, 2 true-count pick, 3 rediffused speed factor,
//   4 + 2i / 5 + 2i   Box-Muller pair of true secondary i's normal (stainless),
//   4 + i             true secondary i's energy draw (simple),
//   20 + i            true secondary i's energy when the log-normal energies exceed e0,
//   28 + 2i / 29 + 2i theta / phi of true secondary i (stainless).
// The reference's normal draws are GSL's ziggurat; Box-Muller gives the same distribution.
#pragma once
#include "common.h"

#define SECEM_SLOTS 44
#define SECEM_MAXTS 8u
#define PURPOSE_SECEM 7u

namespace secem {

constexpr double kPi = 3.141592653589793;  // constants.pyx:9
constexpr double kQe = 1.602176565e-19;    // constants.pyx:5
constexpr double kMe = 9.10938291e-31;     // particleEmitter.pyx:512 (SimpleSecEmitter's own mass)

struct Rng {
    const double *table;
    uint32_t item, step, s0, s1;
    __device__ __forceinline__ double u(int k) const
    {
        if (table) return table[(uint64_t)item * SECEM_SLOTS + k];
        return philox_uniform(item, (uint32_t)k, step, PURPOSE_SECEM, s0, s1);
    }
};

// Yields and counts for one impact (particleEmitter.pyx:383-418 / :528-550).  Returns the
// secondary count and writes impact[0] = e0, impact[1] = nE + 2 nR + 4 nTS (exact in a double).
template <typename T, typename R>
__device__ inline uint32_t pass1(const oecl_secem_t &P, const T *a, const T *nv, const R &rng, double *impact)
{
    const double vx = a[2], vy = a[3], vz = a[4];
    const double v0 = sqrt(vx * vx + vy * vy + vz * vz);
    double e0, deltaE, deltaR, deltaTS;
    if (P.model == OECL_SECEM_SIMPLE) {
        e0 = (0.5 * kMe / kQe) * v0 * v0;
        const double s = 1.813, eHat = 310;
        deltaE = 0.05;
        deltaR = 0.5;
        deltaTS = 1.4 * s * e0 / eHat / (s - 1 + pow(e0 / eHat, s));
    } else {
        e0 = (0.5 * P.particleMass / kQe) * v0 * v0;
        double theta0 = acos((vx * (double)nv[0] + vy * (double)nv[1]) / v0);
        if (theta0 > 0.5 * kPi) theta0 = kPi - theta0;
        // clip(theta0, 0, theta0max) as particleEmitter.pyx:33-39 (NaN -> theta0max)
        if (!(theta0 <= P.theta0max && theta0 >= 0.)) theta0 = theta0 < 0. ? 0. : P.theta0max;
        const double ct = cos(theta0);
        const double deltaE0 = P.p1EInf + (P.p1Ehat - P.p1EInf) * exp(-pow(fabs(e0 - P.eEHat) / P.w, P.p) / P.p);
        deltaE = deltaE0 * (1 + P.e1 * (1 - pow(ct, P.e2)));
        const double deltaR0 = P.p1RInf * (1 - exp(-pow(e0 / P.eR, P.r)));
        deltaR = deltaR0 * (1 + P.r1 * (1 - pow(ct, P.r2)));
        const double eHat = P.eHat0 * (1 + P.t3 * (1 - pow(ct, P.t4)));
        const double deltaTS0 = P.deltaTSHat * P.s * e0 / eHat / (P.s - 1 + pow(e0 / eHat, P.s));
        deltaTS = deltaTS0 * (1 + P.t1 * (1 - pow(ct, P.t2)));
        const double temp = deltaE + deltaR;
        if (temp >= P.deltaERmax) {
            deltaE = deltaE / temp * P.deltaERmax;
            deltaR = deltaR / temp * P.deltaERmax;
        }
    }
    const uint32_t nE = rng.u(0) <= deltaE ? 1u : 0u;
    const uint32_t nR = rng.u(1) <= deltaR ? 1u : 0u;
    const uint32_t fl = (uint32_t)deltaTS;
    uint32_t nTS = rng.u(2) <= (deltaTS - fl) ? fl + 1 : fl;
    if (nTS > SECEM_MAXTS) nTS = SECEM_MAXTS;  // unreachable for the yields the host accepts
    impact[0] = e0;
    impact[1] = (double)(nE + 2 * nR + 4 * nTS);
    return nE + nR + nTS;
}

// Writes the secondaries of one impact (particleEmitter.pyx:420-480 / :552-594) through
// out(k) (a T* to 6 values, or nullptr to stop), k = 0 .. count-1.
template <typename T, typename R, typename OUT>
__device__ inline void pass2(const oecl_secem_t &P, const T *a, const T *nv, const double *impact, const R &rng,
                             OUT out)
{
    const double e0 = impact[0];
    const uint32_t code = (uint32_t)impact[1];
    const uint32_t nE = code & 1u, nR = (code >> 1) & 1u, nTS = code >> 2;
    uint32_t k = 0;
    if (nE) {
        T *o = out(k++);
        if (!o) return;
        o[0] = a[0];
        o[1] = a[1];
        o[5] = a[5];
        o[2] = -a[2];
        o[3] = -a[3];
        o[4] = -a[4];
    }
    if (nR) {
        const double vFac = P.model == OECL_SECEM_SIMPLE ? sqrt(rng.u(3)) : sqrt(pow(rng.u(3), 1. / (1. + P.q)));
        T *o = out(k++);
        if (!o) return;
        o[0] = a[0];
        o[1] = a[1];
        o[5] = a[5];
        o[2] = (T)(-(double)a[2] * vFac);
        o[3] = (T)(-(double)a[3] * vFac);
        o[4] = (T)(-(double)a[4] * vFac);
    }
    if (nTS == 0) return;
    if (P.model == OECL_SECEM_SIMPLE) {
        for (uint32_t ii = 0; ii < nTS; ++ii) {
            const double eSec = rng.u(4 + ii) * 5;
            const double vFac = sqrt(eSec / e0);
            T *o = out(k++);
            if (!o) return;
            o[0] = a[0];
            o[1] = a[1];
            o[5] = a[5];
            o[2] = (T)(-(double)a[2] * vFac);
            o[3] = (T)(-(double)a[3] * vFac);
            o[4] = (T)(-(double)a[4] * vFac);
        }
        return;
    }
    double eSec[SECEM_MAXTS];
    double eSum = 0.;
    for (uint32_t ii = 0; ii < nTS; ++ii) {
        const double z = sqrt(-2. * log(1. - rng.u(4 + 2 * ii))) * cos(2. * kPi * rng.u(5 + 2 * ii));
        eSec[ii] = exp(P.sigmaETS * z + P.muETS);
        eSum += eSec[ii];
    }
    if (eSum > e0)
        for (uint32_t ii = 0; ii < nTS; ++ii) eSec[ii] = rng.u(20 + ii) * e0 / nTS;
    const double eFactorInv = 1. / (0.5 * P.particleMass / kQe);
    const double n0x = nv[0], n0y = nv[1];
    for (uint32_t ii = 0; ii < nTS; ++ii) {
        T *o = out(k++);
        if (!o) return;
        const double vSec = sqrt(eFactorInv * eSec[ii]);
        const double th = asin(rng.u(28 + 2 * ii));
        const double ct = cos(th), st = sin(th);
        const double ph = rng.u(29 + 2 * ii) * kPi * 2.;
        const double cp = cos(ph), sp = sin(ph);
        const double t1 = ct * n0x - st * n0y;
        const double t2 = st * n0x + ct * n0y;
        o[0] = a[0];
        o[1] = a[1];
        o[5] = a[5];
        o[2] = (T)(vSec * ((cp + n0x * n0x * (1 - cp)) * t1 + n0x * n0y * (1 - cp) * t2));
        o[3] = (T)(vSec * ((cp + n0y * n0y * (1 - cp)) * t2 + n0x * n0y * (1 - cp) * t1));
        o[4] = (T)(vSec * (-sp * n0y * t1 + sp * n0x * t2));
    }
}

}  // namespace secem
