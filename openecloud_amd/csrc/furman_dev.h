// Furman-Pivi emission for one impact (particleEmitter.pyx:178-305), shared by the emitter
// kernels and the fused step.  RNG is any functor with double u(int slot); the slot of every
// draw is fixed per impact:
//   0 nSec pick, 1 branch pick, 2-3 Box-Muller pair, 4 the further nSec == 1 draw,
//   5..13 thetaK, 14 y, 15+2i / 16+2i theta / phi of secondary i.
#pragma once
#include "common.h"
#include "specfun.h"

#define FURMAN_SLOTS 35
#define PURPOSE_FURMAN 1u

namespace furman {

constexpr double kPi = 3.141592653589793;  // constants.pyx:9
constexpr double kQe = 1.602176565e-19;    // constants.pyx:5

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

// First loop (particleEmitter.pyx:178-229) for one impact: returns nSec and writes
// impact[4] = (e0, deltaE, deltaR, pN[1]) after the reference's clamps.
template <typename T, typename R>
__device__ inline uint32_t pass1(const oecl_furman_t &P, const T *a, const T *nv, const R &rng,
                                 double *impact)
{
    const int m = 10;
    double vx = a[2], vy = a[3], vz = a[4];
    double nx0 = nv[0], ny0 = nv[1];
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
    double u = rng.u(0);
    double cum = 0., pN1 = 0.;
    uint32_t ns = m;
    const sf::Binom10 bin(pr);
    for (int ii = 0; ii < m; ++ii) {
        double pn = bin.pmf(ii) * (1 - deltaE - deltaR);
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
    if (ns == 0) pN1 = bin.pmf(1) * (1 - deltaE - deltaR) + deltaE + deltaR;
    impact[0] = e0;
    impact[1] = deltaE;
    impact[2] = deltaR;
    impact[3] = pN1;
    return ns;
}

// Second loop (particleEmitter.pyx:236-305) for one impact with ns >= 1 secondaries.  `last`
// holds (e0, deltaE, deltaR, pN1) of the LAST impact of the first loop: the nSec == 1 branch
// takes its deltaE, deltaR and pN[1] from there, as the reference does (:245-246).  Writes
// secondary ii through out(ii) (a T* to 6 values, or nullptr to skip).
template <typename T, typename R, typename OUT>
__device__ inline void pass2(const oecl_furman_t &P, const T *a, const T *nv, uint32_t ns,
                             double e0, const double *last, const R &rng, OUT out)
{
    double eSec[10];
    if (ns == 1) {
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
            double ap = P.pSmallN[0];
            eSec[0] = P.epsN[0] * sf::gammaincinv(ap, rng.u(4) * sf::gammainc(ap, e0 / P.epsN[0]));
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
    double n0x = nv[0], n0y = nv[1];
    for (uint32_t ii = 0; ii < ns; ++ii) {
        T *o = out(ii);
        if (!o) return;
        double vSec = sqrt(eFactorInv * eSec[ii]);
        double th = asin(rng.u(15 + 2 * ii));
        double ct = cos(th), st = sin(th);
        double ph = rng.u(16 + 2 * ii) * kPi * 2.;
        double cp = cos(ph), sp = sin(ph);
        double t1 = ct * n0x - st * n0y;
        double t2 = st * n0x + ct * n0y;
        o[0] = a[0];
        o[1] = a[1];
        o[5] = a[5];
        o[2] = (T)(vSec * ((cp + n0x * n0x * (1 - cp)) * t1 + n0x * n0y * (1 - cp) * t2));
        o[3] = (T)(vSec * ((cp + n0y * n0y * (1 - cp)) * t2 + n0x * n0y * (1 - cp) * t1));
        o[4] = (T)(vSec * (-sp * n0y * t1 + sp * n0x * t2));
    }
}


// Second loop with the secondaries of one impact spread over a group of MAXSEC lanes of one
// wave: lane ii of the group draws the angle thetaK[ii] (ii < ns - 1) or the radius y
// (ii = ns - 1) at the same time as the other lanes, the group exchanges them by shuffles,
// and lane ii assembles secondary ii with exactly the arithmetic of pass2 (same values, same
// order).  All 64 lanes of the wave call it; lanes without an impact pass ns = 0.  `base` is
// the wave lane of the group's first lane; o points to this lane's output row (or nullptr).
template <typename T, typename R>
__device__ inline void pass2_lane(const oecl_furman_t &P, const T *a, const T *nv, uint32_t ns, double e0,
                                  const double *last, const R &rng, uint32_t ii, int base, T *o)
{
    double th = 0., yv = 0., eS = 0.;
    double en = 0.;
    if (ns == 1) {
        if (ii == 0) {
            double aE = last[1] / last[3];
            double aR = last[2] / last[3];
            double u = rng.u(1);
            if (u < aE) {
                double z = sqrt(-2. * log(1. - rng.u(2))) * cos(2. * kPi * rng.u(3));
                if (P.scaleSigmaE == 0) eS = e0 - fabs(P.sigmaE * z);
                else eS = e0 - fabs(P.sigmaE * sqrt(e0 / 300) * z);
                if (eS < 0.) eS = rng.u(4) * e0;
            } else if (u < aE + aR) {
                eS = e0 * pow(rng.u(4), 1. / (1. + P.q));
            } else {
                double ap = P.pSmallN[0];
                eS = P.epsN[0] * sf::gammaincinv(ap, rng.u(4) * sf::gammainc(ap, e0 / P.epsN[0]));
            }
        }
    } else if (ns >= 2) {
        double pn = P.pSmallN[ns - 1];
        en = P.epsN[ns - 1];
        if (ii < ns - 1) {
            th = asin(sqrt(sf::betaincinv(pn * (ns - ii - 1), pn, rng.u(5 + ii))));
        } else if (ii == ns - 1) {
            double p0 = sf::gammainc(ns * pn, e0 / en);
            yv = sqrt(sf::gammaincinv(ns * pn, rng.u(14) * p0));
        }
    }
    const double y = __shfl(yv, base + (ns >= 2 ? (int)ns - 1 : 0), 64);
    double thK[9];
#pragma unroll
    for (int kk = 0; kk < 9; ++kk) thK[kk] = __shfl(th, base + kk, 64);
    if (ns >= 2 && ii < ns) {
        double yk;
        if (ii == 0) {
            yk = y * cos(thK[0]);
        } else if (ii == ns - 1) {
            yk = y;
            for (uint32_t kk = 0; kk < ns - 1; ++kk) yk *= sin(thK[kk]);
        } else {
            yk = y * cos(thK[ii]);
            for (uint32_t kk = 0; kk < ii; ++kk) yk *= sin(thK[kk]);
        }
        eS = en * (yk * yk);
    }
    if (ii >= ns || !o) return;
    double eFactorInv = 1. / (0.5 * P.particleMass / kQe);
    double n0x = nv[0], n0y = nv[1];
    double vSec = sqrt(eFactorInv * eS);
    double thv = asin(rng.u(15 + 2 * ii));
    double ct = cos(thv), st = sin(thv);
    double ph = rng.u(16 + 2 * ii) * kPi * 2.;
    double cp = cos(ph), sp = sin(ph);
    double t1 = ct * n0x - st * n0y;
    double t2 = st * n0x + ct * n0y;
    o[0] = a[0];
    o[1] = a[1];
    o[5] = a[5];
    o[2] = (T)(vSec * ((cp + n0x * n0x * (1 - cp)) * t1 + n0x * n0y * (1 - cp) * t2));
    o[3] = (T)(vSec * ((cp + n0y * n0y * (1 - cp)) * t2 + n0x * n0y * (1 - cp) * t1));
    o[4] = (T)(vSec * (-sp * n0y * t1 + sp * n0x * t2));
}

}  // namespace furman
