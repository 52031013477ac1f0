"""CPU ORACLE -- test infrastructure only (never imported by the product).

Restatement of openECLOUD's Furman-Pivi secondary emission (particleEmitter.pyx:77-308) with
INJECTED uniforms instead of the reference's global GSL MT19937 stream.  The special
functions come from scipy.special (library routines), which pin the reference's GSL calls
(specFun.pyx:34-41) and its hand-written inverses (specFun.pyx:53-341, accurate to 1e-12 and
1e-9 by their own stopping rules) to within those accuracies.

Uniform slots per impact j, U[j, k] (shared with the device emitter):
  k = 0        pass-1 draw that picks nSec                         (particleEmitter.pyx:224)
  k = 1        pass-2 elastic / rediffused / true branch draw        (:250)
  k = 2, 3     Box-Muller pair -> the elastic normal               (:254-256)
  k = 4        the one further draw of the nSec == 1 branches      (:258, :260, :262-263)
  k = 5 .. 13  thetaK draws, ii = 0 .. nSec-2                       (:268-269)
  k = 14       the y draw                                         (:270)
  k = 15 + 2i  theta draw of secondary i, k = 16 + 2i its phi draw   (:289, :292)
The normal is z = sqrt(-2 ln(1 - U[2])) * cos(2 pi U[3]).
"""
import math

import numpy as np
from scipy import special as sps

ELEMENTARY_CHARGE = 1.602176565e-19  # constants.pyx:5
M_E = 9.10938291e-31  # constants.pyx:6
PI = 3.141592653589793  # constants.pyx:9
N_UNIFORMS = 35


def material(name='copper', seyMax=-1., reflec=-1., eHat0=None):
    """particleEmitter.pyx:107-152 (stainlessMod / copperMod) as a dict; eHat0 optional."""
    if name == 'stainless':
        p = dict(p1EInf=0.07, p1Ehat=0.5, eEHat=0., w=100., p=0.9, e1=0.26, e2=2., q=0.4,
                 sigmaE=1.9, p1RInf=0.74, eR=40., r=1., r1=0.26, r2=2., deltaTSHat=1.22,
                 s=1.813, eHat0=310., t3=0.7, t4=1., t1=0.66, t2=0.8,
                 epsN=[3.9, 6.2, 13., 8.8, 6.25, 2.25, 9.2, 5.3, 17.8, 10.],
                 pSmallN=[1.6, 2., 1.8, 4.7, 1.8, 2.4, 1.8, 1.8, 2.3, 1.8])
    elif name == 'copper':
        p = dict(p1EInf=0.02, p1Ehat=0.496, eEHat=0., w=60.86, p=1., e1=0.26, e2=2., sigmaE=2.,
                 p1RInf=0.2, eR=0.041, r=0.104, q=0.5, r1=0.26, r2=2., deltaTSHat=1.8848,
                 s=1.54, eHat0=276.8, t3=0.7, t4=1., t1=0.66, t2=0.8,
                 epsN=[2.5, 3.3, 2.5, 2.5, 2.8, 1.3, 1.5, 1.5, 1.5, 1.5],
                 pSmallN=[1.5, 1.75, 1., 3.75, 8.5, 11.5, 2.5, 3., 2.5, 3.])
    else:
        raise ValueError('Material "%s" not implemented.' % name)
    p['theta0max'] = 84. * PI / 180
    p['deltaERmax'] = 0.99
    p['m'] = 10
    if seyMax >= 0.5:
        p['deltaTSHat'] = seyMax - p['p1RInf'] - p['p1EInf']
    if reflec >= 0.:
        p['p1Ehat'] = reflec
    if eHat0 is not None:
        p['eHat0'] = float(eHat0)
    return p


def binom_pmf(k, pr, n):
    """gsl_ran_binomial_pdf (specFun.pyx:34-35): C(n, k) p^k (1-p)^(n-k), the powers formed by
    repeated multiplication in a fixed order (the device code's order)."""
    pk = 1.
    for _ in range(k):
        pk *= pr
    qk = 1.
    q = 1. - pr
    for _ in range(n - k):
        qk *= q
    return math.comb(n, k) * pk * qk


def yields(e0, theta0, P):
    """particleEmitter.pyx:185-229: delta_e, delta_r, delta_ts' (after clamps) and pN."""
    m = P['m']
    deltaE0 = P['p1EInf'] + (P['p1Ehat'] - P['p1EInf']) * math.exp(
        -((abs(e0 - P['eEHat']) / P['w']) ** P['p']) / P['p'])
    deltaE = deltaE0 * (1 + P['e1'] * (1 - math.cos(theta0) ** P['e2']))
    deltaR0 = P['p1RInf'] * (1 - math.exp(-(e0 / P['eR']) ** P['r']))
    deltaR = deltaR0 * (1 + P['r1'] * (1 - math.cos(theta0) ** P['r2']))
    eHat = P['eHat0'] * (1 + P['t3'] * (1 - math.cos(theta0) ** P['t4']))
    deltaTS0 = P['deltaTSHat'] * P['s'] * e0 / eHat / (P['s'] - 1 + (e0 / eHat) ** P['s'])
    deltaTS = deltaTS0 * (1 + P['t1'] * (1 - math.cos(theta0) ** P['t2']))
    temp = deltaE + deltaR
    if temp >= P['deltaERmax']:
        deltaE = deltaE / temp * P['deltaERmax']
        deltaR = deltaR / temp * P['deltaERmax']
    dtsp = deltaTS / (1 - deltaE - deltaR)
    if dtsp >= m:
        dtsp = m
    elif dtsp < 0:
        dtsp = 0
    pr = dtsp / m
    pN = [binom_pmf(ii, pr, m) * (1 - deltaE - deltaR) for ii in range(m + 1)]
    pN[1] += deltaE + deltaR
    return deltaE, deltaR, deltaTS, pN


def generate_secondaries(absorbed, normals, U, P, particle_mass=M_E, scaleSigmaE=0):
    """particleEmitter.pyx:154-308 (generateSecondaries) with uniforms U[n0, 35].

    Returns (secondaries[nSecSum, 6], nSec[n0], per-impact e0)."""
    absorbed = np.asarray(absorbed, dtype=np.float64)
    normals = np.asarray(normals, dtype=np.float64)
    n0 = absorbed.shape[0]
    m = P['m']
    eFactor = 0.5 * particle_mass / ELEMENTARY_CHARGE
    eFactorInv = 1. / eFactor
    e0 = np.empty(n0)
    nSec = np.empty(n0, dtype=np.int64)
    deltaE = deltaR = 0.
    pN = [0.] * (m + 1)
    for jj in range(n0):
        vx, vy, vz = absorbed[jj, 2:5]
        v0 = math.sqrt(vx * vx + vy * vy + vz * vz)
        e0[jj] = eFactor * v0 * v0
        theta0 = math.acos((vx * normals[jj, 0] + vy * normals[jj, 1]) / v0)
        if theta0 > 0.5 * PI:
            theta0 = PI - theta0
        theta0 = min(max(theta0, 0.), P['theta0max'])
        deltaE, deltaR, _, pN = yields(e0[jj], theta0, P)
        cum = 0.
        nSec[jj] = m
        u = U[jj, 0]
        for ii in range(m):
            cum += pN[ii]
            if u <= cum:
                nSec[jj] = ii
                break
    total = int(nSec.sum())
    out = np.empty((total, 6))
    if total == 0:
        return out, nSec, e0
    # The nSec == 1 branch below uses deltaE, deltaR and pN of the LAST impact of the
    # first loop (particleEmitter.pyx:245-246), as the reference does.
    cur = 0
    for jj in range(n0):
        ns = int(nSec[jj])
        eSec = [0.] * max(ns, 1)
        if ns == 0:
            continue
        if ns == 1:
            aE = deltaE / pN[1]
            aR = deltaR / pN[1]
            u = U[jj, 1]
            if u < aE:
                z = math.sqrt(-2. * math.log(1. - U[jj, 2])) * math.cos(2. * PI * U[jj, 3])
                if scaleSigmaE == 0:
                    eSec[0] = e0[jj] - abs(P['sigmaE'] * z)
                else:
                    eSec[0] = e0[jj] - abs(P['sigmaE'] * math.sqrt(e0[jj] / 300) * z)
                if eSec[0] < 0.:
                    eSec[0] = U[jj, 4] * e0[jj]
            elif u < aE + aR:
                eSec[0] = e0[jj] * U[jj, 4] ** (1. / (1. + P['q']))
            else:
                a = P['pSmallN'][0]
                eSec[0] = P['epsN'][0] * sps.gammaincinv(
                    a, U[jj, 4] * sps.gammainc(a, e0[jj] / P['epsN'][0]))
        else:
            pn = P['pSmallN'][ns - 1]
            en = P['epsN'][ns - 1]
            p0 = sps.gammainc(ns * pn, e0[jj] / en)
            thetaK = [math.asin(math.sqrt(sps.betaincinv(pn * (ns - ii - 1), pn, U[jj, 5 + ii])))
                      for ii in range(ns - 1)]
            y = math.sqrt(sps.gammaincinv(ns * pn, U[jj, 14] * p0))
            yK = [0.] * ns
            yK[0] = y * math.cos(thetaK[0])
            yK[ns - 1] = y
            for ii in range(ns - 1):
                yK[ns - 1] *= math.sin(thetaK[ii])
            for ii in range(1, ns - 1):
                yK[ii] = y * math.cos(thetaK[ii])
                for kk in range(ii):
                    yK[ii] *= math.sin(thetaK[kk])
            eSec = [en * (yK[ii] * yK[ii]) for ii in range(ns)]
        n0x, n0y = normals[jj]
        for ii in range(ns):
            vSec = math.sqrt(eFactorInv * eSec[ii])
            th = math.asin(U[jj, 15 + 2 * ii])
            ct, st = math.cos(th), math.sin(th)
            ph = U[jj, 16 + 2 * ii] * PI * 2.
            cp, sp = math.cos(ph), math.sin(ph)
            t1 = ct * n0x - st * n0y
            t2 = st * n0x + ct * n0y
            out[cur, 0] = absorbed[jj, 0]
            out[cur, 1] = absorbed[jj, 1]
            out[cur, 5] = absorbed[jj, 5]
            out[cur, 2] = vSec * ((cp + n0x * n0x * (1 - cp)) * t1 + n0x * n0y * (1 - cp) * t2)
            out[cur, 3] = vSec * ((cp + n0y * n0y * (1 - cp)) * t2 + n0x * n0y * (1 - cp) * t1)
            out[cur, 4] = vSec * (-sp * n0y * t1 + sp * n0x * t2)
            cur += 1
    return out, nSec, e0
