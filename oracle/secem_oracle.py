"""CPU ORACLE -- test infrastructure only (never imported by the product).

Restatement of openECLOUD's two simpler emitters with INJECTED uniforms instead of the
reference's global GSL MT19937 stream:
  SecElecEmitter    particleEmitter.pyx:316-482 (stainlessMod :334-355)
  SimpleSecEmitter  particleEmitter.pyx:489-596
The reference draws its true-secondary normals with GSL's ziggurat (randomGen.pyx:32); here a
normal is z = sqrt(-2 ln(1 - U[a])) cos(2 pi U[b]) (Box-Muller, the same distribution).

Uniform slots per impact j, U[j, k] (shared with the device emitters):
  k = 0            elastic pick                     (:410 / :542)
  k = 1            rediffused pick                  (:412 / :544)
  k = 2            true-count pick                  (:415 / :547)
  k = 3            rediffused speed factor          (:440 / :572)
  k = 4+2i, 5+2i   Box-Muller pair of true secondary i   (stainless, :452)
  k = 4+i          energy of true secondary i            (simple, :583)
  k = 20+i         fallback energy of true secondary i   (stainless, :457)
  k = 28+2i, 29+2i theta / phi of true secondary i       (stainless, :464, :467)

Pinned by the tests through what the model fixes: the yields at their defining points
(deltaE(eEHat, 0) = p1Ehat, deltaR -> p1RInf, deltaTS(eHat0, 0) = deltaTSHat, the
deltaERmax rescale), mean counts = deltaE + deltaR + deltaTS over many impacts, the rediffused
energy law P(E/e0 <= x) = x^(1+q), log-normal true energies (mean and spread of ln E), the
fallback range, energy and direction of every emitted velocity (|v|^2 = 2 e E / m, cosine
law about the normal), exact mirroring of the elastic ones.
"""
import numpy as np

ELEMENTARY_CHARGE = 1.602176565e-19  # constants.pyx:5
M_E = 9.10938291e-31                 # constants.pyx:6, particleEmitter.pyx:512
PI = 3.141592653589793               # constants.pyx:9
N_UNIFORMS = 44
MAX_TS = 8


def stainless(seyMax=-1., reflec=-1.):
    """stainlessMod (particleEmitter.pyx:334-355) as a dict."""
    p = dict(p1EInf=0.07, p1Ehat=0.5, eEHat=0., w=100., p=0.9, e1=0.26, e2=2., q=0.4,
             p1RInf=0.74, eR=40., r=1., r1=0.26, r2=2., deltaTSHat=1.22, s=1.813, eHat0=310.,
             t3=0.7, t4=1., t1=0.66, t2=0.8, sigmaETS=1.041, muETS=2.456,
             theta0max=84. * PI / 180., deltaERmax=0.99)
    if seyMax >= 0.5:
        p['deltaTSHat'] = seyMax - p['p1RInf'] - p['p1EInf']
        if p['deltaTSHat'] < 0.:
            p['deltaTSHat'] = 1e-6
    if reflec >= 0.:
        p['p1Ehat'] = reflec
    return p


def stainless_yields(e0, theta0, P):
    """(deltaE, deltaR, deltaTS) of particleEmitter.pyx:395-403 for energies e0 [eV] and
    incidence angles theta0 (already folded and clipped); the deltaERmax rescale of :405-408."""
    e0 = np.asarray(e0, np.float64)
    ct = np.cos(theta0)
    dE0 = P['p1EInf'] + (P['p1Ehat'] - P['p1EInf']) * np.exp(
        -(np.abs(e0 - P['eEHat']) / P['w']) ** P['p'] / P['p'])
    dE = dE0 * (1 + P['e1'] * (1 - ct ** P['e2']))
    dR0 = P['p1RInf'] * (1 - np.exp(-(e0 / P['eR']) ** P['r']))
    dR = dR0 * (1 + P['r1'] * (1 - ct ** P['r2']))
    eHat = P['eHat0'] * (1 + P['t3'] * (1 - ct ** P['t4']))
    dTS0 = P['deltaTSHat'] * P['s'] * e0 / eHat / (P['s'] - 1 + (e0 / eHat) ** P['s'])
    dTS = dTS0 * (1 + P['t1'] * (1 - ct ** P['t2']))
    tmp = dE + dR
    big = tmp >= P['deltaERmax']
    dE = np.where(big, dE / np.where(big, tmp, 1.) * P['deltaERmax'], dE)
    dR = np.where(big, dR / np.where(big, tmp, 1.) * P['deltaERmax'], dR)
    return dE, dR, dTS


def simple_yields(e0):
    """(0.05, 0.5, deltaTS) of particleEmitter.pyx:536-540."""
    s, eHat = 1.813, 310.
    e0 = np.asarray(e0, np.float64)
    return (np.full_like(e0, 0.05), np.full_like(e0, 0.5),
            1.4 * s * e0 / eHat / (s - 1 + (e0 / eHat) ** s))


def _theta0(ap, nv):
    v0 = np.sqrt(ap[:, 2] ** 2 + ap[:, 3] ** 2 + ap[:, 4] ** 2)
    with np.errstate(invalid='ignore', divide='ignore'):
        th = np.arccos((ap[:, 2] * nv[:, 0] + ap[:, 3] * nv[:, 1]) / v0)
    th = np.where(th > 0.5 * PI, PI - th, th)
    return v0, th


def counts(ap, nv, U, model, P=None, particle_mass=M_E):
    """First loop: (nE, nR, nTS, e0) per impact.  model 'stainless' or 'simple'."""
    ap = np.asarray(ap, np.float64)
    nv = np.asarray(nv, np.float64)
    v0, th = _theta0(ap, nv)
    if model == 'simple':
        e0 = 0.5 * M_E / ELEMENTARY_CHARGE * v0 * v0
        dE, dR, dTS = simple_yields(e0)
    else:
        e0 = 0.5 * particle_mass / ELEMENTARY_CHARGE * v0 * v0
        # clip (particleEmitter.pyx:33-39): NaN -> theta0max
        thc = np.where((th <= P['theta0max']) & (th >= 0.), th, np.where(th < 0., 0., P['theta0max']))
        dE, dR, dTS = stainless_yields(e0, thc, P)
    nE = (U[:, 0] <= dE).astype(np.int64)
    nR = (U[:, 1] <= dR).astype(np.int64)
    fl = np.floor(dTS).astype(np.int64)
    nTS = np.where(U[:, 2] <= dTS - fl, fl + 1, fl)
    if np.any(nTS > MAX_TS):
        raise ValueError('more than %d true secondaries' % MAX_TS)
    return nE, nR, nTS, e0


def generate_secondaries(ap, nv, U, model, P=None, particle_mass=M_E):
    """(secondaries (nSec, 6), nSec per impact) of generateSecondaries."""
    ap = np.asarray(ap, np.float64)
    nv = np.asarray(nv, np.float64)
    nE, nR, nTS, e0 = counts(ap, nv, U, model, P, particle_mass)
    out = []
    eFactorInv = 1. / (0.5 * particle_mass / ELEMENTARY_CHARGE)
    for j in range(ap.shape[0]):
        a, u = ap[j], U[j]
        if nE[j]:
            out.append([a[0], a[1], -a[2], -a[3], -a[4], a[5]])
        if nR[j]:
            vf = np.sqrt(u[3]) if model == 'simple' else np.sqrt(u[3] ** (1. / (1. + P['q'])))
            out.append([a[0], a[1], -a[2] * vf, -a[3] * vf, -a[4] * vf, a[5]])
        n = int(nTS[j])
        if n == 0:
            continue
        if model == 'simple':
            for i in range(n):
                vf = np.sqrt(u[4 + i] * 5 / e0[j])
                out.append([a[0], a[1], -a[2] * vf, -a[3] * vf, -a[4] * vf, a[5]])
            continue
        z = np.sqrt(-2. * np.log(1. - u[4:4 + 2 * n:2])) * np.cos(2. * PI * u[5:5 + 2 * n:2])
        eSec = np.exp(P['sigmaETS'] * z + P['muETS'])
        eSum = 0.
        for i in range(n):
            eSum += eSec[i]
        if eSum > e0[j]:
            eSec = u[20:20 + n] * e0[j] / n
        nx, ny = nv[j]
        for i in range(n):
            vSec = np.sqrt(eFactorInv * eSec[i])
            th = np.arcsin(u[28 + 2 * i])
            ct, st = np.cos(th), np.sin(th)
            ph = u[29 + 2 * i] * PI * 2.
            cp, sp = np.cos(ph), np.sin(ph)
            t1 = ct * nx - st * ny
            t2 = st * nx + ct * ny
            out.append([a[0], a[1],
                        vSec * ((cp + nx * nx * (1 - cp)) * t1 + nx * ny * (1 - cp) * t2),
                        vSec * ((cp + ny * ny * (1 - cp)) * t2 + nx * ny * (1 - cp) * t1),
                        vSec * (-sp * ny * t1 + sp * nx * t2), a[5]])
    sec = np.asarray(out, np.float64).reshape(-1, 6)
    return sec, nE + nR + nTS
