"""Secondary emission and loading (particleEmitter.pyx:41-596) on the GPU.

FurmanEmitter, SecElecEmitter and SimpleSecEmitter keep the reference's constructors and
generateSecondaries(); FurmanEmitter additionally accepts eHat0 (the energy of maximum
true-secondary yield, 'E_max'), which the reference hard-codes per material.  Random draws
come from Philox (see randomGen) or, for testing, from an injected uniform table of shape
(nAbs, 35) for FurmanEmitter and (nAbs, 44) for the other two.
"""
import ctypes

import numpy as np
import torch

from . import _lib as L
from . import randomGen
from .constants import doubleMinVal, pi


def _material(name, seyMax, reflec, eHat0):
    """particleEmitter.pyx:107-152 (stainlessMod / copperMod)."""
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
    p['theta0max'] = 84. * pi / 180
    p['deltaERmax'] = 0.99
    if seyMax >= 0.5:
        p['deltaTSHat'] = seyMax - p['p1RInf'] - p['p1EInf']
    if reflec >= 0.:
        p['p1Ehat'] = reflec
    if eHat0 is not None:
        p['eHat0'] = float(eHat0)
    return p


class FurmanEmitter:
    """particleEmitter.pyx:77-311."""

    def __init__(self, particleBoundaryObj, particlesObj, material='copper', seyMax=-1.,
                 reflec=-1., scaleSigmaE=0, eHat0=None):
        self.particlesObj = particlesObj
        self.particleBoundaryObj = particleBoundaryObj
        self.params = _material(material, seyMax, reflec, eHat0)
        self.particleMass = particlesObj.getParticleMass()
        self.scaleSigmaE = int(scaleSigmaE)
        self.m = 10
        self.step = 0
        self.desc = self._make_desc()
        dev = L.device()
        self.secondaries = torch.empty((0, 6), dtype=particlesObj.dtype, device=dev)
        self._nsec = torch.empty(1, dtype=torch.int32, device=dev)
        self._off = torch.empty(2, dtype=torch.int32, device=dev)
        self._impact = torch.empty(4, dtype=torch.float64, device=dev)
        self.warnings = {'deltaERmax': 0, 'deltaTSPrime': 0}

    def _make_desc(self):
        d = L.FurmanDesc()
        P = self.params
        for k in ('p1EInf', 'p1Ehat', 'eEHat', 'w', 'p', 'e1', 'e2', 'sigmaE', 'p1RInf', 'eR',
                  'r', 'q', 'r1', 'r2', 'deltaTSHat', 's', 'eHat0', 't1', 't2', 't3', 't4',
                  'theta0max', 'deltaERmax'):
            setattr(d, k, float(P[k]))
        d.epsN = (ctypes.c_double * 10)(*P['epsN'])
        d.pSmallN = (ctypes.c_double * 10)(*P['pSmallN'])
        d.particleMass = self.particleMass
        d.scaleSigmaE = self.scaleSigmaE
        return d

    def _buffers(self, na):
        dev = L.device()
        if self._nsec.shape[0] < na:
            cap = int(na * 1.25) + 64
            self._nsec = torch.empty(cap, dtype=torch.int32, device=dev)
            self._off = torch.empty(cap + 1, dtype=torch.int32, device=dev)
            self._impact = torch.empty(4 * cap, dtype=torch.float64, device=dev)

    def _generate(self, uniforms=None, step=None):
        """particleEmitter.pyx:154-308 -> (nSecSum, 6) device tensor.

        uniforms: optional injected table (nAbs, 35) of FP64 uniforms; step: the Philox
        stream index (defaults to a counter advanced on every call)."""
        pb = self.particleBoundaryObj
        p = self.particlesObj
        na = pb.getAbsorbedMacroParticleCount()
        if step is None:
            step = self.step
            self.step += 1
        if na == 0:
            self.secondaries = torch.empty((0, 6), dtype=p.dtype, device=L.device())
            return self.secondaries
        self._buffers(na)
        u = None
        if uniforms is not None:
            u = L.to_dev(np.asarray(uniforms, np.float64).reshape(na, L.FURMAN_SLOTS))
        k0, k1 = randomGen.key()
        ap = pb.absorbedParticles
        nv = pb.normalVectors
        L.check(L.lib().oecl_furman_pass1(
            ctypes.byref(self.desc), L.i32(p.code), L.ptr(ap), L.ptr(nv), L.u32(na), L.ptr(u),
            L.u32(k0), L.u32(k1), L.u32(step), L.ptr(self._nsec), L.ptr(self._impact),
            L.stream()), 'furman pass 1')
        work = torch.empty(int(L.lib().oecl_scan_work_bytes(na)) // 4 + 8, dtype=torch.int32,
                           device=L.device())
        L.check(L.lib().oecl_scan_u32(L.ptr(self._nsec), L.u32(na), L.ptr(self._off),
                                      L.ptr(work), L.stream()), 'furman scan')
        total = int(self._off[na].item())
        self.secondaries = torch.empty((total, 6), dtype=p.dtype, device=L.device())
        if total > 0:
            L.check(L.lib().oecl_furman_pass2(
                ctypes.byref(self.desc), L.i32(p.code), L.ptr(ap), L.ptr(nv), L.u32(na),
                L.ptr(self._nsec), L.ptr(self._off), L.ptr(self._impact), L.ptr(u), L.u32(k0),
                L.u32(k1), L.u32(step), L.ptr(self.secondaries), L.u32(total), L.stream()),
                'furman pass 2')
        return self.secondaries

    @L.host_out
    def generateSecondaries(self, uniforms=None, step=None):
        """The secondaries of the absorbed particles (see _generate)."""
        return self._generate(uniforms, step)

    @L.host_out
    def getEmittedParticles(self):
        return self.secondaries

    @L.host_out
    def getNSec(self):
        na = self.particleBoundaryObj.getAbsorbedMacroParticleCount()
        return self._nsec[:na]


class _SimpleModelEmitter:
    """Shared driver of the two simpler emitters: first loop (counts), scan, second loop."""
    model = None

    def __init__(self, particleBoundaryObj, particlesObj):
        self.particlesObj = particlesObj
        self.particleBoundaryObj = particleBoundaryObj
        self.particleMass = particlesObj.getParticleMass()
        self.step = 0
        dev = L.device()
        self.secondaries = torch.empty((0, 6), dtype=particlesObj.dtype, device=dev)
        self._nsec = torch.empty(1, dtype=torch.int32, device=dev)
        self._off = torch.empty(2, dtype=torch.int32, device=dev)
        self._impact = torch.empty(4, dtype=torch.float64, device=dev)

    def _make_desc(self, P):
        d = L.SecemDesc()
        d.model = self.model
        for k, v in P.items():
            setattr(d, k, float(v))
        d.particleMass = self.particleMass
        return d

    def _generate(self, uniforms=None, step=None):
        """-> (nSecSum, 6) device tensor, the secondaries of each impact in the reference's
        order (elastic, rediffused, true ones).  uniforms: optional injected table (nAbs, 44)
        of FP64 uniforms; step: the Philox stream index (defaults to a call counter)."""
        pb = self.particleBoundaryObj
        p = self.particlesObj
        na = pb.getAbsorbedMacroParticleCount()
        if step is None:
            step = self.step
            self.step += 1
        dev = L.device()
        if na == 0:
            self.secondaries = torch.empty((0, 6), dtype=p.dtype, device=dev)
            return self.secondaries
        if self._nsec.shape[0] < na:
            cap = int(na * 1.25) + 64
            self._nsec = torch.empty(cap, dtype=torch.int32, device=dev)
            self._off = torch.empty(cap + 1, dtype=torch.int32, device=dev)
            self._impact = torch.empty(4 * cap, dtype=torch.float64, device=dev)
        u = None
        if uniforms is not None:
            u = L.to_dev(np.asarray(uniforms, np.float64).reshape(na, L.SECEM_SLOTS))
        k0, k1 = randomGen.key()
        ap, nv = pb.absorbedParticles, pb.normalVectors
        L.check(L.lib().oecl_secem_pass1(
            ctypes.byref(self.secemDesc), L.i32(p.code), L.ptr(ap), L.ptr(nv), L.u32(na), L.ptr(u),
            L.u32(k0), L.u32(k1), L.u32(step), L.ptr(self._nsec), L.ptr(self._impact), L.stream()),
            'secondary emission pass 1')
        work = torch.empty(int(L.lib().oecl_scan_work_bytes(na)) // 4 + 8, dtype=torch.int32, device=dev)
        L.check(L.lib().oecl_scan_u32(L.ptr(self._nsec), L.u32(na), L.ptr(self._off), L.ptr(work),
                                      L.stream()), 'secondary emission scan')
        total = int(self._off[na].item())
        self.secondaries = torch.empty((total, 6), dtype=p.dtype, device=dev)
        if total > 0:
            L.check(L.lib().oecl_secem_pass2(
                ctypes.byref(self.secemDesc), L.i32(p.code), L.ptr(ap), L.ptr(nv), L.u32(na),
                L.ptr(self._off), L.ptr(self._impact), L.ptr(u), L.u32(k0), L.u32(k1), L.u32(step),
                L.ptr(self.secondaries), L.u32(total), L.stream()), 'secondary emission pass 2')
        return self.secondaries

    @L.host_out
    def generateSecondaries(self, uniforms=None, step=None):
        """The secondaries of the absorbed particles (see _generate)."""
        return self._generate(uniforms, step)

    @L.host_out
    def getEmittedParticles(self):
        return self.secondaries

    @L.host_out
    def getNSec(self):
        na = self.particleBoundaryObj.getAbsorbedMacroParticleCount()
        return self._nsec[:na]


class SecElecEmitter(_SimpleModelEmitter):
    """particleEmitter.pyx:316-482: the stainless-steel Furman-Pivi yields with Bernoulli elastic
    and rediffused secondaries and floor(deltaTS) or floor(deltaTS) + 1 true secondaries of
    log-normal energy (sigmaETS 1.041, muETS 2.456).  Device kernels oecl_secem_pass1/2.

    The true-secondary count is limited to 8 per impact (10 secondaries in all): yields with
    deltaTSHat * (1 + t1) >= 8, i.e. seyMax above about 5.6, are rejected."""
    model = L.SECEM_STAINLESS

    def __init__(self, particleBoundaryObj, particlesObj, seyMax=-1., reflec=-1.):
        super().__init__(particleBoundaryObj, particlesObj)
        self.params = self.stainlessMod(seyMax, reflec)
        if self.params['deltaTSHat'] * (1. + self.params['t1']) >= 8.:
            raise ValueError('SecElecEmitter: seyMax %g allows more than 8 true secondaries per impact'
                             % seyMax)
        self.secemDesc = self._make_desc(self.params)

    @staticmethod
    def stainlessMod(seyMax, reflec):
        """particleEmitter.pyx:334-355."""
        p = dict(p1EInf=0.07, p1Ehat=0.5, eEHat=0., w=100., p=0.9, e1=0.26, e2=2., q=0.4,
                 p1RInf=0.74, eR=40., r=1., r1=0.26, r2=2., deltaTSHat=1.22, s=1.813, eHat0=310.,
                 t3=0.7, t4=1., t1=0.66, t2=0.8, sigmaETS=1.041, muETS=2.456,
                 theta0max=84. * pi / 180., deltaERmax=0.99)
        if seyMax >= 0.5:
            p['deltaTSHat'] = seyMax - p['p1RInf'] - p['p1EInf']
            if p['deltaTSHat'] < 0.:
                p['deltaTSHat'] = 1.e-6
                print('WARNING: Maximum SEY too small. Setting true SEY to zero.')
        if reflec >= 0.:
            p['p1Ehat'] = reflec
        return p


class SimpleSecEmitter(_SimpleModelEmitter):
    """particleEmitter.pyx:489-596 (marked there as for testing only): elastic yield 0.05,
    rediffused 0.5, true-secondary yield 1.4 s x / (s - 1 + x^s) with x = e0 / 310 eV and
    s = 1.813, true-secondary energies uniform in [0, 5) eV, and every secondary leaving along
    the reversed incident velocity.  seyMax and reflec are accepted and ignored, as there."""
    model = L.SECEM_SIMPLE

    def __init__(self, particleBoundaryObj, particlesObj, seyMax=-1., reflec=-1.):
        super().__init__(particleBoundaryObj, particlesObj)
        self.secemDesc = self._make_desc({})


def homoLoader(gridObj, particlesObj, partBoundObj, macroParticleCount, weightInit,
               vsigma=10000. * doubleMinVal, randomizeWeight=0):
    """particleEmitter.pyx:41-72: uniform positions inside the wall, normal velocities."""
    n = int(macroParticleCount)
    p = particlesObj
    cap = max(int(n * 1.25) + 1024, 1024)
    p._data = torch.empty((6, cap), dtype=p.dtype, device=L.device())
    p._spare = None
    p._alloc_scratch(cap)
    failed = torch.zeros(1, dtype=torch.int32, device=L.device())
    k0, k1 = randomGen.key()
    L.check(L.lib().oecl_homo_loader(
        ctypes.byref(partBoundObj.desc), L.i32(p.code), L.ptr(p._data), L.u64(cap), L.u32(n),
        L.f64(weightInit), L.f64(vsigma), L.i32(randomizeWeight), L.u32(k0), L.u32(k1),
        L.ptr(failed), L.stream()), 'homoLoader')
    if int(failed.item()) != 0:
        raise L.OeclError('homoLoader: %d particles found no inside point in 64 draws'
                          % int(failed.item()))
    p.macroParticleCount = n
    p._refCapacity = n


class PhotoEmitter:
    """Primary photoelectrons from the chamber wall while a bunch passes.

    The reference has no photoelectron source (its Readme lists photoemission as an outlook);
    this one is defined as follows.  During the step starting at t a fraction f(t) =
    beam.bunchFraction(t) of a bunch passes, which frees photoYield * nParticles * f(t)
    photoelectrons per metre.  They are carried by macro-particles at a rate of macroPerBunch
    per bunch: the fractional macro-particle count and the charge are carried over from step
    to step, so the total charge is exact and every macro-particle of a step has the same
    weight.  Each photoelectron starts on the wall (inset by 1e-3 cell; rectangle uniform in
    arc length, ellipse uniform in the parametric angle) with energy |energyMean +
    energySigma z| eV and the reference emitter's secondary angles about the inward normal
    (oecl_photo_emit).  Rectangular and elliptical walls only.
    """

    def __init__(self, partBoundObj, particlesObj, beamObj, photoYield=1e-3, macroPerBunch=10000,
                 energyMean=7., energySigma=5.):
        if partBoundObj.desc.kind not in (0, 1):
            raise NotImplementedError('photoelectrons: rectangular or elliptical wall only')
        self.pb = partBoundObj
        self.p = particlesObj
        self.beam = beamObj
        self.photoYield = float(photoYield)
        self.macroPerBunch = float(macroPerBunch)
        self.energyMean = float(energyMean)
        self.energySigma = float(energySigma)
        self._counts = np.zeros(0, np.uint32)
        self._weights = np.zeros(0, np.float64)
        self._carry = (0., 0.)

    def tables(self, nSteps):
        """Macro-particle count and weight of the steps 0 .. nSteps-1 (numpy arrays)."""
        k0 = self._counts.size
        if nSteps > k0:
            cnt = np.zeros(nSteps - k0, np.uint32)
            wts = np.zeros(nSteps - k0, np.float64)
            mAcc, qAcc = self._carry
            dt = self.beam.dt
            for k in range(k0, nSteps):
                f = self.beam.bunchFraction(k * dt)
                if f <= 0.:
                    continue
                mAcc += self.macroPerBunch * f
                qAcc += self.photoYield * self.beam.nParticles * f
                c = int(mAcc)
                if c > 0:
                    cnt[k - k0] = c
                    wts[k - k0] = qAcc / c
                    mAcc -= c
                    qAcc = 0.
            self._carry = (mAcc, qAcc)
            self._counts = np.concatenate([self._counts, cnt])
            self._weights = np.concatenate([self._weights, wts])
        return self._counts[:nSteps], self._weights[:nSteps]

    def generate(self, step):
        """The photoelectrons of time step `step` as a device (n, 6) array."""
        cnt, wts = self.tables(step + 1)
        n = int(cnt[step])
        out = torch.empty((n, 6), dtype=self.p.dtype, device=L.device())
        if n:
            k0, k1 = randomGen.key()
            L.check(L.lib().oecl_photo_emit(
                ctypes.byref(self.pb.desc), L.i32(self.p.code), L.ptr(out), L.u32(n), L.u32(step),
                L.u32(k0), L.u32(k1), L.f64(self.energyMean), L.f64(self.energySigma),
                L.f64(float(wts[step])), L.f64(self.p.particleMass), L.stream()), 'photoelectrons')
        return out
