"""Rigid Gaussian proton beams (beam.py:13-136 of the reference) as charge sources.

The transverse profile is computed once on the host exactly as the reference does (products
of normal-CDF differences over each dual cell within 5 sigma + max(dx, dy), normalised to
sum 1) and kept on the device; getCharge(t) returns the device vector scale(t) * profile or
0 like the reference.  The fused step uses getScale(t) and getProfile() directly.

Constants: the reference's beam.py takes e and c from scipy.constants, whose e at the time
(SciPy 0.12, CODATA 2010) equals constants.pyx's; that value is used here.
"""
import math

import numpy as np
import torch

from . import _lib as L
from .constants import c as C_LIGHT
from .constants import elementary_charge


def _norm_cdf(x):
    return 0.5 * (1. + np.vectorize(math.erf)(np.asarray(x) / math.sqrt(2.)))


def _profile(gridObj, radiusSigma, radiusLimitSigma, xc, yc):
    """beam.py:47-62 / :108-124."""
    nx, ny, npx = gridObj.getNxExt(), gridObj.getNyExt(), gridObj.getNpExt()
    dx, dy = gridObj.getDx(), gridObj.getDy()
    xm = np.asarray(gridObj.getXMesh())
    ym = np.asarray(gridObj.getYMesh())
    X = np.tile(xm, ny)
    Y = np.reshape(np.tile(ym, (nx, 1)).transpose(), npx)
    pts = ((radiusLimitSigma * radiusSigma + max(dx, dy)) ** 2 - X ** 2 - Y ** 2) > 0
    lim = radiusLimitSigma
    prof = np.zeros(npx)
    fx = (_norm_cdf(np.clip((X - xc + dx / 2.) / radiusSigma, -lim, lim)) -
          _norm_cdf(np.clip((X - xc - dx / 2.) / radiusSigma, -lim, lim)))
    fy = (_norm_cdf(np.clip((Y - yc + dy / 2.) / radiusSigma, -lim, lim)) -
          _norm_cdf(np.clip((Y - yc - dy / 2.) / radiusSigma, -lim, lim)))
    prof[pts] = (fx * fy)[pts]
    prof /= np.sum(prof)
    return prof


class _BeamBase:
    bunchLengthSigma = 0.1
    bunchLengthLimitSigma = 5

    def _init(self, gridObj, dt, nParticles, tBunchSpacing, radiusSigma):
        self.gridObj = gridObj
        self.dt = dt
        self.nParticles = nParticles
        self.charge = elementary_charge
        self.beamVelocity = C_LIGHT
        self.radiusSigma = radiusSigma
        self.radiusLimitSigma = 5
        self.xBeamCenter = 0.
        self.yBeamCenter = 0.
        self.tBunchSpacing = tBunchSpacing
        self.tBunchLengthSigma = self.bunchLengthSigma / self.beamVelocity
        self.qTransversalProfile = _profile(gridObj, radiusSigma, self.radiusLimitSigma,
                                            self.xBeamCenter, self.yBeamCenter)
        self._profile_dev = None

    def getTransversalProfile(self):
        return self.qTransversalProfile

    def getProfile(self):
        """Device copy of the transverse profile."""
        if self._profile_dev is None:
            self._profile_dev = L.to_dev(self.qTransversalProfile)
        return self._profile_dev

    def _long(self, tRed):
        lim = self.bunchLengthLimitSigma
        s = self.tBunchLengthSigma
        a = float(_norm_cdf(np.clip((tRed + self.dt / 2.) / s - lim, -lim, lim)))
        b = float(_norm_cdf(np.clip((tRed - self.dt / 2.) / s - lim, -lim, lim)))
        return (a - b) * self.nParticles / self.dt / self.beamVelocity * self.charge

    def bunchFraction(self, t):
        """Fraction of one bunch's protons that passes during the step starting at t (0 outside
        a bunch window): getScale(t) without the charge-per-length factor."""
        s = self.getScale(t)
        if s == 0.:
            return 0.
        return s * self.dt * self.beamVelocity / (self.nParticles * self.charge)

    @L.host_out
    def getCharge(self, t):
        s = self.getScale(t)
        if s == 0.:
            return 0
        return s * self.getProfile()


class LHCBeam(_BeamBase):
    """beam.py:13-79: LHC filling pattern (4 batches of 72 bunches, 8 empty), sigma 1 mm."""

    def __init__(self, gridObj, dt, nParticles=1.e11, tBunchSpacing=25.e-9):
        self._init(gridObj, dt, nParticles, 25.e-9, 0.001)
        self.circumference = 6900
        self.tRevolution = self.circumference / self.beamVelocity
        self.nBatches = 4
        self.nBunchesPerBatch = 72
        self.tBatchLength = self.nBunchesPerBatch * self.tBunchSpacing
        self.nEmptyBunches = 8
        self.tBatchSpacing = self.nEmptyBunches * self.tBunchSpacing
        self.tTrainLength = (self.nBatches * self.nBunchesPerBatch +
                             (self.nBatches - 1) * self.nEmptyBunches) * self.tBunchSpacing

    def getScale(self, t):
        cond = (np.mod(t, self.tBunchSpacing) < 2 * self.bunchLengthLimitSigma * self.tBunchLengthSigma
                and np.mod(t, self.tBatchLength + self.tBatchSpacing) < self.tBatchLength
                and np.mod(t, self.tRevolution) < self.tTrainLength)
        if not cond:
            return 0.
        return self._long(np.mod(np.mod(t, self.tRevolution), self.tBunchSpacing))


class LHCBeamEndless(_BeamBase):
    """beam.py:81-136: an endless train at tBunchSpacing, sigma 2 mm."""

    def __init__(self, gridObj, dt, nParticles=1.e11, tBunchSpacing=25.e-9):
        self._init(gridObj, dt, nParticles, tBunchSpacing, 0.002)

    def getScale(self, t):
        if not (np.mod(t, self.tBunchSpacing) < 2 * self.bunchLengthLimitSigma * self.tBunchLengthSigma):
            return 0.
        return self._long(np.mod(t, self.tBunchSpacing))


class GaussianBunchTrain(_BeamBase):
    """A train of nBunches Gaussian bunches (the reference hard-codes sigma and the filling
    pattern; this exposes them): transverse sigma radiusSigma, nParticles per bunch,
    spacing tBunchSpacing, rms length bunchLengthSigma (m)."""

    def __init__(self, gridObj, dt, nParticles=1.15e11, tBunchSpacing=25.e-9, radiusSigma=1.5e-3,
                 nBunches=4, bunchLengthSigma=0.1):
        self.bunchLengthSigma = bunchLengthSigma
        self._init(gridObj, dt, nParticles, tBunchSpacing, radiusSigma)
        self.nBunches = nBunches

    def getScale(self, t):
        if t >= self.nBunches * self.tBunchSpacing:
            return 0.
        if not (np.mod(t, self.tBunchSpacing) < 2 * self.bunchLengthLimitSigma * self.tBunchLengthSigma):
            return 0.
        return self._long(np.mod(t, self.tBunchSpacing))
