"""Particle walls (particleBoundary.pyx:19-481) on the GPU.

AbsorbRectangular, AbsorbElliptical and AbsorbArbitraryCutCell keep the reference's methods:
indexInside, saveAbsorbed, calculateInteractionPoint, calculateNormalVectors, isInside and the
getters.  Absorbed particles are an (nAbs, 6) device tensor in the particles' dtype.
"""
import ctypes

import numpy as np
import torch

from . import _lib as L


def _wall_arrays(gridObj):
    """Device copies of the cut-cell arrays the wall kernels read (cached on the grid)."""
    cache = getattr(gridObj, '_oecl_wall_arrays', None)
    if cache is None:
        dev = L.device()
        cache = dict(
            insideFaces=torch.as_tensor(np.ascontiguousarray(gridObj.getInsideFaces(), np.uint16)
                                        .view(np.int16)).to(dev),
            boundaryPoints=torch.as_tensor(np.ascontiguousarray(
                np.asarray(gridObj.getBoundaryPoints(), np.float64).reshape(-1, 2))).to(dev),
            cutCellPointsInd=torch.as_tensor(np.ascontiguousarray(
                gridObj.getCutCellPointsInd(), np.int32)).to(dev),
            cutCellNormalVectors=torch.as_tensor(np.ascontiguousarray(
                gridObj.getCutCellNormalVectors(), np.float64)).to(dev))
        if cache['boundaryPoints'].numel() == 0:
            cache['boundaryPoints'] = torch.zeros((1, 2), dtype=torch.float64, device=dev)
        gridObj._oecl_wall_arrays = cache
    return cache


class ParticleBoundary:
    """particleBoundary.pyx:19-74."""
    kind = None

    def __init__(self, gridObj, particlesObj):
        self.gridObj = gridObj
        self.particlesObj = particlesObj
        self.nx = gridObj.getNxExt()
        self.ny = gridObj.getNyExt()
        self.np = gridObj.getNpExt()
        self.lx = gridObj.getLx()
        self.ly = gridObj.getLy()
        self.dx = gridObj.getDx()
        self.dy = gridObj.getDy()
        dev = L.device()
        dt = particlesObj.dtype
        self.absorbedParticles = torch.empty((1, 6), dtype=dt, device=dev)
        self.normalVectors = torch.empty((1, 2), dtype=dt, device=dev)
        self.remainingTimeStep = torch.empty(1, dtype=dt, device=dev)
        self.absorbedMacroParticleCount = 0
        self.clipWarnings = 0
        self.notFoundWarnings = 0
        self._dev = torch.zeros(4, dtype=torch.int32, device=dev)
        self._arrays = _wall_arrays(gridObj) if self.kind == L.WALL_CUTCELL else None
        self.desc = self._make_desc()

    def _make_desc(self):
        g = self.gridObj
        d = L.WallDesc()
        d.kind = self.kind
        d.lx, d.ly, d.dx, d.dy = g.getLx(), g.getLy(), g.getDx(), g.getDy()
        d.lxHalfExt, d.lyHalfExt = 0.5 * g.getLxExt(), 0.5 * g.getLyExt()
        d.nx = g.getNxExt()
        if self._arrays is not None:
            a = self._arrays
            d.insideFaces = a['insideFaces'].data_ptr()
            d.boundaryPoints = a['boundaryPoints'].data_ptr()
            d.cutCellPointsInd = a['cutCellPointsInd'].data_ptr()
            d.cutCellNormalVectors = a['cutCellNormalVectors'].data_ptr()
        return d

    def indexInside(self):
        """indexInside of the concrete wall (writes particlesObj.isInside, 1 = keep)."""
        p = self.particlesObj
        self._dev.zero_()
        L.check(L.lib().oecl_index_inside(
            ctypes.byref(self.desc), L.i32(p.code), L.ptr(p._data), L.u64(p.capacity),
            L.u32(p.macroParticleCount), L.ptr(p.isInside), L.ptr(self._dev[1:2]), L.stream()),
            'indexInside')

    def saveAbsorbed(self):
        """particleBoundary.pyx:38-65: count the absorbed particles (a host read, as the
        reference's getAbsorbedMacroParticleCount is) and copy them in index order."""
        p = self.particlesObj
        n = p.macroParticleCount
        work = torch.empty(int(L.lib().oecl_scan_work_bytes(n)) // 4 + 8, dtype=torch.int32,
                           device=L.device())
        cap = self.absorbedParticles.shape[0]
        L.check(L.lib().oecl_save_absorbed(
            L.i32(p.code), L.ptr(p._data), L.u64(p.capacity), L.u32(n), L.ptr(p.isInside),
            L.ptr(self.absorbedParticles), L.u32(cap), L.ptr(self._dev), L.ptr(work), L.stream()),
            'saveAbsorbed')
        na = int(self._dev[0].item())
        if na > cap:
            self.absorbedParticles = torch.empty((int(na * 1.1) + 1, 6), dtype=p.dtype,
                                                 device=L.device())
            L.check(L.lib().oecl_save_absorbed(
                L.i32(p.code), L.ptr(p._data), L.u64(p.capacity), L.u32(n), L.ptr(p.isInside),
                L.ptr(self.absorbedParticles), L.u32(self.absorbedParticles.shape[0]),
                L.ptr(self._dev), L.ptr(work), L.stream()), 'saveAbsorbed')
        self.absorbedMacroParticleCount = na
        if self.kind == L.WALL_CUTCELL and int(self._dev[1].item()) != 0:
            self.clipWarnings += 1

    def _interaction(self):
        p = self.particlesObj
        na = self.absorbedMacroParticleCount
        if self.remainingTimeStep.shape[0] < na:
            self.remainingTimeStep = torch.empty(int(na * 1.1) + 1, dtype=p.dtype, device=L.device())
        if self.normalVectors.shape[0] < na:
            self.normalVectors = torch.empty((int(na * 1.1) + 1, 2), dtype=p.dtype,
                                             device=L.device())
        self._dev[2:3].zero_()
        L.check(L.lib().oecl_wall_interaction(
            ctypes.byref(self.desc), L.i32(p.code), L.ptr(self.absorbedParticles), L.u32(na),
            L.f64(p.getDt()), L.ptr(self.remainingTimeStep), L.ptr(self.normalVectors),
            L.ptr(self._dev[2:3]), L.stream()), 'calculateInteractionPoint')
        self._normals_done = True

    def calculateInteractionPoint(self):
        """Moves the absorbed particles back to the wall.  The device kernel also writes the
        wall normal at the impact point, which calculateNormalVectors then returns."""
        self._interaction()

    def calculateNormalVectors(self):
        """Normals at the impact points (computed with the interaction point)."""
        if not getattr(self, '_normals_done', False):
            self._interaction()
        self._normals_done = False

    @L.host_out
    def getAbsorbedParticles(self):
        return self.absorbedParticles[:self.absorbedMacroParticleCount]

    def getAbsorbedMacroParticleCount(self):
        return self.absorbedMacroParticleCount

    @L.host_out
    def getNormalVectors(self):
        return self.normalVectors[:self.absorbedMacroParticleCount]

    @L.host_out
    def getRemainingTimeStep(self):
        return self.remainingTimeStep[:self.absorbedMacroParticleCount]


class AbsorbRectangular(ParticleBoundary):
    """particleBoundary.pyx:77-221.

    cornerFix: where a particle left through a corner region (both coordinates out), the
    reference gives four of the eight cases the normal of the other wall
    (particleBoundary.pyx:154-187).  cornerFix=False keeps that, for parity; cornerFix=True
    takes the normal of the wall crossed first (the one whose crossing time is the larger,
    which is also the time the interaction point is moved back by)."""
    kind = L.WALL_RECT

    def __init__(self, gridObj, particlesObj, cornerFix=False):
        self.cornerFix = bool(cornerFix)
        super().__init__(gridObj, particlesObj)

    def _make_desc(self):
        d = super()._make_desc()
        d.rectCornerFix = 1 if self.cornerFix else 0
        return d

    def isInside(self, x, y):
        lxh, lyh = self.lx * 0.5, self.ly * 0.5
        return 0 if (x < -lxh or x > lxh or y < -lyh or y > lyh) else 1


class AbsorbElliptical(ParticleBoundary):
    """particleBoundary.pyx:225-297."""
    kind = L.WALL_ELLIP

    def isInside(self, x, y):
        return 0 if (x ** 2 * (4. / self.lx ** 2) + y ** 2 * (4. / self.ly ** 2)) >= 1. else 1


class AbsorbArbitraryCutCell(ParticleBoundary):
    """particleBoundary.pyx:302-481."""
    kind = L.WALL_CUTCELL

    def isInside(self, x, y):
        g = self.gridObj
        ix = int((x + 0.5 * g.getLxExt()) * (1. / g.getDx()))
        iy = int((y + 0.5 * g.getLyExt()) * (1. / g.getDy()))
        c = ix + iy * g.getNxExt()
        f = g.getInsideFaces()[c]
        if f == 8:
            return 1
        if f == 0:
            return 0
        bp = g.getBoundaryPoints()
        a, b = g.getCutCellPointsInd()[c]
        return 1 if ((bp[a, 0] - x) * (bp[b, 1] - y) > (bp[b, 0] - x) * (bp[a, 1] - y)) else 0
