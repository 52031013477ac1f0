"""Particles: the reference's Particles class (particles.pyx:14-197) on the GPU.

Storage is one (6, capacity) structure-of-arrays tensor in HBM -- rows x, y, vx, vy, vz, w --
in FP64 (default, the reference's precision) or FP32.  getParticleData() returns the
(N, 6) transposed VIEW of that block, so callers that scale weights through it, as
particleManagement does in the reference (particleManagement.pyx:46,51), still alias the
stored data.  Every compute method calls a HIP kernel of liboecl; no CPU fallback exists.
"""
import ctypes

import numpy as np
import torch

from . import _lib as L
from .constants import elementary_charge, m_e, m_p


def _as_dev(a, dtype):
    if isinstance(a, torch.Tensor):
        t = a
        if t.device.type != 'cuda':
            t = t.to(L.device())
        if t.dtype != dtype:
            t = t.to(dtype)
        return t.contiguous()
    return L.to_dev(np.asarray(a), dtype).contiguous()


class Particles:
    """particles.pyx:16-35.  dtype: torch.float64 (reference precision) or torch.float32.

    ordering: 'reference' reproduces the reference's particle order after
    addAndRemoveParticles exactly (including its capacity-dependent choice between the
    in-place and the new-array rule); 'fast' always uses the hole-filling rule, which moves
    only the particles that must move and yields the same particle set.

    orderedDeposit: chargeToGrid sums every node's contributions in particle index order, as
    the reference's sequential loop does (particles.pyx:336-351), so the charge on the grid is
    bit-identical to it; the default atomic deposit agrees to rounding (1e-13 relative)."""

    def __init__(self, particleSpecies, particleType, dtype=torch.float64, ordering='reference',
                 orderedDeposit=False):
        L.require_gpu()
        L.lib()
        self.macroParticleCount = 0
        if particleSpecies == 'electrons':
            self.particleMass = m_e
            self.particleCharge = -elementary_charge
        elif particleSpecies == 'protons':
            self.particleMass = m_p
            self.particleCharge = elementary_charge
        else:
            raise NotImplementedError("Not yet implemented.")
        if particleType == 'variableWeight':
            self.nCoords = 6
        else:
            raise NotImplementedError("Not yet implemented.")
        if ordering not in ('reference', 'fast'):
            raise ValueError("ordering must be 'reference' or 'fast'")
        self.dtype = dtype
        self.code = L.dtype_code(dtype)
        self.ordering = ordering
        self.orderedDeposit = bool(orderedDeposit)
        self._ordWork = None
        self.dt = 0.
        self.bConst = np.zeros(3)
        self._refCapacity = 0
        dev = L.device()
        self._data = torch.empty((6, 0), dtype=dtype, device=dev)
        self._spare = None
        self.chargeOnGrid = torch.empty(1, dtype=torch.float64, device=dev)
        self._alloc_scratch(0)
        self._count = torch.zeros(4, dtype=torch.int32, device=dev)

    # ---------------------------------------------------------------- storage
    @property
    def capacity(self):
        return self._data.shape[1]

    def _alloc_scratch(self, cap):
        dev = L.device()
        self.inCell = torch.empty(cap, dtype=torch.int32, device=dev)
        self._w4 = torch.empty((4, cap), dtype=self.dtype, device=dev)
        self.eAtParticles = torch.empty((cap, 2), dtype=self.dtype, device=dev)
        self.bAtParticles = torch.empty((cap, 3), dtype=self.dtype, device=dev)
        self.isInside = torch.empty(cap, dtype=torch.uint8, device=dev)

    def reserve(self, cap):
        """Grow the device capacity to at least cap particles (keeps the data)."""
        if cap <= self.capacity:
            return
        new = torch.empty((6, int(cap)), dtype=self.dtype, device=L.device())
        n = self.macroParticleCount
        if n:
            new[:, :n].copy_(self._data[:, :n])
        self._data = new
        self._spare = None
        self._alloc_scratch(int(cap))

    def _spare_block(self):
        if self._spare is None or self._spare.shape != self._data.shape:
            self._spare = torch.empty_like(self._data)
        return self._spare

    # ---------------------------------------------------------------- reference API
    def setBConst(self, bConst):
        """particles.pyx:123-124."""
        self.bConst = np.asarray(bConst, dtype=np.float64).reshape(3).copy()

    def setParticleData(self, particleData):
        """particles.pyx:126-140: the data becomes the particle set; its row count is the
        reference's capacity."""
        pd = _as_dev(particleData, self.dtype)
        if pd.ndim != 2 or pd.shape[1] != 6:
            raise ValueError('particleData must have shape (N, 6)')
        n = pd.shape[0]
        cap = max(int(n * 1.25) + 1024, 1024)
        self._data = torch.empty((6, cap), dtype=self.dtype, device=L.device())
        self._data[:, :n].copy_(pd.t())
        self._spare = None
        self.macroParticleCount = n
        self._refCapacity = n
        self._alloc_scratch(cap)

    @L.host_out
    def getParticleData(self):
        """(N, 6) view aliasing the device storage (particles.pyx:142-143)."""
        return self._data[:, :self.macroParticleCount].t()

    @L.host_out
    def getFullParticleData(self):
        return self._data.t()

    def getNCoords(self):
        return self.nCoords

    def getMacroParticleCount(self):
        return self.macroParticleCount

    def setMacroParticleCount(self, macroParticleCount):
        self.macroParticleCount = int(macroParticleCount)

    @L.host_out
    def getChargeOnGrid(self):
        return self.chargeOnGrid

    def setChargeOnGrid(self, chargeOnGrid):
        self.chargeOnGrid = _as_dev(chargeOnGrid, torch.float64)

    def getParticleCharge(self):
        return self.particleCharge

    def getParticleMass(self):
        return self.particleMass

    @L.host_out
    def getInCell(self):
        return self.inCell[:self.macroParticleCount]

    @L.host_out
    def getFullIsInside(self):
        return self.isInside

    @L.host_out
    def getIsInside(self):
        return self.isInside[:self.macroParticleCount]

    @L.host_out
    def getWeightLowXLowY(self):
        return self._w4[0, :self.macroParticleCount]

    @L.host_out
    def getWeightUpXLowY(self):
        return self._w4[1, :self.macroParticleCount]

    @L.host_out
    def getWeightLowXUpY(self):
        return self._w4[2, :self.macroParticleCount]

    @L.host_out
    def getWeightUpXUpY(self):
        return self._w4[3, :self.macroParticleCount]

    @L.host_out
    def getEAtParticles(self):
        return self.eAtParticles[:self.macroParticleCount]

    @L.host_out
    def getBAtParticles(self):
        return self.bAtParticles

    def getDt(self):
        return self.dt

    def getMeanWeight(self):
        """particles.pyx:39-47 (FP64 sum of the weights divided by N)."""
        from .diagnostics import weight_sum
        return weight_sum(self) / self.macroParticleCount

    def calcGridWeights(self, gridObj):
        """particles.pyx:52-56."""
        g = L.grid_desc(gridObj)
        n = self.macroParticleCount
        L.check(L.lib().oecl_calc_grid_weights(
            ctypes.byref(g), L.i32(self.code), L.ptr(self._data), L.u64(self.capacity), L.u32(n),
            L.ptr(self.inCell), L.ptr(self._w4), L.u64(self.capacity), L.stream()),
            'calcGridWeights')

    def chargeToGrid(self, gridObj):
        """particles.pyx:70-77."""
        npx = gridObj.getNpExt()
        if self.chargeOnGrid.shape[0] < npx:
            self.chargeOnGrid = torch.empty(npx, dtype=torch.float64, device=L.device())
        self._charge(gridObj, self.chargeOnGrid)

    def _charge(self, gridObj, rho):
        """The charge of the particles on the grid into rho, from the weights of the last
        calcGridWeights, atomic or in the reference's order (orderedDeposit)."""
        g = L.grid_desc(gridObj)
        n = self.macroParticleCount
        if self.orderedDeposit:
            need = int(L.lib().oecl_charge_ordered_work_bytes(L.u32(n), L.u32(gridObj.getNpExt()))) // 8 + 8
            if self._ordWork is None or self._ordWork.numel() < need:
                self._ordWork = torch.empty(int(need * 1.25), dtype=torch.float64, device=L.device())
            L.check(L.lib().oecl_charge_to_grid_ordered(
                ctypes.byref(g), L.i32(self.code), L.ptr(self._data), L.u64(self.capacity), L.u32(n),
                L.ptr(self.inCell), L.ptr(self._w4), L.u64(self.capacity), L.f64(self.particleCharge),
                L.ptr(rho), L.ptr(self._ordWork), L.stream()), 'chargeToGrid (ordered)')
            return
        L.check(L.lib().oecl_charge_to_grid(
            ctypes.byref(g), L.i32(self.code), L.ptr(self._data), L.u64(self.capacity),
            L.u32(n), L.ptr(self.inCell), L.ptr(self._w4),
            L.u64(self.capacity), L.f64(self.particleCharge), L.ptr(rho),
            L.stream()), 'chargeToGrid')

    def _field(self, gridObj, field, dims, out):
        f = _as_dev(field, torch.float64)
        if f.numel() < dims * gridObj.getNpExt():
            raise ValueError('field has %d values, need %d' % (f.numel(), dims * gridObj.getNpExt()))
        g = L.grid_desc(gridObj)
        L.check(L.lib().oecl_field_to_particles(
            ctypes.byref(g), L.i32(self.code), L.ptr(f), L.ptr(self.inCell), L.ptr(self._w4),
            L.u64(self.capacity), L.u32(self.macroParticleCount), L.i32(dims), L.ptr(out),
            L.stream()), 'fieldToParticles')

    def eFieldToParticles(self, gridObj, eAtGridPoints):
        """particles.pyx:58-62."""
        self._field(gridObj, eAtGridPoints, 2, self.eAtParticles)

    def bFieldToParticles(self, gridObj, bAtGridPoints):
        """particles.pyx:64-68."""
        self._field(gridObj, bAtGridPoints, 3, self.bAtParticles)

    def borisPush(self, dt, typeBField=2):
        """particles.pyx:79-92."""
        self.dt = dt
        b = (ctypes.c_double * 3)(*self.bConst.tolist())
        L.check(L.lib().oecl_boris_push(
            L.i32(self.code), L.ptr(self._data), L.u64(self.capacity),
            L.u32(self.macroParticleCount), L.ptr(self.eAtParticles), L.ptr(self.bAtParticles),
            b, L.i32(typeBField), L.f64(dt), L.f64(self.particleCharge / self.particleMass),
            L.stream()), 'borisPush')

    def count_removed(self, keep, n):
        """Number of zero flags in keep[:n] (device scan, one host read)."""
        work = torch.empty(int(L.lib().oecl_scan_work_bytes(n)) // 4 + 8, dtype=torch.int32,
                           device=L.device())
        L.check(L.lib().oecl_save_absorbed(
            L.i32(self.code), L.ptr(self._data), L.u64(self.capacity), L.u32(n), L.ptr(keep),
            ctypes.c_void_p(0), L.u32(0), L.ptr(self._count), L.ptr(work), L.stream()), 'count')
        return int(self._count[0].item())

    def addAndRemoveParticles(self, addParticleData, keepParticles):
        """particles.pyx:94-121."""
        n = self.macroParticleCount
        add = _as_dev(addParticleData, self.dtype).reshape(-1, 6)
        na = add.shape[0]
        keep = keepParticles
        if not (isinstance(keep, torch.Tensor) and keep.dtype == torch.uint8 and keep.is_cuda):
            keep = _as_dev(keepParticles, torch.int64).to(torch.uint8)
        if keep.numel() < n:
            raise ValueError('keepParticles has fewer entries than particles')
        removed = self.count_removed(keep, n)
        kc = n - removed
        new = kc + na
        if na == 0 and kc == n:
            return
        if self.ordering == 'fast':
            mode = 3
        elif na == 0:
            mode = 0
        elif new > self._refCapacity:
            mode = 2
            self._refCapacity = int(new * 1.1)
        else:
            mode = 1
        if new > self.capacity:
            keep = keep.clone()
            self.reserve(int(new * 1.25) + 1024)
        spare = self._spare_block() if mode in (1, 2) else None
        work = torch.empty(int(L.lib().oecl_compact_work_bytes(n, na)) // 4 + 8,
                           dtype=torch.int32, device=L.device())
        addc = add.contiguous()
        L.check(L.lib().oecl_add_and_remove(
            L.i32(self.code), L.ptr(self._data), L.ptr(spare), L.u64(self.capacity), L.u32(n),
            L.ptr(keep), L.ptr(addc), L.u32(na), L.i32(mode), L.ptr(self._count), L.ptr(work),
            L.stream()), 'addAndRemoveParticles')
        if mode == 2:
            self._data, self._spare = self._spare, self._data
        self.macroParticleCount = new

    def sortByColumn(self, sortByColumn):
        """particles.pyx:49-50: order the particles by one column (device radix sort)."""
        from .particleManagement import sort_by_column
        sort_by_column(self, int(sortByColumn))
