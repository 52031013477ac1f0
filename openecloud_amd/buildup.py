"""The buildup time loop (the loop the reference writes out in example01_basic.py:52-97 and
its siblings) in two forms:

  Buildup.step_reference()  calls the reference's public methods one by one, in the
                            reference's order (calcGridWeights, chargeToGrid, solve,
                            eFieldToParticles, [bFieldToParticles], borisPush, indexInside,
                            saveAbsorbed, interaction point, normals, generateSecondaries,
                            addAndRemoveParticles);
  Buildup.step()            the plain fused GPU step: one kernel gathers E (and B), pushes,
                            tests the wall and deposits the charge of the surviving particles
                            for the next step; the absorbed particles are collected, moved to
                            the wall, emitted from, and the compaction moves only particles
                            that must move.  Same physics, same particle set.
Both use the same device objects, so they can be mixed.
"""
import ctypes

import numpy as np
import torch

from . import _lib as L
from . import randomGen

_VP = ctypes.c_void_p


class StepArgs(ctypes.Structure):
    """Mirror of oecl_step_args (include/oecl.h)."""
    _fields_ = [('dtype', ctypes.c_int32), ('data', _VP), ('stride', ctypes.c_uint64),
                ('nDev', _VP), ('isInside', _VP), ('rho', _VP), ('q', _VP), ('profile', _VP),
                ('beamScale', _VP), ('nScale', ctypes.c_uint32), ('phi', _VP), ('E', _VP),
                ('e2nRowPtr', _VP), ('e2nCol', _VP), ('e2nVal', _VP), ('bField', _VP),
                ('b0', ctypes.c_double), ('b1', ctypes.c_double), ('b2', ctypes.c_double),
                ('typeB', ctypes.c_int32), ('dt', ctypes.c_double), ('qm', ctypes.c_double),
                ('qe', ctypes.c_double), ('absorbed', _VP), ('normals', _VP), ('remaining', _VP),
                ('absCap', ctypes.c_uint32), ('nSec', _VP), ('secOff', _VP), ('impact', _VP),
                ('secondaries', _VP), ('secCap', ctypes.c_uint32), ('emit', ctypes.c_int32),
                ('seed0', ctypes.c_uint32), ('seed1', ctypes.c_uint32), ('stepDev', _VP),
                ('stats', _VP), ('countHist', _VP), ('histCap', ctypes.c_uint32), ('work', _VP),
                ('evPushBegin', _VP), ('evPushEnd', _VP), ('photoCount', _VP), ('photoWeight', _VP),
                ('nPhoto', ctypes.c_uint32), ('photoE0', ctypes.c_double),
                ('photoSigmaE', ctypes.c_double), ('photoSlots', _VP), ('photoCap', ctypes.c_uint32),
                ('secem', _VP), ('Eil', _VP), ('externalPhi', ctypes.c_int32), ('Bil', _VP),
                ('e2nKind', _VP), ('e2nIrr', _VP), ('nIrr', ctypes.c_uint32), ('e2nCx', ctypes.c_double),
                ('e2nCy', ctypes.c_double), ('pushArith', ctypes.c_int32), ('Eil32', _VP), ('Bil32', _VP),
                ('nearCells', ctypes.c_double), ('depSpan', ctypes.c_uint32), ('depTile', ctypes.c_int32)]


PUSH_ARITHMETIC = ('fp64', 'mixed')


def push_arithmetic(dtype, mode):
    """The push arithmetic of the one-call step for particles of `dtype`:
      'fp64'   weights, gathers, Boris rotation, wall test and deposit in FP64 registers (the
               only form for FP64 particles);
      'mixed'  FP32 particles: gather and Boris push in FP32 from FP32 copies of E and B, the
               charge split, run sums and rho in FP64, and the FP64 form for every particle
               within `nearCells` cells of the wall before or after its push (emitted electrons
               and grazing orbits), so absorption decisions are the FP64 ones.
    None gives 'mixed' for FP32 particles and 'fp64' for FP64 ones."""
    fp32 = dtype == torch.float32
    if mode is None:
        return 'mixed' if fp32 else 'fp64'
    if mode not in PUSH_ARITHMETIC:
        raise ValueError('pushArithmetic must be one of %s' % (PUSH_ARITHMETIC,))
    if mode == 'mixed' and not fp32:
        raise ValueError('the mixed push needs FP32 particles')
    return mode


def set_push_fields(obj, a, nNodes, mode, nearCells):
    """Allocates on obj the field copies the push of `mode` reads (nNodes: grid nodes of all
    runs) and points the step arguments a at them; obj._bil is the (Bx, By) copy or None."""
    dev = L.device()
    if mode == 'mixed':
        if not nearCells >= 1.:
            raise ValueError('nearCells must be at least 1')
        obj._eil = None
        obj._eil32 = torch.zeros(2 * nNodes, dtype=torch.float32, device=dev)
        obj._bil32 = obj._bil.float().contiguous() if obj._bil is not None else None
        a.Eil = None
        a.Eil32 = obj._eil32.data_ptr()
        a.Bil32 = obj._bil32.data_ptr() if obj._bil32 is not None else None
        a.pushArith = 1
        a.nearCells = float(nearCells)
    else:
        obj._eil = torch.zeros(2 * nNodes, dtype=torch.float64, device=dev)
        obj._eil32 = obj._bil32 = None
        a.Eil = obj._eil.data_ptr()
        a.Eil32 = None
        a.Bil32 = None
        a.pushArith = 0
        a.nearCells = 0.


def set_e2n_classes(a, ps):
    """Step arguments for the row-class evaluation of E = phiToE phi (PoissonSolver.e2nKind)."""
    a.e2nKind = ps.e2nKind.data_ptr()
    a.e2nIrr = ps.e2nIrr.data_ptr() if ps.e2nIrr.numel() else None
    a.nIrr = int(ps.e2nIrr.numel())
    a.e2nCx, a.e2nCy = ps.e2nCx, ps.e2nCy


def b_layout(bField, npx):
    """The (Bx, By)-per-node copy of a grid field whose Bz is zero everywhere (the step's Bil);
    None when Bz is not zero."""
    if bField is None or bool((bField[2 * npx:3 * npx] != 0).any()):
        return None
    return torch.stack([bField[:npx], bField[npx:2 * npx]], dim=1).contiguous()


ST_ABS_OVERFLOW, ST_SEC_OVERFLOW, ST_CAP_OVERFLOW, ST_PHOTO_OVERFLOW = 1, 2, 4, 8


class Buildup:

    def __init__(self, gridObj, particlesObj, boundaryObj, emitterObj, poissonObj, beamObj=None,
                 dt=1e-11, typeBField=0, bAtGridPoints=None, secondaryEmission=True,
                 photoEmitter=None):
        self.g = gridObj
        self.photo = photoEmitter
        self.p = particlesObj
        self.pb = boundaryObj
        self.em = emitterObj
        self.ps = poissonObj
        self.beam = beamObj
        self.dt = float(dt)
        self.typeB = int(typeBField)
        self.secondaryEmission = secondaryEmission
        self.bField = None if bAtGridPoints is None else L.to_dev(np.asarray(bAtGridPoints, np.float64))
        if self.typeB == 2 and self.bField is None:
            raise ValueError('typeBField=2 needs bAtGridPoints')
        self.it = 0
        dev = L.device()
        npx = gridObj.getNpExt()
        self.gd = L.grid_desc(gridObj)
        self.rho = torch.zeros(npx, dtype=torch.float64, device=dev)
        self.q = torch.zeros(npx, dtype=torch.float64, device=dev)
        self._rho_valid = False
        self._cnt = torch.zeros(8, dtype=torch.int32, device=dev)
        self._pin = torch.zeros(8, dtype=torch.int32).pin_memory()
        self._work = None
        self.history = {'macro': [], 'absorbed': [], 'emitted': [], 'photo': [], 'charge': []}

    # ------------------------------------------------------------------ helpers
    def _work_buf(self, n, nadd=0):
        need = int(max(L.lib().oecl_compact_work_bytes(n, nadd), L.lib().oecl_scan_work_bytes(n))) // 4 + 64
        if self._work is None or self._work.numel() < need:
            self._work = torch.empty(int(need * 1.25), dtype=torch.int32, device=L.device())
        return self._work

    def deposit_all(self):
        """rho = charge of all current particles (the reference's chargeToGrid; in its
        summation order when the particles use orderedDeposit)."""
        p = self.p
        if p.orderedDeposit:
            p.calcGridWeights(self.g)
            p._charge(self.g, self.rho)
            self._rho_valid = True
            return
        self.rho.zero_()
        L.check(L.lib().oecl_deposit_add(ctypes.byref(self.gd), L.i32(p.code), L.ptr(p._data),
                                         L.u64(p.capacity), L.u32(p.macroParticleCount),
                                         L.f64(p.particleCharge), L.ptr(self.rho), L.stream()),
                'deposit')
        self._rho_valid = True

    def invalidate_charge(self):
        """Call after changing the particles outside step() (e.g. particle management)."""
        self._rho_valid = False

    def _rhs(self, t):
        if self.beam is None:
            return self.rho
        s = self.beam.getScale(t)
        if s == 0.:
            return self.rho
        L.check(L.lib().oecl_axpy_rhs(L.u32(self.g.getNpExt()), L.ptr(self.rho), L.f64(s),
                                      L.ptr(self.beam.getProfile()), L.ptr(self.q), L.stream()),
                'rhs')
        return self.q

    # ------------------------------------------------------------------ reference order
    def step_reference(self):
        g, p, pb, ps = self.g, self.p, self.pb, self.ps
        t = self.it * self.dt
        p.calcGridWeights(g)
        p.chargeToGrid(g)
        q = p.chargeOnGrid
        if self.beam is not None:
            s = self.beam.getScale(t)
            if s != 0.:
                q = q + s * self.beam.getProfile()
        ps._solve(q)
        p.eFieldToParticles(g, ps.eAtGridPoints)
        if self.typeB == 2:
            p.bFieldToParticles(g, self.bField)
        p.borisPush(self.dt, typeBField=self.typeB)
        pb.indexInside()
        pb.saveAbsorbed()
        na = pb.getAbsorbedMacroParticleCount()
        nsec = 0
        sec = torch.empty((0, 6), dtype=p.dtype, device=L.device())
        if na > 0:
            pb.calculateInteractionPoint()
            pb.calculateNormalVectors()
            if self.secondaryEmission:
                sec = self.em._generate(step=self.it)
            nsec = sec.shape[0]
        ph = self.photo.generate(self.it) if self.photo is not None else sec[:0]
        if na > 0 or ph.shape[0] > 0:
            p.addAndRemoveParticles(torch.cat([sec, ph]) if ph.shape[0] else sec, p.isInside[:p.macroParticleCount])
        self._rho_valid = False
        self._record(na, nsec, ph.shape[0])
        self.it += 1

    # ------------------------------------------------------------------ fused plain step
    def step(self):
        g, p, pb, ps, em = self.g, self.p, self.pb, self.ps, self.em
        t = self.it * self.dt
        if not self._rho_valid or p.orderedDeposit:
            self.deposit_all()
        E = ps._solve(self._rhs(t))
        n = p.macroParticleCount
        s = L.stream()
        self.rho.zero_()
        b = (ctypes.c_double * 3)(*p.bConst.tolist())
        L.check(L.lib().oecl_push_fused(
            ctypes.byref(self.gd), ctypes.byref(pb.desc), L.i32(p.code), L.ptr(p._data),
            L.u64(p.capacity), L.u32(n), L.ptr(E), L.ptr(self.bField), b, L.i32(self.typeB),
            L.f64(self.dt), L.f64(p.particleCharge / p.particleMass), L.f64(p.particleCharge),
            L.ptr(p.isInside), L.ptr(self.rho), s), 'push')
        p.dt = self.dt
        # absorbed particles, in index order
        work = self._work_buf(n)
        cap = pb.absorbedParticles.shape[0]
        L.check(L.lib().oecl_save_absorbed(
            L.i32(p.code), L.ptr(p._data), L.u64(p.capacity), L.u32(n), L.ptr(p.isInside),
            L.ptr(pb.absorbedParticles), L.u32(cap), L.ptr(self._cnt), L.ptr(work), s), 'absorb')
        na = int(self._cnt[0].item())
        if na > cap:
            pb.absorbedParticles = torch.empty((int(na * 1.5) + 1024, 6), dtype=p.dtype,
                                               device=L.device())
            L.check(L.lib().oecl_save_absorbed(
                L.i32(p.code), L.ptr(p._data), L.u64(p.capacity), L.u32(n), L.ptr(p.isInside),
                L.ptr(pb.absorbedParticles), L.u32(pb.absorbedParticles.shape[0]),
                L.ptr(self._cnt), L.ptr(work), s), 'absorb')
        pb.absorbedMacroParticleCount = na
        nsec = 0
        sec = None
        if na > 0:
            pb._interaction()
            if self.secondaryEmission:
                sec = em._generate(step=self.it)
                nsec = sec.shape[0]
        nph = 0
        if self.photo is not None:
            ph = self.photo.generate(self.it)
            nph = ph.shape[0]
            if nph:
                sec = ph if not nsec else torch.cat([sec, ph])
        nadd = nsec + nph
        if nadd:
            L.check(L.lib().oecl_deposit_add_aos(
                ctypes.byref(self.gd), L.i32(p.code), L.ptr(sec), L.u32(nadd),
                L.f64(p.particleCharge), L.ptr(self.rho), s), 'deposit secondaries')
        if na > 0 or nadd > 0:
            nsec = nadd
            new = n - na + nsec
            if new > p.capacity:
                keep = p.isInside[:n].clone()
                p.reserve(int(new * 1.25) + 1024)
            else:
                keep = p.isInside
            work = self._work_buf(n, nsec)
            spare = p._spare_block() if p.ordering == 'reference' else None
            mode = 3
            if p.ordering == 'reference':
                if nsec == 0:
                    mode = 0
                elif new > p._refCapacity:
                    mode = 2
                    p._refCapacity = int(new * 1.1)
                else:
                    mode = 1
            L.check(L.lib().oecl_add_and_remove(
                L.i32(p.code), L.ptr(p._data), L.ptr(spare), L.u64(p.capacity), L.u32(n),
                L.ptr(keep), L.ptr(sec), L.u32(nsec), L.i32(mode), L.ptr(self._cnt[1:2]),
                L.ptr(work), s), 'compact')
            if mode == 2:
                p._data, p._spare = p._spare, p._data
            p.macroParticleCount = new
        self._rho_valid = True
        self._record(na, nsec - nph, nph)
        self.it += 1

    # ------------------------------------------------------------------ native step
    def native_setup(self, absCap=None, secCap=None, maxSteps=100000, pushArithmetic=None, nearCells=2.):
        """Allocates the device state of the one-call step (oecl_step_run).  The beam scale of
        every step up to maxSteps is tabulated now (it depends on the step index only).
        pushArithmetic: 'fp64' or 'mixed' (push_arithmetic; None: 'mixed' for FP32 particles),
        nearCells the wall margin of the mixed push."""
        self.pushArithmetic = push_arithmetic(self.p.dtype, pushArithmetic)
        p, pb, ps = self.p, self.pb, self.ps
        if ps.method != 'direct':
            raise ValueError('the native step needs PoissonSolver(method="direct"); method "%s" runs through '
                             'step() and step_reference()' % ps.method)
        dev = L.device()
        n = p.macroParticleCount
        cap = p.capacity
        absCap = int(absCap or max(65536, cap // 8))
        secCap = int(secCap or 2 * absCap)
        self.absCap, self.secCap, self.maxSteps = absCap, secCap, int(maxSteps)
        pb.absorbedParticles = torch.empty((absCap, 6), dtype=p.dtype, device=dev)
        pb.normalVectors = torch.empty((absCap, 2), dtype=p.dtype, device=dev)
        pb.remainingTimeStep = torch.empty(absCap, dtype=p.dtype, device=dev)
        self._nsec = torch.empty(absCap, dtype=torch.int32, device=dev)
        self._secoff = torch.empty(absCap + 1, dtype=torch.int32, device=dev)
        self._impact = torch.empty(4 * absCap, dtype=torch.float64, device=dev)
        self._sec = torch.empty((secCap, 6), dtype=p.dtype, device=dev)
        self._nDev = torch.tensor([n], dtype=torch.int32, device=dev)
        self._stepDev = torch.tensor([self.it], dtype=torch.int32, device=dev)
        self._stats = torch.zeros(8, dtype=torch.int32, device=dev)
        nTab = self.it + self.maxSteps   # tables are indexed by the absolute step
        self._tabEnd = nTab              # steps at or past this index would read past the tables
        self._hist = torch.zeros(nTab, dtype=torch.int32, device=dev)
        self._stepwork = torch.empty(int(L.lib().oecl_step_work_bytes(L.u64(cap), L.u32(absCap))) // 8 + 64,
                                     dtype=torch.float64, device=dev)
        if self.beam is not None:
            sc = np.array([self.beam.getScale(k * self.dt) for k in range(nTab)])
            self._scale = L.to_dev(sc)
            prof = self.beam.getProfile()
        else:
            self._scale, prof = None, None
        k0, k1 = randomGen.key()
        a = StepArgs()
        a.dtype = p.code
        a.data = p._data.data_ptr()
        a.stride = cap
        a.nDev = self._nDev.data_ptr()
        a.isInside = None          # the one-call step keeps its holes in lists; no per-slot flags
        a.rho = self.rho.data_ptr()
        a.q = self.q.data_ptr()
        a.profile = prof.data_ptr() if prof is not None else None
        a.beamScale = self._scale.data_ptr() if self._scale is not None else None
        a.nScale = nTab if self._scale is not None else 0
        a.phi = ps.phi.data_ptr()
        a.E = ps.eAtGridPoints.data_ptr()
        a.e2nRowPtr = ps.rowPtr.data_ptr()
        a.e2nCol = ps.col.data_ptr()
        a.e2nVal = ps.val.data_ptr()
        set_e2n_classes(a, ps)
        a.bField = self.bField.data_ptr() if self.bField is not None else None
        a.b0, a.b1, a.b2 = [float(v) for v in p.bConst]
        a.typeB = self.typeB
        a.dt = self.dt
        a.qm = p.particleCharge / p.particleMass
        a.qe = p.particleCharge
        a.absorbed = pb.absorbedParticles.data_ptr()
        a.normals = pb.normalVectors.data_ptr()
        a.remaining = pb.remainingTimeStep.data_ptr()
        a.absCap = absCap
        a.nSec = self._nsec.data_ptr()
        a.secOff = self._secoff.data_ptr()
        a.impact = self._impact.data_ptr()
        a.secondaries = self._sec.data_ptr()
        a.secCap = secCap
        a.emit = 1 if self.secondaryEmission else 0
        a.seed0, a.seed1 = k0, k1
        a.stepDev = self._stepDev.data_ptr()
        a.stats = self._stats.data_ptr()
        a.countHist = self._hist.data_ptr()
        a.histCap = nTab
        if self.photo is not None:
            cnt, wts = self.photo.tables(nTab)
            self._photoCount = torch.as_tensor(cnt.astype(np.int32)).to(dev)
            self._photoWeight = L.to_dev(wts)
            self.photoCap = max(1, int(cnt.max()) if cnt.size else 1)
            self._photoSlots = torch.empty((self.photoCap, 6), dtype=p.dtype, device=dev)
            a.photoCount = self._photoCount.data_ptr()
            a.photoWeight = self._photoWeight.data_ptr()
            a.nPhoto = nTab
            a.photoE0 = self.photo.energyMean
            a.photoSigmaE = self.photo.energySigma
            a.photoSlots = self._photoSlots.data_ptr()
            a.photoCap = self.photoCap
        secem = getattr(self.em, 'secemDesc', None)
        if secem is not None:
            a.secem = ctypes.addressof(secem)
            self._prm = L.FurmanDesc()           # only its particleMass is read with secem set
            self._prm.particleMass = p.particleMass
        else:
            self._prm = self.em.desc
        self._bil = b_layout(self.bField, self.g.getNpExt()) if self.typeB == 2 else None
        a.Bil = self._bil.data_ptr() if self._bil is not None else None
        set_push_fields(self, a, self.g.getNpExt(), self.pushArithmetic, nearCells)
        a.work = self._stepwork.data_ptr()
        self._args = a
        if not self._rho_valid:
            self.deposit_all()

    def step_native(self):
        """One full step in one library call; counts stay on the device (call sync())."""
        if self.it >= self._tabEnd:
            raise L.OeclError('native step: maxSteps exhausted; call native_setup again')
        if self.p.orderedDeposit:
            # the charge of the current particles in the reference's order, replacing the
            # push's deposit (needs the host count: synchronises)
            self.sync()
            self.deposit_all()
        L.check(L.lib().oecl_step_run(ctypes.byref(self.gd), ctypes.byref(self.pb.desc),
                                      ctypes.byref(self._prm), self.ps._fp,
                                      ctypes.byref(self._args), L.stream()), 'step')
        self.it += 1

    def capture_native(self, nSteps):
        """Captures nSteps native steps into one HIP graph (torch.cuda.CUDAGraph).  Every
        replay() runs the next nSteps steps: the step index lives on the device.  The graph
        holds the current device pointers, so sort_particles(keepBuffers=True) must be used
        while the graph is in use.  Event timing is switched off inside the graph."""
        self._args.evPushBegin = None
        self._args.evPushEnd = None
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        g = torch.cuda.CUDAGraph()
        it0 = self.it
        with torch.cuda.graph(g, stream=s):
            for _ in range(int(nSteps)):
                self.step_native()
        torch.cuda.current_stream().wait_stream(s)
        self.it = it0   # capturing runs nothing
        self._graph, self._graphSteps = g, int(nSteps)
        return g

    def replay(self):
        """Runs the captured steps (capture_native)."""
        if self.it + self._graphSteps > self._tabEnd:
            raise L.OeclError('native step: maxSteps exhausted; call native_setup again')
        self._graph.replay()
        self.it += self._graphSteps

    def sync(self):
        """Reads the device counts back, updates the Particles object and checks the status
        word (raises on any overflow)."""
        st = self._stats.cpu().tolist()
        self.p.macroParticleCount = int(self._nDev.item())
        self.pb.absorbedMacroParticleCount = min(st[4], self.absCap)
        if st[2] & ST_ABS_OVERFLOW:
            raise L.OeclError('more wall hits in one step than absCap=%d' % self.absCap)
        if st[2] & ST_SEC_OVERFLOW:
            raise L.OeclError('more secondaries in one step than secCap=%d' % self.secCap)
        if st[2] & ST_PHOTO_OVERFLOW:
            raise L.OeclError('more photoelectrons in one step than the tables allow')
        if st[2] & ST_CAP_OVERFLOW:
            raise L.OeclError('particle count exceeded the capacity %d' % self.p.capacity)
        return {'absorbed_total': st[0], 'emitted_total': st[1], 'not_found': st[3],
                'absorbed_last': st[4], 'added_last': st[5], 'photo_total': st[6]}

    def sort_particles(self, keepBuffers=False):
        """Orders the particles by grid cell (device radix sort + permutation).  Needs the host
        count to be current (call sync() first after native steps).  keepBuffers: copy the
        sorted block back so the particle storage keeps its address (captured graphs)."""
        p = self.p
        n = p.macroParticleCount
        if n < 2:
            return
        dev = L.device()
        if getattr(self, '_sortkeys', None) is None or self._sortkeys.numel() < p.capacity:
            self._sortkeys = torch.empty(p.capacity, dtype=torch.int64, device=dev)
            self._sortperm = torch.empty(p.capacity, dtype=torch.int32, device=dev)
            self._sortwork = torch.empty(int(L.lib().oecl_sort_work_bytes(p.capacity)) // 8 + 64,
                                         dtype=torch.float64, device=dev)
        spare = p._spare_block()
        L.check(L.lib().oecl_sort_by_cell(
            ctypes.byref(self.gd), L.i32(p.code), L.ptr(p._data), L.ptr(spare), L.u64(p.capacity),
            L.u32(n), L.ptr(self._sortkeys), L.ptr(self._sortperm), L.ptr(self._sortwork),
            L.stream()), 'sort by cell')
        if keepBuffers:
            p._data[:, :n].copy_(spare[:, :n])
        else:
            p._data, p._spare = spare, p._data
            if getattr(self, '_args', None) is not None:
                self._args.data = p._data.data_ptr()

    def manage(self, targetMacroParticleCount, upper=1.05, lower=0.95, method='local', **kw):
        """Particle management inside a native buildup (example03_kdTree.py:57-59): when the
        macro-particle count is above upper * target or below lower * target, localRandom
        (or globalRandom) brings it back to the target.  The particle block keeps its address
        and capacity (a captured graph stays valid), the device count is updated and the
        charge of the managed particles is deposited again.  Synchronises; returns the
        management's operation counts, or None when the count is within the band."""
        from . import particleManagement
        self.sync()
        p = self.p
        n = p.macroParticleCount
        target = int(targetMacroParticleCount)
        if lower * target <= n <= upper * target:
            return None
        block, cap = p._data, p.capacity
        if method == 'local':
            ops = particleManagement.localRandom(self.g, p, self.pb, target, **kw)
        elif method == 'global':
            particleManagement.globalRandom(self.g, p, target)
            ops = {'method': 'global'}
        else:
            raise ValueError('method must be "local" or "global"')
        if p.capacity != cap:
            raise L.OeclError('particle management grew the block beyond its capacity %d' % cap)
        if p._data is not block:
            m = p.macroParticleCount
            block[:, :m].copy_(p._data[:, :m])
            p._data, p._spare = block, p._data
        self._nDev.fill_(p.macroParticleCount)
        self.deposit_all()
        return ops

    def counts(self, first, last):
        """Particle counts at the start of steps first .. last-1 (native steps)."""
        return self._hist[first:last].cpu().tolist()

    def _record(self, na, nsec, nph=0):
        h = self.history
        h['macro'].append(self.p.macroParticleCount)
        h['absorbed'].append(na)
        h['emitted'].append(nsec)
        h['photo'].append(nph)

    def run(self, nSteps, fused=True):
        for _ in range(int(nSteps)):
            if fused:
                self.step()
            else:
                self.step_reference()
