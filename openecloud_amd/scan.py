"""Batched parameter scans: R independent buildups advanced together.

The reference runs one buildup per process (example01_basic.py); a scan over secondary
emission yield and bunch intensity is R separate runs.  BuildupScan advances R runs that share
the grid, the wall and the Poisson operator with one library call per step for all of them
(oecl_step_run_batch), so the device sees R times the work per launch:
  - every run's particles live in its own segment of one SoA block (a multiple of PB = 256 slots,
    so that each push block belongs to one run);
  - the R Poisson right-hand sides are solved together (the DST products get R times the
    columns);
  - the per-run hit, emission and placement kernels run one block or block row per run;
  - runs keep their own Furman parameters, beam intensity and Philox key (so run r reproduces
    a standalone Buildup with the same seed, particles and parameters);
  - capture(n) records n steps of all runs into one HIP graph, and the cell sort is one radix
    sort over all segments.
Photoelectrons are not available in a scan.
"""
import ctypes
import time

import numpy as np
import torch

from . import _lib as L
from .buildup import (StepArgs, ST_ABS_OVERFLOW, ST_CAP_OVERFLOW, ST_SEC_OVERFLOW, b_layout, push_arithmetic,
                      set_e2n_classes, set_push_fields)

PB = 256


class RunsDesc(ctypes.Structure):
    """Mirror of oecl_runs_t (include/oecl.h)."""
    _fields_ = [('R', ctypes.c_uint32), ('segOff', ctypes.c_void_p), ('blockRun', ctypes.c_void_p),
                ('absOff', ctypes.c_void_p), ('furman', ctypes.c_void_p), ('keys', ctypes.c_void_p)]


def plan_segments(counts, growth=2.0, slack=1024, absFrac=1. / 16, absMin=1024):
    """Segment and hit-row layout for runs of the given particle counts: run r may grow to
    growth * count + slack particles (rounded up to 256) and takes segCap / absFrac hit rows
    (at least absMin).  Returns (segOff[R + 1], absOff[R + 1]) as int64 arrays."""
    counts = np.asarray(counts, dtype=np.int64)
    seg = np.ceil((counts * float(growth) + slack) / PB).astype(np.int64) * PB
    ab = np.maximum(np.ceil(seg * float(absFrac)).astype(np.int64), absMin)
    segOff = np.concatenate([[0], np.cumsum(seg)])
    absOff = np.concatenate([[0], np.cumsum(ab)])
    if segOff[-1] > 0xFFFFFF00 or absOff[-1] > 0xFFFFFF00:
        raise ValueError('scan too large for 32-bit particle indices')
    return segOff, absOff


def seed_key(seed):
    """The Philox key randomGen.seed(seed) sets."""
    s = int(seed) & 0xFFFFFFFFFFFFFFFF
    return s & 0xFFFFFFFF, (s >> 32) & 0xFFFFFFFF


class BuildupScan:
    """R buildups (openecloud_amd.buildup.Buildup objects) stepped together.

    The buildups must share the Grid, the wall kind, the PoissonSolver (method 'direct'), dt,
    the magnetic field and the particle dtype; each keeps its particles, FurmanEmitter (or
    any emitter with the same parameters), beam (its intensity and timing give the run's
    beam-scale table; the transverse profile must be the same) and seed.  The runs' particles
    are copied into the scan's block at construction; read them back with run_particles(r).
    """

    def __init__(self, buildups, seeds, growth=2.0, absFrac=1. / 16, maxSteps=100000, slack=1024,
                 pushArithmetic=None, nearCells=2.):
        if len(buildups) != len(seeds) or not buildups:
            raise ValueError('one seed per buildup')
        b0 = buildups[0]
        for b in buildups[1:]:
            if (b.g is not b0.g or b.ps is not b0.ps or b.dt != b0.dt or b.typeB != b0.typeB
                    or b.p.dtype != b0.p.dtype or type(b.pb) is not type(b0.pb)):
                raise ValueError('the runs of a scan share grid, Poisson solver, dt, field type and dtype')
            if b.photo is not None:
                raise ValueError('photoelectrons are not available in a scan')
        if b0.ps.method != 'direct':
            raise ValueError('a scan needs PoissonSolver(method="direct")')
        if hasattr(b0.em, 'secemDesc'):
            raise ValueError('a scan runs the Furman emitter')
        if int(L.lib().oecl_push_block_size()) != PB:
            raise L.OeclError('the library pushes %d particles per block, the scan lays out %d'
                              % (int(L.lib().oecl_push_block_size()), PB))
        self.runs = list(buildups)
        self.R = R = len(buildups)
        self.g, self.ps, self.pb = b0.g, b0.ps, b0.pb
        self.dt, self.typeB = b0.dt, b0.typeB
        self.dtype = b0.p.dtype
        self.pushArithmetic = push_arithmetic(self.dtype, pushArithmetic)
        self._nearCells = float(nearCells)
        self.code = b0.p.code
        self.qe = b0.p.particleCharge
        self.qm = b0.p.particleCharge / b0.p.particleMass
        self.it = b0.it
        self.maxSteps = int(maxSteps)
        dev = L.device()
        npx = self.g.getNpExt()
        self.np = npx
        counts = [b.p.macroParticleCount for b in buildups]
        segOff, absOff = plan_segments(counts, growth, slack=slack, absFrac=absFrac)
        self.segOff, self.absOff = segOff, absOff
        cap = int(segOff[-1])
        absTot = int(absOff[-1])
        self.capacity, self.absTotal = cap, absTot
        self.data = torch.zeros((6, cap), dtype=self.dtype, device=dev)
        self.spare = torch.zeros((6, cap), dtype=self.dtype, device=dev)
        for r, b in enumerate(buildups):
            n = counts[r]
            self.data[:, int(segOff[r]):int(segOff[r]) + n].copy_(b.p._data[:, :n])
        blockRun = np.repeat(np.arange(R, dtype=np.int64), (segOff[1:] - segOff[:-1]) // PB)
        self._segOff = torch.as_tensor(segOff.astype(np.int32)).to(dev)
        self._absOff = torch.as_tensor(absOff.astype(np.int32)).to(dev)
        self._blockRun = torch.as_tensor(blockRun.astype(np.int32)).to(dev)
        # per-run Furman parameters (oecl_furman_t array) and Philox keys
        fsz = ctypes.sizeof(L.FurmanDesc)
        raw = bytearray(fsz * R)
        for r, b in enumerate(buildups):
            raw[r * fsz:(r + 1) * fsz] = bytes(b.em.desc)
        self._furman = torch.frombuffer(raw, dtype=torch.uint8).to(dev)
        keys = np.array([seed_key(s) for s in seeds], dtype=np.uint32).reshape(-1)
        self.keys = keys.reshape(R, 2)
        self._keys = torch.as_tensor(keys.astype(np.int64).astype(np.int32)).to(dev)
        self._nDev = torch.tensor(counts, dtype=torch.int32, device=dev)
        self._stepDev = torch.tensor([self.it], dtype=torch.int32, device=dev)
        self._stats = torch.zeros(8 * R, dtype=torch.int32, device=dev)
        nTab = self.it + self.maxSteps
        self._nTab = nTab
        self._hist = torch.zeros(R * nTab, dtype=torch.int32, device=dev)
        self.rho = torch.zeros(R * npx, dtype=torch.float64, device=dev)
        self.q = torch.zeros(R * npx, dtype=torch.float64, device=dev)
        self.phi = torch.zeros(R * npx, dtype=torch.float64, device=dev)
        self.E = torch.zeros(R * 2 * npx, dtype=torch.float64, device=dev)
        self.absorbed = torch.zeros((absTot, 6), dtype=self.dtype, device=dev)
        self.normals = torch.zeros((absTot, 2), dtype=self.dtype, device=dev)
        self.remaining = torch.zeros(absTot, dtype=self.dtype, device=dev)
        self._nsec = torch.zeros(absTot, dtype=torch.int32, device=dev)
        self._secoff = torch.zeros(absTot + R, dtype=torch.int32, device=dev)
        self._impact = torch.zeros(4 * absTot, dtype=torch.float64, device=dev)
        self._work = torch.zeros(int(L.lib().oecl_step_batch_work_bytes(L.u64(cap), L.u32(absTot), L.u32(R))) // 8
                                 + 64, dtype=torch.float64, device=dev)
        beams = [b.beam for b in buildups]
        if any(bm is not None for bm in beams):
            if any(bm is None for bm in beams):
                raise ValueError('either every run has a beam or none has')
            prof0 = beams[0].getProfile()
            for bm in beams[1:]:
                if not torch.equal(bm.getProfile(), prof0):
                    raise ValueError('the runs of a scan share the beam profile')
            sc = np.array([[bm.getScale(k * self.dt) for k in range(nTab)] for bm in beams])
            self._scale = L.to_dev(sc.reshape(-1))
            self._scaleHost = sc
            self._prof = prof0
        else:
            self._scale, self._prof, self._scaleHost = None, None, None
        self._bField = b0.bField
        self._bil = b_layout(self._bField, npx) if self.typeB == 2 else None
        self._desc = RunsDesc(R, self._segOff.data_ptr(), self._blockRun.data_ptr(), self._absOff.data_ptr(),
                              self._furman.data_ptr(), self._keys.data_ptr())
        self._args = self._make_args(b0)
        L.check(L.lib().oecl_fastpoisson_reserve(self.ps._fp, L.u32(R)), 'scan Poisson batch')
        self._graph = None
        self._sortkeys = None
        self._it0 = self.it
        self._learned = None
        self.learnedStats = {'iterations': [], 'relres_max': 0., 'fallbacks': 0, 'judge_max_rel_err': 0.,
                             'judged': 0}
        self.deposit_all()

    def set_solver(self, preconditioner=None, tol=1e-8, maxIter=4000, judgeEvery=0):
        """The Poisson solve of the steps.  None: the direct solve inside the batched step (the
        default).  A trained UNetPreconditioner: every run's potential by the learned flexible CG
        to tol, warm-started from the run's previous potential, the direct solve taking over when
        it does not converge; every judgeEvery-th step (0: never) the direct solve of each run
        judges the result (learnedStats['judge_max_rel_err']).  The learned solve reads residual
        norms on the host, so such steps run eagerly (no graph capture)."""
        self._learned = preconditioner
        self._ltol, self._lmax, self._ljudge = float(tol), int(maxIter), int(judgeEvery)
        self._args.externalPhi = 1 if preconditioner is not None else 0
        if preconditioner is not None:
            self._graph = None

    def _solve_learned(self):
        """phi_r for this step's charge rho_r + s_r(step) profile of every run."""
        net, ps, npx = self._learned, self.ps, self.np
        st = self.learnedStats
        judge = self._ljudge > 0 and (self.it - self._it0) % self._ljudge == 0
        for r in range(self.R):
            rho_r = self.rho[r * npx:(r + 1) * npx]
            q_r = self.q[r * npx:(r + 1) * npx]
            sc = float(self._scaleHost[r, self.it]) if self._scaleHost is not None else 0.
            if sc != 0.:
                L.check(L.lib().oecl_axpy_rhs(L.u32(npx), L.ptr(rho_r), L.f64(sc), L.ptr(self._prof), L.ptr(q_r),
                                              L.stream()), 'scan rhs')
            else:
                q_r.copy_(rho_r)
        # every run's flexible CG in one set of launches (each run stops at its own check point)
        its, rels = net.fcg_batch(ps, self.q, self.phi, self.R, tol=self._ltol, maxIter=self._lmax)
        for r in range(self.R):
            q_r = self.q[r * npx:(r + 1) * npx]
            phi_r = self.phi[r * npx:(r + 1) * npx]
            it, rr = int(its[r]), float(rels[r])
            st['iterations'].append(it)
            if not rr <= self._ltol:
                st['fallbacks'] += 1
                ps._direct(q_r, phi_r)
            else:
                st['relres_max'] = max(st['relres_max'], rr)
            if judge:
                ref = torch.empty_like(phi_r)
                ps._direct(q_r, ref)
                den = float(torch.linalg.norm(ref))
                if den > 0.:
                    st['judge_max_rel_err'] = max(st['judge_max_rel_err'], float(torch.linalg.norm(phi_r - ref)) / den)
                st['judged'] += 1

    def _make_args(self, b0):
        p, ps = b0.p, self.ps
        a = StepArgs()
        a.dtype = self.code
        a.data = self.data.data_ptr()
        a.stride = self.capacity
        a.nDev = self._nDev.data_ptr()
        a.isInside = None          # the batched step keeps its holes in lists; no per-slot flags
        a.rho = self.rho.data_ptr()
        a.q = self.q.data_ptr()
        a.profile = self._prof.data_ptr() if self._prof is not None else None
        a.beamScale = self._scale.data_ptr() if self._scale is not None else None
        a.nScale = self._nTab if self._scale is not None else 0
        a.phi = self.phi.data_ptr()
        a.E = self.E.data_ptr()
        a.e2nRowPtr = ps.rowPtr.data_ptr()
        a.e2nCol = ps.col.data_ptr()
        a.e2nVal = ps.val.data_ptr()
        set_e2n_classes(a, ps)
        a.bField = self._bField.data_ptr() if self._bField is not None else None
        a.b0, a.b1, a.b2 = [float(v) for v in p.bConst]
        a.typeB = self.typeB
        a.dt = self.dt
        a.qm = self.qm
        a.qe = self.qe
        a.absorbed = self.absorbed.data_ptr()
        a.normals = self.normals.data_ptr()
        a.remaining = self.remaining.data_ptr()
        a.absCap = self.absTotal
        a.nSec = self._nsec.data_ptr()
        a.secOff = self._secoff.data_ptr()
        a.impact = self._impact.data_ptr()
        a.secondaries = None
        a.secCap = 0
        a.emit = 1 if b0.secondaryEmission else 0
        a.stepDev = self._stepDev.data_ptr()
        a.stats = self._stats.data_ptr()
        a.countHist = self._hist.data_ptr()
        a.histCap = self._nTab
        a.work = self._work.data_ptr()
        a.Bil = self._bil.data_ptr() if self._bil is not None else None
        set_push_fields(self, a, self.R * self.np, self.pushArithmetic, self._nearCells)
        return a

    # ------------------------------------------------------------------ state
    def deposit_all(self):
        """rho_r = charge of run r's particles, for every run."""
        self.rho.zero_()
        counts = self._nDev.cpu().tolist()
        esz = self.data.element_size()
        for r in range(self.R):
            base = self.data.data_ptr() + int(self.segOff[r]) * esz
            L.check(L.lib().oecl_deposit_add(
                ctypes.byref(self.runs[0].gd), L.i32(self.code), ctypes.c_void_p(base), L.u64(self.capacity),
                L.u32(counts[r]), L.f64(self.qe), ctypes.c_void_p(self.rho.data_ptr() + 8 * r * self.np),
                L.stream()), 'scan deposit')

    def counts(self):
        """Live macro-particle count of every run (synchronises)."""
        return self._nDev.cpu().tolist()

    def run_particles(self, r):
        """Run r's particles as an (N, 6) view of the scan's block (synchronises)."""
        n = int(self._nDev[r].item())
        s0 = int(self.segOff[r])
        return self.data[:, s0:s0 + n].t()

    def run_rho(self, r):
        return self.rho[r * self.np:(r + 1) * self.np]

    def history(self, r, first, last):
        """Run r's particle counts at the start of steps first .. last-1."""
        return self._hist[r * self._nTab + first:r * self._nTab + last].cpu().tolist()

    def sync(self):
        """Reads every run's counts and status words; raises on any overflow."""
        st = self._stats.cpu().numpy().reshape(self.R, 8)
        for r in range(self.R):
            if st[r, 2] & ST_ABS_OVERFLOW:
                raise L.OeclError('run %d: more wall hits in one step than its %d hit rows'
                                  % (r, int(self.absOff[r + 1] - self.absOff[r])))
            if st[r, 2] & ST_SEC_OVERFLOW:
                raise L.OeclError('run %d: secondary overflow' % r)
            if st[r, 2] & ST_CAP_OVERFLOW:
                raise L.OeclError('run %d: particle count exceeded its segment of %d'
                                  % (r, int(self.segOff[r + 1] - self.segOff[r])))
        return {'absorbed_total': st[:, 0].tolist(), 'emitted_total': st[:, 1].tolist(),
                'not_found': st[:, 3].tolist(), 'counts': self.counts()}

    # ------------------------------------------------------------------ stepping
    def step(self):
        """One step of every run, one library call (after the runs' learned solves, when set)."""
        if self.it - self._it0 >= self.maxSteps:
            raise L.OeclError('scan: maxSteps exhausted')
        if self._learned is not None:
            self._solve_learned()
        L.check(L.lib().oecl_step_run_batch(ctypes.byref(self.runs[0].gd), ctypes.byref(self.pb.desc),
                                            ctypes.byref(self._desc), self.ps._fp, ctypes.byref(self._args),
                                            L.stream()), 'scan step')
        self.it += 1

    def set_deposit(self, span=0, waveTiles=False):
        """The push's deposit tiles (oecl_step_args.depSpan, depTile): span, the lanes over which
        runs of equal cells are summed (0: the library's default of 4; else a power of two,
        2-64); waveTiles, per-wave LDS tiles only instead of one tile per push block.  Changes
        only the summation order of rho.  A captured graph is dropped (it holds the old
        arguments)."""
        span = int(span)
        if span and (span < 2 or span > 64 or span & (span - 1)):
            raise ValueError('span must be 0 or a power of two in 2..64')
        self._args.depSpan = span
        self._args.depTile = 1 if waveTiles else 0
        self._graph = None

    def capture(self, nSteps):
        """Records nSteps steps of every run into one HIP graph; replay() runs the next
        nSteps.  The scan's buffers keep their addresses (sort() copies back)."""
        if self._learned is not None:
            raise L.OeclError('scan: the learned solve runs eagerly (set_solver(None) to capture)')
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        g = torch.cuda.CUDAGraph()
        it0 = self.it
        with torch.cuda.graph(g, stream=s):
            for _ in range(int(nSteps)):
                self.step()
        torch.cuda.current_stream().wait_stream(s)
        self.it = it0
        self._graph, self._graphSteps = g, int(nSteps)
        return g

    def replay(self):
        if self.it - self._it0 + self._graphSteps > self.maxSteps:
            raise L.OeclError('scan: maxSteps exhausted')
        self._graph.replay()
        self.it += self._graphSteps

    def sort(self, order=0, binShift=0):
        """Orders every run's particles by cell (one radix sort over all segments).  The
        sorted block is copied back, so captured graphs stay valid."""
        dev = L.device()
        if self._sortkeys is None:
            self._sortkeys = torch.empty(self.capacity, dtype=torch.int64, device=dev)
            self._sortperm = torch.empty(self.capacity, dtype=torch.int32, device=dev)
            self._sortwork = torch.empty(int(L.lib().oecl_sort_work_bytes(L.u32(self.capacity))) // 8 + 64,
                                         dtype=torch.float64, device=dev)
        L.check(L.lib().oecl_sort_runs_by_cell(
            ctypes.byref(self.runs[0].gd), L.i32(self.code), L.ptr(self.data), L.ptr(self.spare),
            L.u64(self.capacity), L.u32(self.R), L.ptr(self._segOff), L.ptr(self._blockRun), L.ptr(self._nDev),
            L.i32(order), L.i32(binShift), L.ptr(self._sortkeys), L.ptr(self._sortperm), L.ptr(self._sortwork),
            L.stream()), 'scan sort')
        self.data.copy_(self.spare)

    def run(self, nSteps, sortEvery=50, graphSteps=None, order=0, binShift=0):
        """Advances every run nSteps steps: graphs of graphSteps steps (default: sortEvery, or
        no graph when None and sortEvery is 0), a cell sort every sortEvery steps."""
        done = 0
        nSteps = int(nSteps)
        if graphSteps and (self._graph is None or self._graphSteps != graphSteps):
            self.capture(graphSteps)
        while done < nSteps:
            if sortEvery and self.it % sortEvery == 0:
                self.sort(order, binShift)
            if graphSteps and nSteps - done >= graphSteps and (
                    not sortEvery or (self.it % sortEvery) + graphSteps <= sortEvery):
                self.replay()
                done += graphSteps
            else:
                self.step()
                done += 1
