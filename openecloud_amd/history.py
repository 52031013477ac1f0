"""Step history and data dumps of a buildup.

The reference's driver keeps two histories, the physical particle count numpy.sum(w) and the
macro-particle count. Every nHistoryOut steps it saves them together with the particle data,
as physicalParticleCount.npy and particleData.npy (example01_basic.py:45-78).

`BuildupHistory` keeps those histories, plus absorbed, emitted and photoelectron counts, for
any of the three step forms of `Buildup`. It writes the same files. It can also dump pairs
(q, phi) of the Poisson problem the buildup solves, with q = rho + beam charge and phi the
potential for that q, as the data set a learned preconditioner is trained on.
"""
import os

import numpy as np
import torch

from . import diagnostics


class BuildupHistory:

    def __init__(self, buildup, outDir=None, dumpEvery=0, pairEvery=0):
        """outDir: where save() and the periodic dumps write (None: nothing is written).
        dumpEvery: save the histories and particleData.npy every this many recorded steps.
        pairEvery: keep a (q, phi) pair every this many recorded steps."""
        self.b = buildup
        self.outDir = outDir
        self.dumpEvery = int(dumpEvery)
        self.pairEvery = int(pairEvery)
        self.step = []
        self.physicalParticleCount = []
        self.macroParticleCount = []
        self.absorbed = []
        self.emitted = []
        self.photo = []
        self.pairs = []
        self._last = None
        if outDir is not None:
            os.makedirs(outDir, exist_ok=True)

    def record(self):
        """Records the state after the last step.

        Native steps keep their counts on the device, so call Buildup.sync() before
        recording. For them, the absorbed, emitted and photoelectron counts are the differences
        of the running totals since the previous record.
        """
        b, p = self.b, self.b.p
        k = len(self.step)
        self.step.append(b.it)
        self.physicalParticleCount.append(diagnostics.weight_sum(p))
        self.macroParticleCount.append(p.getMacroParticleCount())
        if getattr(b, '_stats', None) is not None and getattr(b, '_args', None) is not None:
            st = b._stats.cpu().tolist()
            tot = (st[0], st[1], st[6])
            prev = self._last or (0, 0, 0)
            self._last = tot
            self.absorbed.append(tot[0] - prev[0])
            self.emitted.append(tot[1] - prev[1])
            self.photo.append(tot[2] - prev[2])
        else:
            h = b.history
            self.absorbed.append(h['absorbed'][-1] if h['absorbed'] else 0)
            self.emitted.append(h['emitted'][-1] if h['emitted'] else 0)
            self.photo.append(h['photo'][-1] if h['photo'] else 0)
        if self.pairEvery and k % self.pairEvery == 0:
            self.pairs.append(self.poisson_pair())
        if self.dumpEvery and self.outDir is not None and k > 0 and k % self.dumpEvery == 0:
            self.save()

    def poisson_pair(self):
        """(q, phi) for the current particles: q = rho + beam charge at the current time.
        phi = A^-1 q is solved now, so the pair is consistent. Returns float64 host arrays
        shaped (nyExt, nxExt)."""
        b = self.b
        if not b._rho_valid:
            b.deposit_all()
        q = b._rhs(b.it * b.dt).clone()
        b.ps.solvePhi(q)
        shape = (b.g.getNyExt(), b.g.getNxExt())
        return (q.cpu().numpy().reshape(shape).copy(), b.ps.phi.cpu().numpy().reshape(shape).copy())

    def arrays(self):
        return {'step': np.asarray(self.step, np.int64),
                'physicalParticleCount': np.asarray(self.physicalParticleCount, np.float64),
                'macroParticleCount': np.asarray(self.macroParticleCount, np.int64),
                'absorbed': np.asarray(self.absorbed, np.int64),
                'emitted': np.asarray(self.emitted, np.int64),
                'photo': np.asarray(self.photo, np.int64)}

    def save(self, outDir=None):
        """Writes physicalParticleCount.npy, macroParticleCount.npy and particleData.npy (the
        reference's files, particle data AoS (N, 6) float64), history.npz with every
        history, and poissonPairs.npz (q, phi stacked) when pairs were kept."""
        d = outDir or self.outDir
        if d is None:
            raise ValueError('no output directory')
        os.makedirs(d, exist_ok=True)
        a = self.arrays()
        np.save(os.path.join(d, 'physicalParticleCount.npy'), a['physicalParticleCount'])
        np.save(os.path.join(d, 'macroParticleCount.npy'), a['macroParticleCount'])
        np.save(os.path.join(d, 'particleData.npy'),
                self.b.p._data[:, :self.b.p.macroParticleCount].t().to(torch.float64).cpu().numpy())
        np.savez(os.path.join(d, 'history.npz'), **a)
        if self.pairs:
            np.savez(os.path.join(d, 'poissonPairs.npz'), q=np.stack([q for q, _ in self.pairs]),
                     phi=np.stack([f for _, f in self.pairs]))
