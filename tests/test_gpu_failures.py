"""Failure handling of the one-call step and the scan, forced from the host with capacities that
are too small: the device flags the overflow, skips the compaction of an overflowing hit list
(the count stays), and sync() raises with the limit that was hit.  Empty inputs pass through."""
import numpy as np
import pytest
import torch

from conftest import random_particles

pytestmark = pytest.mark.gpu


def _buildup(n=20000, seed=61, sey=1.6, reserve=None, g=None, ps=None):
    from openecloud_amd import randomGen
    from openecloud_amd.beam import GaussianBunchTrain
    from openecloud_amd.buildup import Buildup
    from openecloud_amd.grid import Grid
    from openecloud_amd.particleBoundary import AbsorbElliptical
    from openecloud_amd.particleEmitter import FurmanEmitter
    from openecloud_amd.particles import Particles
    from openecloud_amd.poissonSolver import PoissonSolver
    g = g or Grid(64, 48, 0.09, 0.07, 1)
    pd = random_particles(g, n, seed=seed, wall='ellip', vscale=3e6)
    pd[:, 5] = 5e4
    randomGen.seed(seed)
    p = Particles('electrons', 'variableWeight', ordering='fast')
    p.setParticleData(pd)
    if reserve:
        p.reserve(reserve)
    p.setBConst([0., 0.5, 0.])
    pb = AbsorbElliptical(g, p)
    em = FurmanEmitter(pb, p, 'copper', seyMax=sey, eHat0=300.)
    ps = ps or PoissonSolver(g, method='direct', checkEvery=0)
    beam = GaussianBunchTrain(g, 2e-11, nParticles=1e11, tBunchSpacing=25e-9, radiusSigma=2e-3, nBunches=1,
                              bunchLengthSigma=0.003)
    return Buildup(g, p, pb, em, ps, beam, dt=2e-11, typeBField=1)


def test_hit_overflow_is_flagged_and_leaves_the_count():
    from openecloud_amd import _lib as L
    b = _buildup()
    n0 = b.p.macroParticleCount
    b.native_setup(absCap=8, secCap=64, maxSteps=8)
    b.step_native()
    with pytest.raises(L.OeclError, match='absCap=8'):
        b.sync()
    assert b.p.macroParticleCount == n0


def test_secondary_overflow_is_flagged():
    from openecloud_amd import _lib as L
    b = _buildup(sey=2.4)
    b.native_setup(absCap=8192, secCap=4, maxSteps=8)
    b.step_native()
    with pytest.raises(L.OeclError, match='secCap=4'):
        b.sync()


def test_capacity_overflow_is_flagged():
    from openecloud_amd import _lib as L
    b = _buildup(n=20000, sey=2.4)
    p = b.p
    # a particle block with room for 8 more particles than are loaded
    cap = p.macroParticleCount + 8
    p._data = p._data[:, :cap].clone()
    p._spare = None
    p._alloc_scratch(cap)
    assert p.capacity == cap
    b.native_setup(absCap=8192, secCap=16384, maxSteps=64)
    with pytest.raises(L.OeclError, match='capacity'):
        for _ in range(64):
            b.step_native()
            b.sync()
    assert p.macroParticleCount <= cap


def test_scan_segment_overflow_names_the_run():
    from openecloud_amd import _lib as L
    from openecloud_amd.scan import PB, BuildupScan
    b0 = _buildup(n=3000, seed=70, sey=0.5)
    b1 = _buildup(n=3000, seed=71, sey=3.0, g=b0.g, ps=b0.ps)
    # segments with no room to grow beyond the rounding to a push block: run 1 (many
    # secondaries) overflows
    sc = BuildupScan([b0, b1], [70, 71], growth=1.0, slack=0, maxSteps=200)
    assert int(sc.segOff[2] - sc.segOff[1]) - 3000 < PB
    with pytest.raises(L.OeclError, match='run 1'):
        for _ in range(200):
            sc.step()
            sc.sync()


def test_empty_inputs():
    from openecloud_amd.grid import Grid
    from openecloud_amd.particleBoundary import AbsorbElliptical
    from openecloud_amd.particleEmitter import FurmanEmitter, SecElecEmitter
    from openecloud_amd.particles import Particles
    from openecloud_amd.poissonSolver import PoissonSolver
    g = Grid(32, 32, 0.09, 0.07, 1)
    p = Particles('electrons', 'variableWeight')
    p.setParticleData(np.zeros((0, 6)))
    pb = AbsorbElliptical(g, p)
    pb.absorbedMacroParticleCount = 0
    assert FurmanEmitter(pb, p, 'copper').generateSecondaries().shape == (0, 6)
    assert SecElecEmitter(pb, p).generateSecondaries().shape == (0, 6)
    ps = PoissonSolver(g, method='direct', checkEvery=0)
    phi = ps.solvePhi(torch.zeros(g.getNpExt(), dtype=torch.float64, device='cuda'))
    assert float(phi.abs().max()) == 0.
