"""Host arrays (openecloud_amd.set_host_arrays(True)): a driver written against the reference's
array conventions runs unchanged.  The loop below has the shape of the reference's basic
example (example01_basic.py:52-97): the physical particle count from numpy.sum over the weight
column, the Poisson right-hand side built with numpy.asarray(getChargeOnGrid()) + the beam
charge, numpy.save of getParticleData(), and addAndRemoveParticles fed with the NumPy results of
generateSecondaries() and getIsInside().  Its particle state after the steps equals the same
steps run on device tensors."""
import os
import tempfile

import numpy as np
import pytest
import torch

from conftest import random_particles

pytestmark = pytest.mark.gpu


def _objects():
    from openecloud_amd import randomGen
    from openecloud_amd.beam import LHCBeamEndless
    from openecloud_amd.grid import Grid
    from openecloud_amd.particleBoundary import AbsorbElliptical
    from openecloud_amd.particleEmitter import FurmanEmitter
    from openecloud_amd.particles import Particles
    from openecloud_amd.poissonSolver import PoissonSolver
    randomGen.seed(3)
    g = Grid(32, 32, 0.04, 0.04, 1)
    pd = random_particles(g, 20000, seed=9, wall='ellip', vscale=2e6)
    pd[:, 5] = 1e4
    p = Particles('electrons', 'variableWeight', orderedDeposit=True)   # reproducible rho
    p.setParticleData(pd)
    pb = AbsorbElliptical(g, p)
    em = FurmanEmitter(pb, p)
    ps = PoissonSolver(g)
    beam = LHCBeamEndless(g, 2.155172413795e-11)
    return g, p, pb, em, ps, beam


def _loop(host, nsteps, outdir):
    import openecloud_amd
    openecloud_amd.set_host_arrays(host)
    try:
        g, p, pb, em, ps, beam = _objects()
        dt = 2.155172413795e-11
        phys = np.zeros(nsteps)
        for ii in range(nsteps):
            w = p.getParticleData()[:, 5]
            phys[ii] = np.sum(w) if host else float(w.sum())
            p.calcGridWeights(g)
            p.chargeToGrid(g)
            if host:
                np.save(os.path.join(outdir, 'particleData.npy'), p.getParticleData())
                ps.solve(np.asarray(p.getChargeOnGrid()) + beam.getCharge(ii * dt))
            else:
                ps.solve(p.getChargeOnGrid() + beam.getCharge(ii * dt))
            p.eFieldToParticles(g, ps.getEAtGridPoints())
            p.borisPush(dt, typeBField=0)
            pb.indexInside()
            pb.saveAbsorbed()
            if pb.getAbsorbedMacroParticleCount() > 0:
                pb.calculateInteractionPoint()
                pb.calculateNormalVectors()
                p.addAndRemoveParticles(em.generateSecondaries(), p.getIsInside())
        data = p.getParticleData()
        return phys, (data if host else data.cpu().numpy()).copy()
    finally:
        openecloud_amd.set_host_arrays(False)


def test_reference_style_driver_runs_on_host_arrays():
    with tempfile.TemporaryDirectory() as d:
        phys_h, pd_h = _loop(True, 8, d)
        saved = np.load(os.path.join(d, 'particleData.npy'))
        assert saved.ndim == 2 and saved.shape[1] == 6
    phys_d, pd_d = _loop(False, 8, None)
    assert isinstance(pd_h, np.ndarray)
    assert pd_h.shape == pd_d.shape
    assert np.array_equal(pd_h, pd_d)
    assert np.allclose(phys_h, phys_d, rtol=1e-12)


def test_host_arrays_switch_only_changes_the_return_type():
    import openecloud_amd
    g, p, pb, em, ps, beam = _objects()
    assert isinstance(p.getParticleData(), torch.Tensor)
    openecloud_amd.set_host_arrays(True)
    try:
        assert openecloud_amd.host_arrays()
        a = p.getParticleData()
        assert isinstance(a, np.ndarray) and a.shape == (p.getMacroParticleCount(), 6)
        assert isinstance(p.getChargeOnGrid(), np.ndarray)
        assert isinstance(ps.getPhi(), np.ndarray)
    finally:
        openecloud_amd.set_host_arrays(False)
    assert isinstance(ps.getPhi(), torch.Tensor)
