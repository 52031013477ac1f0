"""The one-call step captured into a HIP graph: every replay matches the eager steps."""
import numpy as np
import pytest
import torch

from conftest import random_particles

pytestmark = pytest.mark.gpu


def _setup(seed_particles, photo):
    from openecloud_amd import randomGen
    from openecloud_amd.beam import GaussianBunchTrain
    from openecloud_amd.buildup import Buildup
    from openecloud_amd.grid import Grid
    from openecloud_amd.particleBoundary import AbsorbElliptical
    from openecloud_amd.particleEmitter import FurmanEmitter, PhotoEmitter
    from openecloud_amd.particles import Particles
    from openecloud_amd.poissonSolver import PoissonSolver
    g = Grid(64, 48, 0.09, 0.07, 1)
    pd0 = random_particles(g, 20000, seed=seed_particles, wall='ellip', vscale=3e6)
    pd0[:, 5] = 5e4
    dt = 2e-11
    randomGen.seed(91)
    p = Particles('electrons', 'variableWeight', ordering='fast')
    p.setParticleData(pd0)
    p.reserve(40000)
    p.setBConst([0., 0.5, 0.])
    pb = AbsorbElliptical(g, p)
    em = FurmanEmitter(pb, p, 'copper', seyMax=1.6, eHat0=300.)
    ps = PoissonSolver(g, method='direct', checkEvery=0)
    beam = GaussianBunchTrain(g, dt, nParticles=1e11, tBunchSpacing=25e-9, radiusSigma=2e-3,
                              nBunches=1, bunchLengthSigma=0.003)
    ph = PhotoEmitter(pb, p, beam, photoYield=1e-3, macroPerBunch=2000) if photo else None
    b = Buildup(g, p, pb, em, ps, beam, dt=dt, typeBField=1, photoEmitter=ph)
    b.native_setup(absCap=8192, secCap=16384, maxSteps=64)
    return b, p


@pytest.mark.parametrize('photo', [False, True])
def test_graph_replays_match_eager_steps(photo):
    """k steps per graph, three replays with host work between them (sync, and a cell sort
    that keeps the buffers) against the same sequence run eagerly."""
    k = 4
    b, p = _setup(21, photo)
    ref = []
    for r in range(3):
        for _ in range(k):
            b.step_native()
        b.sync()
        ref.append(p.getParticleData().cpu().numpy().copy())
        if r == 1:
            b.sort_particles()
    b2, p2 = _setup(21, photo)
    b2.capture_native(k)
    for r in range(3):
        b2.replay()
        st = b2.sync()
        got = p2.getParticleData().cpu().numpy()
        assert got.shape == ref[r].shape, (r, got.shape, ref[r].shape)
        scale = np.abs(ref[r]).max(axis=0)
        assert (np.abs(got - ref[r]) / scale).max() < 1e-12, r
        if r == 1:
            b2.sort_particles(keepBuffers=True)   # the graph's pointers stay valid
    assert st['absorbed_total'] > 0
    if photo:
        assert st['photo_total'] > 0
