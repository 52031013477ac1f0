"""GPU buildup loop against the oracle loop: same initial particles, same Philox draws,
several full time steps (deposit, Poisson, gather, push, wall, emission, add/remove)."""
import numpy as np
import pytest
import torch

from conftest import random_particles

pytestmark = pytest.mark.gpu


def _oracle_loop(orc, g, pd, nsteps, dt, typeB, bconst, profile, beam_scale, P, key, cap,
                 rule='reference', photo=None, secem=None, wall='ellip', ops=None):
    """photo: (counts, weights, e0, sigmaE) of a photoelectron source, added after each
    step's secondaries (oracle.photo_oracle).  secem: (model, parameters) of a simpler emitter
    (oracle.secem_oracle) used in place of the Furman one.  wall: 'ellip' or 'rect'.  ops: an
    OraclePoisson of g to reuse."""
    from oracle.grid_oracle import OraclePoisson
    if ops is None:
        ops = OraclePoisson(g)
    Q = -1.602176565e-19
    qm = Q / 9.10938291e-31
    nx, npx = g.getNxExt(), g.getNpExt()
    out = []
    for it in range(nsteps):
        ic, w = orc.calc_grid_weights(pd, g.getDx(), g.getDy(), g.getLxExt(), g.getLyExt(), nx)
        rho = orc.charge_to_grid(pd, ic, w, Q, nx, npx)
        s = beam_scale(it * dt)
        E = ops.solve(rho + s * profile if s else rho)
        eat = orc.field_to_particles(E, ic, w, nx, npx, 2)
        # typeB 2: bconst is the field on the grid, [Bx | By | Bz] (3 npExt), gathered here
        bat = orc.field_to_particles(bconst, ic, w, nx, npx, 3) if typeB == 2 else bconst
        pd = orc.boris_push(pd, eat, dt, qm, typeB, bat)
        ins, pd = orc.index_inside(pd, wall, g.getLx(), g.getLy())
        ab = orc.save_absorbed(pd, ins)
        sec = np.zeros((0, 6))
        if ab.shape[0]:
            moved, _, nv, _ = orc.interaction_and_normals(ab, wall, g.getLx(), g.getLy(),
                                                          g.getDx(), g.getDy())
            na = ab.shape[0]
            if secem is not None:
                from oracle import secem_oracle as so
                U = np.empty((na, so.N_UNIFORMS))
                for b in range(so.N_UNIFORMS // 2):
                    U[:, 2 * b:2 * b + 2] = orc.words_to_uniforms(
                        orc.philox_words(na, 0, b, it, 7, key[0], key[1]))
                sec, _ = so.generate_secondaries(moved, nv, U, secem[0], secem[1])
            else:
                U = np.empty((na, 36))
                for b in range(18):
                    U[:, 2 * b:2 * b + 2] = orc.words_to_uniforms(
                        orc.philox_words(na, 0, b, it, 1, key[0], key[1]))
                from oracle.furman_oracle import generate_secondaries
                sec, _, _ = generate_secondaries(moved, nv, U[:, :35], P)
        if photo is not None and photo[0][it] > 0:
            from oracle.photo_oracle import photo_emit
            ph = photo_emit(1, g.getLx(), g.getLy(), g.getDx(), g.getDy(), int(photo[0][it]), it,
                            key[0], key[1], photo[2], photo[3], float(photo[1][it]), 9.10938291e-31)
            sec = np.concatenate([sec, ph])
        if ab.shape[0] or sec.shape[0]:
            if rule == 'fast':
                pd = orc.add_and_remove_fast(pd, sec, ins)
            else:
                pd, cap = orc.add_and_remove(pd, sec, ins, cap)
        out.append(pd.copy())
    return out


@pytest.mark.parametrize('fused,ordered', [(False, False), (True, False), (False, True), (True, True)])
def test_buildup_matches_oracle_fp64(orc, fused, ordered):
    from openecloud_amd import randomGen
    from openecloud_amd.beam import GaussianBunchTrain
    from openecloud_amd.buildup import Buildup
    from openecloud_amd.grid import Grid
    from openecloud_amd.particleBoundary import AbsorbElliptical
    from openecloud_amd.particleEmitter import FurmanEmitter
    from openecloud_amd.particles import Particles
    from openecloud_amd.poissonSolver import PoissonSolver
    from oracle.furman_oracle import material
    g = Grid(64, 48, 0.09, 0.07, 1)
    n = 20000
    pd0 = random_particles(g, n, seed=11, wall='ellip', vscale=3e6)
    pd0[:, 5] = 5e4
    dt = 2e-11
    randomGen.seed(2024)
    key = randomGen.key()
    p = Particles('electrons', 'variableWeight', ordering='reference', orderedDeposit=ordered)
    p.setParticleData(pd0)
    p.setBConst([0., 0.5, 0.])
    pb = AbsorbElliptical(g, p)
    em = FurmanEmitter(pb, p, 'copper', seyMax=1.6, eHat0=300.)
    ps = PoissonSolver(g, method='direct', checkEvery=1, tol=1e-12)
    beam = GaussianBunchTrain(g, dt, nParticles=1e11, tBunchSpacing=25e-9, radiusSigma=2e-3,
                              nBunches=1, bunchLengthSigma=0.003)
    b = Buildup(g, p, pb, em, ps, beam, dt=dt, typeBField=1)
    nsteps = 12
    ref = _oracle_loop(orc, g, pd0.copy(), nsteps, dt, 1, np.array([0., 0.5, 0.]),
                       beam.getTransversalProfile(), beam.getScale,
                       material('copper', 1.6, eHat0=300.), key, n)
    emitted = 0
    for it in range(nsteps):
        b.step() if fused else b.step_reference()
        got = p.getParticleData().cpu().numpy()
        assert got.shape == ref[it].shape, (it, got.shape, ref[it].shape)
        scale = np.abs(ref[it]).max(axis=0)
        err = (np.abs(got - ref[it]) / scale).max()
        assert err < 1e-12, (it, err)
        emitted += b.history['emitted'][-1]
    assert emitted > 0 and sum(b.history['absorbed']) > 0


@pytest.mark.parametrize('dtype', ['f64', 'f32'])
def test_native_step_matches_oracle(orc, dtype):
    """The one-call native step (device-resident counts) against the oracle loop with the
    same hole-filling order; FP64 to 1e-12 of each column's scale; FP32 particles (the mixed
    push) to 1e-5 after 10 steps on the particles that no compaction has moved."""
    from openecloud_amd import randomGen
    from openecloud_amd.beam import GaussianBunchTrain
    from openecloud_amd.buildup import Buildup
    from openecloud_amd.grid import Grid
    from openecloud_amd.particleBoundary import AbsorbElliptical
    from openecloud_amd.particleEmitter import FurmanEmitter
    from openecloud_amd.particles import Particles
    from openecloud_amd.poissonSolver import PoissonSolver
    from oracle.furman_oracle import material
    g = Grid(64, 48, 0.09, 0.07, 1)
    n = 20000
    pd0 = random_particles(g, n, seed=12, wall='ellip', vscale=3e6)
    pd0[:, 5] = 5e4
    dt = 2e-11
    randomGen.seed(77)
    key = randomGen.key()
    tdt = torch.float64 if dtype == 'f64' else torch.float32
    if dtype == 'f32':
        pd0 = pd0.astype(np.float32).astype(np.float64)
    p = Particles('electrons', 'variableWeight', dtype=tdt, ordering='fast')
    p.setParticleData(pd0)
    p.setBConst([0., 0.5, 0.])
    pb = AbsorbElliptical(g, p)
    em = FurmanEmitter(pb, p, 'copper', seyMax=1.6, eHat0=300.)
    ps = PoissonSolver(g, method='direct', checkEvery=0)
    beam = GaussianBunchTrain(g, dt, nParticles=1e11, tBunchSpacing=25e-9, radiusSigma=2e-3,
                              nBunches=1, bunchLengthSigma=0.003)
    b = Buildup(g, p, pb, em, ps, beam, dt=dt, typeBField=1)
    b.native_setup(absCap=8192, secCap=16384, maxSteps=64)
    nsteps = 10
    ref = _oracle_loop(orc, g, pd0.copy(), nsteps, dt, 1, np.array([0., 0.5, 0.]),
                       beam.getTransversalProfile(), beam.getScale,
                       material('copper', 1.6, eHat0=300.), key, n, rule='fast')
    for it in range(nsteps):
        b.step_native()
        st = b.sync()
        got = p.getParticleData().cpu().numpy().astype(np.float64)
        if dtype == 'f64':
            assert got.shape == ref[it].shape, (it, got.shape, ref[it].shape)
            scale = np.abs(ref[it]).max(axis=0)
            assert (np.abs(got - ref[it]) / scale).max() < 1e-12, it
        else:
            # FP32 push: same particle count within a few wall-test flips, and the FP64
            # oracle's state to FP32 accuracy on the particles both still hold
            assert abs(got.shape[0] - ref[it].shape[0]) <= 0.002 * n, it
    assert st['absorbed_total'] > 0 and st['emitted_total'] > 0
    if dtype == 'f32':
        m = min(got.shape[0], ref[-1].shape[0])
        # compare the first particles, which no compaction has moved in either run
        idx = np.nonzero(np.all(np.abs(got[:2000, :2] - ref[-1][:2000, :2]) < 1e-6, axis=1))[0]
        assert idx.size > 1900
        scale = np.abs(ref[-1]).max(axis=0)
        err = (np.abs(got[idx] - ref[-1][idx]) / scale).max(axis=0)
        print('FP32 particles, %s push, after %d steps: %d of the first 2000 in place, max deviation per column %s'
              % (b.pushArithmetic, nsteps, idx.size, err), flush=True)
        assert err.max() < 1e-5


@pytest.mark.parametrize('native', [False, True])
def test_photoelectrons_in_the_step_match_oracle(orc, native):
    """A bunch passage with the photoelectron source: the Python fused step and the one-call
    native step against the oracle loop (photoelectrons after the secondaries, fast rule)."""
    from openecloud_amd import randomGen
    from openecloud_amd.beam import GaussianBunchTrain
    from openecloud_amd.buildup import Buildup
    from openecloud_amd.grid import Grid
    from openecloud_amd.particleBoundary import AbsorbElliptical
    from openecloud_amd.particleEmitter import FurmanEmitter, PhotoEmitter
    from openecloud_amd.particles import Particles
    from openecloud_amd.poissonSolver import PoissonSolver
    from oracle.furman_oracle import material
    g = Grid(64, 48, 0.09, 0.07, 1)
    n = 20000
    pd0 = random_particles(g, n, seed=13, wall='ellip', vscale=3e6)
    pd0[:, 5] = 5e4
    dt = 2e-11
    randomGen.seed(78)
    key = randomGen.key()
    p = Particles('electrons', 'variableWeight', ordering='fast')
    p.setParticleData(pd0)
    p.setBConst([0., 0.5, 0.])
    pb = AbsorbElliptical(g, p)
    em = FurmanEmitter(pb, p, 'copper', seyMax=1.6, eHat0=300.)
    ps = PoissonSolver(g, method='direct', checkEvery=0)
    beam = GaussianBunchTrain(g, dt, nParticles=1e11, tBunchSpacing=25e-9, radiusSigma=2e-3,
                              nBunches=1, bunchLengthSigma=0.003)
    photo = PhotoEmitter(pb, p, beam, photoYield=1e-3, macroPerBunch=2000)
    b = Buildup(g, p, pb, em, ps, beam, dt=dt, typeBField=1, photoEmitter=photo)
    nsteps = 8
    cnt, wts = photo.tables(nsteps)
    assert cnt.sum() > 1000
    # the weights carry the exact photoelectron charge of the steps with a bunch
    frac = np.array([beam.bunchFraction(k * dt) for k in range(nsteps)])
    assert abs((cnt * wts).sum() - 1e-3 * 1e11 * frac.sum()) <= 1e-12 * (cnt * wts).sum() + \
        1e-3 * 1e11 * max(photo._carry[1], 0.)
    ref = _oracle_loop(orc, g, pd0.copy(), nsteps, dt, 1, np.array([0., 0.5, 0.]),
                       beam.getTransversalProfile(), beam.getScale,
                       material('copper', 1.6, eHat0=300.), key, n, rule='fast',
                       photo=(cnt, wts, photo.energyMean, photo.energySigma))
    if native:
        b.native_setup(absCap=8192, secCap=16384, maxSteps=64)
    for it in range(nsteps):
        if native:
            b.step_native()
            st = b.sync()
        else:
            b.step()
        got = p.getParticleData().cpu().numpy()
        assert got.shape == ref[it].shape, (it, got.shape, ref[it].shape)
        scale = np.abs(ref[it]).max(axis=0)
        assert (np.abs(got - ref[it]) / scale).max() < 1e-12, it
    if native:
        assert st['photo_total'] == int(cnt.sum())
    else:
        assert sum(b.history['photo']) == int(cnt.sum())


@pytest.mark.parametrize('model', ['stainless', 'simple'])
@pytest.mark.parametrize('form', ['reference', 'fused', 'native'])
def test_simple_emitters_in_the_step_match_oracle(orc, model, form):
    """SecElecEmitter / SimpleSecEmitter as the step's emitter in the three step forms."""
    from openecloud_amd import randomGen
    from openecloud_amd.beam import GaussianBunchTrain
    from openecloud_amd.buildup import Buildup
    from openecloud_amd.grid import Grid
    from openecloud_amd.particleBoundary import AbsorbElliptical
    from openecloud_amd.particleEmitter import SecElecEmitter, SimpleSecEmitter
    from openecloud_amd.particles import Particles
    from openecloud_amd.poissonSolver import PoissonSolver
    from oracle import secem_oracle as so
    g = Grid(64, 48, 0.09, 0.07, 1)
    n = 20000
    pd0 = random_particles(g, n, seed=14, wall='ellip', vscale=3e6)
    pd0[:, 5] = 5e4
    dt = 2e-11
    randomGen.seed(91)
    key = randomGen.key()
    rule = 'fast' if form == 'native' else 'reference'
    p = Particles('electrons', 'variableWeight', ordering=rule)
    p.setParticleData(pd0)
    p.setBConst([0., 0.5, 0.])
    pb = AbsorbElliptical(g, p)
    em = (SecElecEmitter if model == 'stainless' else SimpleSecEmitter)(pb, p, seyMax=2.0)
    ps = PoissonSolver(g, method='direct', checkEvery=0)
    beam = GaussianBunchTrain(g, dt, nParticles=1e11, tBunchSpacing=25e-9, radiusSigma=2e-3,
                              nBunches=1, bunchLengthSigma=0.003)
    b = Buildup(g, p, pb, em, ps, beam, dt=dt, typeBField=1)
    if form == 'native':
        b.native_setup(absCap=8192, secCap=16384, maxSteps=64)
    nsteps = 8
    ref = _oracle_loop(orc, g, pd0.copy(), nsteps, dt, 1, np.array([0., 0.5, 0.]),
                       beam.getTransversalProfile(), beam.getScale, None, key, n, rule=rule,
                       secem=(model, so.stainless(seyMax=2.0)))
    for it in range(nsteps):
        if form == 'native':
            b.step_native()
            b.sync()
        elif form == 'fused':
            b.step()
        else:
            b.step_reference()
        got = p.getParticleData().cpu().numpy()
        assert got.shape == ref[it].shape, (it, got.shape, ref[it].shape)
        scale = np.abs(ref[it]).max(axis=0)
        err = (np.abs(got - ref[it]) / scale).max()
        assert err < 1e-12, (it, err)
    assert ref[-1].shape[0] != n


def test_buildup_with_learned_solve_matches_oracle(orc):
    """The reference-order step with PoissonSolver(method='learned'): the flexible CG with a
    network trained on this grid, warm-started from the previous potential, to 1e-11; the
    direct solve judges every solve.  The particle state follows the oracle's FP64 loop (whose
    solve is exact) to the accuracy the tolerance allows, with no fallback.  (Measured: 52-92
    iterations a step; the bunch changes the charge too much from one step to the next for the
    warm start to save much here.)"""
    from openecloud_amd import randomGen
    from openecloud_amd.beam import GaussianBunchTrain
    from openecloud_amd.buildup import Buildup
    from openecloud_amd.grid import Grid
    from openecloud_amd.learnedPreconditioner import UNetPreconditioner
    from openecloud_amd.particleBoundary import AbsorbElliptical
    from openecloud_amd.particleEmitter import FurmanEmitter
    from openecloud_amd.particles import Particles
    from openecloud_amd.poissonSolver import PoissonSolver
    from oracle.furman_oracle import material
    g = Grid(64, 48, 0.09, 0.07, 1)
    n = 20000
    pd0 = random_particles(g, n, seed=11, wall='ellip', vscale=3e6)
    pd0[:, 5] = 5e4
    dt = 2e-11
    randomGen.seed(2024)
    key = randomGen.key()
    p = Particles('electrons', 'variableWeight', ordering='reference')
    p.setParticleData(pd0)
    p.setBConst([0., 0.5, 0.])
    pb = AbsorbElliptical(g, p)
    em = FurmanEmitter(pb, p, 'copper', seyMax=1.6, eHat0=300.)
    ps = PoissonSolver(g, method='learned', tol=1e-11, maxIter=2000, checkEvery=1)
    net = UNetPreconditioner(g, seed=0)
    net.train(ps, net.smoothed_pairs(ps, net.operator_pairs(ps, 32, seed=1)), steps=300, lr=5e-4, batch=2)
    ps.setPreconditioner(net)
    beam = GaussianBunchTrain(g, dt, nParticles=1e11, tBunchSpacing=25e-9, radiusSigma=2e-3,
                              nBunches=1, bunchLengthSigma=0.003)
    b = Buildup(g, p, pb, em, ps, beam, dt=dt, typeBField=1)
    nsteps = 12
    ref = _oracle_loop(orc, g, pd0.copy(), nsteps, dt, 1, np.array([0., 0.5, 0.]),
                       beam.getTransversalProfile(), beam.getScale,
                       material('copper', 1.6, eHat0=300.), key, n)
    its = []
    for it in range(nsteps):
        b.step_reference()
        its.append(ps.lastIterations)
        got = p.getParticleData().cpu().numpy()
        assert got.shape == ref[it].shape, (it, got.shape, ref[it].shape)
        scale = np.abs(ref[it]).max(axis=0)
        err = (np.abs(got - ref[it]) / scale).max()
        assert err < 1e-8, (it, err, its)
    print('learned solve in the buildup: iterations', its, 'judge max rel err', ps.judgeMaxErr)
    assert ps.fallbacks == 0 and ps.judgeCount == nsteps and ps.judgeMaxErr < 1e-8, (its, ps.judgeMaxErr)


def test_native_step_deposit_paths_match_oracle(orc):
    """The push's deposit merges run totals in one LDS tile per block when the block's cells are
    close, in per-wave tiles when each wave's cells are close but the block's are not, and adds
    them to rho directly otherwise.  Particles laid out to take each path (4 blocks of 256 each:
    compact blocks, blocks of four far-apart compact waves, scattered blocks), one native step
    with no wall hits, and the deposited charge against the oracle's deposit of the stepped
    particles."""
    from openecloud_amd import randomGen
    from openecloud_amd.beam import GaussianBunchTrain
    from openecloud_amd.buildup import Buildup
    from openecloud_amd.grid import Grid
    from openecloud_amd.particleBoundary import AbsorbElliptical
    from openecloud_amd.particleEmitter import FurmanEmitter
    from openecloud_amd.particles import Particles
    from openecloud_amd.poissonSolver import PoissonSolver
    g = Grid(128, 96, 0.09, 0.07, 1)
    rng = np.random.default_rng(21)
    dx, dy = g.getDx(), g.getDy()

    def cloud(cx, cy, m):      # m particles within one cell of (cx, cy)
        return np.column_stack([cx + rng.uniform(-dx, dx, m), cy + rng.uniform(-dy, dy, m)])
    pos = []
    for k in range(4):         # compact blocks: all 4 waves around one point
        pos.append(cloud(-0.02 + 0.01 * k, 0.005 * k, 256))
    for k in range(4):         # far-apart waves, each compact (a block box of ~76 x 42 cells)
        for cx, cy in ((-0.025, -0.015), (0.025, 0.015), (-0.025, 0.015), (0.025, -0.015)):
            pos.append(cloud(cx + 0.001 * k, cy, 64))
    for k in range(4):         # scattered
        r = np.sqrt(rng.uniform(0, 1, 256)) * 0.8
        t = rng.uniform(0, 2 * np.pi, 256)
        pos.append(np.column_stack([0.045 * r * np.cos(t), 0.035 * r * np.sin(t)]))
    xy = np.concatenate(pos)
    n = xy.shape[0]
    pd0 = np.zeros((n, 6))
    pd0[:, :2] = xy
    pd0[:, 2:5] = rng.normal(0., 1e5, (n, 3))
    pd0[:, 5] = rng.uniform(0.5, 1.5, n) * 1e5
    randomGen.seed(5)
    p = Particles('electrons', 'variableWeight', ordering='fast')
    p.setParticleData(pd0)
    p.setBConst([0., 0.5, 0.])
    pb = AbsorbElliptical(g, p)
    em = FurmanEmitter(pb, p, 'copper', seyMax=1.6, eHat0=300.)
    ps = PoissonSolver(g, method='direct', checkEvery=0)
    beam = GaussianBunchTrain(g, 1e-12, nParticles=1e11, tBunchSpacing=25e-9, radiusSigma=2e-3, nBunches=1)
    b = Buildup(g, p, pb, em, ps, beam, dt=1e-12, typeBField=1)
    b.native_setup(absCap=4096, secCap=8192, maxSteps=8)
    b.step_native()
    st = b.sync()
    assert st['absorbed_total'] == 0 and p.getMacroParticleCount() == n
    pd = p.getParticleData().cpu().numpy()
    ic, w = orc.calc_grid_weights(pd, dx, dy, g.getLxExt(), g.getLyExt(), g.getNxExt())
    rho_o = orc.charge_to_grid(pd, ic, w, p.getParticleCharge(), g.getNxExt(), g.getNpExt())
    rho = b.rho.cpu().numpy()
    assert np.allclose(rho, rho_o, rtol=1e-13, atol=1e-13 * np.abs(rho_o).max())


@pytest.mark.parametrize('layout', ['eb', 'separate'])
def test_native_step_grid_b_matches_oracle(orc, layout):
    """B on the grid (quadrupole plus a skew term, typeBField 2, gathered at each particle like
    E): the one-call native step against the oracle loop, FP64 to 1e-12 of each column's
    scale (a swapped or dropped B component, with fields of ~2 T, would differ at order one).
    Bz is zero, so the step gathers B from the (Bx, By) layout and E from its interleaved copy
    ('eb'); 'separate' takes that layout away and gathers B from the three B arrays."""
    from openecloud_amd import randomGen
    from openecloud_amd.beam import GaussianBunchTrain
    from openecloud_amd.buildup import Buildup
    from openecloud_amd.grid import Grid
    from openecloud_amd.magneticField import multipoleExpansion
    from openecloud_amd.particleBoundary import AbsorbElliptical
    from openecloud_amd.particleEmitter import FurmanEmitter
    from openecloud_amd.particles import Particles
    from openecloud_amd.poissonSolver import PoissonSolver
    from oracle.furman_oracle import material
    g = Grid(64, 48, 0.09, 0.07, 1)
    bgrid = np.asarray(multipoleExpansion(g, [0., 50.], [0., 3.]), np.float64)
    n = 20000
    pd0 = random_particles(g, n, seed=21, wall='ellip', vscale=3e6)
    pd0[:, 5] = 5e4
    dt = 2e-11
    randomGen.seed(91)
    key = randomGen.key()
    p = Particles('electrons', 'variableWeight', ordering='fast')
    p.setParticleData(pd0)
    pb = AbsorbElliptical(g, p)
    em = FurmanEmitter(pb, p, 'copper', seyMax=1.6, eHat0=300.)
    ps = PoissonSolver(g, method='direct', checkEvery=0)
    beam = GaussianBunchTrain(g, dt, nParticles=1e11, tBunchSpacing=25e-9, radiusSigma=2e-3,
                              nBunches=1, bunchLengthSigma=0.003)
    b = Buildup(g, p, pb, em, ps, beam, dt=dt, typeBField=2, bAtGridPoints=bgrid)
    b.native_setup(absCap=8192, secCap=16384, maxSteps=64)
    assert b._bil is not None
    if layout == 'separate':
        b._args.Bil = None
    nsteps = 10
    ref = _oracle_loop(orc, g, pd0.copy(), nsteps, dt, 2, bgrid, beam.getTransversalProfile(),
                       beam.getScale, material('copper', 1.6, eHat0=300.), key, n, rule='fast')
    for it in range(nsteps):
        b.step_native()
        st = b.sync()
        got = p.getParticleData().cpu().numpy()
        assert got.shape == ref[it].shape, (it, got.shape, ref[it].shape)
        scale = np.abs(ref[it]).max(axis=0)
        err = (np.abs(got - ref[it]) / scale).max()
        assert err < 1e-12, (it, err)
    assert st['absorbed_total'] > 0 and st['emitted_total'] > 0


@pytest.mark.parametrize('typeB,path', [(0, 'classes'), (2, 'classes'), (2, 'csr')])
def test_native_step_field_equals_the_csr_product(typeB, path):
    """E written by the step (row classes of phiToE, or the plain CSR product when the classes are
    taken away) equals the library's CSR product of the step's phi bit for bit, and the
    interleaved (Ex, Ey) copy the push reads holds the same values (the (Bx, By) copy of the grid
    field untouched)."""
    from openecloud_amd import randomGen
    from openecloud_amd.beam import GaussianBunchTrain
    from openecloud_amd.buildup import Buildup
    from openecloud_amd.grid import Grid
    from openecloud_amd.magneticField import multipoleExpansion
    from openecloud_amd.particleBoundary import AbsorbElliptical
    from openecloud_amd.particleEmitter import FurmanEmitter
    from openecloud_amd.particles import Particles
    from openecloud_amd.poissonSolver import PoissonSolver
    g = Grid(96, 72, 0.09, 0.07, 1)
    npx = g.getNpExt()
    bgrid = np.asarray(multipoleExpansion(g, [0., 50.]), np.float64) if typeB == 2 else None
    pd0 = random_particles(g, 30000, seed=4, wall='ellip', vscale=3e6)
    pd0[:, 5] = 5e4
    randomGen.seed(7)
    p = Particles('electrons', 'variableWeight', ordering='fast')
    p.setParticleData(pd0)
    pb = AbsorbElliptical(g, p)
    em = FurmanEmitter(pb, p, 'copper', seyMax=1.6)
    ps = PoissonSolver(g, method='direct', checkEvery=0)
    assert 0 < int(ps.e2nIrr.numel()) < 2 * npx // 10
    beam = GaussianBunchTrain(g, 2e-11, nParticles=1e11, tBunchSpacing=25e-9, radiusSigma=2e-3, nBunches=1,
                              bunchLengthSigma=0.003)
    b = Buildup(g, p, pb, em, ps, beam, dt=2e-11, typeBField=typeB, bAtGridPoints=bgrid)
    b.native_setup(absCap=8192, secCap=16384, maxSteps=8)
    if path == 'csr':
        b._args.e2nKind = None
    for _ in range(3):
        b.step_native()
    b.sync()
    E = ps.eAtGridPoints.clone()
    ref = ps.phiToE().clone()
    assert torch.equal(E, ref)
    il = b._eil.view(npx, 2)
    assert torch.equal(il[:, 0], E[:npx]) and torch.equal(il[:, 1], E[npx:])
    if typeB == 2:
        B = torch.as_tensor(bgrid).cuda()
        assert torch.equal(b._bil[:, 0], B[:npx]) and torch.equal(b._bil[:, 1], B[npx:2 * npx])
