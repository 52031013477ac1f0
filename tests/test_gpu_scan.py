"""Batched parameter scans (BuildupScan, oecl_step_run_batch): every run of a scan against the
same run stepped alone by the one-call native step, graph replays, and the batched sort."""
import numpy as np
import pytest
import torch

from conftest import random_particles

pytestmark = pytest.mark.gpu

SEEDS = (101, 202, 303)
NS = (5000, 12000, 20000)
SEY = (1.3, 1.6, 2.0)
PPB = (5e10, 1e11, 2e11)


def _shared(dtype=torch.float64, grid=(64, 48, 0.09, 0.07, 1)):
    from openecloud_amd.grid import Grid
    from openecloud_amd.poissonSolver import PoissonSolver
    g = Grid(*grid)
    ps = PoissonSolver(g, method='direct', checkEvery=0)
    return g, ps


def _run(g, ps, r, dtype=torch.float64):
    from openecloud_amd.beam import GaussianBunchTrain
    from openecloud_amd.buildup import Buildup
    from openecloud_amd.particleBoundary import AbsorbElliptical
    from openecloud_amd.particleEmitter import FurmanEmitter
    from openecloud_amd.particles import Particles
    pd = random_particles(g, NS[r], seed=40 + r, wall='ellip', vscale=3e6)
    pd[:, 5] = 5e4
    if dtype == torch.float32:
        pd = pd.astype(np.float32).astype(np.float64)
    dt = 2e-11
    p = Particles('electrons', 'variableWeight', dtype=dtype, ordering='fast')
    p.setParticleData(pd)
    p.setBConst([0., 0.3, 0.])
    pb = AbsorbElliptical(g, p)
    em = FurmanEmitter(pb, p, 'copper', seyMax=SEY[r], eHat0=300.)
    beam = GaussianBunchTrain(g, dt, nParticles=PPB[r], tBunchSpacing=25e-9, radiusSigma=2e-3,
                              nBunches=1, bunchLengthSigma=0.003)
    return Buildup(g, p, pb, em, ps, beam, dt=dt, typeBField=1)


def _alone(g, ps, nsteps, sort_at=(), dtype=torch.float64):
    from openecloud_amd import randomGen
    out = []
    for r in range(3):
        b = _run(g, ps, r, dtype)
        randomGen.seed(SEEDS[r])
        b.native_setup(absCap=4096, secCap=8192, maxSteps=64)
        states = []
        for it in range(nsteps):
            if it in sort_at:
                b.sync()
                b.sort_particles(keepBuffers=True)
            b.step_native()
            b.sync()
            states.append(b.p.getParticleData().cpu().numpy().astype(np.float64).copy())
        out.append(states)
    return out


def _scan(g, ps, dtype=torch.float64):
    from openecloud_amd.scan import BuildupScan
    return BuildupScan([_run(g, ps, r, dtype) for r in range(3)], SEEDS, maxSteps=64)


def _close(got, ref, tol):
    assert got.shape == ref.shape, (got.shape, ref.shape)
    scale = np.abs(ref).max(axis=0)
    return (np.abs(got - ref) / scale).max() < tol


def test_scan_runs_match_standalone_native_steps():
    g, ps = _shared()
    nsteps, sort_at = 10, (5,)
    ref = _alone(g, ps, nsteps, sort_at)
    sc = _scan(g, ps)
    for it in range(nsteps):
        if it in sort_at:
            sc.sort()
        sc.step()
        st = sc.sync()
        for r in range(3):
            got = sc.run_particles(r).cpu().numpy()
            assert _close(got, ref[r][it], 1e-12), (it, r)
    assert all(a > 0 for a in st['absorbed_total']) and sum(st['emitted_total']) > 0
    assert st['counts'] == [ref[r][-1].shape[0] for r in range(3)]
    # the per-run histories hold each step's starting count
    assert sc.history(2, 1, nsteps) == [ref[2][k].shape[0] for k in range(nsteps - 1)]


def test_scan_runs_match_standalone_on_a_2048_wide_box():
    """The same on a flat 2048 x 32 ellipse, whose 2047-column Poisson box takes the persistent
    2048-point DST passes: the three runs' lines in one launch of each pass."""
    g, ps = _shared(grid=(2048, 32, 0.09, 0.0015, 1))
    nsteps = 3
    ref = _alone(g, ps, nsteps)
    sc = _scan(g, ps)
    for it in range(nsteps):
        sc.step()
        sc.sync()
        for r in range(3):
            got = sc.run_particles(r).cpu().numpy()
            assert _close(got, ref[r][it], 1e-12), (it, r)


def test_scan_graph_replays_match_eager_steps():
    g, ps = _shared()
    eager = _scan(g, ps)
    for _ in range(8):
        eager.step()
    eager.sync()
    sc = _scan(g, ps)
    sc.step()             # first step eagerly (allocations happen outside the capture)
    sc.capture(7)
    sc.replay()
    st = sc.sync()
    assert sc.it == eager.it == 8
    for r in range(3):
        assert _close(sc.run_particles(r).cpu().numpy(), eager.run_particles(r).cpu().numpy(), 1e-12), r
    assert st['counts'] == eager.counts()


@pytest.mark.parametrize('order,binShift', [(0, 0), (1, 0), (0, 2), (1, 1)])
def test_scan_sort_orders_each_run_in_its_segment(order, binShift):
    g, ps = _shared()
    sc = _scan(g, ps)
    for _ in range(3):
        sc.step()
    sc.sync()
    before = [sc.run_particles(r).cpu().numpy().copy() for r in range(3)]
    sc.sort(order, binShift)
    nx = g.getNxExt()
    for r in range(3):
        after = sc.run_particles(r).cpu().numpy()
        # same particles (as a multiset of rows), now in bin order, stable within a bin
        key = lambda a: np.lexsort(a.T[::-1])
        assert np.array_equal(before[r][key(before[r])], after[key(after)])
        ix = np.floor((after[:, 0] + 0.5 * g.getLxExt()) / g.getDx()).astype(np.int64) >> binShift
        iy = np.floor((after[:, 1] + 0.5 * g.getLyExt()) / g.getDy()).astype(np.int64) >> binShift
        if order == 0:
            b = ix + iy * (((nx - 1) >> binShift) + 1)
        else:
            b = np.zeros_like(ix)
            for k in range(16):
                b |= ((ix >> k) & 1) << (2 * k) | ((iy >> k) & 1) << (2 * k + 1)
        assert np.all(np.diff(b) >= 0), r


def test_scan_fp32_runs_match_standalone():
    g, ps = _shared()
    nsteps = 5
    ref = _alone(g, ps, nsteps, dtype=torch.float32)
    sc = _scan(g, ps, dtype=torch.float32)
    for it in range(nsteps):
        sc.step()
    sc.sync()
    for r in range(3):
        got = sc.run_particles(r).cpu().numpy().astype(np.float64)
        assert abs(got.shape[0] - ref[r][-1].shape[0]) <= 2
        m = min(got.shape[0], ref[r][-1].shape[0])
        idx = np.nonzero(np.all(np.abs(got[:m, :2] - ref[r][-1][:m, :2]) < 1e-7, axis=1))[0]
        assert idx.size > 0.95 * m
        scale = np.abs(ref[r][-1]).max(axis=0)
        assert (np.abs(got[idx] - ref[r][-1][idx]) / scale).max() < 1e-5


def test_scan_with_learned_solve_follows_the_direct_scan():
    """The batched step with every run's potential from the learned flexible CG (the network
    trained briefly on this grid) against the same scan with the direct solve: the particles
    agree to the solve tolerance's effect, no run falls back, and the judge confirms the
    potentials."""
    from openecloud_amd.learnedPreconditioner import UNetPreconditioner
    g, ps = _shared()
    net = UNetPreconditioner(g, seed=0)
    net.train(ps, net.smoothed_pairs(ps, net.operator_pairs(ps, 16, seed=1)), steps=300, lr=5e-4, batch=2)
    ref, lrn = _scan(g, ps), _scan(g, ps)
    lrn.set_solver(net, tol=1e-12, maxIter=4000, judgeEvery=2)
    for it in range(6):
        ref.step()
        lrn.step()
    ref.sync()
    lrn.sync()
    st = lrn.learnedStats
    assert st['fallbacks'] == 0 and st['relres_max'] <= 1e-12 and st['judged'] > 0
    assert st['judge_max_rel_err'] < 1e-9
    for r in range(3):
        assert _close(lrn.run_particles(r).cpu().numpy(), ref.run_particles(r).cpu().numpy(), 1e-9), r


@pytest.mark.parametrize('span,waveTiles', [(16, False), (64, False), (4, True), (2, True)])
def test_scan_deposit_tiles_change_only_the_summation_order(span, waveTiles):
    """The autotuner's deposit tile choices (run-sum span, per-wave tiles): every run of the scan
    follows the default-tile scan to 1e-12 (FP64) over 8 steps, with the same counts."""
    g, ps = _shared()
    ref = _scan(g, ps)
    sc = _scan(g, ps)
    sc.set_deposit(span, waveTiles)
    for it in range(8):
        ref.step()
        sc.step()
    st, st0 = sc.sync(), ref.sync()
    assert st['counts'] == st0['counts']
    for r in range(3):
        assert _close(sc.run_particles(r).cpu().numpy(), ref.run_particles(r).cpu().numpy(), 1e-12), r
    with pytest.raises(ValueError):
        sc.set_deposit(3)


def test_autotuner_picks_a_sort_interval_and_tiles():
    """tune_scan on a small scan: both stages are timed, the choice is one of the candidates and
    the scan is left set to its tiles; the cache returns the same choice without timing."""
    import os
    import tempfile
    from openecloud_amd.autotune import tune_scan
    g, ps = _shared(torch.float32)
    sc = _scan(g, ps, torch.float32)
    cands = [dict(sortEvery=2, order=0, binShift=0), dict(sortEvery=3, order=1, binShift=1)]
    tiles = [dict(span=4, waveTiles=False), dict(span=16, waveTiles=True)]
    with tempfile.TemporaryDirectory() as d:
        cache = os.path.join(d, 'tune.json')
        best, results = tune_scan(sc, candidates=cands, rounds=1, cacheFile=cache, tiles=tiles)
        assert len(results) == 4
        assert best['sortEvery'] in (2, 3) and best['graphSteps'] == best['sortEvery']
        assert (best['span'], best['waveTiles']) in [(4, False), (16, True)]
        assert (int(sc._args.depSpan), bool(sc._args.depTile)) == (best['span'], best['waveTiles'])
        sc.set_deposit(0, False)
        again, res2 = tune_scan(sc, candidates=cands, rounds=1, cacheFile=cache, tiles=tiles)
        assert res2 == [] and again == best
        assert (int(sc._args.depSpan), bool(sc._args.depTile)) == (best['span'], best['waveTiles'])
