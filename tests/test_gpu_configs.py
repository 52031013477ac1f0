"""GPU parity at the sizes of the five benchmark configurations, through the same classes and
C ABI as the smaller tests:

  C1  128x64 rectangular chamber, 1e5 macro-electrons, no magnetic field, endless LHC beam,
      Furman delta_max 1.6 / E_max 300 eV: whole buildup steps against the oracle loop
      (example01_basic.py:52-97), method by method and as the one-call step, FP64 to 1e-12;
  C2  512x384 elliptical chamber, 2e6 macro-electrons, By 0.5 T, photoelectrons: the one-call
      step against the oracle loop, FP64 to 1e-12; the potential against SuperLU to 1e-10;
      FP32 storage against the FP64 oracle;
  C3  64 buildups on a 256x256 rectangular chamber stepped together: two runs against the
      oracle loop, FP64 to 1e-12;
  C4  1024x1024 cut-cell chamber with a conductor: the direct solve against SuperLU to 1e-10,
      and the learned flexible CG to 1e-8;
  C5  2048x2048 elliptical chamber, 2e7 FP32 macro-electrons, 50 T/m quadrupole on the grid:
      the direct solve against its CG-refined solution to 1e-10, a sample of the push against
      the oracle, and the one-call step's particle and charge bookkeeping.
Each test bounds its own steps; the oracle runs on samples where the whole set would be slow."""
import numpy as np
import pytest
import torch

from test_gpu_buildup import _oracle_loop

pytestmark = pytest.mark.gpu

Q = -1.602176565e-19
M_E = 9.10938291e-31


def _cmp(got, ref):
    assert got.shape == ref.shape, (got.shape, ref.shape)
    scale = np.abs(ref).max(axis=0)
    return float((np.abs(got - ref) / scale).max())


# ------------------------------------------------------------------------------------------ C1

@pytest.mark.parametrize('mode', ['reference', 'native'])
def test_c1_buildup_matches_oracle(orc, mode):
    from openecloud_amd import randomGen
    from openecloud_amd.beam import LHCBeamEndless
    from openecloud_amd.buildup import Buildup
    from openecloud_amd.grid import Grid
    from openecloud_amd.particleBoundary import AbsorbRectangular
    from openecloud_amd.particleEmitter import FurmanEmitter, homoLoader
    from openecloud_amd.particles import Particles
    from openecloud_amd.poissonSolver import PoissonSolver
    from oracle.furman_oracle import material
    print('C1 buildup, mode %s' % mode, flush=True)
    g = Grid(128, 64, 0.1, 0.05, 0)
    n = 100000
    dt = 1e-11
    randomGen.seed(0)
    key = randomGen.key()
    ref_order = mode == 'reference'
    p = Particles('electrons', 'variableWeight', ordering='reference' if ref_order else 'fast',
                  orderedDeposit=True)
    pb = AbsorbRectangular(g, p)
    homoLoader(g, p, pb, n, 1e9 / n, 4.194e5)
    pd0 = p.getParticleData().cpu().numpy().copy()
    em = FurmanEmitter(pb, p, 'copper', seyMax=1.6, eHat0=300.)
    ps = PoissonSolver(g, method='direct', checkEvery=0)
    beam = LHCBeamEndless(g, dt)
    b = Buildup(g, p, pb, em, ps, beam, dt=dt, typeBField=0)
    nsteps = 6
    ref = _oracle_loop(orc, g, pd0.copy(), nsteps, dt, 0, np.zeros(3), beam.getTransversalProfile(),
                       beam.getScale, material('copper', 1.6, eHat0=300.), key, n,
                       rule='reference' if ref_order else 'fast', wall='rect')
    if not ref_order:
        b.native_setup(absCap=16384, secCap=32768, maxSteps=16)
    for it in range(nsteps):
        if ref_order:
            b.step_reference()
        else:
            b.step_native()
            b.sync()
        err = _cmp(p.getParticleData().cpu().numpy(), ref[it])
        assert err < 1e-12, (it, err)
    assert beam.getScale(0.) > 0.


# ------------------------------------------------------------------------------------------ C2

def _c2(dtype, n=2000000):
    import bench
    return bench.build_dipole(n, dtype)


def test_c2_full_size_step_matches_oracle(orc):
    from openecloud_amd import randomGen
    from oracle.furman_oracle import material
    from oracle.grid_oracle import OraclePoisson
    print('C2 full size, FP64', flush=True)
    g, p, pb, em, ps, b = _c2('f64')
    key = randomGen.key()
    n = p.getMacroParticleCount()
    pd0 = p.getParticleData().cpu().numpy().copy()
    b.native_setup(absCap=65536, secCap=131072, maxSteps=16)
    nsteps = 3
    cnt, wts = b.photo.tables(nsteps)
    ops = OraclePoisson(g)
    ref = _oracle_loop(orc, g, pd0.copy(), nsteps, b.dt, 1, np.array([0., 0.5, 0.]),
                       b.beam.getTransversalProfile(), b.beam.getScale, material('copper', 1.3), key, n,
                       rule='fast', photo=(cnt, wts, 7., 5.), ops=ops)
    added = 0
    for it in range(nsteps):
        b.step_native()
        st = b.sync()
        added += st['added_last']
        err = _cmp(p.getParticleData().cpu().numpy(), ref[it])
        assert err < 1e-12, (it, err)
    assert st['absorbed_total'] > 0 and added > 0
    # the potential of the last step's charge against SuperLU (poissonSolver.pyx:83-88)
    q = (b.rho + b.beam.getScale(nsteps * b.dt) * b.beam.getProfile()).cpu().numpy()
    e_ref = ops.solve(q)
    e = ps.solve(torch.as_tensor(q).cuda()).cpu().numpy()
    phi = ps.getPhi().cpu().numpy()
    assert np.abs(phi - ops.phi).max() <= 1e-10 * np.abs(ops.phi).max()
    assert np.abs(e - e_ref).max() <= 1e-10 * np.abs(e_ref).max()


def test_c2_fp32_storage_against_fp64_oracle(orc):
    """FP32 particle storage at full size: after 3 steps, the particles that neither run has
    moved by compaction agree with the FP64 oracle to 1e-5 of each column's scale, and the
    counts agree within the wall-test flips that FP32 positions cause."""
    from openecloud_amd import randomGen
    from oracle.furman_oracle import material
    print('C2 full size, FP32 storage', flush=True)
    g, p, pb, em, ps, b = _c2('f32')
    key = randomGen.key()
    n = p.getMacroParticleCount()
    pd0 = p.getParticleData().cpu().numpy().astype(np.float64)
    b.native_setup(absCap=65536, secCap=131072, maxSteps=16)
    nsteps = 3
    cnt, wts = b.photo.tables(nsteps)
    ref = _oracle_loop(orc, g, pd0.copy(), nsteps, b.dt, 1, np.array([0., 0.5, 0.]),
                       b.beam.getTransversalProfile(), b.beam.getScale, material('copper', 1.3), key, n,
                       rule='fast', photo=(cnt, wts, 7., 5.))
    for it in range(nsteps):
        b.step_native()
        b.sync()
    got = p.getParticleData().cpu().numpy().astype(np.float64)
    assert abs(got.shape[0] - ref[-1].shape[0]) <= 1e-4 * n
    m = min(got.shape[0], ref[-1].shape[0]) // 2
    scale = np.abs(ref[-1]).max(axis=0)
    rel = np.abs(got[:m] - ref[-1][:m]) / scale
    # the same particle in both runs: all six columns within 1e-3 (secondaries of one impact
    # share its position, so position alone does not identify a particle); those must then
    # agree to 1e-5
    same = np.all(rel < 1e-3, axis=1)
    assert same.mean() > 0.99
    err = rel[same].max()
    print('C2 FP32 storage: %.2e of the particles matched, max error %.2e' % (same.mean(), err), flush=True)
    assert err < 1e-5, err


# ------------------------------------------------------------------------------------------ C3

def test_c3_scan_runs_match_oracle(orc):
    """The 64-run scan (256x256 rectangular chamber, By 0.3 T, delta_max over [1.1, 2.0],
    log-uniform intensities and counts), FP64: every run advanced by the batched step; the
    smallest and a middle-sized run against the oracle loop with their own parameters."""
    import bench
    from oracle.furman_oracle import material
    from openecloud_amd.scan import seed_key
    print('C3 scan of 64 runs', flush=True)
    g, ps, runs, scan, prm = bench.build_scan(64, 'f64', 16)
    counts = np.asarray(prm['n'])
    pick = [int(np.argmin(counts)), int(np.argsort(counts)[len(counts) // 3])]
    pd0 = {r: scan.run_particles(r).cpu().numpy().copy() for r in pick}
    nsteps = 4
    states = {r: [] for r in pick}
    for it in range(nsteps):
        scan.step()
        scan.sync()
        for r in pick:
            states[r].append(scan.run_particles(r).cpu().numpy().copy())
    from oracle.grid_oracle import OraclePoisson
    ops = OraclePoisson(g)
    for r in pick:
        b = runs[r]
        ref = _oracle_loop(orc, g, pd0[r].copy(), nsteps, b.dt, 1, np.array([0., 0.3, 0.]),
                           b.beam.getTransversalProfile(), b.beam.getScale,
                           material('copper', float(prm['sey'][r]), eHat0=300.), seed_key(1000 + r),
                           pd0[r].shape[0], rule='fast', wall='rect', ops=ops)
        for it in range(nsteps):
            err = _cmp(states[r][it], ref[it])
            assert err < 1e-12, (r, it, err)


# ------------------------------------------------------------------------------------------ C4

def _c4_grid():
    from openecloud_amd.grid import Grid
    ax, ay, xc, rc = 0.045, 0.035, 0.02, 0.0025

    def bf(x, y):
        x = np.asarray(x, np.float64)
        y = np.asarray(y, np.float64)
        return ((x * x / (ax * ax) + y * y / (ay * ay) < 1.) & ((x - xc) ** 2 + y * y > rc * rc)).astype(np.int64)
    return Grid(1024, 1024, 2 * ax, 2 * ay, 3, boundFunc=bf)


def test_c4_cutcell_1024_direct_and_learned_solves():
    from openecloud_amd.learnedPreconditioner import UNetPreconditioner
    from openecloud_amd.poissonSolver import PoissonSolver
    from oracle.grid_oracle import OraclePoisson
    print('C4 1024^2 cut cell: direct solve against SuperLU', flush=True)
    g = _c4_grid()
    ps = PoissonSolver(g, method='direct', checkEvery=0)
    rng = np.random.default_rng(4)
    npx = g.getNpExt()
    # a cloud-like charge: smooth random field plus a Gaussian beam term, on the inside nodes
    xs = (np.arange(g.getNxExt()) - g.getNxExt() / 2) * g.getDx()
    ys = (np.arange(g.getNyExt()) - g.getNyExt() / 2) * g.getDy()
    X, Y = np.meshgrid(xs, ys)
    q = (np.exp(-(X ** 2 + Y ** 2) / (2 * 0.0015 ** 2)) * 1e-9 +
         rng.standard_normal((g.getNyExt(), g.getNxExt())) * 1e-12).ravel()
    ops = OraclePoisson(g)
    e_ref = ops.solve(q)
    e = ps.solve(torch.as_tensor(q).cuda()).cpu().numpy()
    phi = ps.getPhi().cpu().numpy()
    assert np.abs(phi - ops.phi).max() <= 1e-10 * np.abs(ops.phi).max()
    assert np.abs(e - e_ref).max() <= 1e-10 * np.abs(e_ref).max()
    print('C4 learned flexible CG to 1e-8', flush=True)
    net = UNetPreconditioner(g, seed=0)
    train = net.smoothed_pairs(ps, net.operator_pairs(ps, 16, seed=1))
    net.train(ps, train, steps=1500, lr=5e-4, batch=2)
    qd = torch.as_tensor(q).cuda()
    z, it, rr, ok = net.fcg(ps, qd, tol=1e-8, maxIter=20000)
    print('C4 learned flexible CG: %d iterations, relative residual %.2e' % (it, rr), flush=True)
    assert ok and rr <= 1e-8, (it, rr)
    assert float(torch.linalg.norm(z.cpu() - torch.as_tensor(ops.phi)) / np.linalg.norm(ops.phi)) < 1e-6


# ------------------------------------------------------------------------------------------ C5

@pytest.fixture(scope='module')
def c5():
    import bench
    print('C5 2048^2 set-up', flush=True)
    log = lambda m: print(m, flush=True)
    return bench.build_full(20000000, log)


def test_c5_direct_solve_against_refined_solution(c5):
    g, bq, p, pb, em, ps, b = c5
    print('C5 direct solve', flush=True)
    b.deposit_all()
    q = b.rho + b.beam.getScale(0.) * b.beam.getProfile()
    ps.solvePhi(q)
    phi = ps.getPhi().clone()
    e = ps.phiToE().clone()
    # refinement: Jacobi CG from the direct solution to a 1e-14 relative residual
    tol0, ps.tol = ps.tol, 1e-14
    ps._pcg(ps._rhs(q), 20000)
    ps.tol = tol0
    phi_r = ps.getPhi().clone()
    e_r = ps.phiToE().clone()
    err_phi = float((phi - phi_r).abs().max() / phi_r.abs().max())
    err_e = float((e - e_r).abs().max() / e_r.abs().max())
    print('C5 direct solve: phi %.2e, E %.2e relative to the refined solution' % (err_phi, err_e), flush=True)
    assert err_phi <= 1e-10 and err_e <= 1e-10


def test_c5_native_step_bookkeeping(c5):
    """Three one-call steps of the 2e7-particle buildup: the count is the previous count minus
    the step's wall hits plus its secondaries, the weight sum changes by exactly the hits'
    and secondaries' weights, every particle is inside the chamber, and the charge the step
    leaves on the grid equals a fresh deposit of the particles."""
    from openecloud_amd.buildup import Buildup
    from openecloud_amd.diagnostics import moments
    g, bq, p, pb, em, ps, b0 = c5
    print('C5 native steps', flush=True)
    b = Buildup(g, p, pb, em, ps, b0.beam, dt=2e-12, typeBField=2, bAtGridPoints=bq)
    n = p.getMacroParticleCount()
    b.native_setup(absCap=n // 32, secCap=n // 16, maxSteps=8)
    for _ in range(3):
        n0 = p.getMacroParticleCount()
        w0 = moments(p)[0]
        b.step_native()
        st = b.sync()
        na, nadd = st['absorbed_last'], st['added_last']
        assert p.getMacroParticleCount() == n0 - na + nadd
        wa = float(pb.absorbedParticles[:na, 5].to(torch.float64).sum())
        ws = float(b._sec[:nadd, 5].to(torch.float64).sum())
        assert abs(moments(p)[0] - (w0 - wa + ws)) <= 1e-12 * w0
    pdn = p.getParticleData()
    x, y = pdn[:, 0].double(), pdn[:, 1].double()
    assert bool(torch.isfinite(pdn).all())
    s = x * x * (4. / g.getLx() ** 2) + y * y * (4. / g.getLy() ** 2)
    assert int((s >= 1.).sum()) == 0, int((s >= 1.).sum())
    rho = b.rho.clone()
    b.deposit_all()
    assert float((rho - b.rho).abs().max()) <= 1e-12 * float(b.rho.abs().max())


def test_c5_push_sample_matches_oracle(c5, orc):
    """Gather E and the quadrupole B, Boris push and wall test of all 2e7 FP32 particles by the
    method-by-method kernels; a sample of 2e5 against the oracle (FP64 arithmetic on the same
    FP32 inputs) to 1e-5 of each column's scale (FP32 arithmetic on the device)."""
    g, bq, p, pb, em, ps, b = c5
    print('C5 push sample', flush=True)
    n = p.getMacroParticleCount()
    rng = np.random.default_rng(5)
    idx = np.sort(rng.choice(n, 200000, replace=False))
    pd0 = p.getParticleData().cpu().numpy().astype(np.float64)[idx]
    b.deposit_all()
    E = ps.solve(b.rho).clone()
    p.calcGridWeights(g)
    p.eFieldToParticles(g, E)
    p.bFieldToParticles(g, b.bField)
    p.borisPush(2e-12, typeBField=2)
    pb.indexInside()
    got = p.getParticleData().cpu().numpy().astype(np.float64)[idx]
    ins = p.getIsInside().cpu().numpy()[idx]
    nx, npx = g.getNxExt(), g.getNpExt()
    ic, w = orc.calc_grid_weights(pd0, g.getDx(), g.getDy(), g.getLxExt(), g.getLyExt(), nx)
    eat = orc.field_to_particles(E.cpu().numpy(), ic, w, nx, npx, 2)
    bat = orc.field_to_particles(np.asarray(bq, np.float64), ic, w, nx, npx, 3)
    ref = orc.boris_push(pd0, eat, 2e-12, Q / M_E, 2, bat)
    ins_ref, ref = orc.index_inside(ref, 'ellip', g.getLx(), g.getLy())
    scale = np.abs(ref).max(axis=0)
    assert (np.abs(got - ref) / scale).max() < 1e-5
    assert np.mean(ins == ins_ref) > 1 - 1e-4


def test_c5_merge_event_reaches_target(c5):
    """localRandom at the 1.5x trigger on the 2e7-particle buildup, merging groups of up to 8
    particles as the full workload does: the count falls to the target exactly and the total
    weight is unchanged."""
    group = 8
    from openecloud_amd.diagnostics import moments
    from openecloud_amd.particleManagement import localRandom
    g, bq, p, pb, em, ps, b = c5
    print('C5 merge event', flush=True)
    n = p.getMacroParticleCount()
    target = int(n / 1.5)
    w0 = moments(p)[0]
    info = localRandom(g, p, pb, target, mergeGroup=group)
    print('C5 merge, groups of up to %d: %d -> %d in %d rounds' % (group, n, p.getMacroParticleCount(), info['rounds']),
          flush=True)
    assert p.getMacroParticleCount() == target
    assert abs(moments(p)[0] - w0) <= 1e-12 * w0
