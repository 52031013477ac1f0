"""Buildup benchmark of the plain GPU PIC step on one MI355X.

Workload (the elliptical-chamber dipole buildup): elliptical cut-cell chamber with semi-axes
45 mm x 35 mm on a 512 x 384 grid, uniform dipole By = 0.5 T, 2e6 macro-electrons loaded
uniformly (1 eV Maxwellian, 1e9 electrons per metre in total), Gaussian bunches with
sigma 1.5 mm, 1.15e11 protons per bunch, 25 ns spacing, 4 bunches, dt = 5e-12 s, Furman
copper wall with delta_max = 1.3.  A 'step' is one full time step of the buildup loop:
Poisson solve with the beam, gather + Boris push + wall test + deposit, absorption, wall
interaction, Furman emission and compaction.

Prints ONE JSON line: particle-steps per second (sum over timed steps of the live macro-
particle count divided by the wall time), with the fused push kernel's roofline and the
CPU baseline beside it.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, 'profiles', 'push_hbm_traffic_ellip512x384_2e6_f64.md')


def measured_traffic(dtype, n):
    """HBM bytes per push launch from the committed counter summary (FP64, 2e6 particles), or
    None when the summary does not cover this run's particle type and count."""
    if dtype != 'f64' or n != 2000000 or not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        if line.startswith('traffic_bytes_per_launch_f64:'):
            return float(line.split(':')[1])
    return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=200)
    ap.add_argument('--warmup', type=int, default=20)
    ap.add_argument('--particles', type=float, default=2e6)
    ap.add_argument('--dtype', default='f64', choices=['f64', 'f32'])
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--calib-copy', action='store_true',
                    help='run one plain 1 GiB device copy first (HBM counter calibration)')
    ap.add_argument('--cpu-seconds', type=float, default=12.0)
    ap.add_argument('--eager', action='store_true',
                    help='dipole: launch every step on its own (default: replay the steps between two '
                         'sorts as one captured HIP graph)')
    ap.add_argument('--sort-every', type=int, default=50,
                    help='re-order particles by cell every this many steps (0: never)')
    ap.add_argument('--train-steps', type=int, default=8000, help='learned: Adam steps')
    ap.add_argument('--pairs', type=int, default=192, help='learned: training pairs')
    ap.add_argument('--train-pairs', choices=('operator', 'buildup'), default='operator',
                    help='learned: train on (A phi, phi) with phi random at 0-64 cell scales, or on the '
                         'buildup dumps (half) and white-noise charges (half) with phi from the direct solve')
    ap.add_argument('--grid', type=int, default=1024, help='learned: grid cells per side')
    ap.add_argument('--workload', default='dipole', choices=['dipole', 'scan', 'learned', 'full'],
                    help='dipole: the elliptical dipole buildup (default); scan: 64 buildups '
                         'stepped together (delta_max and intensity scan)')
    ap.add_argument('--runs', type=int, default=64, help='scan: number of buildups')
    ap.add_argument('--scan-start', type=int, default=700,
                    help='scan: the timed steps start at this step of the simulation (the first bunch passage, '
                         'once multipacting has set in), whatever the autotuner ran before')
    ap.add_argument('--tune-cache', default=None, help='scan: autotuner cache file (JSON)')
    ap.add_argument('--scan-config', default=None,
                    help='scan: sortEvery,order,binShift to use instead of autotuning')
    ap.add_argument('--no-f32-compare', action='store_true',
                    help='dipole: skip the FP32-storage run compared against the FP64 one')
    return ap.parse_args()


def build_case(n, dtype):
    import torch
    from openecloud_amd import randomGen
    from openecloud_amd.beam import GaussianBunchTrain
    from openecloud_amd.buildup import Buildup
    from openecloud_amd.grid import Grid
    from openecloud_amd.particleBoundary import AbsorbElliptical
    from openecloud_amd.particleEmitter import FurmanEmitter, PhotoEmitter, homoLoader
    from openecloud_amd.particles import Particles
    from openecloud_amd.poissonSolver import PoissonSolver
    randomGen.seed(0)
    dt = 5e-12
    g = Grid(512, 384, 0.09, 0.07, 1)
    p = Particles('electrons', 'variableWeight', dtype=torch.float64 if dtype == 'f64' else torch.float32,
                  ordering='fast')
    p.setBConst([0., 0.5, 0.])
    pb = AbsorbElliptical(g, p)
    em = FurmanEmitter(pb, p, 'copper', seyMax=1.3)
    ps = PoissonSolver(g, method='direct', checkEvery=0)
    beam = GaussianBunchTrain(g, dt, nParticles=1.15e11, tBunchSpacing=25e-9, radiusSigma=1.5e-3,
                              nBunches=4)
    homoLoader(g, p, pb, int(n), 1e9 / n, 4.194e5)
    p.reserve(int(n * 2))
    # primary photoelectrons, 1e-3 per proton per metre, 1e5 macro-particles per bunch
    photo = PhotoEmitter(pb, p, beam, photoYield=1e-3, macroPerBunch=1e5, energyMean=7., energySigma=5.)
    b = Buildup(g, p, pb, em, ps, beam, dt=dt, typeBField=1, photoEmitter=photo)
    return g, p, pb, em, ps, b


def cpu_baseline(seconds):
    """The oracle's C restatement of the reference loops (AoS, one core) on a bounded sample:
    particle loop (weights, deposit, gather, constant-B Boris push, elliptical wall test) over
    a 2e5-particle sample of the same grid, plus the reference's SuperLU back-substitution on
    the same 512 x 384 system; extrapolated to 2e6 particles per step."""
    import ctypes
    import numpy as np
    import oracle
    from oracle.grid_oracle import OraclePoisson
    from openecloud_amd.grid import Grid
    g = Grid(512, 384, 0.09, 0.07, 1)
    rng = np.random.default_rng(0)
    ns = 200000
    pd = np.empty((ns, 6))
    k = 0
    while k < ns:
        x = (rng.random(ns) - 0.5) * 0.09
        y = (rng.random(ns) - 0.5) * 0.07
        ok = x * x * 4 / 0.09 ** 2 + y * y * 4 / 0.07 ** 2 < 1
        m = min(ns - k, ok.sum())
        pd[k:k + m, 0] = x[ok][:m]
        pd[k:k + m, 1] = y[ok][:m]
        k += m
    pd[:, 2:5] = rng.standard_normal((ns, 3)) * 4.194e5
    pd[:, 5] = 500.
    npx = g.npExt
    E = rng.standard_normal(2 * npx) * 1e3
    lib = oracle.lib()
    D = ctypes.POINTER(ctypes.c_double)
    ic = np.empty(ns, np.uint32)
    w4 = np.empty(4 * ns)
    eat = np.empty(2 * ns)
    bat = np.empty(3 * ns)
    ins = np.empty(ns, np.uint16)
    rho = np.empty(npx)
    bc = np.array([0., 0.5, 0.])
    qm = -1.602176565e-19 / 9.10938291e-31
    args = (pd.ctypes.data_as(D), ctypes.c_uint32(ns), E.ctypes.data_as(D), ctypes.c_void_p(0),
            ctypes.c_int(1), bc.ctypes.data_as(D), ctypes.c_double(1e-14), ctypes.c_double(qm),
            ctypes.c_double(1 / g.dx), ctypes.c_double(1 / g.dy), ctypes.c_double(g.lxExt / 2),
            ctypes.c_double(g.lyExt / 2), ctypes.c_uint32(g.nxExt), ctypes.c_uint32(npx),
            ctypes.c_int(1), ctypes.c_double(0.09), ctypes.c_double(0.07),
            ic.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), w4.ctypes.data_as(D),
            eat.ctypes.data_as(D), bat.ctypes.data_as(D),
            ins.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)), rho.ctypes.data_as(D),
            ctypes.c_double(-1.602176565e-19))
    lib.orc_particle_step(*args)
    t0 = time.perf_counter()
    reps = 0
    while time.perf_counter() - t0 < seconds * 0.6:
        lib.orc_particle_step(*args)
        reps += 1
    t_part = (time.perf_counter() - t0) / reps / ns          # s per particle-step
    ps = OraclePoisson(g)
    q = rng.standard_normal(npx) * 1e-12
    ps.solve(q)
    t0 = time.perf_counter()
    nsol = 0
    while time.perf_counter() - t0 < seconds * 0.4:
        ps.solve(q)
        nsol += 1
    t_lu = (time.perf_counter() - t0) / nsol
    n_full = 2e6
    step_time = t_part * n_full + t_lu
    return dict(value=n_full / step_time, unit='particle-steps/s', cores=1, kind='port',
                sample='%d-particle sample x %d reps of the C particle loop (%.3g s/particle-step) '
                       '+ %d SuperLU solves of the 512x384 system (%.3g s each), extrapolated '
                       'to 2e6 particles per step' % (ns, reps, t_part, nsol, t_lu),
                poisson_solves_per_s=1. / t_lu)


def build_scan(R, dtype, maxSteps):
    """The scan workload: R buildups on a shared 256 x 256 rectangular 100 x 50 mm chamber,
    By = 0.3 T, dt = 1e-11 s, Gaussian bunches sigma 2 mm at 25 ns, 8 bunches, Furman copper
    with E_max 300 eV.  delta_max runs over linspace(1.1, 2.0, R); bunch intensity and
    macro-particle count are log-uniform over [5e10, 2e11] ppb and [1e4, 2e6] (seed 0); each
    run carries 1e9 electrons per metre, loaded uniformly with a 1 eV Maxwellian."""
    import numpy as np
    import torch
    from openecloud_amd import randomGen
    from openecloud_amd.beam import GaussianBunchTrain
    from openecloud_amd.buildup import Buildup
    from openecloud_amd.grid import Grid
    from openecloud_amd.particleBoundary import AbsorbRectangular
    from openecloud_amd.particleEmitter import FurmanEmitter, homoLoader
    from openecloud_amd.particles import Particles
    from openecloud_amd.poissonSolver import PoissonSolver
    from openecloud_amd.scan import BuildupScan
    rng = np.random.default_rng(0)
    sey = np.linspace(1.1, 2.0, R)
    ppb = np.exp(rng.uniform(np.log(5e10), np.log(2e11), R))
    nmp = np.exp(rng.uniform(np.log(1e4), np.log(2e6), R)).astype(np.int64)
    dt = 1e-11
    g = Grid(256, 256, 0.1, 0.05, 0)
    ps = PoissonSolver(g, method='direct', checkEvery=0)
    tdt = torch.float64 if dtype == 'f64' else torch.float32
    runs, seeds = [], []
    for r in range(R):
        seeds.append(1000 + r)
        randomGen.seed(1000 + r)
        p = Particles('electrons', 'variableWeight', dtype=tdt, ordering='fast')
        p.setBConst([0., 0.3, 0.])
        pb = AbsorbRectangular(g, p)
        em = FurmanEmitter(pb, p, 'copper', seyMax=float(sey[r]), eHat0=300.)
        beam = GaussianBunchTrain(g, dt, nParticles=float(ppb[r]), tBunchSpacing=25e-9, radiusSigma=2e-3,
                                  nBunches=8)
        homoLoader(g, p, pb, int(nmp[r]), 1e9 / float(nmp[r]), 4.194e5)
        runs.append(Buildup(g, p, pb, em, ps, beam, dt=dt, typeBField=1))
    scan = BuildupScan(runs, seeds, growth=1.5, maxSteps=maxSteps)
    return g, ps, runs, scan, dict(sey=sey, ppb=ppb, n=nmp)


def compare_f32(a, n, b64, extra):
    """The same seeded workload with FP32 particle storage (arithmetic stays FP64), stepped
    exactly like the FP64 run (warm-up, the timed steps as graph replays or eager launches,
    sorts, then `extra` eager steps without a sort); its throughput and the differences of the
    final state from the FP64 run: cloud density (charge on the grid, relative L2), total charge
    and macro-particle count."""
    import torch
    g, p, pb, em, ps, b = build_case(n, 'f32')
    b.native_setup(absCap=max(65536, n // 8), secCap=max(131072, n // 4),
                   maxSteps=a.warmup + a.steps + extra + 40)
    b.sort_particles()
    for k in range(a.warmup):
        b.step_native()
        if a.sort_every and k % a.sort_every == a.sort_every - 1:
            b.sync()
            b.sort_particles()
    b.sync()
    i0 = b.it
    if not a.eager:
        per = a.sort_every if a.sort_every else a.steps
        b.capture_native(per)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if not a.eager:
        for k in range(a.steps // per):
            b.replay()
            if a.sort_every:
                b.sync()
                b.sort_particles(keepBuffers=True)
    else:
        for k in range(a.steps):
            b.step_native()
            if a.sort_every and k % a.sort_every == a.sort_every - 1:
                b.sync()
                b.sort_particles()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    b.sync()
    rate = float(sum(b.counts(i0, i0 + a.steps))) / wall
    for k in range(extra):
        b.step_native()
    b.sync()
    b64.deposit_all()
    b.deposit_all()
    r64, r32 = b64.rho, b.rho
    q64 = float(b64.p.getParticleData()[:, 5].to(torch.float64).sum())
    q32 = float(b.p.getParticleData()[:, 5].to(torch.float64).sum())
    return {'value': rate, 'unit': 'particle-steps/s',
            'cloud_density_rel_l2': float(torch.linalg.norm(r32 - r64) / torch.linalg.norm(r64)),
            'total_charge_rel': abs(q32 - q64) / abs(q64),
            'macro_particles': [b64.p.macroParticleCount, b.p.macroParticleCount],
            'note': 'FP32 particle storage, FP64 arithmetic, same seeds and steps; differences from the '
                    'FP64 run after the timed steps'}


def main_learned(a):
    """The learned-preconditioner workload: a cut-cell grid of the elliptical chamber (45 x 35
    mm semi-axes) with a circular conductor of radius 2.5 mm centred at x = 20 mm.  Test pairs
    (q, phi): buildup clouds dumped every 50 steps (charge plus the beam term at that time) and
    white-noise charges, phi from the direct solve.  Training pairs: (A phi, phi) with phi
    random at length scales of 0-64 cells (default), or more buildup dumps and white-noise
    charges (--train-pairs buildup).  The U-Net is trained on the device; then Poisson solves
    to 1e-8 relative residual by the learned flexible CG, by Jacobi PCG and by the direct
    solve."""
    import numpy as np
    import torch
    from openecloud_amd import _lib as L
    from openecloud_amd import randomGen
    from openecloud_amd.beam import GaussianBunchTrain
    from openecloud_amd.buildup import Buildup
    from openecloud_amd.grid import Grid
    from openecloud_amd.learnedPreconditioner import UNetPreconditioner
    from openecloud_amd.particleBoundary import AbsorbArbitraryCutCell
    from openecloud_amd.particleEmitter import FurmanEmitter, homoLoader
    from openecloud_amd.particles import Particles
    from openecloud_amd.poissonSolver import PoissonSolver
    L.require_gpu()
    log = lambda m: print(m, file=sys.stderr, flush=True)
    ax, ay, xc, rc = 0.045, 0.035, 0.02, 0.0025

    def bf(x, y):
        x = np.asarray(x, np.float64)
        y = np.asarray(y, np.float64)
        return ((x * x / (ax * ax) + y * y / (ay * ay) < 1.) & ((x - xc) ** 2 + y * y > rc * rc)).astype(np.int64)
    n = a.grid
    t0 = time.perf_counter()
    g = Grid(n, n, 2 * ax, 2 * ay, 3, boundFunc=bf)
    ps = PoissonSolver(g, method='direct', checkEvery=0)
    pj = PoissonSolver(g, method='pcg', tol=1e-8, maxIter=200000)
    log('grid and direct solver: %.1f s, %d capacitance rows' % (time.perf_counter() - t0, ps.capacitanceRows))
    # training pairs from a buildup on this grid: every 50 steps q = rho + beam(t)
    randomGen.seed(0)
    dt = 5e-12
    p = Particles('electrons', 'variableWeight', dtype=torch.float32, ordering='fast')
    p.setBConst([0., 0.5, 0.])
    pb = AbsorbArbitraryCutCell(g, p)
    em = FurmanEmitter(pb, p, 'copper', seyMax=1.6, eHat0=300.)
    beam = GaussianBunchTrain(g, dt, nParticles=1.15e11, tBunchSpacing=25e-9, radiusSigma=1.5e-3, nBunches=4)
    npart = 200000
    homoLoader(g, p, pb, npart, 1e9 / npart, 4.194e5)
    p.reserve(3 * npart)
    b = Buildup(g, p, pb, em, ps, beam, dt=dt, typeBField=1)
    ntest = 4   # test pairs of each kind; one more of each is the validation set
    fromBuildup = a.train_pairs == 'buildup'
    ncl = (a.pairs // 2 if fromBuildup else 0) + ntest + 1
    nwh = (a.pairs - a.pairs // 2 if fromBuildup else 0) + ntest + 1
    b.native_setup(absCap=65536, secCap=131072, maxSteps=50 * ncl + 10)
    clouds, noise = [], []
    t0 = time.perf_counter()
    for k in range(ncl):
        for _ in range(50):
            b.step_native()
        b.sync()
        b.deposit_all()
        q = b._rhs(b.it * dt).clone()
        clouds.append((q, ps.solvePhi(q).clone()))
    rng = np.random.default_rng(1)
    mask = ps.mask.to(torch.float64)
    for k in range(nwh):
        q = torch.as_tensor(rng.standard_normal(g.getNpExt())).cuda() * mask * 1e-9
        noise.append((q, ps.solvePhi(q).clone()))
    test = clouds[-ntest:] + noise[-ntest:]
    val = [clouds[-ntest - 1][0], noise[-ntest - 1][0]]
    net = UNetPreconditioner(g, seed=0)
    raw = (clouds[:-ntest - 1] + noise[:-ntest - 1]) if fromBuildup else net.operator_pairs(ps, a.pairs, seed=1)
    pairs = raw
    log('%d training pairs (%s), %d test pairs: %.1f s' % (len(raw), a.train_pairs, len(test),
                                                          time.perf_counter() - t0))
    train = net.smoothed_pairs(ps, raw)
    del clouds, noise
    # the same flexible CG with the untrained network (the ablation)
    base = [net.fcg(ps, q, tol=1e-8, maxIter=5000)[1] for q, _ in test]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = net.train(ps, train, steps=a.train_steps, lr=5e-4, lrEnd=5e-5, batch=4,
                     logEvery=max(1, a.train_steps // 10), log=log, validate=val,
                     validateEvery=max(1, a.train_steps // 6), validateIter=4000)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    # U-Net apply time (the preconditioner's cost per iteration)
    r = test[0][0]
    for _ in range(3):
        net.precondition(ps, r)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(20):
        net.precondition(ps, r)
    torch.cuda.synchronize()
    t_apply = (time.perf_counter() - t0) / 20
    H, W = net.H, net.W
    flops = 0
    for lv in range(4):
        convs = {0: 4, 1: 4, 2: 4, 3: 2}[lv]
        flops += convs * 2 * 9 * 16 * 16 * (H >> lv) * (W >> lv)
    res = {'learned': [], 'jacobi': [], 'direct': []}
    for q, phi in test:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        z, it, rr, ok = net.fcg(ps, q, tol=1e-8, maxIter=5000)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        err = float(torch.linalg.norm(z - phi) / torch.linalg.norm(phi))
        res['learned'].append((t1 - t0, it, rr, ok, err))
        t0 = time.perf_counter()
        pj.solvePhi(q)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        res['jacobi'].append((t1 - t0, pj.lastIterations, None, True, None))
        t0 = time.perf_counter()
        for _ in range(10):
            ps.solvePhi(q)
        torch.cuda.synchronize()
        res['direct'].append(((time.perf_counter() - t0) / 10, 0, None, True, None))
    # the buildup continued in the reference order with the learned solve (warm-started from
    # the previous potential, the direct solve judging every 10th solve), against the same steps
    # with the direct solve
    psl = PoissonSolver(g, method='learned', tol=1e-8, maxIter=4000, checkEvery=10, preconditioner=net)
    steps_b = 20
    bsteps = {}
    for name, solver in (('direct', ps), ('learned', psl)):
        b.ps = solver
        its, pst = [], 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps_b):
            pst += p.getMacroParticleCount()
            b.step_reference()
            its.append(solver.lastIterations)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        bsteps[name] = {'ms_per_step': 1e3 * el / steps_b, 'particle_steps_per_s': pst / el,
                        'iterations_mean': float(np.mean(its))}
    b.ps = ps
    bsteps['learned'].update(judge_max_rel_err=psl.judgeMaxErr, judge_count=psl.judgeCount,
                             fallbacks=psl.fallbacks)
    log('buildup steps: %s' % json.dumps(bsteps))
    summ = {}
    for k, v in res.items():
        tm = [x[0] for x in v]
        summ[k] = {'solves_per_s': len(tm) / sum(tm), 'iterations_mean': float(np.mean([x[1] for x in v]))}
    summ['learned']['relres_max'] = max(x[2] for x in res['learned'])
    summ['learned']['converged'] = all(x[3] for x in res['learned'])
    summ['learned']['solution_rel_err_vs_direct_max'] = max(x[4] for x in res['learned'])
    summ['learned']['iterations_untrained_network'] = float(np.mean(base))
    summ['learned']['iterations_per_test'] = [int(x[1]) for x in res['learned']]
    summ['learned']['iterations_untrained_per_test'] = [int(x) for x in base]
    summ['jacobi']['iterations_per_test'] = [int(x[1]) for x in res['jacobi']]
    out = {
        'metric': 'Poisson solves/sec to 1e-8 relative residual, learned-preconditioner flexible CG',
        'value': summ['learned']['solves_per_s'],
        'unit': 'solves/s',
        'n_gpus': 1,
        'steps': len(test),
        'warmup': 0,
        'ms_per_step': 1e3 / summ['learned']['solves_per_s'],
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'bf16 network (FP32 accumulation), FP64 CG, damped Jacobi (omega 2/3) around the network',
        'data': 'synthetic: test pairs from a seeded buildup on this grid (4) and white-noise charges (4); '
                'validation right-hand sides: one more of each; training pairs ' +
                ('(A phi, phi), phi random at 0-64 cell scales' if not fromBuildup else
                 'from the same buildup and white-noise charges'),
        'config': {'workload': '%dx%d cut-cell elliptical chamber 45x35 mm with a conductor of radius 2.5 mm at '
                               'x = 20 mm; U-Net 4 levels, 16 channels' % (n, n),
                   'grid': [n, n], 'inside_nodes': int(ps.mask.sum().item()),
                   'capacitance_rows': ps.capacitanceRows, 'pairs': len(pairs), 'train_pairs': a.train_pairs,
                   'fcg_directions': net.mDirs, 'smoothing_omega': net.omega,
                   'train_steps': a.train_steps, 'validation_iterations': net.validation,
                   'train_seconds': t_train, 'loss_history': hist},
        'solvers': summ,
        'unet_apply_ms': t_apply * 1e3,
        'unet_apply_gflop': flops / 1e9,
        'roofline': {'bound': 'mfma', 'kernel': 'U-Net apply (14 bf16 convolutions)',
                     'achieved': flops / t_apply / 1e12, 'peak': 2500.0, 'unit': 'TFLOP/s',
                     'frac': flops / t_apply / 1e12 / 2500.0, 'traffic': None},
        'learned_vs_jacobi_time': summ['learned']['solves_per_s'] / summ['jacobi']['solves_per_s'],
        'learned_vs_jacobi_iterations': summ['jacobi']['iterations_mean'] / max(1., summ['learned']['iterations_mean']),
        'buildup_reference_steps': dict(bsteps, steps=steps_b, particles=p.getMacroParticleCount()),
    }
    print(json.dumps(out))


def main_full(a):
    """The combined workload: a 2048 x 2048 grid of the elliptical 90 x 70 mm chamber, a 50 T/m
    quadrupole (the grid field path, typeBField 2), a.particles (default 2e7) FP32
    macro-electrons, 20 bunches of 1.15e11 with sigma 1 mm, dt 2e-12 s, Furman copper with
    delta_max 1.5, the one-call step with the direct Poisson solve.  After the timed steps one
    merge event of localRandom reduces the count by the 1.5x trigger's factor (target = count
    / 1.5), timed, with its charge change."""
    import ctypes
    import torch
    from openecloud_amd import _lib as L
    from openecloud_amd import particleManagement, randomGen
    from openecloud_amd.beam import GaussianBunchTrain
    from openecloud_amd.buildup import Buildup
    from openecloud_amd.diagnostics import weight_sum
    from openecloud_amd.grid import Grid
    from openecloud_amd.magneticField import multipoleExpansion
    from openecloud_amd.particleBoundary import AbsorbElliptical
    from openecloud_amd.particleEmitter import FurmanEmitter, homoLoader
    from openecloud_amd.particles import Particles
    from openecloud_amd.poissonSolver import PoissonSolver
    L.require_gpu()
    log = lambda m: print(m, file=sys.stderr, flush=True)
    n = int(a.particles) if '--particles' in sys.argv else 20000000
    dt = 2e-12
    t0 = time.perf_counter()
    randomGen.seed(0)
    g = Grid(2048, 2048, 0.09, 0.07, 1)
    bq = multipoleExpansion(g, [0., 50.])
    ps = PoissonSolver(g, method='direct', checkEvery=0)
    log('grid, field and direct solver: %.1f s (%d capacitance rows)' % (time.perf_counter() - t0,
                                                                        ps.capacitanceRows))
    p = Particles('electrons', 'variableWeight', dtype=torch.float32, ordering='fast')
    pb = AbsorbElliptical(g, p)
    em = FurmanEmitter(pb, p, 'copper', seyMax=1.5)
    beam = GaussianBunchTrain(g, dt, nParticles=1.15e11, tBunchSpacing=25e-9, radiusSigma=1e-3, nBunches=20)
    homoLoader(g, p, pb, n, 1e9 / n, 4.194e5)
    p.reserve(int(1.6 * n))
    b = Buildup(g, p, pb, em, ps, beam, dt=dt, typeBField=2, bAtGridPoints=bq)
    b.native_setup(absCap=max(65536, n // 32), secCap=max(131072, n // 16), maxSteps=a.warmup + a.steps + 20)
    b.sort_particles()
    se = a.sort_every
    for k in range(a.warmup):
        b.step_native()
        if se and k % se == se - 1:
            b.sync()
            b.sort_particles()
    b.sync()
    evs = []
    for _ in range(a.steps):
        e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
        L.check(L.lib().oecl_event_create(ctypes.byref(e0)))
        L.check(L.lib().oecl_event_create(ctypes.byref(e1)))
        evs.append((e0, e1))
    it0 = b.it
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(a.steps):
        b._args.evPushBegin, b._args.evPushEnd = evs[k][0].value, evs[k][1].value
        b.step_native()
        if se and k % se == se - 1:
            b.sync()
            b.sort_particles()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    b._args.evPushBegin = b._args.evPushEnd = None
    st = b.sync()
    counts = b.counts(it0, it0 + a.steps)
    push = []
    for e0, e1 in evs:
        ms = ctypes.c_float()
        L.check(L.lib().oecl_event_elapsed_ms(e0, e1, ctypes.byref(ms)))
        push.append(ms.value * 1e3)
        L.lib().oecl_event_destroy(e0)
        L.lib().oecl_event_destroy(e1)
    avg_us = sum(push) / len(push)
    bytes_per = sum(11 * 4 * c + 48 * g.getNpExt() for c in counts) / len(counts)
    achieved = bytes_per / (avg_us * 1e-6) / 1e9
    # one merge event of the 1.5x trigger
    nBefore = p.macroParticleCount
    q0 = weight_sum(p)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    mg = particleManagement.localRandom(g, p, pb, int(nBefore / 1.5))
    torch.cuda.synchronize()
    t_merge = time.perf_counter() - t1
    q1 = weight_sum(p)
    out = {
        'metric': 'buildup particle-steps/sec (2048x2048 quadrupole, one-call step)',
        'value': float(sum(counts)) / wall,
        'unit': 'particle-steps/s',
        'n_gpus': 1,
        'steps': a.steps,
        'warmup': a.warmup,
        'ms_per_step': wall / a.steps * 1e3,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32',
        'data': 'synthetic (seeded uniform load, Philox draws)',
        'config': {'workload': '2048x2048 elliptical 90x70 mm chamber, quadrupole 50 T/m (grid field), '
                               '%d FP32 macro-electrons, Gaussian bunches sigma 1 mm 1.15e11 ppb x20, dt 2e-12 s, '
                               'Furman copper delta_max 1.5, direct Poisson solve' % n,
                   'grid': [2048, 2048], 'particles': n, 'capacitance_rows': ps.capacitanceRows,
                   'absorbed_total': st['absorbed_total'], 'emitted_total': st['emitted_total']},
        'roofline': {'bound': 'hbm', 'kernel': 'k_push_dev', 'achieved': achieved, 'peak': HBM_PEAK_GBS,
                     'unit': 'GB/s', 'frac': achieved / HBM_PEAK_GBS, 'traffic': None, 'avg_launch_us': avg_us,
                     'bytes_per_launch': bytes_per},
        'merge_event': {'count_before': nBefore, 'count_after': p.macroParticleCount, 'seconds': t_merge,
                        'charge_rel_change': abs(q1 - q0) / abs(q0), 'ops': mg},
    }
    print(json.dumps(out))


def main_scan(a):
    import ctypes
    import torch
    from openecloud_amd import _lib as L
    from openecloud_amd import randomGen
    from openecloud_amd.autotune import tune_scan
    L.require_gpu()
    R = a.runs
    start = max(int(a.scan_start), a.warmup)
    maxSteps = start + a.steps + 2000
    g, ps, runs, scan, prm = build_scan(R, a.dtype, maxSteps)
    log = lambda m: print(m, file=sys.stderr)
    scan.sort()
    for _ in range(a.warmup):
        scan.step()
    scan.sync()
    if a.scan_config:
        se_, o_, b_ = [int(v) for v in a.scan_config.split(',')]
        best, results = dict(sortEvery=se_, order=o_, binShift=b_, graphSteps=se_), []
    else:
        best, results = tune_scan(scan, cacheFile=a.tune_cache, log=log)
    se = int(best['sortEvery'])
    steps = (a.steps + se - 1) // se * se
    # the timed window is steps [start, start + steps) of a fresh scan from the same seeds, so
    # that it does not depend on how many steps the autotuner ran (or whether it ran at all)
    del scan, runs
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    g, ps, runs, scan, prm = build_scan(R, a.dtype, maxSteps)
    scan.sort(best['order'], best['binShift'])
    for _ in range(a.warmup):
        scan.step()
    scan.sync()
    scan.capture(se)
    while scan.it + se <= start:
        scan.sort(best['order'], best['binShift'])
        scan.replay()
    while scan.it < start:
        scan.step()
    st0 = scan.sync()
    it0 = scan.it
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps // se):
        scan.sort(best['order'], best['binShift'])
        scan.replay()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    st = scan.sync()
    psteps = float(sum(sum(scan.history(r, it0, it0 + steps)) for r in range(R)))
    rate = psteps / wall
    # push kernel timing: one more sort interval of eager steps with HIP events around each push
    evs = []
    scan.sort(best['order'], best['binShift'])
    i1 = scan.it
    for k in range(se):
        e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
        L.check(L.lib().oecl_event_create(ctypes.byref(e0)))
        L.check(L.lib().oecl_event_create(ctypes.byref(e1)))
        evs.append((e0, e1))
        scan._args.evPushBegin, scan._args.evPushEnd = e0.value, e1.value
        scan.step()
    scan._args.evPushBegin = scan._args.evPushEnd = None
    torch.cuda.synchronize()
    push_us = []
    for e0, e1 in evs:
        ms = ctypes.c_float()
        L.check(L.lib().oecl_event_elapsed_ms(e0, e1, ctypes.byref(ms)))
        push_us.append(ms.value * 1e3)
        L.lib().oecl_event_destroy(e0)
        L.lib().oecl_event_destroy(e1)
    tot = [sum(scan.history(r, i1 + k, i1 + k + 1)[0] for r in range(R)) for k in range(se)]
    s_b = 8 if a.dtype == 'f64' else 4
    by = [11 * s_b * tot[k] + 24 * g.getNpExt() * R for k in range(se)]
    avg_us = sum(push_us) / len(push_us)
    achieved = (sum(by) / len(by)) / (avg_us * 1e-6) / 1e9
    print('scan push time (us) by steps since the last sort: ' +
          ' '.join('%d:%.0f' % (k, t) for k, t in enumerate(push_us)), file=sys.stderr)
    # the same runs one at a time with the single-run one-call step (3 steps each)
    seq_ps, seq_t = 0., 0.
    for r, b in enumerate(runs):
        randomGen.seed(1000 + r)
        b.native_setup(absCap=max(4096, b.p.capacity // 16), secCap=max(8192, b.p.capacity // 8), maxSteps=8)
        b.step_native()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(3):
            b.step_native()
        torch.cuda.synchronize()
        seq_t += time.perf_counter() - t
        seq_ps += float(sum(b.counts(b.it - 3, b.it)))
        b.sync()
    # the reference-order method-by-method GPU step on 4 runs of the scan
    ref_ps, ref_t = 0., 0.
    for r in (0, R // 3, 2 * R // 3, R - 1):
        b = runs[r]
        b.step_reference()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(2):
            n = b.p.macroParticleCount
            b.step_reference()
            ref_ps += n
        torch.cuda.synchronize()
        ref_t += time.perf_counter() - t
    seq_rate, ref_rate = seq_ps / seq_t, ref_ps / ref_t
    n0 = [int(x) for x in prm['n']]
    out = {
        'metric': 'batched scan buildup particle-steps/sec (%d buildups per step call)' % R,
        'value': rate,
        'unit': 'particle-steps/s',
        'n_gpus': 1,
        'steps': steps,
        'warmup': a.warmup,
        'ms_per_step': wall / steps * 1e3,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': a.dtype,
        'data': 'synthetic (seeded uniform loads, Philox draws)',
        'config': {'workload': '%d buildups on a shared 256x256 rectangular 100x50 mm chamber, By 0.3 T, '
                               'dt 1e-11 s, Gaussian bunches sigma 2 mm x8, delta_max linspace(1.1, 2.0), '
                               'ppb log-uniform [5e10, 2e11], macro-particles log-uniform [1e4, 2e6]' % R,
                   'runs': R, 'grid': [256, 256], 'macro_particles_total_start': int(sum(n0)),
                   'macro_particles_total_end': int(sum(st['counts'])),
                   'absorbed_total': int(sum(st['absorbed_total'])), 'emitted_total': int(sum(st['emitted_total'])),
                   'timed_steps': [it0, it0 + steps],
                   'absorbed_in_timed_steps': int(sum(st['absorbed_total']) - sum(st0['absorbed_total'])),
                   'emitted_in_timed_steps': int(sum(st['emitted_total']) - sum(st0['emitted_total']))},
        'poisson_solves_per_s': R * steps / wall,
        'roofline': {'bound': 'hbm', 'kernel': 'k_push_dev', 'achieved': achieved, 'peak': HBM_PEAK_GBS,
                     'unit': 'GB/s', 'frac': achieved / HBM_PEAK_GBS, 'traffic': None, 'avg_launch_us': avg_us,
                     'bytes_per_launch': sum(by) / len(by)},
        'autotune': {'best': best, 'candidates_us_per_step': [[c, 1e6 * t] for c, t in results]},
        'plain_one_run_per_call': {'value': seq_rate, 'unit': 'particle-steps/s',
                                   'note': 'the same runs stepped one at a time with the single-run one-call step'},
        'plain_reference_order': {'value': ref_rate, 'unit': 'particle-steps/s',
                                  'note': 'method-by-method GPU step in the reference order, 4 of the runs'},
        'speedup_vs_one_run_per_call': rate / seq_rate,
        'speedup_vs_reference_order': rate / ref_rate,
    }
    print(json.dumps(out))


def main():
    a = parse()
    if a.gpus != 1:
        print('bench.py runs on one GPU only: this build has no multi-GPU path '
              '(see the sharding design note); got --gpus %d' % a.gpus, file=sys.stderr)
        sys.exit(2)
    if a.workload == 'learned':
        return main_learned(a)
    if a.workload == 'full':
        return main_full(a)
    if a.workload == 'scan':
        if a.dtype == 'f64' and '--dtype' not in sys.argv:
            a.dtype = 'f32'   # the scan's particles are FP32 unless asked otherwise
        return main_scan(a)
    import torch
    from openecloud_amd import _lib as L
    L.require_gpu()
    n = int(a.particles)
    g, p, pb, em, ps, b = build_case(n, a.dtype)
    import ctypes
    b.native_setup(absCap=max(65536, n // 8), secCap=max(131072, n // 4),
                   maxSteps=a.warmup + a.steps + 40)
    b.sort_particles()
    if a.calib_copy:
        # one plain 1 GiB device copy (beyond the 256 MB last-level cache) for the counters
        src_t = torch.ones(1 << 27, dtype=torch.float64, device='cuda')
        dst_t = torch.empty_like(src_t)
        torch.cuda.synchronize()
        L.check(L.lib().oecl_copy_bytes(L.ptr(dst_t), L.ptr(src_t), L.u64(src_t.numel() * 8), L.stream()),
                'calibration copy')
        torch.cuda.synchronize()
        del src_t, dst_t
    for k in range(a.warmup):
        b.step_native()
        if a.sort_every and k % a.sort_every == a.sort_every - 1:
            b.sync()
            b.sort_particles()
    torch.cuda.synchronize()
    b.sync()
    # timed region: K full steps; the push kernel of each step is bracketed by HIP events
    # recorded on the stream the step runs on
    evs = []
    for _ in range(a.steps):
        e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
        L.check(L.lib().oecl_event_create(ctypes.byref(e0)))
        L.check(L.lib().oecl_event_create(ctypes.byref(e1)))
        evs.append((e0, e1))
    it0 = b.it
    if not a.eager:
        # steps between two sorts as one HIP graph; the sort keeps the particle buffers
        per = a.sort_every if a.sort_every else a.steps
        if a.steps % per:
            print('graph replay needs --steps to be a multiple of --sort-every (or --eager)', file=sys.stderr)
            sys.exit(2)
        b.capture_native(per)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(a.steps // per):
            b.replay()
            if a.sort_every:
                b.sync()
                b.sort_particles(keepBuffers=True)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        # push timing of the same kernel: a few eager steps with events after the timed region
        nev = min(20, len(evs))
        for k in range(nev):
            b._args.evPushBegin = evs[k][0].value
            b._args.evPushEnd = evs[k][1].value
            b.step_native()
        evs = evs[:nev]
    else:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(a.steps):
            b._args.evPushBegin = evs[k][0].value
            b._args.evPushEnd = evs[k][1].value
            b.step_native()
            if a.sort_every and k % a.sort_every == a.sort_every - 1:
                b.sync()              # host count for the sort launch, and the status check
                b.sort_particles()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
    b._args.evPushBegin = None
    b._args.evPushEnd = None
    st = b.sync()
    wall = t1 - t0
    counts = b.counts(it0, it0 + a.steps)
    psteps = float(sum(counts))
    push_ms = []
    for e0, e1 in evs:
        ms = ctypes.c_float()
        L.check(L.lib().oecl_event_elapsed_ms(e0, e1, ctypes.byref(ms)))
        push_ms.append(ms.value)
        L.lib().oecl_event_destroy(e0)
        L.lib().oecl_event_destroy(e1)
    if a.sort_every:
        by_age = {}
        for k, ms in enumerate(push_ms):
            by_age.setdefault(k % a.sort_every, []).append(ms * 1e3)
        print('push time (us) by steps since the last sort: ' +
              ' '.join('%d:%.1f' % (k, sum(v) / len(v)) for k, v in sorted(by_age.items())),
              file=sys.stderr)
    s = 8 if a.dtype == 'f64' else 4
    npx = g.getNpExt()
    bytes_per = [11 * s * c + 24 * npx for c in counts]
    avg_ms = sum(push_ms) / len(push_ms)
    avg_bytes = sum(bytes_per) / len(bytes_per)
    achieved = avg_bytes / (avg_ms * 1e-3) / 1e9
    # Poisson solves/s of the direct device solve (phi only, then phi -> E)
    q = b.rho.clone()
    torch.cuda.synchronize()
    tp = time.perf_counter()
    nsol = 100
    for _ in range(nsol):
        ps.solve(q)
    torch.cuda.synchronize()
    poisson_rate = nsol / (time.perf_counter() - tp)
    res = ps.residual(q)
    out = {
        'metric': 'buildup particle-steps/sec (plain GPU step)',
        'value': psteps / wall,
        'unit': 'particle-steps/s',
        'n_gpus': 1,
        'steps': a.steps,
        'warmup': a.warmup,
        'ms_per_step': wall / a.steps * 1e3,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': a.dtype,
        'data': 'synthetic (seeded uniform load, Philox draws)',
        'config': {'workload': 'elliptical 45x35 mm cut-cell chamber, 512x384 grid, By 0.5 T, '
                               '%d macro-electrons, Gaussian bunches sigma 1.5 mm 1.15e11 ppb x4, '
                               'dt 5e-12 s, Furman copper delta_max 1.3, primary photoelectrons '
                               '1e-3 per proton per m' % n,
                   'grid': [512, 384], 'particles': n, 'dt': 5e-12,
                   'macro_particles_start_end': [counts[0], p.macroParticleCount],
                   'absorbed_total': st['absorbed_total'], 'emitted_total': st['emitted_total'],
                   'photoelectrons_total': st['photo_total']},
        'poisson_solves_per_s': poisson_rate,
        'poisson_relres': res,
        'poisson_capacitance_rows': ps.capacitanceRows,
        'sort_every': a.sort_every,
        'roofline': {'bound': 'hbm', 'kernel': 'k_push_dev', 'achieved': achieved,
                     'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': achieved / HBM_PEAK_GBS,
                     'traffic': measured_traffic(a.dtype, n), 'avg_launch_us': avg_ms * 1e3,
                     'bytes_per_launch': avg_bytes},
    }
    if not a.no_f32_compare and a.dtype == 'f64':
        out['f32_vs_f64'] = compare_f32(a, n, b, b.it - it0 - a.steps)
    if not a.no_cpu_baseline:
        cb = cpu_baseline(a.cpu_seconds)
        out['cpu_baseline'] = {k: cb[k] for k in ('value', 'unit', 'cores', 'kind', 'sample')}
        out['cpu_baseline_poisson_solves_per_s'] = cb['poisson_solves_per_s']
    print(json.dumps(out))


if __name__ == '__main__':
    main()
