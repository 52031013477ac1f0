"""Buildup benchmark of the GPU PIC step on one MI355X.

Default workload (the largest single-GPU configuration): a 2048 x 2048 grid of an elliptical
90 x 70 mm chamber with a 50 T/m quadrupole on the grid (typeBField 2), 2e7 FP32
macro-electrons (1 eV Maxwellian, 1e9 electrons per metre), 20 Gaussian bunches of 1.15e11
protons with sigma 1 mm at 25 ns, dt = 2e-12 s, Furman copper wall with delta_max = 1.5, and
localRandom particle management whenever the count leaves [0.95, 1.5] x 2e7 (checked every
sort interval, as the reference checks every 50 steps).  A 'step' is one full time step of the
buildup loop: Poisson solve with the beam, gather + Boris push + wall test + deposit,
absorption, wall interaction, Furman emission and compaction.

Beside it, the same line reports the plain GPU step on the elliptical dipole buildup (512 x 384
grid, 2e6 FP64 macro-electrons, By = 0.5 T) and the Poisson solves/s of both grids.

Prints ONE JSON line: particle-steps per second (sum over the timed steps of the live
macro-particle count divided by the wall time), with the push kernel's roofline and the CPU
baseline beside it.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
WORKLOADS = ('full', 'dipole', 'scan', 'learned', 'northstar', 'push')


def profile_traffic(name):
    """HBM bytes per push launch from a committed counter summary under profiles/ (the line
    'traffic_bytes_per_launch: <bytes>'), or None when there is none for this workload."""
    if not name:
        return None
    path = os.path.join(ROOT, 'profiles', name)
    if not os.path.isfile(path):
        return None
    for line in open(path):
        if line.startswith('traffic_bytes_per_launch:'):
            return float(line.split(':')[1])
    return None


def push_model(dtype, arith, gridB):
    """(kernel, algorithmic bytes per particle, per grid node, description) of one push launch:
    the particle reads x, y, vx, vy, vz, w and writes x, y, vx, vy, vz (11 values); per node
    the push reads E (FP64 (Ex, Ey), or its FP32 copy for the mixed push), a grid B field's
    (Bx, By) likewise, and adds into rho (an FP64 read and write)."""
    s = 8 if dtype == 'f64' else 4
    mixed = dtype == 'f32' and arith != 'fp64'
    fe = 8 if mixed else 16
    node = fe + (fe if gridB else 0) + 16
    kern = 'k_push_mixed' if mixed else 'k_push_dev'
    desc = ('%d B per %s macro-particle + %d B per grid node (E%s %s, rho FP64 read and write)'
            % (11 * s, 'FP64' if dtype == 'f64' else 'FP32', node, ', B' if gridB else '', 'FP32' if mixed else 'FP64'))
    return kern, 11 * s, node, desc


TRAFFIC_PROFILES = {('dipole', 'f64', 2000000): 'push_hbm_traffic_ellip512x384_2e6_f64.md',
                    ('scan', 'mixed', 64): 'push_hbm_traffic_scan_64runs_256sq_f32.md',
                    ('push', 'fp64', 20000000): 'push_hbm_traffic_rect2049sq_2e7_f32_quadrupole.md',
                    ('push', 'mixed', 20000000): 'push_mixed_hbm_traffic_rect2049sq_2e7_f32_quadrupole.md'}


def parse(argv=None):
    """Command line, checked completely before anything touches the GPU."""
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=5)
    ap.add_argument('--particles', type=float, default=None,
                    help='macro-particles (full: 2e7, dipole: 2e6)')
    ap.add_argument('--dtype', default=None, choices=['f64', 'f32'],
                    help='particle storage (full and scan: f32, dipole: f64)')
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--no-dipole', action='store_true', help='full: skip the dipole line beside it')
    ap.add_argument('--calib-copy', action='store_true',
                    help='run one plain 1 GiB device copy first (HBM counter calibration)')
    ap.add_argument('--cpu-seconds', type=float, default=15.0)
    ap.add_argument('--eager', action='store_true',
                    help='dipole: launch every step on its own (default: replay the steps between two '
                         'sorts as one captured HIP graph)')
    ap.add_argument('--sort-every', type=int, default=50,
                    help='re-order particles by cell every this many steps (0: never)')
    ap.add_argument('--target', type=float, default=None,
                    help='full: particle-management target count (default: the loaded count)')
    ap.add_argument('--train-steps', type=int, default=None,
                    help='learned, northstar: Adam steps (default 8000 for learned, 16000 for northstar)')
    ap.add_argument('--pairs', type=int, default=192, help='learned: training pairs')
    ap.add_argument('--train-pairs', choices=('operator', 'buildup'), default='operator',
                    help='learned: train on (A phi, phi) with phi random at 0-64 cell scales, or on the '
                         'buildup dumps (half) and white-noise charges (half) with phi from the direct solve')
    ap.add_argument('--grid', type=int, default=1024, help='learned: grid cells per side')
    ap.add_argument('--fcg-dirs', type=int, default=8,
                    help='learned, northstar: directions the flexible CG keeps (1-8)')
    ap.add_argument('--workload', default='full', choices=WORKLOADS,
                    help='full: the 2048^2 quadrupole buildup with particle management (default); '
                         'dipole: the 512x384 dipole buildup; scan: 64 buildups stepped together; '
                         'learned: the learned-preconditioner Poisson solve on a 1024^2 cut-cell grid; '
                         'northstar: the 64-run scan with every run solved by the learned flexible CG, against '
                         'the plain method; push: the full workload\'s push at its size (2e7 FP32 particles, '
                         'quadrupole on a 2048^2-cell grid) on a rectangular chamber whose Poisson box needs no '
                         'capacitance rows, for counter passes on the push alone')
    ap.add_argument('--runs', type=int, default=64, help='scan: number of buildups')
    ap.add_argument('--scan-start', type=int, default=700,
                    help='scan: the timed steps start at this step of the simulation (the first bunch passage, '
                         'once multipacting has set in), whatever the autotuner ran before')
    ap.add_argument('--tune-cache', default=None, help='scan: autotuner cache file (JSON)')
    ap.add_argument('--scan-config', default=None,
                    help='scan: sortEvery,order,binShift to use instead of autotuning')
    ap.add_argument('--no-f32-compare', action='store_true',
                    help='dipole: skip the FP32-storage run compared against the FP64 one')
    ap.add_argument('--push-arith', default=None, choices=('mixed', 'fp64'),
                    help='push arithmetic of FP32 particles: mixed (FP32 gather and Boris push, FP64 charge '
                         'sums and the FP64 push within 2 cells of the wall; the default) or fp64')
    a = ap.parse_args(argv)
    if a.gpus != 1:
        ap.exit(2, 'bench.py runs on one GPU only: this build has no multi-GPU path (the sharding '
                   'design is described with the build); got --gpus %d\n' % a.gpus)
    if a.steps < 1:
        ap.error('--steps must be at least 1')
    if a.warmup < 0 or a.sort_every < 0:
        ap.error('--warmup and --sort-every must not be negative')
    if a.particles is not None and a.particles < 1:
        ap.error('--particles must be positive')
    if a.scan_config is not None:
        try:
            v = [int(x) for x in a.scan_config.split(',')]
        except ValueError:
            v = []
        if len(v) != 3 or v[0] < 1 or v[1] not in (0, 1) or not 0 <= v[2] <= 8:
            ap.error('--scan-config is sortEvery,order,binShift (sortEvery >= 1, order 0 or 1, binShift 0-8)')
    if a.dtype is None:
        a.dtype = 'f64' if a.workload == 'dipole' else 'f32'
    if a.push_arith == 'mixed' and a.dtype == 'f64':
        ap.error('--push-arith mixed needs FP32 particles')
    if a.particles is None:
        a.particles = 2e6 if a.workload == 'dipole' else 2e7
    if not 1 <= a.fcg_dirs <= 8:
        ap.error('--fcg-dirs is 1-8')
    if a.train_steps is None:
        a.train_steps = 16000 if a.workload == 'northstar' else 8000
    return a


def step_plan(steps, sortEvery, graph):
    """The order of work for `steps` buildup steps with a cell sort after every sortEvery-th
    step (0: never).  Entries: ('graph', k) replays a captured graph of k steps, ('step',) runs
    one step on its own, ('sort',) synchronises and sorts.  With graph, whole sort intervals
    are graph replays and the remaining steps run on their own; the sorts fall after the same
    steps either way, so both plans advance the buildup identically."""
    steps, sortEvery = int(steps), int(sortEvery)
    plan = []
    if graph and sortEvery and steps >= sortEvery:
        for _ in range(steps // sortEvery):
            plan += [('graph', sortEvery), ('sort',)]
        plan += [('step',)] * (steps % sortEvery)
    elif graph and not sortEvery:
        plan = [('graph', steps)]
    else:
        for k in range(steps):
            plan.append(('step',))
            if sortEvery and k % sortEvery == sortEvery - 1:
                plan.append(('sort',))
    return plan


def plan_graph_steps(plan):
    """Steps per captured graph in a plan (0 when it replays none)."""
    return max([e[1] for e in plan if e[0] == 'graph'], default=0)


def plan_steps(plan):
    return sum(e[1] if e[0] == 'graph' else 1 for e in plan if e[0] != 'sort')


def run_plan(b, plan, events=None, onSort=None):
    """Runs a step plan on a native Buildup.  events: a list of (begin, end) HIP events, one
    pair per eager step, recorded around that step's push kernel.  onSort: called after each
    synchronisation, before the sort (particle management).  Returns the eager steps run."""
    keep = plan_graph_steps(plan) > 0
    k = 0
    for e in plan:
        if e[0] == 'graph':
            b.replay()
        elif e[0] == 'step':
            if events is not None and k < len(events):
                b._args.evPushBegin, b._args.evPushEnd = events[k][0].value, events[k][1].value
            b.step_native()
            b._args.evPushBegin = b._args.evPushEnd = None
            k += 1
        else:
            b.sync()
            if onSort is not None:
                onSort()
            b.sort_particles(keepBuffers=keep)
    return k


def make_events(n):
    import ctypes
    from openecloud_amd import _lib as L
    evs = []
    for _ in range(n):
        e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
        L.check(L.lib().oecl_event_create(ctypes.byref(e0)))
        L.check(L.lib().oecl_event_create(ctypes.byref(e1)))
        evs.append((e0, e1))
    return evs


def event_us(evs):
    """Elapsed microseconds of each event pair (and releases them)."""
    import ctypes
    from openecloud_amd import _lib as L
    out = []
    for e0, e1 in evs:
        ms = ctypes.c_float()
        L.check(L.lib().oecl_event_elapsed_ms(e0, e1, ctypes.byref(ms)))
        out.append(ms.value * 1e3)
        L.lib().oecl_event_destroy(e0)
        L.lib().oecl_event_destroy(e1)
    return out


def event_free(evs):
    """Releases event pairs that were never recorded."""
    from openecloud_amd import _lib as L
    for e0, e1 in evs:
        L.lib().oecl_event_destroy(e0)
        L.lib().oecl_event_destroy(e1)


def host_info():
    model = platform.processor() or ''
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                model = line.split(':', 1)[1].strip()
                break
    except OSError:
        pass
    try:
        cpus = len(os.sched_getaffinity(0))
    except AttributeError:
        cpus = os.cpu_count()
    return {'nproc': cpus, 'cpu_model': model}


# ---------------------------------------------------------------------------------------------
# workloads

def build_dipole(n, dtype):
    """The elliptical dipole buildup: 45 x 35 mm semi-axes on a 512 x 384 grid, By = 0.5 T,
    n macro-electrons, Gaussian bunches sigma 1.5 mm, 1.15e11 ppb, 4 bunches, dt 5e-12 s,
    Furman copper delta_max 1.3, primary photoelectrons 1e-3 per proton per metre."""
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
    photo = PhotoEmitter(pb, p, beam, photoYield=1e-3, macroPerBunch=1e5, energyMean=7., energySigma=5.)
    b = Buildup(g, p, pb, em, ps, beam, dt=dt, typeBField=1, photoEmitter=photo)
    return g, p, pb, em, ps, b


def build_full(n, log):
    """The 2048 x 2048 quadrupole buildup (see the module docstring)."""
    import torch
    from openecloud_amd import randomGen
    from openecloud_amd.beam import GaussianBunchTrain
    from openecloud_amd.buildup import Buildup
    from openecloud_amd.grid import Grid
    from openecloud_amd.magneticField import multipoleExpansion
    from openecloud_amd.particleBoundary import AbsorbElliptical
    from openecloud_amd.particleEmitter import FurmanEmitter, homoLoader
    from openecloud_amd.particles import Particles
    from openecloud_amd.poissonSolver import PoissonSolver
    dt = 2e-12
    t0 = time.perf_counter()
    randomGen.seed(0)
    g = Grid(2048, 2048, 0.09, 0.07, 1)
    bq = multipoleExpansion(g, [0., 50.])
    log('full: grid and field %.1f s' % (time.perf_counter() - t0))
    ps = PoissonSolver(g, method='direct', checkEvery=0)
    log('full: direct solver %.1f s (%d capacitance rows)' % (time.perf_counter() - t0, ps.capacitanceRows))
    p = Particles('electrons', 'variableWeight', dtype=torch.float32, ordering='fast')
    pb = AbsorbElliptical(g, p)
    em = FurmanEmitter(pb, p, 'copper', seyMax=1.5)
    beam = GaussianBunchTrain(g, dt, nParticles=1.15e11, tBunchSpacing=25e-9, radiusSigma=1e-3, nBunches=20)
    homoLoader(g, p, pb, n, 1e9 / n, 4.194e5)
    p.reserve(int(1.6 * n) + 4096)
    b = Buildup(g, p, pb, em, ps, beam, dt=dt, typeBField=2, bAtGridPoints=bq)
    return g, bq, p, pb, em, ps, b


def poisson_rate(ps, q, nsol):
    import torch
    ps.solve(q)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(nsol):
        ps.solve(q)
    torch.cuda.synchronize()
    return nsol / (time.perf_counter() - t0)


def measure_full(a, log):
    """Times a.steps steps of the 2048^2 quadrupole buildup (see the module docstring) after
    a.warmup steps; HIP events around every push kernel."""
    import torch
    from openecloud_amd import _lib as L
    from openecloud_amd.diagnostics import weight_sum
    target = int(a.target) if a.target else int(a.particles)
    # loaded just above the 1.5x trigger: the first management check (step 0) merges down to the
    # target, later checks run at every sort
    n = int(1.525 * target) if not a.target else int(a.particles)
    g, bq, p, pb, em, ps, b = build_full(n, log)
    b.native_setup(absCap=max(65536, n // 32), secCap=max(131072, n // 16),
                   maxSteps=a.warmup + a.steps + 20, pushArithmetic=a.push_arith)
    kern, bpp, bpn, bdesc = push_model('f32', b.pushArithmetic, True)
    mgmt = []

    def manage():
        import torch
        torch.cuda.synchronize()
        t = time.perf_counter()
        nb = p.macroParticleCount
        ops = b.manage(target, upper=1.5, lower=0.95, mergeGroup=8)
        torch.cuda.synchronize()
        if ops is not None:
            mgmt.append(dict(ops, step=b.it, count_before=nb, count_after=p.macroParticleCount,
                             seconds=time.perf_counter() - t))

    manage()
    b.sort_particles()
    log('full: management at step 0: %s' % mgmt)
    run_plan(b, step_plan(a.warmup, a.sort_every, False), onSort=manage)
    b.sync()
    plan = step_plan(a.steps, a.sort_every, False)
    evs = make_events(a.steps)
    it0 = b.it
    st0 = b.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run_plan(b, plan, events=evs, onSort=manage)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    st = b.sync()
    counts = b.counts(it0, it0 + a.steps)
    push = event_us(evs)
    avg_us = sum(push) / len(push)
    npx = g.getNpExt()
    bytes_per = sum(bpp * c + bpn * npx for c in counts) / len(counts)
    achieved = bytes_per / (avg_us * 1e-6) / 1e9
    hits = (st['absorbed_total'] - st0['absorbed_total']) / a.steps
    # one forced management event (the 1.5x trigger's reduction from the current count), timed
    nBefore = p.macroParticleCount
    q0 = weight_sum(p)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ops = b.manage(int(nBefore / 1.5), upper=1.0, lower=0., mergeGroup=8)
    torch.cuda.synchronize()
    t_merge = time.perf_counter() - t1
    q1 = weight_sum(p)
    solves = poisson_rate(ps, b.rho.clone(), 20)
    res = ps.residual(b.rho.clone())
    # counters are taken on the same push kernel at this size through --workload push (a
    # rectangular chamber of 2049^2 cells: the same particle count, grid size and field)
    traffic, traffic_src = None, 'not measured for this workload'
    if target == 20000000:
        name = TRAFFIC_PROFILES.get(('push', b.pushArithmetic, 20000000), '')
        traffic = profile_traffic(name)
        if traffic is not None:
            traffic_src = ('profiles/%s: the same %s at this size (2e7 FP32, quadrupole on the grid) on the '
                           'rectangular 2049^2-cell chamber of --workload push' % (name, kern))
    out = {
        'value': float(sum(counts)) / wall,
        'ms_per_step': wall / a.steps * 1e3,
        'config': {'workload': '2048x2048 elliptical 90x70 mm chamber, quadrupole 50 T/m on the grid, '
                               '%d FP32 macro-electrons, Gaussian bunches sigma 1 mm 1.15e11 ppb x20, dt 2e-12 s, '
                               'Furman copper delta_max 1.5, direct Poisson solve, localRandom management (groups of up to 8) '
                               'outside [0.95, 1.5] x %d (loaded at %d, merged at the step-0 check)'
                               % (target, target, n),
                   'grid': [2048, 2048], 'particles': target, 'loaded': n, 'management_target': target,
                   'capacitance_rows': ps.capacitanceRows, 'sort_every': a.sort_every,
                   'macro_particles_start_end': [counts[0], p.macroParticleCount if not ops else nBefore],
                   'hits_per_step': hits, 'absorbed_total': st['absorbed_total'],
                   'emitted_total': st['emitted_total']},
        'poisson_solves_per_s': solves,
        'poisson_relres': res,
        'push_arithmetic': b.pushArithmetic,
        'roofline': {'bound': 'hbm', 'kernel': kern, 'achieved': achieved, 'peak': HBM_PEAK_GBS,
                     'unit': 'GB/s', 'frac': achieved / HBM_PEAK_GBS, 'traffic': traffic,
                     'traffic_source': traffic_src, 'avg_launch_us': avg_us,
                     'bytes_per_launch': bytes_per, 'bytes_model': bdesc},
        'management_events': mgmt,
        'management_event': {'count_before': nBefore, 'count_after': p.macroParticleCount, 'seconds': t_merge,
                             'charge_rel_change': abs(q1 - q0) / abs(q0), 'ops': ops},
    }
    return out, g, bq, hits, counts


def measure_dipole(a, log, steps, warmup, n, dtype, eager, compare=False):
    """Times `steps` steps of the dipole buildup (graph replays between sorts unless eager);
    push kernel times from eager steps with HIP events after the timed region."""
    import torch
    g, p, pb, em, ps, b = build_dipole(n, dtype)
    b.native_setup(absCap=max(65536, n // 8), secCap=max(131072, n // 4),
                   maxSteps=warmup + steps + 40, pushArithmetic=a.push_arith if dtype == 'f32' else None)
    kern, bpp, bpn, bdesc = push_model(dtype, b.pushArithmetic, False)
    b.sort_particles()
    if a.calib_copy:
        from openecloud_amd import _lib as L
        # one plain 1 GiB device copy (beyond the 256 MB last-level cache) for the counters
        src_t = torch.ones(1 << 27, dtype=torch.float64, device='cuda')
        dst_t = torch.empty_like(src_t)
        torch.cuda.synchronize()
        L.check(L.lib().oecl_copy_bytes(L.ptr(dst_t), L.ptr(src_t), L.u64(src_t.numel() * 8), L.stream()),
                'calibration copy')
        torch.cuda.synchronize()
        del src_t, dst_t
    run_plan(b, step_plan(warmup, a.sort_every, False))
    b.sync()
    plan = step_plan(steps, a.sort_every, not eager)
    gs = plan_graph_steps(plan)
    if gs:
        b.capture_native(gs)
    it0 = b.it
    evs = make_events(steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ran = run_plan(b, plan, events=evs)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    evs_used, evs_left = evs[:ran], evs[ran:]
    # push kernel timing: a few more eager steps with events when the timed steps were replays
    extra = 0
    if len(evs_used) < min(20, steps):
        extra = min(20, steps) - len(evs_used)
        run_plan(b, [('step',)] * extra, events=evs_left[:extra])
        evs_used = evs_used + evs_left[:extra]
        evs_left = evs_left[extra:]
    st = b.sync()
    counts = b.counts(it0, it0 + steps)
    push_us = event_us(evs_used)
    event_free(evs_left)
    npx = g.getNpExt()
    avg_bytes = sum(bpp * c + bpn * npx for c in counts) / len(counts)
    avg_us = sum(push_us) / len(push_us)
    achieved = avg_bytes / (avg_us * 1e-6) / 1e9
    q = b.rho.clone()
    tr = profile_traffic(TRAFFIC_PROFILES.get(('dipole', dtype, n), ''))
    out = {
        'value': float(sum(counts)) / wall,
        'unit': 'particle-steps/s',
        'steps': steps,
        'warmup': warmup,
        'ms_per_step': wall / steps * 1e3,
        'dtype': dtype,
        'config': {'workload': 'elliptical 45x35 mm cut-cell chamber, 512x384 grid, By 0.5 T, '
                               '%d macro-electrons, Gaussian bunches sigma 1.5 mm 1.15e11 ppb x4, '
                               'dt 5e-12 s, Furman copper delta_max 1.3, primary photoelectrons '
                               '1e-3 per proton per m' % n,
                   'grid': [512, 384], 'particles': n, 'dt': 5e-12, 'graph_steps': gs,
                   'macro_particles_start_end': [counts[0], p.macroParticleCount],
                   'absorbed_total': st['absorbed_total'], 'emitted_total': st['emitted_total'],
                   'photoelectrons_total': st['photo_total']},
        'poisson_solves_per_s': poisson_rate(ps, q, 100),
        'poisson_relres': ps.residual(q),
        'poisson_capacitance_rows': ps.capacitanceRows,
        'push_arithmetic': b.pushArithmetic,
        'roofline': {'bound': 'hbm', 'kernel': kern, 'achieved': achieved,
                     'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': achieved / HBM_PEAK_GBS,
                     'traffic': tr,
                     'traffic_source': ('profiles/' + TRAFFIC_PROFILES[('dipole', dtype, n)]) if tr else
                     'not measured for this workload',
                     'avg_launch_us': avg_us, 'bytes_per_launch': avg_bytes, 'bytes_model': bdesc,
                     'note': 'the 96 MB FP64 particle state fits the 256 MB last-level cache'},
    }
    if compare:
        out['f32_vs_f64'] = compare_f32(a, n, b, warmup, steps, eager, extra)
    return out


def compare_f32(a, n, b64, warmup, steps, eager, extra):
    """The same seeded workload with FP32 particle storage, stepped exactly like the FP64 run
    (warm-up, the timed steps, `extra` eager steps); its throughput and the differences of the
    final state from the FP64 run: cloud density (charge on the grid, relative L2), total charge
    and macro-particle count."""
    import torch
    g, p, pb, em, ps, b = build_dipole(n, 'f32')
    b.native_setup(absCap=max(65536, n // 8), secCap=max(131072, n // 4), maxSteps=warmup + steps + 40,
                   pushArithmetic=a.push_arith)
    b.sort_particles()
    run_plan(b, step_plan(warmup, a.sort_every, False))
    b.sync()
    plan = step_plan(steps, a.sort_every, not eager)
    gs = plan_graph_steps(plan)
    if gs:
        b.capture_native(gs)
    i0 = b.it
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run_plan(b, plan)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    b.sync()
    rate = float(sum(b.counts(i0, i0 + steps))) / wall
    run_plan(b, [('step',)] * extra)
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
            'push_arithmetic': b.pushArithmetic,
            'note': 'FP32 particle storage, same seeds and steps; differences from the FP64 run after '
                    'the timed steps'}


# ---------------------------------------------------------------------------------------------
# CPU baseline: the oracle's C restatement of the reference loops, one core

def _sample_in_ellipse(rng, ns, lx, ly):
    import numpy as np
    pd = np.empty((ns, 6))
    k = 0
    while k < ns:
        x = (rng.random(ns) - 0.5) * lx
        y = (rng.random(ns) - 0.5) * ly
        ok = x * x * 4 / lx ** 2 + y * y * 4 / ly ** 2 < 1
        m = min(ns - k, int(ok.sum()))
        pd[k:k + m, 0] = x[ok][:m]
        pd[k:k + m, 1] = y[ok][:m]
        k += m
    pd[:, 2:5] = rng.standard_normal((ns, 3)) * 4.194e5
    pd[:, 5] = 500.
    return pd


def cpu_particle_rate(g, seconds, typeB, bConst=None, B=None, wall=1, ns=200000):
    """Seconds per particle-step of the oracle's C particle loop (weights, deposit, E gather,
    [B gather], Boris push, wall test, saveAbsorbed, compaction) on a sample of ns particles
    uniform in the elliptical chamber of grid g, with a random E on the grid."""
    import ctypes
    import numpy as np
    import oracle
    rng = np.random.default_rng(0)
    pd = _sample_in_ellipse(rng, ns, g.getLx(), g.getLy())
    npx = g.getNpExt()
    E = rng.standard_normal(2 * npx) * 1e3
    lib = oracle.lib()
    D = ctypes.POINTER(ctypes.c_double)
    U16 = ctypes.POINTER(ctypes.c_uint16)
    ic = np.empty(ns, np.uint32)
    w4 = np.empty(4 * ns)
    eat = np.empty(2 * ns)
    bat = np.empty(3 * ns)
    ins = np.empty(ns, np.uint16)
    rho = np.empty(npx)
    absb = np.empty((ns, 6))
    bc = np.asarray(bConst if bConst is not None else [0., 0., 0.], np.float64)
    Bp = np.ascontiguousarray(B, np.float64) if B is not None else None
    qm = -1.602176565e-19 / 9.10938291e-31
    args = (pd.ctypes.data_as(D), ctypes.c_uint32(ns), E.ctypes.data_as(D),
            Bp.ctypes.data_as(D) if Bp is not None else ctypes.c_void_p(0),
            ctypes.c_int(typeB), bc.ctypes.data_as(D), ctypes.c_double(2e-12), ctypes.c_double(qm),
            ctypes.c_double(1 / g.dx), ctypes.c_double(1 / g.dy), ctypes.c_double(g.lxExt / 2),
            ctypes.c_double(g.lyExt / 2), ctypes.c_uint32(g.nxExt), ctypes.c_uint32(npx),
            ctypes.c_int(wall), ctypes.c_double(g.getLx()), ctypes.c_double(g.getLy()),
            ic.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), w4.ctypes.data_as(D),
            eat.ctypes.data_as(D), bat.ctypes.data_as(D), ins.ctypes.data_as(U16), rho.ctypes.data_as(D),
            ctypes.c_double(-1.602176565e-19))
    lib.orc_save_absorbed.restype = ctypes.c_uint32

    def once():
        lib.orc_particle_step(*args)
        na = lib.orc_save_absorbed(pd.ctypes.data_as(D), ctypes.c_uint32(ns), ins.ctypes.data_as(U16),
                                   absb.ctypes.data_as(D))
        if na:
            ins[:] = 1   # keep the sample's size: the compaction below moves nothing
        lib.orc_remove(pd.ctypes.data_as(D), ins.ctypes.data_as(U16), ctypes.c_uint32(ns), ctypes.c_uint32(ns))

    once()
    t0 = time.perf_counter()
    reps = 0
    while time.perf_counter() - t0 < seconds:
        once()
        reps += 1
    return (time.perf_counter() - t0) / reps / ns, ns, reps


def cpu_emission_rate(seconds, seyMax, eHat0=None):
    """Seconds per wall impact of the Furman emission restatement (particleEmitter.pyx:154-308,
    Python with scipy.special) on impacts at 1-400 eV and random incidence on an ellipse."""
    import numpy as np
    from oracle import furman_oracle as F
    rng = np.random.default_rng(1)
    P = F.material('copper', seyMax=seyMax, eHat0=eHat0)
    n0 = 256
    th = rng.random(n0) * 2 * np.pi
    nrm = np.stack([-np.cos(th), -np.sin(th)], 1)
    e = rng.uniform(1., 400., n0)
    v = np.sqrt(2 * e * 1.602176565e-19 / 9.10938291e-31)
    ang = rng.uniform(-1.4, 1.4, n0)
    ab = np.zeros((n0, 6))
    ab[:, 2] = -v * (np.cos(ang) * nrm[:, 0] - np.sin(ang) * nrm[:, 1])
    ab[:, 3] = -v * (np.sin(ang) * nrm[:, 0] + np.cos(ang) * nrm[:, 1])
    ab[:, 5] = 500.
    t0 = time.perf_counter()
    done = 0
    while time.perf_counter() - t0 < seconds or done == 0:
        U = rng.random((n0, F.N_UNIFORMS))
        F.generate_secondaries(ab, nrm, U, P)
        done += n0
    return (time.perf_counter() - t0) / done, done


def cpu_superlu_rate(g, seconds):
    import numpy as np
    from oracle.grid_oracle import OraclePoisson
    rng = np.random.default_rng(2)
    ps = OraclePoisson(g)
    q = rng.standard_normal(g.getNpExt()) * 1e-12
    ps.solve(q)
    t0 = time.perf_counter()
    nsol = 0
    while time.perf_counter() - t0 < seconds or nsol == 0:
        ps.solve(q)
        nsol += 1
    return (time.perf_counter() - t0) / nsol, nsol


def cpu_baseline_full(g, bq, n_avg, hits, seconds):
    """The 2048^2 quadrupole step on one core: the C particle loop with B on the grid on a
    2e5-particle sample, plus the Furman emission per impact times the GPU run's impacts per
    step.  The reference's SuperLU factorisation of this 3.3e6-unknown system does not fit a
    bounded sample, so the Poisson solve is left out of this figure (which favours the CPU);
    SuperLU solves/s of the 512 x 384 system are reported beside it."""
    from openecloud_amd.grid import Grid
    t_part, ns, reps = cpu_particle_rate(g, seconds * 0.55, 2, B=bq)
    t_imp, nimp = cpu_emission_rate(seconds * 0.2, 1.5)
    g2 = Grid(512, 384, 0.09, 0.07, 1)
    t_lu, nsol = cpu_superlu_rate(g2, seconds * 0.25)
    step = t_part * n_avg + t_imp * hits
    info = host_info()
    return {'value': n_avg / step, 'unit': 'particle-steps/s', 'cores': 1, 'kind': 'port',
            'sample': '%d-particle sample x %d reps of the C particle loop with the quadrupole on the 2048^2 grid '
                      '(%.3g s/particle-step) plus %d Furman impacts of the Python/scipy restatement (%.3g s each) '
                      'x %.0f impacts per step, scaled to %.3g particles per step; Poisson solve not included '
                      '(SuperLU at 2048^2 exceeds the sample bound)' % (ns, reps, t_part, nimp, t_imp, hits, n_avg),
            'nproc': info['nproc'], 'cpu_model': info['cpu_model'],
            'superlu_solves_per_s_512x384': 1. / t_lu}


def cpu_baseline_dipole(g, n, hits, seconds):
    """The dipole step on one core: the C particle loop (constant By) on a 2e5-particle sample
    of the 512 x 384 chamber, the Furman emission per impact times the GPU run's impacts per
    step, and the reference's SuperLU back-substitution of the same system."""
    t_part, ns, reps = cpu_particle_rate(g, seconds * 0.5, 1, bConst=[0., 0.5, 0.])
    t_imp, nimp = cpu_emission_rate(seconds * 0.2, 1.3)
    t_lu, nsol = cpu_superlu_rate(g, seconds * 0.3)
    step = t_part * n + t_imp * hits + t_lu
    info = host_info()
    return {'value': n / step, 'unit': 'particle-steps/s', 'cores': 1, 'kind': 'port',
            'sample': '%d-particle sample x %d reps of the C particle loop (%.3g s/particle-step) + %d Furman '
                      'impacts (Python/scipy restatement, %.3g s each) x %.0f per step + %d SuperLU solves of the '
                      '512x384 system (%.3g s each), scaled to %d particles per step'
                      % (ns, reps, t_part, nimp, t_imp, hits, nsol, t_lu, n),
            'nproc': info['nproc'], 'cpu_model': info['cpu_model'], 'poisson_solves_per_s': 1. / t_lu}


def build_scan(R, dtype, maxSteps, pushArithmetic=None):
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
    scan = BuildupScan(runs, seeds, growth=1.5, maxSteps=maxSteps, pushArithmetic=pushArithmetic)
    return g, ps, runs, scan, dict(sey=sey, ppb=ppb, n=nmp)


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
    net = UNetPreconditioner(g, seed=0, mDirs=a.fcg_dirs)
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


def main_scan(a):
    import ctypes
    import torch
    from openecloud_amd import _lib as L
    from openecloud_amd import randomGen
    from openecloud_amd.autotune import tune_scan
    L.require_gpu()
    R = a.runs
    log = lambda m: print(m, file=sys.stderr, flush=True)
    start = max(int(a.scan_start), a.warmup)
    maxSteps = start + a.steps + 2000
    g, ps, runs, scan, prm = build_scan(R, a.dtype, maxSteps, a.push_arith)
    scan.sort()
    for _ in range(a.warmup):
        scan.step()
    scan.sync()
    if not a.scan_config:
        # tune where the timed window runs: the state at step `start` (multipacting under way),
        # reached with a sort every 25 steps
        scan.capture(25)
        while scan.it + 25 <= start:
            scan.sort()
            scan.replay()
        while scan.it < start:
            scan.step()
        scan.sync()
    if a.scan_config:
        se_, o_, b_ = [int(v) for v in a.scan_config.split(',')]
        best, results = dict(sortEvery=se_, order=o_, binShift=b_, graphSteps=se_), []
    else:
        best, results = tune_scan(scan, cacheFile=a.tune_cache, log=log)
    se = int(best['sortEvery'])
    steps = a.steps
    # the timed window is steps [start, start + steps) of a fresh scan from the same seeds, so
    # that it does not depend on how many steps the autotuner ran (or whether it ran at all)
    del scan, runs
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    g, ps, runs, scan, prm = build_scan(R, a.dtype, maxSteps, a.push_arith)
    scan.set_deposit(int(best.get('span', 0)), bool(best.get('waveTiles', False)))
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
    # whole sort intervals as graph replays, the remaining steps of a last interval one by one
    for _ in range(steps // se):
        scan.sort(best['order'], best['binShift'])
        scan.replay()
    if steps % se:
        scan.sort(best['order'], best['binShift'])
        for _ in range(steps % se):
            scan.step()
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
    kern, bpp, bpn, bdesc = push_model(a.dtype, scan.pushArithmetic, False)
    by = [bpp * tot[k] + bpn * g.getNpExt() * R for k in range(se)]
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
    tr = profile_traffic(TRAFFIC_PROFILES.get(('scan', scan.pushArithmetic, R), ''))
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
        'push_arithmetic': scan.pushArithmetic,
        'roofline': {'bound': 'hbm', 'kernel': kern, 'achieved': achieved, 'peak': HBM_PEAK_GBS,
                     'unit': 'GB/s', 'frac': achieved / HBM_PEAK_GBS, 'traffic': tr,
                     'traffic_source': ('profiles/' + TRAFFIC_PROFILES[('scan', scan.pushArithmetic, R)]) if tr else
                     'not measured for this workload',
                     'avg_launch_us': avg_us, 'bytes_per_launch': sum(by) / len(by), 'bytes_model': bdesc},
        'autotune': {'best': best, 'candidates_us_per_step': [[c, 1e6 * t] for c, t in results]},
        'plain_one_run_per_call': {'value': seq_rate, 'unit': 'particle-steps/s',
                                   'note': 'the same runs stepped one at a time with the single-run one-call step'},
        'plain_reference_order': {'value': ref_rate, 'unit': 'particle-steps/s',
                                  'note': 'method-by-method GPU step in the reference order, 4 of the runs'},
        'speedup_vs_one_run_per_call': rate / seq_rate,
        'speedup_vs_reference_order': rate / ref_rate,
    }
    print(json.dumps(out))



def main_full(a):
    """The default line: the 2048^2 quadrupole buildup, with the dipole buildup and the CPU
    baseline beside it."""
    import torch
    from openecloud_amd import _lib as L
    L.require_gpu()
    log = lambda m: print(m, file=sys.stderr, flush=True)
    res, g, bq, hits, counts = measure_full(a, log)
    log('full: %.4g particle-steps/s, %.3f ms/step' % (res['value'], res['ms_per_step']))
    out = {
        'metric': 'buildup particle-steps/sec and Poisson solves/sec, learned-PCG batched vs plain',
        'value': res['value'],
        'unit': 'particle-steps/s',
        'n_gpus': 1,
        'steps': a.steps,
        'warmup': a.warmup,
        'ms_per_step': res['ms_per_step'],
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32',
        'data': 'synthetic (seeded uniform load, Philox draws)',
        'config': res['config'],
        'solver': 'direct (capacitance-matrix DST solve, the reference LU as its oracle)',
        'poisson_solves_per_s': res['poisson_solves_per_s'],
        'poisson_relres': res['poisson_relres'],
        'roofline': res['roofline'],
        'management_events': res['management_events'],
        'management_event': res['management_event'],
    }
    if not a.no_dipole:
        torch.cuda.synchronize()
        d = measure_dipole(a, log, a.steps, a.warmup, 2000000, 'f64', a.eager)
        log('dipole: %.4g particle-steps/s, %.3f ms/step' % (d['value'], d['ms_per_step']))
        out['plain_dipole'] = d
    if not a.no_cpu_baseline:
        n_avg = sum(counts) / len(counts)
        cb = cpu_baseline_full(g, bq, n_avg, hits, a.cpu_seconds)
        out['cpu_baseline'] = cb
    print(json.dumps(out), flush=True)


def main_dipole(a):
    from openecloud_amd import _lib as L
    L.require_gpu()
    log = lambda m: print(m, file=sys.stderr, flush=True)
    n = int(a.particles)
    d = measure_dipole(a, log, a.steps, a.warmup, n, a.dtype, a.eager,
                       compare=not a.no_f32_compare and a.dtype == 'f64')
    out = {'metric': 'buildup particle-steps/sec and Poisson solves/sec, learned-PCG batched vs plain',
           'n_gpus': 1, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
           'data': 'synthetic (seeded uniform load, Philox draws)', 'solver': 'direct'}
    out.update(d)
    if not a.no_cpu_baseline:
        from openecloud_amd.grid import Grid
        g = Grid(512, 384, 0.09, 0.07, 1)
        hits = d['config']['absorbed_total'] / max(1, a.warmup + a.steps)
        out['cpu_baseline'] = cpu_baseline_dipole(g, n, hits, a.cpu_seconds)
    print(json.dumps(out), flush=True)


def _scan_rate(scan, steps, sortEvery):
    """Particle-steps/s of `steps` eager steps of every run of a scan (a cell sort every
    sortEvery steps, counted from the first)."""
    import torch
    it0 = scan.it
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        if sortEvery and k % sortEvery == 0:
            scan.sort()
        scan.step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    scan.sync()
    ps_ = float(sum(sum(scan.history(r, it0, it0 + steps)) for r in range(scan.R)))
    return ps_ / wall, wall, ps_


def main_northstar(a):
    """The combined form: the 64-run scan (256x256 rectangular 100x50 mm chamber, By 0.3 T,
    delta_max over [1.1, 2.0], log-uniform intensities and counts, FP32 particles with FP64
    charge and Poisson arithmetic) with every run's Poisson problem solved by the learned
    flexible CG to 1e-8 relative residual, warm-started from the run's previous potential; the
    U-Net is trained on the device on (q, phi) pairs dumped from the scan's own runs, and the
    direct solve is fallback and judge.  Beside it: the same runs with the batched direct
    solve, the plain GPU method (the reference-order method-by-method step, a drop-in for the
    reference's loop) on four of the runs, and the cloud density of the learned FP32 runs
    against the same runs stepped in FP64 with the direct solve."""
    import numpy as np
    import torch
    from openecloud_amd import _lib as L
    from openecloud_amd.learnedPreconditioner import UNetPreconditioner
    L.require_gpu()
    log = lambda m: print(m, file=sys.stderr, flush=True)
    R, W, K = a.runs, a.warmup, a.steps
    dumpEvery, nDumps = 10, 4
    maxSteps = W + K + dumpEvery * nDumps + 40
    # training pairs from the scan itself: every run's (q, phi) every dumpEvery steps
    t0 = time.perf_counter()
    g, ps, runs, scan, prm = build_scan(R, 'f32', maxSteps)
    scan.sort()
    pairs = []
    for d in range(nDumps):
        for _ in range(dumpEvery):
            scan.step()
        for r in range(R):
            sc = float(scan._scaleHost[r, scan.it])
            q = (scan.run_rho(r) + sc * scan._prof).clone()
            pairs.append((q, ps.solvePhi(q).clone()))
    scan.sync()
    net = UNetPreconditioner(g, seed=0, mDirs=a.fcg_dirs)
    val = [pairs[-1][0], pairs[-R // 2][0]]   # held out: the model with the fewest iterations on them is kept
    # the scan's own dumps, and as many operator pairs (A phi, phi) with phi random at 0-64 cell
    # scales: cloud potentials alone leave the network's short reach poorly trained
    train = net.smoothed_pairs(ps, pairs[:-1] + net.operator_pairs(ps, len(pairs), seed=1))
    t1 = time.perf_counter()
    hist = net.train(ps, train, steps=a.train_steps, lr=5e-4, lrEnd=5e-5, batch=4,
                     logEvery=max(1, a.train_steps // 10), validate=val,
                     validateEvery=max(1, a.train_steps // 6), validateIter=4000)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    log('northstar: %d pairs, trained in %.1f s (set-up %.1f s)' % (len(pairs), t_train, t1 - t0))
    del scan, runs, pairs, train
    torch.cuda.empty_cache()
    # the learned batched form
    g, ps, runs, scan, prm = build_scan(R, 'f32', maxSteps)
    scan.sort()
    scan.set_solver(net, tol=1e-8, maxIter=4000, judgeEvery=max(1, K // 4))
    for _ in range(W):
        scan.step()
    scan.sync()
    nIt0 = len(scan.learnedStats['iterations'])
    rate_l, wall_l, psteps_l = _scan_rate(scan, K, a.sort_every)
    its = scan.learnedStats['iterations'][nIt0:]
    lstats = dict(scan.learnedStats)
    lstats['iterations'] = None
    rho_l = [scan.run_rho(r).clone() for r in range(R)]
    counts_l = scan.counts()
    log('northstar: learned batched %.4g particle-steps/s, %.1f FCG iterations per solve' % (rate_l, np.mean(its)))
    del scan, runs
    torch.cuda.empty_cache()
    # the same runs, batched direct solve
    g, ps, runs, scan, prm = build_scan(R, 'f32', maxSteps)
    scan.sort()
    for _ in range(W):
        scan.step()
    scan.sync()
    rate_d, wall_d, _ = _scan_rate(scan, K, a.sort_every)
    # the plain GPU method: the reference-order method-by-method step on four of the runs
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
    rate_plain = ref_ps / ref_t
    del scan, runs
    torch.cuda.empty_cache()
    # cloud density: the learned FP32 runs against the same runs in FP64 with the direct solve
    g, ps, runs, scan, prm = build_scan(R, 'f64', maxSteps)
    scan.sort()
    for k in range(W + K):
        if k >= W and a.sort_every and (k - W) % a.sort_every == 0:
            scan.sort()
        scan.step()
    scan.sync()
    errs = []
    for r in range(R):
        r64 = scan.run_rho(r)
        den = float(torch.linalg.norm(r64))
        errs.append(float(torch.linalg.norm(rho_l[r] - r64)) / den if den > 0. else 0.)
    counts64 = scan.counts()
    out = {
        'metric': 'buildup particle-steps/sec and Poisson solves/sec, learned-PCG batched vs plain',
        'value': rate_l,
        'unit': 'particle-steps/s',
        'n_gpus': 1,
        'steps': K,
        'warmup': W,
        'ms_per_step': wall_l / K * 1e3,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32 particles, f64 Poisson (learned FCG: bf16 network with f32 accumulation, f64 vectors)',
        'data': 'synthetic (seeded uniform loads, Philox draws); the network trained on the scan\'s own dumps',
        'config': {'workload': '%d buildups on a shared 256x256 rectangular 100x50 mm chamber, By 0.3 T, dt 1e-11 s, '
                               'Gaussian bunches sigma 2 mm x8, delta_max linspace(1.1, 2.0), ppb log-uniform '
                               '[5e10, 2e11], macro-particles log-uniform [1e4, 2e6]; every run\'s Poisson problem by '
                               'the learned flexible CG to 1e-8, warm-started, direct solve as fallback and judge' % R,
                   'runs': R, 'grid': [256, 256], 'train_pairs': '%d scan dumps + %d operator pairs' % (R * nDumps - 1,
                                                                                            R * nDumps),
                   'train_steps': a.train_steps, 'fcg_dirs': a.fcg_dirs,
                   'train_seconds': t_train, 'loss_first_last': [hist[0][1], hist[-1][1]] if hist else None,
                   'validation_iterations': net.validation},
        'poisson_solves_per_s': R * K / wall_l,
        'learned': {'fcg_iterations_mean': float(np.mean(its)), 'fcg_iterations_max': int(np.max(its)),
                    'relres_max': lstats['relres_max'], 'fallbacks': lstats['fallbacks'],
                    'judge_max_rel_err': lstats['judge_max_rel_err'], 'judged': lstats['judged']},
        'batched_direct': {'value': rate_d, 'unit': 'particle-steps/s', 'ms_per_step': wall_d / K * 1e3,
                           'poisson_solves_per_s': R * K / wall_d},
        'plain_reference_order': {'value': rate_plain, 'unit': 'particle-steps/s',
                                  'note': 'method-by-method GPU step in the reference order, 4 of the runs'},
        'ratio_learned_batched_vs_plain': rate_l / rate_plain,
        'ratio_direct_batched_vs_plain': rate_d / rate_plain,
        'cloud_density_vs_fp64_direct': {'rel_l2_max': max(errs), 'rel_l2_mean': float(np.mean(errs)),
                                         'macro_particles_f32_f64': [int(sum(counts_l)), int(sum(counts64))],
                                         'after_steps': W + K},
    }
    print(json.dumps(out), flush=True)


def main_push(a):
    """The push of the full workload at its size: 2e7 FP32 macro-electrons (uniform load, 1 eV
    Maxwellian), the 50 T/m quadrupole on the grid, a 2049 x 2049-cell rectangular 90 x 70 mm
    chamber (2051 extended nodes: the Poisson box is 2047 = 2^11 - 1 columns wide with no
    capacitance rows, so the set-up has no dense factorisation), cell-sorted once, then
    warm-up and timed one-call steps; HIP events around every push kernel.  For counter passes
    on the push alone (rocprofv3 --kernel-include-regex k_push_dev)."""
    import torch
    from openecloud_amd import _lib as L
    from openecloud_amd import randomGen
    from openecloud_amd.beam import GaussianBunchTrain
    from openecloud_amd.buildup import Buildup
    from openecloud_amd.grid import Grid
    from openecloud_amd.magneticField import multipoleExpansion
    from openecloud_amd.particleBoundary import AbsorbRectangular
    from openecloud_amd.particleEmitter import FurmanEmitter, homoLoader
    from openecloud_amd.particles import Particles
    from openecloud_amd.poissonSolver import PoissonSolver
    L.require_gpu()
    log = lambda m: print(m, file=sys.stderr, flush=True)
    n = int(a.particles)
    randomGen.seed(0)
    g = Grid(2049, 2049, 0.09, 0.07, 0)
    ps = PoissonSolver(g, method='direct', checkEvery=0)
    log('push: grid %d^2 extended, %d capacitance rows' % (g.getNxExt(), ps.capacitanceRows))
    bq = multipoleExpansion(g, [0., 50.])
    p = Particles('electrons', 'variableWeight', dtype=torch.float32, ordering='fast')
    pb = AbsorbRectangular(g, p)
    em = FurmanEmitter(pb, p, 'copper', seyMax=1.5)
    beam = GaussianBunchTrain(g, 2e-12, nParticles=1.15e11, tBunchSpacing=25e-9, radiusSigma=1e-3, nBunches=20)
    homoLoader(g, p, pb, n, 1e9 / n, 4.194e5)
    p.reserve(int(1.6 * n) + 4096)
    b = Buildup(g, p, pb, em, ps, beam, dt=2e-12, typeBField=2, bAtGridPoints=bq)
    b.native_setup(absCap=max(65536, n // 32), secCap=max(131072, n // 16), maxSteps=a.warmup + a.steps + 4,
                   pushArithmetic=a.push_arith)
    kern, bpp, bpn, bdesc = push_model('f32', b.pushArithmetic, True)
    b.sort_particles()
    run_plan(b, [('step',)] * a.warmup)
    b.sync()
    evs = make_events(a.steps)
    it0 = b.it
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run_plan(b, [('step',)] * a.steps, events=evs)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    b.sync()
    counts = b.counts(it0, it0 + a.steps)
    push = event_us(evs)
    avg_us = sum(push) / len(push)
    npx = g.getNpExt()
    bytes_per = sum(bpp * c + bpn * npx for c in counts) / len(counts)
    achieved = bytes_per / (avg_us * 1e-6) / 1e9
    tname = TRAFFIC_PROFILES.get(('push', b.pushArithmetic, 20000000), '')
    tb = profile_traffic(tname) if n == 20000000 else None
    out = {
        'metric': 'push kernel of the full workload: particle-steps/sec',
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
        'config': {'workload': '2049x2049-cell rectangular 90x70 mm chamber (no capacitance rows), quadrupole '
                               '50 T/m on the grid, %d FP32 macro-electrons, dt 2e-12 s, direct Poisson solve' % n,
                   'grid': [2049, 2049], 'particles': n, 'push_us': push},
        'push_arithmetic': b.pushArithmetic,
        'roofline': {'bound': 'hbm', 'kernel': kern, 'achieved': achieved, 'peak': HBM_PEAK_GBS,
                     'unit': 'GB/s', 'frac': achieved / HBM_PEAK_GBS, 'traffic': tb,
                     'traffic_source': ('profiles/' + tname) if tb else None,
                     'avg_launch_us': avg_us, 'bytes_per_launch': bytes_per, 'bytes_model': bdesc},
    }
    log('push: %.1f us per launch, %.3f of 8 TB/s' % (avg_us, achieved / HBM_PEAK_GBS))
    print(json.dumps(out), flush=True)


def main(argv=None):
    a = parse(argv)
    if a.workload == 'push':
        return main_push(a)
    if a.workload == 'northstar':
        return main_northstar(a)
    if a.workload == 'learned':
        return main_learned(a)
    if a.workload == 'scan':
        return main_scan(a)
    if a.workload == 'dipole':
        return main_dipole(a)
    return main_full(a)


if __name__ == '__main__':
    main()
