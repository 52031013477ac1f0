"""CPU: bench.py's argument checks and step plans (no GPU work)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def sorts_after(plan):
    """Step counts after which each sort falls."""
    k, out = 0, []
    for e in plan:
        if e[0] == 'graph':
            k += e[1]
        elif e[0] == 'step':
            k += 1
        else:
            out.append(k)
    return out


@pytest.mark.parametrize('steps,sortEvery', [(20, 50), (20, 5), (50, 50), (120, 50), (1, 1), (7, 3),
                                              (200, 50), (20, 0), (3, 0), (99, 10)])
def test_plans_run_exactly_the_steps_with_the_same_sorts(steps, sortEvery):
    eager = bench.step_plan(steps, sortEvery, False)
    graph = bench.step_plan(steps, sortEvery, True)
    assert bench.plan_steps(eager) == steps
    assert bench.plan_steps(graph) == steps
    assert sorts_after(eager) == sorts_after(graph)
    if sortEvery:
        assert sorts_after(eager) == [sortEvery * (i + 1) for i in range(steps // sortEvery)]
    else:
        assert sorts_after(eager) == []
    gs = bench.plan_graph_steps(graph)
    assert all(e[1] == gs for e in graph if e[0] == 'graph')
    assert bench.plan_graph_steps(eager) == 0


def test_driver_command_plans():
    # the driver's default: 20 timed steps, 5 warm-up steps, sorts every 50
    a = bench.parse(['--gpus', '1', '--steps', '20', '--warmup', '5'])
    assert a.workload == 'full' and a.dtype == 'f32' and a.particles == 2e7
    assert bench.step_plan(a.steps, a.sort_every, True) == [('step',)] * 20
    assert bench.step_plan(a.warmup, a.sort_every, False) == [('step',)] * 5
    assert bench.step_plan(120, 50, True) == [('graph', 50), ('sort',), ('graph', 50), ('sort',)] + [('step',)] * 20


def test_dipole_defaults():
    a = bench.parse(['--workload', 'dipole'])
    assert a.dtype == 'f64' and a.particles == 2e6


@pytest.mark.parametrize('argv', [['--steps', '0'], ['--warmup', '-1'], ['--sort-every', '-2'],
                                  ['--scan-config', '5,2,0'], ['--scan-config', 'x'], ['--particles', '0']])
def test_bad_arguments_exit_2(argv):
    with pytest.raises(SystemExit) as e:
        bench.parse(argv)
    assert e.value.code == 2


def test_more_than_one_gpu_exits_before_any_gpu_work():
    # a fresh interpreter: the refusal comes from the argument check, before torch is imported
    code = ('import sys; sys.path.insert(0, %r); import bench\n'
            'try:\n    bench.main(["--gpus", "2"])\n'
            'except SystemExit as e:\n    print("rc", e.code, "torch" in sys.modules)\n') % ROOT
    r = subprocess.run([sys.executable, '-c', code], capture_output=True, text=True, timeout=120)
    assert 'rc 2 False' in r.stdout, r.stdout + r.stderr
    assert 'one GPU only' in r.stderr


def test_traffic_profiles_resolve():
    """Every traffic profile the bench names exists and holds its per-launch bytes; a workload
    without one gets None (not an error)."""
    assert bench.profile_traffic('') is None
    assert bench.profile_traffic('no_such_profile.md') is None
    for key, name in bench.TRAFFIC_PROFILES.items():
        v = bench.profile_traffic(name)
        assert v is not None and v > 1e6, (key, name)
    k, bpp, bpn, _ = bench.push_model('f32', 'mixed', True)
    assert (k, bpp, bpn) == ('k_push_mixed', 44, 32)
    k, bpp, bpn, _ = bench.push_model('f32', 'fp64', True)
    assert (k, bpp, bpn) == ('k_push_dev', 44, 48)
    k, bpp, bpn, _ = bench.push_model('f64', None, False)
    assert (k, bpp, bpn) == ('k_push_dev', 88, 32)
