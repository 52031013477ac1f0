"""CPU: the particle-loop oracle pinned by closed forms, invariants and known answers."""
import math

import numpy as np
import pytest

from conftest import random_particles
from openecloud_amd.grid import Grid

Q = -1.602176565e-19
M = 9.10938291e-31


@pytest.fixture(scope='module')
def g():
    return Grid(128, 64, 0.1, 0.05, 0)


def test_philox_known_answers(orc):
    """Random123 kat_vectors for philox4x32-10."""
    assert [hex(v) for v in orc.philox_words(1, 0, 0, 0, 0, 0, 0)[0]] == \
        ['0x6627e8d5', '0xe169c58d', '0xbc57ac4c', '0x9b00dbd8']
    w = orc.philox_words(1, 0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff)
    assert [hex(v) for v in w[0]] == ['0x408f276d', '0x41c83b0e', '0xa20bc7c6', '0x6d5451fd']
    w = orc.philox_words(1, 0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344, 0xa4093822, 0x299f31d0)
    assert [hex(v) for v in w[0]] == ['0xd16cfe09', '0x94fdcceb', '0x5001e420', '0x24126ea1']


def test_weights_partition_of_unity_and_linear_field(orc, g):
    pd = random_particles(g, 3000, seed=1)
    ic, w = orc.calc_grid_weights(pd, g.dx, g.dy, g.lxExt, g.lyExt, g.nxExt)
    assert np.allclose(w.sum(axis=0), 1., atol=1e-15)
    assert np.all(w >= 0)
    X, Y = np.meshgrid(g.xMesh, g.yMesh)
    a, b = 3.0, -7.0
    f = np.concatenate([(a * X + b * Y).ravel(), (b * X - a * Y).ravel()])
    e = orc.field_to_particles(f, ic, w, g.nxExt, g.npExt, 2)
    assert np.allclose(e[:, 0], a * pd[:, 0] + b * pd[:, 1], atol=1e-14)
    assert np.allclose(e[:, 1], b * pd[:, 0] - a * pd[:, 1], atol=1e-14)


def test_deposit_total_charge_and_node_particle(orc, g):
    pd = random_particles(g, 3000, seed=2)
    ic, w = orc.calc_grid_weights(pd, g.dx, g.dy, g.lxExt, g.lyExt, g.nxExt)
    rho = orc.charge_to_grid(pd, ic, w, Q, g.nxExt, g.npExt)
    assert math.isclose(rho.sum(), Q * pd[:, 5].sum(), rel_tol=1e-13)
    one = np.array([[g.xMesh[40], g.yMesh[20], 0, 0, 0, 2.5]])
    ic, w = orc.calc_grid_weights(one, g.dx, g.dy, g.lxExt, g.lyExt, g.nxExt)
    rho = orc.charge_to_grid(one, ic, w, Q, g.nxExt, g.npExt)
    k = 40 + 20 * g.nxExt
    # the node coordinate round-trips through (x + lxExt/2)/dx only to rounding
    assert np.count_nonzero(np.abs(rho) > 1e-9 * abs(2.5 * Q)) == 1
    assert math.isclose(rho[k], 2.5 * Q, rel_tol=1e-9)


def test_boris_uniform_e_no_b_exact_acceleration(orc):
    pd = np.zeros((4, 6))
    pd[:, 2] = [0, 1e5, -2e5, 3e5]
    e = np.tile([1e4, -2e4], (4, 1))
    dt = 1e-11
    qm = Q / M
    out = orc.boris_push(pd, e, dt, qm, 0)
    assert np.allclose(out[:, 2], pd[:, 2] + qm * dt * 1e4, rtol=1e-15)
    assert np.allclose(out[:, 0], dt * out[:, 2], rtol=1e-15)


def test_boris_uniform_b_rotation_angle_and_energy(orc):
    n = 200
    rng = np.random.default_rng(3)
    pd = np.zeros((n, 6))
    pd[:, 2:5] = rng.standard_normal((n, 3)) * 1e6
    bz = 0.5
    dt = 5e-12
    qm = Q / M
    out = orc.boris_push(pd, np.zeros((n, 2)), dt, qm, 1, np.array([0., 0., bz]))
    v0 = np.linalg.norm(pd[:, 2:5], axis=1)
    v1 = np.linalg.norm(out[:, 2:5], axis=1)
    assert np.allclose(v0, v1, rtol=1e-14)
    ang = np.arctan2(out[:, 3], out[:, 2]) - np.arctan2(pd[:, 3], pd[:, 2])
    ang = (ang + np.pi) % (2 * np.pi) - np.pi
    expect = -2 * math.atan(qm * bz * dt / 2)  # textbook Boris rotation angle
    assert np.allclose(ang, expect, rtol=1e-9)
    assert np.array_equal(out[:, 4], pd[:, 4])  # vz untouched by Bz


def test_boris_e_cross_b_drift(orc):
    """In crossed E, B the guiding centre drifts at E x B / B^2."""
    e = np.array([[1e3, 0.]])
    b = np.array([0., 0., 0.1])
    pd = np.zeros((1, 6))
    pd[0, 3] = -1e3 / 0.1  # start on the drift velocity: no gyration
    dt = 1e-12
    out = pd
    for _ in range(100):
        out = orc.boris_push(out, e, dt, Q / M, 1, b)
    assert np.allclose(out[0, 2:4], [0., -1e4], atol=1e-6)


def test_compaction_rules_hand_cases(orc):
    base = np.arange(8, dtype=float)[:, None] * np.ones((1, 6))
    keep = np.array([1, 0, 1, 1, 0, 1, 0, 1], np.uint16)
    add = 100 + np.arange(2, dtype=float)[:, None] * np.ones((1, 6))
    # _removeParticles: holes below the new count filled from the end
    out, _ = orc.add_and_remove(base, np.zeros((0, 6)), keep, 8)
    assert list(out[:, 0]) == [0, 7, 2, 3, 5]
    # _addAndRemoveParticles (capacity suffices): 2 additions into holes 1 and 4, then the
    # kept tail after hole 4 is shifted down: [0, 100, 2, 3, 101, 5, 7]
    out, _ = orc.add_and_remove(base, add, keep, 8)
    assert list(out[:, 0]) == [0, 100, 2, 3, 101, 5, 7]
    # _addAndRemoveParticlesNewArray (capacity exceeded): stable keep then append
    out, cap = orc.add_and_remove(base, np.vstack([add] * 3), keep, 8)
    assert list(out[:, 0]) == [0, 2, 3, 5, 7, 100, 101, 100, 101, 100, 101]
    assert cap == int(11 * 1.1)


@pytest.mark.parametrize('nadd', [0, 30, 600])
def test_fast_compaction_rule_agrees_with_reference_where_defined(orc, nadd):
    """The fast rule equals _removeParticles when nothing is added and
    _addAndRemoveParticles when the additions cover every hole; otherwise the same set."""
    rng = np.random.default_rng(nadd)
    n = 5000
    pd = rng.standard_normal((n, 6))
    keep = (rng.random(n) > 0.1).astype(np.uint16)
    add = rng.standard_normal((nadd, 6))
    fast = orc.add_and_remove_fast(pd, add, keep)
    ref, _ = orc.add_and_remove(pd, add, keep, n + nadd)   # roomy: the in-place rules
    h = n - int(keep.sum())
    if nadd == 0 or nadd >= h:
        assert np.array_equal(fast, ref)
    key = lambda a: a[np.lexsort(a.T[::-1])]
    assert np.array_equal(key(fast), key(ref))


def test_interaction_on_wall_and_analytic_normals(orc):
    gE = Grid(64, 48, 0.09, 0.07, 1)
    rng = np.random.default_rng(4)
    n = 500
    ang = rng.random(n) * 2 * np.pi
    inner = np.stack([0.0449 * np.cos(ang), 0.0349 * np.sin(ang)], 1)
    v = rng.standard_normal((n, 2)) * 1e6
    v[(v * inner).sum(1) < 0] *= -1  # moving outwards
    dt = 1e-10
    ap = np.zeros((n, 6))
    ap[:, :2] = inner + v * dt
    ap[:, 2:4] = v
    ins, _ = orc.index_inside(ap, 'ellip', 0.09, 0.07)
    ap = ap[ins == 0]
    moved, rts, nv, _ = orc.interaction_and_normals(ap, 'ellip', 0.09, 0.07, gE.dx, gE.dy)
    a = (0.09 - 1e-6 * gE.dx) / 2
    b = (0.07 - 1e-6 * gE.dy) / 2
    assert np.allclose(moved[:, 0] ** 2 / a ** 2 + moved[:, 1] ** 2 / b ** 2, 1., atol=1e-12)
    grad = -np.stack([moved[:, 0] / 0.045 ** 2, moved[:, 1] / 0.035 ** 2], 1)
    grad /= np.linalg.norm(grad, axis=1)[:, None]
    assert np.allclose(nv, grad, atol=1e-12)
    assert np.all((rts > 0) & (rts <= dt * (1 + 1e-9)))
