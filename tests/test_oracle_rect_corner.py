"""Rectangular wall corner exits (particleBoundary.pyx:108-221): the oracle's labels with and
without the corner fix.  Pinned by geometry: with the fix, the moved point lies on the wall
its normal names, for every exit; the reference's labels break that in exactly the four corner
cases it gives the other wall's normal."""
import numpy as np

import oracle as orc

LX, LY, DX, DY = 0.1, 0.06, 0.1 / 63, 0.06 / 47   # Grid(64, 48, LX, LY): dx = lx / (nx - 1)


def _corner_exits(n, seed):
    """Particles that crossed both half-lines within one step, at each of the four corners."""
    rng = np.random.default_rng(seed)
    sx = rng.choice([-1., 1.], n)
    sy = rng.choice([-1., 1.], n)
    ang = rng.random(n) * 0.5 * np.pi * 0.9 + 0.05 * np.pi * 0.5   # away from the axes
    v = 1e7 * (0.5 + rng.random(n))
    ap = np.zeros((n, 6))
    ap[:, 2] = sx * v * np.cos(ang)
    ap[:, 3] = sy * v * np.sin(ang)
    dt = 1e-11
    # a start point inside, close to the corner, then one step of free flight
    x0 = sx * (0.5 * LX - rng.random(n) * 0.5 * np.abs(ap[:, 2]) * dt)
    y0 = sy * (0.5 * LY - rng.random(n) * 0.5 * np.abs(ap[:, 3]) * dt)
    ap[:, 0] = x0 + ap[:, 2] * dt
    ap[:, 1] = y0 + ap[:, 3] * dt
    ap[:, 5] = 1.
    return ap


def _on_named_wall(moved, nv):
    lxh, lyh = 0.5 * LX - 1e-6 * DX, 0.5 * LY - 1e-6 * DY
    tol = 1e-12
    wx = np.where(nv[:, 0] > 0, -lxh, lxh)
    wy = np.where(nv[:, 1] > 0, -lyh, lyh)
    onx = (nv[:, 0] != 0) & (np.abs(moved[:, 0] - wx) < tol * LX) & (np.abs(moved[:, 1]) <= lyh * (1 + 1e-9))
    ony = (nv[:, 1] != 0) & (np.abs(moved[:, 1] - wy) < tol * LY) & (np.abs(moved[:, 0]) <= lxh * (1 + 1e-9))
    return onx | ony


def test_corner_fix_moves_to_the_wall_its_normal_names():
    ap = _corner_exits(20000, 3)
    lxh, lyh = 0.5 * LX - 1e-6 * DX, 0.5 * LY - 1e-6 * DY
    assert np.all(np.abs(ap[:, 0]) > lxh) and np.all(np.abs(ap[:, 1]) > lyh)
    m0, r0, n0, _ = orc.interaction_and_normals(ap, 'rect', LX, LY, DX, DY)
    m1, r1, n1, _ = orc.interaction_and_normals(ap, 'rect', LX, LY, DX, DY, cornerFix=True)
    # the flag changes the normal only
    assert np.array_equal(m0, m1) and np.array_equal(r0, r1)
    assert np.all(_on_named_wall(m1, n1))
    # the reference's labels: x wall crossed first at the lower-left and upper-left corners
    # keeps its normal; right-side x exits and lower y exits get the other wall's normal
    bad = ~_on_named_wall(m0, n0)
    xFirst = n1[:, 0] != 0
    left, low = ap[:, 0] < 0, ap[:, 1] < 0
    expect_bad = (xFirst & ~left) | (~xFirst & low)
    assert np.array_equal(bad, expect_bad)
    assert 0.2 < bad.mean() < 0.8


def test_corner_fix_leaves_side_exits_alone():
    rng = np.random.default_rng(5)
    n = 5000
    ap = np.zeros((n, 6))
    ap[:, 0] = rng.uniform(-0.4 * LX, 0.4 * LX, n)
    ap[:, 1] = np.where(rng.random(n) < 0.5, -1, 1) * (0.5 * LY + 1e-5)
    ap[:, 2] = rng.standard_normal(n) * 1e6
    ap[:, 3] = np.sign(ap[:, 1]) * 1e7
    m0, r0, n0, _ = orc.interaction_and_normals(ap, 'rect', LX, LY, DX, DY)
    m1, r1, n1, _ = orc.interaction_and_normals(ap, 'rect', LX, LY, DX, DY, cornerFix=True)
    assert np.array_equal(n0, n1) and np.array_equal(m0, m1)
    assert np.all(_on_named_wall(m0, n0))
