"""CPU: the oracle's cut-cell wall search (particleBoundary.pyx:373-457, orc_interaction_cutcell)
pinned by geometry on a circular chamber, and by a brute-force search over every chord of the
cut-cell boundary.

A particle that crossed the wall within one cell is moved back along its last step, from the
new position p towards the old one p + r (r = -dt v), to the crossing with the chord of one of
the four cells around it, and 1e-6 of the step beyond it; remainingTimeStep = t (1 + 1e-6) dt.
So the moved point must lie on the segment from p to p + r, on the cut-cell boundary polyline
(which lies within (dx^2 + dy^2) / (8 R) of the circle), and coincide with the crossing that a search
over all chords finds."""
import numpy as np
import pytest

from openecloud_amd.grid import Grid

R = 0.018
SE = 1e-6


@pytest.fixture(scope='module')
def grid():
    return Grid(72, 72, 0.04, 0.04, 3, boundFunc=lambda x, y: 1 if x * x + y * y < R * R else 0)


def _crossings(g, n, seed, reach):
    """n particles that start inside, within `reach` cells of the wall, and end outside after
    one step of dt = 1e-11 s with at most a cell's travel."""
    import oracle
    rng = np.random.default_rng(seed)
    dt = 1e-11
    dx = g.dx
    out = []
    while sum(len(o) for o in out) < n:
        m = 4 * n
        phi = rng.random(m) * 2 * np.pi
        r0 = R - rng.uniform(0.02, reach) * dx
        old = np.stack([r0 * np.cos(phi), r0 * np.sin(phi)], 1)
        ang = phi + rng.uniform(-1.0, 1.0, m)
        step = rng.uniform(0.1, 0.7) * dx
        d = np.stack([np.cos(ang), np.sin(ang)], 1) * step
        new = old + d
        pd = np.zeros((m, 6))
        pd[:, :2] = new
        pd[:, 2:4] = d / dt
        pd[:, 4] = 1e5
        pd[:, 5] = 1.
        ins_old, _ = oracle.index_inside(np.c_[old, np.zeros((m, 4))], 'cutcell', g.getLx(), g.getLy(), grid=g)
        ins_new, _ = oracle.index_inside(pd, 'cutcell', g.getLx(), g.getLy(), grid=g)
        ok = (ins_old == 1) & (ins_new == 0)
        out.append(pd[ok])
    return np.concatenate(out)[:n], dt


def _chords(g):
    bp = np.asarray(g.boundaryPoints, np.float64)
    cpi = np.asarray(g.cutCellPointsInd, np.int64).reshape(-1, 2)
    f = np.asarray(g.insideFaces)
    cells = np.nonzero((f != 0) & (f != 8))[0]
    a = bp[cpi[cells, 0]]
    b = bp[cpi[cells, 1]]
    return a, b


def test_cut_points_lie_on_the_circle(grid):
    a, b = _chords(grid)
    for q in (a, b):
        assert np.abs(np.hypot(q[:, 0], q[:, 1]) - R).max() < 1e-12 * R + 2 ** -41 * grid.dx


@pytest.mark.parametrize('seed', [0, 1, 2])
def test_cutcell_interaction_point_by_geometry(grid, seed):
    import oracle
    g = grid
    pd, dt = _crossings(g, 400, seed, 0.6)
    moved, rts, nv, nf = oracle.interaction_and_normals(pd, 'cutcell', g.getLx(), g.getLy(), g.dx, g.dy,
                                                        grid=g, dt=dt)
    assert nf == 0
    p = pd[:, :2]
    rvec = -dt * pd[:, 2:4]
    # on the segment from p towards p + r, at the fraction remainingTime / dt
    lam = rts / dt
    assert np.all(lam >= -SE * (1 + SE) - 1e-15) and np.all(lam <= (1 + SE) ** 2 + 1e-15)
    exp = p + lam[:, None] * rvec
    assert np.abs(moved[:, :2] - exp).max() <= 1e-15 + 4e-16 * np.abs(p).max()
    # the step's crossing with the cut-cell boundary, found by a search over every chord
    a, b = _chords(g)
    s = b - a
    for k in range(len(pd)):
        rs = rvec[k, 0] * s[:, 1] - rvec[k, 1] * s[:, 0]
        q = a - p[k]
        with np.errstate(divide='ignore', invalid='ignore'):
            t = (q[:, 0] * s[:, 1] - q[:, 1] * s[:, 0]) / rs
            u = (q[:, 0] * rvec[k, 1] - q[:, 1] * rvec[k, 0]) / rs
        hit = (rs != 0) & (t >= -SE) & (t <= 1 + SE) & (u >= -SE) & (u <= 1 + SE)
        assert hit.any()
        tk = lam[k] / (1 + SE)
        assert np.abs(t[hit] - tk).min() < 1e-9
    # the moved point is 1e-6 of the step beyond the polyline; a chord spans at most a cell's
    # diagonal L, so the polyline is within the sagitta L^2 / (8 R) of the circle
    rad = np.hypot(moved[:, 0], moved[:, 1])
    tol = (g.dx ** 2 + g.dy ** 2) / (8 * R) * (1 + 1e-3) + SE * np.hypot(rvec[:, 0], rvec[:, 1]) + 1e-15
    assert np.all(np.abs(rad - R) <= tol)
    # the cell's normal points into the chamber, close to -x/|x| of the circle
    inward = -moved[:, :2] / rad[:, None]
    assert np.all((nv * inward).sum(1) > np.cos(0.5))
    assert np.allclose(np.hypot(nv[:, 0], nv[:, 1]), 1., atol=1e-12)


def test_a_planted_sign_error_is_caught(grid):
    """The same geometric checks reject a search that steps forward instead of back (r = +dt v)."""
    import oracle
    g = grid
    pd, dt = _crossings(g, 200, 5, 0.6)
    flipped = pd.copy()
    flipped[:, 2:4] *= -1
    moved, rts, nv, nf = oracle.interaction_and_normals(flipped, 'cutcell', g.getLx(), g.getLy(), g.dx, g.dy,
                                                        grid=g, dt=dt)
    rad = np.hypot(moved[:, 0], moved[:, 1])
    found = rts > 0
    bad = (~found) | (np.abs(rad - R) > (g.dx ** 2 + g.dy ** 2) / (8 * R) + 1e-9)
    assert nf > 0 or bad.mean() > 0.5
