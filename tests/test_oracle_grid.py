"""CPU: the grid oracle against the reference's own golden vectors, closed forms, and the
product's vectorised Grid against the oracle (bit for bit)."""
import os

import numpy as np
import pytest
import scipy.sparse as spsp

from oracle.grid_oracle import OracleGrid
from openecloud_amd.grid import Grid

GOLD = os.path.join(os.path.dirname(__file__), 'golden', 'reference_grid_vectors.npz')
ARRAYS = ['insidePoints', 'insideEdges', 'insideFaces', 'ds', 'dsi', 'da', 'dai', 'dst', 'dsti',
          'dat', 'dati', 'boundaryPoints', 'cutCellPointsInd', 'cutCellNormalVectors',
          'cutCellCenter']


def _gold_cases():
    z = np.load(GOLD)
    return z, sorted(set(k.split('/')[0] for k in z.files))


@pytest.mark.parametrize('case', _gold_cases()[1])
def test_oracle_grid_matches_reference_vectors(case):
    z, _ = _gold_cases()
    a = z[case + '/args']
    g = OracleGrid(int(a[0]), int(a[1]), a[2], a[3], int(a[4]))
    n = 0
    for f in z.files:
        if not f.startswith(case + '/') or f.split('/')[1] in ('args', 'e2n_row', 'e2n_col', 'e2n_val'):
            continue
        k = f.split('/')[1]
        attr = k[0].lower() + k[1:]
        assert np.array_equal(np.asarray(getattr(g, attr)), z[f]), (case, k)
        n += 1
    assert n >= 8
    npx = g.npExt
    R = spsp.coo_matrix((z[case + '/e2n_val'], (z[case + '/e2n_row'], z[case + '/e2n_col'])),
                        shape=(2 * npx, 2 * npx)).tocsr()
    D = R - g.edgeToNode.tocsr()
    assert D.nnz == 0 or abs(D).max() == 0.


def _circ(x, y):
    return 1 if x * x + y * y < 0.015 ** 2 else 0


def _cases():
    return [((16, 12, 0.1, 0.05, 0), {}), ((20, 17, 0.09, 0.07, 1), {}),
            ((41, 33, 0.09, 0.07, 1), {}), ((30, 30, 0.04, 0.04, 3), {'boundFunc': _circ}),
            ((30, 26, 0.04, 0.04, 2), {'boundFunc': _circ}), ((26, 26, 0.04, 0.04, 1), {})]


@pytest.mark.parametrize('args,kw', _cases())
def test_vectorised_grid_equals_oracle(args, kw):
    a = Grid(*args, **kw)
    b = OracleGrid(*args, **kw)
    for k in ARRAYS:
        x, y = np.asarray(getattr(a, k)), np.asarray(getattr(b, k))
        assert x.shape == y.shape and np.array_equal(x, y), k
    D = (a.edgeToNode - b.edgeToNode).tocsr()
    assert D.nnz == 0 or abs(D).max() == 0.


def test_rectangular_counts_closed_form():
    g = Grid(128, 64, 0.1, 0.05, 0)
    assert (g.nxExt, g.nyExt, g.npExt) == (130, 66, 8580)
    assert int(g.insidePoints.sum()) == 126 * 62 == 7812
    assert g.dx == 0.1 / 127 and g.dy == 0.05 / 63


def test_ellipse_boundary_points_on_wall():
    g = Grid(96, 72, 0.09, 0.07, 1)
    bp = g.boundaryPoints
    r = bp[:, 0] ** 2 * 4 / 0.09 ** 2 + bp[:, 1] ** 2 * 4 / 0.07 ** 2
    # bisection to cutCellAcc = 1e-12 of the mesh coordinates (41 halvings of a cell)
    assert np.all(np.abs(r - 1) < 1e-9)
    ds = g.ds
    assert np.all(ds[g.insideEdges == 1] > 0)
    assert np.all(ds[:g.npExt][g.insideEdges[:g.npExt] == 1] <= g.dx * (1 + 1e-12))
    nv = g.cutCellNormalVectors
    cut = (g.insideFaces != 0) & (g.insideFaces != 8)
    assert np.allclose(np.hypot(nv[cut, 0], nv[cut, 1]), 1.)
