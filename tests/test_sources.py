"""CPU: beam and multipole field sources against closed forms."""
import math

import numpy as np
import pytest

from openecloud_amd.grid import Grid


def test_quadrupole_field_is_linear():
    from openecloud_amd.magneticField import multipoleExpansion
    g = Grid(64, 48, 0.09, 0.07, 1)
    b = multipoleExpansion(g, [0., 50.])
    X, Y = np.meshgrid(g.getXMesh(), g.getYMesh())
    npx = g.getNpExt()
    # normal quadrupole b_1 = 50 T/m: Bx = g y, By = g x
    assert np.allclose(b[:npx], 50. * Y.ravel(), atol=1e-12)
    assert np.allclose(b[npx:2 * npx], 50. * X.ravel(), atol=1e-12)
    assert np.all(b[2 * npx:] == 0.)


def test_dipole_and_skew_terms():
    from openecloud_amd.magneticField import multipoleExpansion
    g = Grid(16, 16, 0.04, 0.04, 1)
    npx = g.getNpExt()
    b = multipoleExpansion(g, [0.5])
    assert np.allclose(b[:npx], 0., atol=1e-15) and np.allclose(b[npx:2 * npx], 0.5)
    b = multipoleExpansion(g, [0.], [0.3])            # skew dipole: Bx = -a0... -> horizontal
    assert np.allclose(np.hypot(b[:npx], b[npx:2 * npx]), 0.3)


@pytest.mark.parametrize('cls', ['LHCBeam', 'LHCBeamEndless', 'GaussianBunchTrain'])
def test_beam_profile_normalised_symmetric(cls):
    from openecloud_amd import beam
    g = Grid(128, 96, 0.09, 0.07, 1)
    b = getattr(beam, cls)(g, 5e-12)
    prof = b.getTransversalProfile().reshape(g.getNyExt(), g.getNxExt())
    assert math.isclose(prof.sum(), 1., rel_tol=1e-12)
    assert np.allclose(prof, prof[::-1, ::-1], atol=1e-15)
    assert math.isclose(prof.max(), prof[g.getNyExt() // 2, g.getNxExt() // 2], rel_tol=1e-12)


def test_beam_line_charge_per_bunch():
    """Sum over one bunch passage of scale * dt * c equals N * e (the longitudinal CDF
    differences telescope)."""
    from openecloud_amd import beam
    from openecloud_amd.constants import c, elementary_charge
    g = Grid(32, 32, 0.09, 0.07, 1)
    dt = 5e-12
    b = beam.GaussianBunchTrain(g, dt, nParticles=1.15e11, nBunches=1)
    tot = sum(b.getScale(k * dt) for k in range(int(25e-9 / dt)))
    assert math.isclose(tot * dt * c, 1.15e11 * elementary_charge, rel_tol=1e-6)
    assert b.getScale(25e-9 + 3 * dt) == 0.   # one bunch only


def test_b_layout_interleaves_a_transverse_grid_field():
    """The step's (Bx, By) layout of a grid field: Bx and By of each node side by side; a field
    with any nonzero Bz keeps the separate arrays (None)."""
    import torch
    from openecloud_amd.buildup import b_layout
    npx = 7
    b = torch.arange(3 * npx, dtype=torch.float64) + 1.
    assert b_layout(b, npx) is None
    b[2 * npx:] = 0.
    bl = b_layout(b, npx)
    assert bl.shape == (npx, 2) and bl.is_contiguous()
    assert torch.equal(bl[:, 0], b[:npx]) and torch.equal(bl[:, 1], b[npx:2 * npx])
    assert b_layout(None, npx) is None
