"""CPU: the photoelectron source's restatement pinned by geometry, kinematics and the energy
and angle laws of its definition (the reference has no photoelectron source)."""
import math

import numpy as np
import pytest

from oracle.photo_oracle import photo_emit

LX, LY = 0.09, 0.07
DX, DY = LX / 511., LY / 383.
ME = 9.10938291e-31
QE = 1.602176565e-19


@pytest.mark.parametrize('kind', [0, 1])
def test_points_on_inset_wall_with_inward_unit_normals(kind):
    out, nv, e = photo_emit(kind, LX, LY, DX, DY, 50000, 3, 5, 6, 7., 5., 2e4, ME, with_normals=True)
    a, b = 0.5 * LX - 1e-3 * DX, 0.5 * LY - 1e-3 * DY
    x, y = out[:, 0], out[:, 1]
    assert np.allclose(np.hypot(nv[:, 0], nv[:, 1]), 1., atol=1e-15)
    if kind == 1:
        assert np.abs((x / a) ** 2 + (y / b) ** 2 - 1.).max() < 1e-14
        # inward: the normal points against the position's outward gradient
        assert np.all(nv[:, 0] * x / a ** 2 + nv[:, 1] * y / b ** 2 < 0)
    else:
        on_x = np.isclose(np.abs(x), a, rtol=0, atol=1e-15)
        on_y = np.isclose(np.abs(y), b, rtol=0, atol=1e-15)
        assert np.all(on_x | on_y)
        assert np.all((np.abs(x) <= a + 1e-15) & (np.abs(y) <= b + 1e-15))
        assert np.all(nv[:, 0] * x + nv[:, 1] * y < 0)
        # uniform in arc length: each side carries its share of the perimeter
        frac_y = np.mean(on_y & ~on_x)
        assert abs(frac_y - a / (a + b)) < 0.01
    assert np.all(out[:, 5] == 2e4)


@pytest.mark.parametrize('kind', [0, 1])
def test_speed_energy_and_angle_law(kind):
    n = 200000
    e0, sig = 7., 5.
    out, nv, e = photo_emit(kind, LX, LY, DX, DY, n, 11, 1, 2, e0, sig, 1., ME, with_normals=True)
    v = out[:, 2:5]
    sp2 = (v ** 2).sum(1)
    assert np.allclose(sp2 * 0.5 * ME / QE, e, rtol=1e-12)
    # |E0 + sigma z|: folded normal mean and second moment
    mu = sig * math.sqrt(2 / math.pi) * math.exp(-e0 ** 2 / (2 * sig ** 2)) + e0 * math.erf(e0 / (sig * math.sqrt(2)))
    assert abs(e.mean() - mu) < 4 * e.std() / math.sqrt(n)
    assert abs((e ** 2).mean() - (e0 ** 2 + sig ** 2)) < 4 * (e ** 2).std() / math.sqrt(n)
    # angle to the inward normal: sin(theta) uniform on [0, 1) (theta = asin(u)); emission inward
    cosang = (v[:, 0] * nv[:, 0] + v[:, 1] * nv[:, 1]) / np.sqrt(sp2)
    s = np.sqrt(np.clip(1 - cosang ** 2, 0, 1))
    assert np.all(cosang > 0)
    hist = np.histogram(s, bins=10, range=(0, 1))[0] / n
    assert np.abs(hist - 0.1).max() < 0.005
    # azimuth about the normal is uniform: mean vz vanishes
    assert abs(v[:, 2].mean()) < 4 * v[:, 2].std() / math.sqrt(n)


def test_counter_streams_are_per_step():
    a = photo_emit(1, LX, LY, DX, DY, 100, 4, 1, 2, 7., 5., 1., ME)
    b = photo_emit(1, LX, LY, DX, DY, 100, 5, 1, 2, 7., 5., 1., ME)
    c = photo_emit(1, LX, LY, DX, DY, 60, 4, 1, 2, 7., 5., 1., ME)
    assert not np.allclose(a, b)
    assert np.array_equal(a[:60], c)
