import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs a ROCm GPU (MI355X) and the built liboecl.so')


@pytest.fixture(scope='session')
def orc():
    import oracle
    oracle.build()
    return oracle


def random_particles(g, n, seed=0, vscale=4.2e5, margin=0.999, wall='rect'):
    """(n, 6) particles uniform inside the un-extended chamber of grid g."""
    rng = np.random.default_rng(seed)
    lx, ly = g.getLx(), g.getLy()
    pd = np.empty((n, 6))
    k = 0
    while k < n:
        x = (rng.random(4 * n) - 0.5) * lx * margin
        y = (rng.random(4 * n) - 0.5) * ly * margin
        if wall == 'ellip':
            ok = x * x * 4 / lx ** 2 + y * y * 4 / ly ** 2 < 1
        else:
            ok = np.ones_like(x, bool)
        x, y = x[ok], y[ok]
        m = min(n - k, x.size)
        pd[k:k + m, 0] = x[:m]
        pd[k:k + m, 1] = y[:m]
        k += m
    pd[:, 2:5] = rng.standard_normal((n, 3)) * vscale
    pd[:, 5] = 1e4 * (0.5 + rng.random(n))
    return pd
