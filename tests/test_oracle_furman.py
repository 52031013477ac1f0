"""CPU: the Furman-Pivi oracle pinned by the model's closed forms and conservation bounds."""
import math

import numpy as np
import pytest
from scipy import special as sps

from oracle.furman_oracle import (ELEMENTARY_CHARGE, M_E, generate_secondaries, material,
                                  yields)


@pytest.mark.parametrize('e0,th', [(5., 0.), (40., 0.3), (300., 0.), (300., 1.2), (1500., 0.7)])
def test_mean_yield_equals_sum_of_components(e0, th):
    """sum_n n P(n) = delta_e + delta_r + delta_ts (Furman & Pivi 2002) when no clamp acts."""
    P = material('copper', 1.6, eHat0=300.)
    dE, dR, dTS, pN = yields(e0, th, P)
    assert math.isclose(sum(pN), 1., rel_tol=1e-12)
    mean = sum(n * pN[n] for n in range(11))
    assert math.isclose(mean, dE + dR + dTS, rel_tol=1e-12)


def test_peak_true_secondary_yield_at_ehat0():
    P = material('copper', 1.6, eHat0=300.)
    _, _, dTS, _ = yields(300., 0., P)
    assert math.isclose(dTS, P['deltaTSHat'], rel_tol=1e-12)
    assert math.isclose(P['deltaTSHat'] + P['p1RInf'] + P['p1EInf'], 1.6, rel_tol=1e-12)


def _impacts(n, e, seed):
    rng = np.random.default_rng(seed)
    v = math.sqrt(2 * e * ELEMENTARY_CHARGE / M_E)
    ap = np.zeros((n, 6))
    ap[:, 0] = 0.05
    ap[:, 2] = -v
    ap[:, 5] = 1.
    nv = np.tile([-1., 0.], (n, 1))
    return ap, nv, rng.random((n, 35))


def test_sampled_yield_energy_bound_and_directions():
    P = material('copper', 1.6, eHat0=300.)
    n = 4000
    ap, nv, U = _impacts(n, 300., 1)
    sec, nsec, e0 = generate_secondaries(ap, nv, U, P)
    dE, dR, dTS, _ = yields(300., 0., P)
    mean = nsec.mean()
    sd = nsec.std() / math.sqrt(n)
    assert abs(mean - (dE + dR + dTS)) < 5 * sd
    eFactor = 0.5 * M_E / ELEMENTARY_CHARGE
    es = eFactor * (sec[:, 2:5] ** 2).sum(1)
    # emitted into the chamber: along the inward normal (-1, 0)
    assert np.all(sec[:, 2] * -1. >= -1e-9 * np.sqrt((sec[:, 2:5] ** 2).sum(1)))
    # true secondaries of one impact share at most its energy
    k = 0
    for j in range(n):
        if nsec[j] >= 2:
            assert es[k:k + nsec[j]].sum() <= e0[j] * (1 + 1e-9)
        k += nsec[j]


def test_inverse_special_functions_inverse_property():
    rng = np.random.default_rng(2)
    for a in (1.5, 3.5, 17.0, 57.5):
        p = rng.random(20)
        x = sps.gammaincinv(a, p)
        assert np.allclose(sps.gammainc(a, x), p, rtol=1e-12, atol=1e-14)
    for a, b in ((1.5, 1.5), (10.5, 3.5), (92., 11.5), (1., 3.)):
        p = rng.random(20)
        x = sps.betaincinv(a, b, p)
        assert np.allclose(sps.betainc(a, b, x), p, rtol=1e-12, atol=1e-14)


def test_binomial_pmf_matches_library():
    """The emitter's binomial probabilities (powers by repeated multiplication) against
    scipy.stats.binom.pmf, and summing to 1."""
    from scipy.stats import binom
    from oracle.furman_oracle import binom_pmf
    for p in (0., 1e-6, 0.013, 0.25, 0.5, 0.8, 0.999, 1.):
        got = np.array([binom_pmf(k, p, 10) for k in range(11)])
        ref = binom.pmf(np.arange(11), 10, p)
        assert np.allclose(got, ref, rtol=1e-13, atol=1e-300)
        assert abs(got.sum() - 1.) < 1e-14
