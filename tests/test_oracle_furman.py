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


# ---- delta_e and delta_r at the points that define them (Furman & Pivi, PRST-AB 5, 124404
# (2002), eqs. 25-29: P1e(inf), P^1e at E^e, P1r(inf), E_r), independent of the restatement's
# arithmetic: each check is a value the model fixes by definition.

def _de_dr(e0, th, P):
    dE, dR, _, _ = yields(e0, th, P)
    return dE, dR


@pytest.mark.parametrize('name', ['copper', 'stainless'])
def test_elastic_yield_defining_points(name):
    P = material(name)
    # peak P^1e at E = E^e (both materials have E^e = 0), normal incidence
    dE, _ = _de_dr(P['eEHat'], 0., P)
    assert math.isclose(dE, P['p1Ehat'], rel_tol=1e-14)
    # P1e(inf) far above the peak (exp(-(E/W)^p / p) underflows to 0 at 1e30 eV)
    dE, _ = _de_dr(1e30, 0., P)
    assert math.isclose(dE, P['p1EInf'], rel_tol=1e-12)
    # one width above the peak the excess over P1e(inf) has fallen by exp(-1/p)
    dE, _ = _de_dr(P['eEHat'] + P['w'], 0., P)
    assert math.isclose(dE, P['p1EInf'] + (P['p1Ehat'] - P['p1EInf']) * math.exp(-1. / P['p']), rel_tol=1e-13)
    # decreasing in E away from the peak
    es = np.linspace(1., 2000., 200)
    d = [_de_dr(e, 0., P)[0] for e in es]
    assert np.all(np.diff(d) < 0)


@pytest.mark.parametrize('name', ['copper', 'stainless'])
def test_rediffused_yield_defining_points(name):
    P = material(name)
    # no rediffusion at zero energy, P1r(inf) at high energy (copper's r = 0.104 makes the
    # approach slow: exp(-(E/E_r)^r) is 1e-2 at 1e5 eV and underflows only near 1e30 eV)
    assert _de_dr(0., 0., P)[1] == 0.
    assert math.isclose(_de_dr(1e30, 0., P)[1], P['p1RInf'], rel_tol=1e-12)
    # at E = E_r the yield is (1 - 1/e) of its limit, for any exponent r
    assert math.isclose(_de_dr(P['eR'], 0., P)[1], P['p1RInf'] * (1. - math.exp(-1.)), rel_tol=1e-13)
    es = np.linspace(0.01, 2000., 200)
    d = [_de_dr(e, 0., P)[1] for e in es]
    assert np.all(np.diff(d) >= 0) and d[1] > d[0]   # rises, then saturates (stainless: to rounding)


@pytest.mark.parametrize('th', [0.2, 0.7, 1.2])
def test_angular_factors(th):
    """delta(theta) / delta(0) = 1 + e1 (1 - cos^e2 theta) for the elastic part and 1 + r1 (1 -
    cos^r2 theta) for the rediffused part (Furman & Pivi eqs. 47-48); the true-secondary
    peak energy moves to E^(theta) = E^0 (1 + t3 (1 - cos^t4 theta)) and its peak value scales
    by 1 + t1 (1 - cos^t2 theta)."""
    P = material('copper', 1.6, eHat0=300.)
    for e in (20., 300.):
        e0, r0 = _de_dr(e, 0., P)
        e1, r1 = _de_dr(e, th, P)
        assert math.isclose(e1 / e0, 1 + P['e1'] * (1 - math.cos(th) ** P['e2']), rel_tol=1e-13)
        assert math.isclose(r1 / r0, 1 + P['r1'] * (1 - math.cos(th) ** P['r2']), rel_tol=1e-13)
    eh = P['eHat0'] * (1 + P['t3'] * (1 - math.cos(th) ** P['t4']))
    pk = yields(eh, th, P)[2]
    assert math.isclose(pk, P['deltaTSHat'] * (1 + P['t1'] * (1 - math.cos(th) ** P['t2'])), rel_tol=1e-13)
    for h in (1e-3, 1.):
        assert yields(eh + h, th, P)[2] < pk and yields(eh - h, th, P)[2] < pk


def test_elastic_rediffused_clamp():
    """delta_e + delta_r is scaled down to deltaERmax = 0.99 when it reaches it
    (particleEmitter.pyx:199-206); the mean number of secondaries is then 0.99 + delta_ts."""
    P = material('copper', 1.6, reflec=1.0)
    dE, dR, dTS, pN = yields(1e-3, 0., P)
    assert math.isclose(dE + dR, 0.99, rel_tol=1e-14)
    # the clamp scales both parts alike: the ratio of the unclamped yields is kept
    uE, uR, _, _ = yields(1e-3, 0., dict(P, deltaERmax=10.))
    assert uE + uR > 0.99
    assert math.isclose(dE / dR, uE / uR, rel_tol=1e-13)
    assert math.isclose(sum(n * pN[n] for n in range(11)), 0.99 + dTS, rel_tol=1e-12)


# ---- the true-secondary energy law (Furman & Pivi sec. III.D; particleEmitter.pyx:262-278):
# for n secondaries the scaled total energy S = sum E_k / eps_n is gamma(n p_n) distributed,
# truncated at E0 / eps_n, and the shares E_k / sum E_k are Dirichlet(p_n, ..., p_n), so each
# share is Beta(p_n, (n - 1) p_n).

def _true_secondaries(n_imp, e, seed):
    P = material('copper', 1.6, eHat0=300.)
    ap, nv, U = _impacts(n_imp, e, seed)
    sec, nsec, e0 = generate_secondaries(ap, nv, U, P)
    eFactor = 0.5 * M_E / ELEMENTARY_CHARGE
    es = eFactor * (sec[:, 2:5] ** 2).sum(1)
    start = np.concatenate([[0], np.cumsum(nsec)])
    return P, nsec, start, es


@pytest.mark.parametrize('ns', [2, 3, 4])
def test_true_secondary_total_energy_is_truncated_gamma(ns):
    from scipy import stats
    e = 150.
    P, nsec, start, es = _true_secondaries(30000, e, 10 + ns)
    sel = np.nonzero(nsec == ns)[0]
    assert sel.size > 300
    en, pn = P['epsN'][ns - 1], P['pSmallN'][ns - 1]
    S = np.array([es[start[j]:start[j] + ns].sum() for j in sel]) / en
    cut = e / en
    assert np.all(S <= cut * (1 + 1e-9))
    cdf = lambda s: sps.gammainc(ns * pn, np.minimum(s, cut)) / sps.gammainc(ns * pn, cut)
    assert stats.kstest(S, cdf).pvalue > 1e-3
    # a wrong shape (n p_n off by one p_n) is rejected by the same test
    cdf_wrong = lambda s: sps.gammainc((ns + 1) * pn, np.minimum(s, cut)) / sps.gammainc((ns + 1) * pn, cut)
    assert stats.kstest(S, cdf_wrong).pvalue < 1e-6


@pytest.mark.parametrize('ns', [2, 3])
def test_true_secondary_energy_shares_are_dirichlet(ns):
    from scipy import stats
    P, nsec, start, es = _true_secondaries(30000, 60., 20 + ns)
    sel = np.nonzero(nsec == ns)[0]
    pn = P['pSmallN'][ns - 1]
    tot = np.array([es[start[j]:start[j] + ns].sum() for j in sel])
    for k in range(ns):
        share = np.array([es[start[j] + k] for j in sel]) / tot
        assert stats.kstest(share, stats.beta(pn, (ns - 1) * pn).cdf).pvalue > 1e-3
    # the shares of one impact add up to one
    assert np.allclose(tot, np.array([es[start[j]:start[j] + ns].sum() for j in sel]))
