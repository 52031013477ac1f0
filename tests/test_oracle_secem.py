"""The simpler emitters' oracle (SecElecEmitter, SimpleSecEmitter) pinned by what the models fix:
yields at their defining points, Bernoulli and floor-plus-Bernoulli counts, the rediffused
energy law, log-normal true energies, the fallback range, the cosine law about the normal."""
import numpy as np
import pytest
from scipy import stats

from oracle import secem_oracle as so


INC = 0.2   # incidence angle of the test impacts [rad]; at exactly 0 rounding can make the
            # reference's acos argument < -1, whose NaN its clip turns into theta0max


def _normal_impacts(n, e0, rng):
    """n impacts at energy e0 [eV] and incidence INC on walls with random inward normals."""
    ang = rng.random(n) * 2 * np.pi
    nrm = np.stack([np.cos(ang), np.sin(ang)], axis=1)
    v = np.sqrt(2 * e0 * so.ELEMENTARY_CHARGE / so.M_E)
    ap = np.zeros((n, 6))
    ap[:, 0] = rng.random(n)
    ap[:, 1] = rng.random(n)
    ap[:, 2] = -v * (np.cos(INC) * nrm[:, 0] - np.sin(INC) * nrm[:, 1])
    ap[:, 3] = -v * (np.cos(INC) * nrm[:, 1] + np.sin(INC) * nrm[:, 0])
    ap[:, 5] = 1 + rng.random(n)
    return ap, nrm


def test_stainless_yields_at_their_defining_points():
    P = so.stainless()
    dE, dR, dTS = so.stainless_yields(np.array([0., 1e7]), 0., P)
    assert dE[0] == P['p1Ehat'] and dR[0] == 0. and dTS[0] == 0.
    assert abs(dR[1] - P['p1RInf']) < 1e-12
    assert abs(dE[1] - P['p1EInf']) < 1e-12
    # the true yield peaks at eHat0 with the value deltaTSHat (normal incidence)
    e = np.linspace(1., 2000., 200001)
    _, _, ts = so.stainless_yields(e, 0., P)
    assert abs(e[np.argmax(ts)] - P['eHat0']) < 0.05
    assert ts.max() <= P['deltaTSHat'] * (1 + 1e-15)
    assert abs(so.stainless_yields(np.array([P['eHat0']]), 0., P)[2][0] - P['deltaTSHat']) < 1e-15
    # seyMax is split into the asymptotic elastic and rediffused yields plus deltaTSHat
    for sey in (1.2, 1.7, 2.4):
        Q = so.stainless(seyMax=sey)
        assert abs(Q['p1EInf'] + Q['p1RInf'] + Q['deltaTSHat'] - sey) < 1e-15
    assert so.stainless(seyMax=0.6)['deltaTSHat'] == 1e-6
    assert so.stainless(reflec=0.3)['p1Ehat'] == 0.3


def test_stainless_deltaERmax_rescale_keeps_the_ratio():
    P = so.stainless()
    Pinf = dict(P, deltaERmax=np.inf)
    th = P['theta0max']
    e = np.array([500., 2000., 1e4])
    dE, dR, _ = so.stainless_yields(e, th, P)
    dE0, dR0, _ = so.stainless_yields(e, th, Pinf)
    assert np.all(dE0 + dR0 >= 0.99)
    assert np.allclose(dE + dR, 0.99, rtol=0, atol=1e-15)
    assert np.allclose(dE / dR, dE0 / dR0, rtol=1e-14)
    # below the limit nothing changes
    dE, dR, _ = so.stainless_yields(np.array([50.]), 0., P)
    dE0, dR0, _ = so.stainless_yields(np.array([50.]), 0., Pinf)
    assert dE[0] == dE0[0] and dR[0] == dR0[0]


@pytest.mark.parametrize('model', ['stainless', 'simple'])
@pytest.mark.parametrize('e0', [40., 310., 900.])
def test_mean_counts_equal_the_yields(model, e0):
    rng = np.random.default_rng(int(e0))
    n = 200000
    ap, nv = _normal_impacts(n, e0, rng)
    U = rng.random((n, so.N_UNIFORMS))
    P = so.stainless(seyMax=2.2)
    nE, nR, nTS, e = so.counts(ap, nv, U, model, P)
    assert np.allclose(e, e0, rtol=1e-12)
    if model == 'simple':
        dE, dR, dTS = so.simple_yields(np.array([e0]))
    else:
        dE, dR, dTS = so.stainless_yields(np.array([e0]), INC, P)
    for cnt, d in ((nE, dE[0]), (nR, dR[0]), (nTS, dTS[0])):
        f = d - np.floor(d)
        sd = np.sqrt(max(f * (1 - f), 1e-4) / n)
        assert abs(cnt.mean() - d) < 5 * sd, (cnt.mean(), d)
    assert set(np.unique(nTS)) <= {int(np.floor(dTS[0])), int(np.floor(dTS[0])) + 1}
    assert set(np.unique(nE)) <= {0, 1} and set(np.unique(nR)) <= {0, 1}


def test_normal_incidence_nan_angle_takes_theta0max():
    """acos of a rounded -1 - eps is NaN; clip (particleEmitter.pyx:33-39) makes it theta0max."""
    P = so.stainless()
    ap = np.zeros((1, 6))
    nv = np.array([[-0.6520162635843662, -0.7582049802141122]])   # (cos, sin) of 4.002148315014479
    ap[0, 2:4] = -3e6 * nv[0]
    U = np.full((1, so.N_UNIFORMS), 0.5)
    v0, th = so._theta0(ap, nv)
    assert np.isnan(th[0])
    e0 = 0.5 * so.M_E * np.sum(ap[0, 2:5] ** 2) / so.ELEMENTARY_CHARGE
    dE, _, _ = so.stainless_yields(np.array([e0]), P['theta0max'], P)
    dE0, _, _ = so.stainless_yields(np.array([e0]), 0., P)
    U[0, 0] = 0.5 * (dE[0] + dE0[0])       # between the two: elastic only at theta0max
    nE, _, _, _ = so.counts(ap, nv, U, 'stainless', P)
    assert dE[0] > dE0[0] and nE[0] == 1


def test_rediffused_energy_law_and_elastic_mirror():
    rng = np.random.default_rng(7)
    n = 60000
    P = so.stainless()
    ap, nv = _normal_impacts(n, 200., rng)
    U = rng.random((n, so.N_UNIFORMS))
    U[:, 0] = 0.            # every impact emits an elastic one
    U[:, 1] = 0.            # and a rediffused one
    U[:, 2] = 1.            # floor(deltaTS) true ones (1 at 200 eV)
    sec, ns = so.generate_secondaries(ap, nv, U, 'stainless', P)
    assert np.all(ns == 3)
    el, rd = sec[0::3], sec[1::3]
    assert np.array_equal(el[:, 2:5], -ap[:, 2:5]) and np.array_equal(el[:, [0, 1, 5]], ap[:, [0, 1, 5]])
    ratio = np.sum(rd[:, 2:5] ** 2, 1) / np.sum(ap[:, 2:5] ** 2, 1)
    ks = stats.kstest(ratio, lambda x: np.clip(x, 0, 1) ** (1 + P['q']))
    assert ks.pvalue > 1e-3
    # rediffused ones leave along the reversed incident direction
    cosang = -np.sum(rd[:, 2:5] * ap[:, 2:5], 1) / np.sqrt(np.sum(rd[:, 2:5] ** 2, 1) * np.sum(ap[:, 2:5] ** 2, 1))
    assert np.allclose(cosang, 1., atol=1e-12)
    sd, nd = so.generate_secondaries(ap, nv, U, 'simple')
    ratio = np.sum(sd[1::3, 2:5] ** 2, 1) / np.sum(ap[:, 2:5] ** 2, 1)
    assert stats.kstest(ratio, 'uniform').pvalue > 1e-3


def test_true_secondaries_log_normal_and_cosine_law():
    rng = np.random.default_rng(11)
    n = 40000
    P = so.stainless(seyMax=3.0)
    ap, nv = _normal_impacts(n, 1000., rng)   # deltaTS about 1.4; sums above e0 ~ 1e-5
    U = rng.random((n, so.N_UNIFORMS))
    U[:, 0] = 1.
    U[:, 1] = 1.            # true secondaries only
    sec, ns = so.generate_secondaries(ap, nv, U, 'stainless', P)
    assert np.all(ns >= 1)
    rep = np.repeat(np.arange(n), ns)
    k = np.arange(sec.shape[0]) - np.repeat(np.cumsum(ns) - ns, ns)   # index within the impact
    v2 = np.sum(sec[:, 2:5] ** 2, 1)
    e = 0.5 * so.M_E * v2 / so.ELEMENTARY_CHARGE
    z = (np.log(e) - P['muETS']) / P['sigmaETS']
    assert stats.kstest(z, 'norm').pvalue > 1e-3
    # angle to the inward normal: cos = sqrt(1 - u^2) with u the theta draw
    cosang = (sec[:, 2] * nv[rep, 0] + sec[:, 3] * nv[rep, 1]) / np.sqrt(v2)
    u = U[rep, 28 + 2 * k]
    assert np.allclose(cosang, np.sqrt(1 - u * u), atol=1e-12)


def test_true_secondaries_fallback_below_their_log_normal_sum():
    rng = np.random.default_rng(13)
    n = 30000
    P = so.stainless(seyMax=2.2)
    ap, nv = _normal_impacts(n, 0.5, rng)    # log-normal energies sum above e0 but ~1e-3 of the time
    U = rng.random((n, so.N_UNIFORMS))
    U[:, 0] = 1.
    U[:, 1] = 1.
    U[:, 2] = 0.                              # floor + 1 true secondaries
    sec, ns = so.generate_secondaries(ap, nv, U, 'stainless', P)
    assert np.all(ns >= 1)
    rep = np.repeat(np.arange(n), ns)
    e = 0.5 * so.M_E * np.sum(sec[:, 2:5] ** 2, 1) / so.ELEMENTARY_CHARGE
    frac = e * ns[rep] / 0.5
    assert frac.max() < 1. + 1e-12
    assert stats.kstest(frac, 'uniform').pvalue > 1e-3


def test_simple_true_secondaries_uniform_energy_reversed():
    rng = np.random.default_rng(17)
    n = 50000
    ap, nv = _normal_impacts(n, 300., rng)
    U = rng.random((n, so.N_UNIFORMS))
    U[:, 0] = 1.
    U[:, 1] = 1.
    sec, ns = so.generate_secondaries(ap, nv, U, 'simple')
    rep = np.repeat(np.arange(n), ns)
    e = 0.5 * so.M_E * np.sum(sec[:, 2:5] ** 2, 1) / so.ELEMENTARY_CHARGE
    assert stats.kstest(e / 5., 'uniform').pvalue > 1e-3
    a = ap[rep, 2:5]
    cosang = -np.sum(sec[:, 2:5] * a, 1) / np.sqrt(np.sum(sec[:, 2:5] ** 2, 1) * np.sum(a * a, 1))
    assert np.allclose(cosang, 1., atol=1e-12)
