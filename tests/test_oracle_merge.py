"""CPU: the oracle's restatement of localRandom's serial merge loop (particleManagement.pyx:
166-201) pinned against a brute-force loop that searches every remaining particle (no kd-tree)
on small sets, and by its invariants."""
import numpy as np
import pytest

from oracle.merge_oracle import greedy_merge


def brute(pd, nc, pw, disLimit, toMerge, U):
    X = pd[:nc, :5] * pw
    inTree = np.ones(nc, bool)
    keep = np.ones(nc, bool)
    w = pd[:nc, 5].copy()
    pairs = []
    ii = jj = 0
    while jj < toMerge and ii < nc - 1:
        if keep[ii] and inTree[ii]:
            inTree[ii] = False
            cand = np.nonzero(inTree)[0]
            if cand.size:
                d2 = ((X[cand] - X[ii]) ** 2).sum(1)
                k = int(np.argmin(d2))
                f = int(cand[k])
                if d2[k] < disLimit:
                    if U[jj] < w[ii] / (w[f] + w[ii]):
                        w[ii] += w[f]
                        keep[f] = False
                        pairs.append((ii, f, d2[k]))
                    else:
                        w[f] += w[ii]
                        keep[ii] = False
                        pairs.append((f, ii, d2[k]))
                    inTree[f] = False
                    jj += 1
        ii += 1
    return keep, w, pairs


@pytest.mark.parametrize('seed,lim', [(0, 1.), (1, 0.02), (2, 1e-4)])
def test_greedy_merge_matches_brute_force(seed, lim):
    rng = np.random.default_rng(seed)
    n = 300
    pd = np.empty((n, 6))
    pd[:, :2] = rng.random((n, 2)) * 0.1
    pd[:, 2:5] = rng.standard_normal((n, 3)) * 1e5
    pd[:, 5] = np.sort(rng.random(n) + 0.5)
    pw = np.array([1., 1., 1e-7, 1e-7, 1e-7])
    U = rng.random(n)
    a = greedy_merge(pd, n - 20, pw, lim, 120, U)
    b = brute(pd, n - 20, pw, lim, 120, U)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert [(x, y) for x, y, _ in a[2]] == [(x, y) for x, y, _ in b[2]]
    # every merge obeys the limit, the total weight is kept, a removed particle is gone
    assert all(d2 < lim for _, _, d2 in a[2])
    assert np.isclose(a[1][a[0]].sum(), pd[:n - 20, 5].sum(), rtol=1e-14)
    assert all(not a[0][r] for _, r, _ in a[2])
