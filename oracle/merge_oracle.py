"""CPU ORACLE -- test infrastructure only (never imported by the product).

Restatement of localRandom's serial merge loop (particleManagement.pyx:166-201, mergeScheme 0)
over particles already sorted by weight: each particle, lightest first, that is still in the
kd-tree is removed from it and merged with its nearest remaining neighbour in the weighted
5-D space (x, y, lambda vx, lambda vy, lambda vz) when the SQUARED distance is below
factorDisLimit * ds (:185); the survivor is the particle itself with probability
w_ii / (w_found + w_ii) and takes the sum of the weights; the neighbour leaves the tree.
scipy.spatial.cKDTree stands in for the reference's kd-tree (kdTree.pyx, "heavily inspired by
the scipy cKDTree implementation"), with removal emulated by skipping removed points.
Uniforms U[j] for the survivor draw of merge j are injected.
"""
import numpy as np


def greedy_merge(pd, nc, pw, disLimit, toMerge, U):
    """pd: (n, 6) particles sorted by weight; merges among [0, nc).  Returns (keep[nc] bool,
    weights[nc], pairs: list of (survivor, removed, squared distance))."""
    from scipy.spatial import cKDTree
    X = np.asarray(pd[:nc, :5], np.float64) * np.asarray(pw, np.float64)
    tree = cKDTree(X)
    inTree = np.ones(nc, bool)
    keep = np.ones(nc, bool)
    w = np.asarray(pd[:nc, 5], np.float64).copy()
    pairs = []
    ii = jj = 0
    while jj < toMerge and ii < nc - 1:
        if not keep[ii] or not inTree[ii]:
            ii += 1
            continue
        inTree[ii] = False
        found = None
        k = 16
        while found is None:
            kk = min(k, nc)
            d, idx = tree.query(X[ii], k=kk)
            for dd, j in zip(np.atleast_1d(d), np.atleast_1d(idx)):
                if j < nc and inTree[j]:
                    found = (int(j), float(dd) ** 2)
                    break
            if found is None and kk >= nc:
                break
            k *= 8
        if found is not None and found[1] < disLimit:
            f, d2 = found
            if U[jj] < w[ii] / (w[f] + w[ii]):
                w[ii] += w[f]
                keep[f] = False
                pairs.append((ii, f, d2))
            else:
                w[f] += w[ii]
                keep[ii] = False
                pairs.append((f, ii, d2))
            inTree[f] = False
            jj += 1
        ii += 1
    return keep, w, pairs
