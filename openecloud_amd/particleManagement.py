"""Macro-particle management (particleManagement.pyx:35-262) on the GPU.

globalRandom keeps the reference's rule exactly (weight rescale by N/target, then a uniformly
random kill set or uniformly random clones); only the random stream differs.  localRandom
keeps the reference's thresholds, operation counts, velocity scale, split rule, merge rule
(nearest neighbour in the weighted 5-D phase space, the squared distance below
factorDisLimit * ds) and survivor rule; the serial kd-tree loop becomes rounds of parallel
matching on the device (oecl_nn_merge): exact nearest neighbours through a Morton-ordered tree,
pairs taken lightest first.  mergeGroup > 2 lets a survivor merge again until its group holds that many
particles (2 = the reference's pairs).  The total weight is conserved exactly: the rounding of
the merged weights in the particle dtype is measured and given back to one particle.
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib as L
from . import randomGen
from .constants import c as C_LIGHT


def _work(nbytes):
    return torch.empty(int(nbytes) // 8 + 64, dtype=torch.float64, device=L.device())


def sort_by_column(p, column):
    """Particles.sortByColumn: stable ascending device radix sort by one column."""
    n = p.macroParticleCount
    if n < 2:
        return
    dev = L.device()
    keys = torch.empty(n, dtype=torch.int64, device=dev)
    perm = torch.empty(n, dtype=torch.int32, device=dev)
    spare = p._spare_block()
    L.check(L.lib().oecl_sort_by_column(
        L.i32(p.code), L.ptr(p._data), L.ptr(spare), L.u64(p.capacity), L.u32(n), L.i32(column),
        L.ptr(keys), L.ptr(perm), L.ptr(_work(L.lib().oecl_sort_work_bytes(n))), L.stream()),
        'sortByColumn')
    # rows beyond n are not used; swap the blocks
    p._data, p._spare = spare, p._data


def globalRandom(gridObj, particlesObj, targetMacroParticleCount):
    """particleManagement.pyx:35-67."""
    p = particlesObj
    n = p.macroParticleCount
    target = int(targetMacroParticleCount)
    diff = n - target
    f = n / target
    k0, k1 = randomGen.key()
    sidx = randomGen.next_stream()
    dev = L.device()
    if diff > 0:
        L.check(L.lib().oecl_scale_weights(L.i32(p.code), L.ptr(p._data), L.u64(p.capacity),
                                           L.u32(n), L.f64(f), L.stream()), 'scale')
        keep = torch.empty(n, dtype=torch.uint8, device=dev)
        keys = torch.empty(n, dtype=torch.int64, device=dev)
        perm = torch.empty(n, dtype=torch.int32, device=dev)
        L.check(L.lib().oecl_random_kill(L.u32(n), L.u32(diff), L.u32(sidx), L.u32(k0), L.u32(k1),
                                         L.ptr(keep), L.ptr(keys), L.ptr(perm),
                                         L.ptr(_work(L.lib().oecl_sort_work_bytes(n))), L.stream()),
                'globalRandom kill')
        p.addAndRemoveParticles(torch.zeros((0, 6), dtype=p.dtype, device=dev), keep)
    elif target > n:
        L.check(L.lib().oecl_scale_weights(L.i32(p.code), L.ptr(p._data), L.u64(p.capacity),
                                           L.u32(n), L.f64(f), L.stream()), 'scale')
        add = torch.empty((-diff, 6), dtype=p.dtype, device=dev)
        L.check(L.lib().oecl_random_clones(L.i32(p.code), L.ptr(p._data), L.u64(p.capacity),
                                           L.u32(n), L.u32(-diff), L.u32(sidx), L.u32(k0),
                                           L.u32(k1), L.ptr(add), L.stream()), 'globalRandom clone')
        keep = torch.ones(n, dtype=torch.uint8, device=dev)
        p.addAndRemoveParticles(add, keep)


def localRandom(gridObj, particlesObj, particleBoundaryObj, targetMacroParticleCount,
                phaseSpaceWeights=None, thresholdFactor=2., lambdavFactor=1., factorDisLimit=1.,
                mergeScheme=0, mergeGroup=2, pairs=None):
    """particleManagement.pyx:74-238.  Returns a dict with the operation counts.  pairs: an
    optional device int32 tensor of at least 2 * toMerge entries that receives (survivor,
    removed) indices, in the weight-sorted order, for every merge."""
    from .diagnostics import moments
    p = particlesObj
    n = p.macroParticleCount
    target = int(targetMacroParticleCount)
    m = moments(p)
    meanWeight = m[0] / n
    goalWeight = meanWeight * n / target
    maxThreshold = thresholdFactor * goalWeight
    minThreshold = goalWeight / thresholdFactor
    dx, dy = gridObj.getDx(), gridObj.getDy()
    ds = min(dx, dy)
    sort_by_column(p, 5)
    dev = L.device()
    out = torch.empty(4, dtype=torch.int32, device=dev)
    L.check(L.lib().oecl_search_sorted_weights(L.i32(p.code), L.ptr(p._data), L.u64(p.capacity),
                                               L.u32(n), L.f64(minThreshold), L.f64(maxThreshold),
                                               L.f64(meanWeight), L.ptr(out), L.stream()), 'search')
    o = out.cpu().tolist()
    toMerge = o[0]
    toSplit = n - o[1]
    # particleManagement.pyx:106-122
    if 2 * toMerge + toSplit >= n:
        fl = 0.95 * n / (2 * toMerge + toSplit)
        toMerge = int(toMerge * fl)
        toSplit = int(toSplit * fl)
    newN = n + toSplit - toMerge
    if newN > target:
        toMerge += newN - target
    elif newN < target:
        toSplit += target - newN
    if 2 * toMerge + toSplit >= n:
        fl = 0.95 * n / (2 * toMerge + toSplit)
        toMerge = int(toMerge * fl)
        toSplit = int(toSplit * fl)
    # particleManagement.pyx:124-128 (sum of v^2 w^2 from the same particle set)
    vTypical = math.sqrt(m[2]) / (n * meanWeight)
    vTypical = max(vTypical, 1.e-12 * C_LIGHT)
    k0, k1 = randomGen.key()
    sidx = randomGen.next_stream()
    add = torch.empty((toSplit, 6), dtype=p.dtype, device=dev)
    if toSplit:
        L.check(L.lib().oecl_split_heaviest(
            ctypes.byref(particleBoundaryObj.desc), L.i32(p.code), L.ptr(p._data), L.u64(p.capacity),
            L.u32(n - toSplit), L.u32(toSplit), L.f64(dx), L.f64(dy), L.f64(gridObj.getLx()),
            L.f64(gridObj.getLy()), L.u32(sidx), L.u32(k0), L.u32(k1), L.ptr(add), L.stream()),
            'localRandom split')
    if phaseSpaceWeights is None or len(phaseSpaceWeights) == 0:
        lam = lambdavFactor * ds / vTypical
        pw = np.array([1., 1., lam, lam, lam])
    else:
        pw = np.asarray(phaseSpaceWeights, np.float64)[:5]
    pwd = L.to_dev(pw)
    nc = n - toSplit
    keep = torch.ones(n, dtype=torch.uint8, device=dev)
    merged = torch.zeros(1, dtype=torch.int32, device=dev)
    nMerged, rounds = ctypes.c_uint32(0), ctypes.c_uint32(0)
    if pairs is not None and pairs.numel() < 2 * toMerge:
        raise ValueError('pairs must hold 2 * toMerge = %d entries' % (2 * toMerge))
    if toMerge and nc > 1:
        L.check(L.lib().oecl_nn_merge(
            L.i32(p.code), L.ptr(p._data), L.u64(p.capacity), L.u32(nc), L.ptr(pwd),
            L.i32(mergeGroup), L.f64(factorDisLimit * ds), L.u32(toMerge), L.i32(mergeScheme),
            L.u32(sidx), L.u32(k0), L.u32(k1), L.ptr(keep), L.ptr(merged), L.ptr(pairs),
            ctypes.byref(nMerged), ctypes.byref(rounds), L.ptr(_work(L.lib().oecl_merge_work_bytes(nc))),
            L.stream()), 'localRandom merge')
    p.addAndRemoveParticles(add, keep)
    # The merged weights were rounded to the particle dtype: the measured difference of the
    # total (FP64, fixed summation order) goes back to the particle in the first slot, so the
    # total charge changes by at most that one weight's final rounding.
    if p.macroParticleCount:
        lost = m[0] - moments(p)[0]
        if lost != 0.:
            w0 = float(p._data[5, 0].item())
            p._data[5, 0] = w0 + lost
    return {'toSplit': toSplit, 'toMerge': toMerge, 'merged': nMerged.value, 'rounds': rounds.value,
            'vTypical': vTypical}
