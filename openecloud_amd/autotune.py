"""Autotuner for the batched scan's bin and interval sizes.

The step's own kernels have fixed shapes; what changes its speed is how well the particles
stay ordered: the push gathers E and deposits rho through per-wave LDS tiles that cover the
bounding box of a wave's cells, so it is fastest right after a cell sort and slows as the
particles drift.  The knobs are
  sortEvery   steps between two cell sorts (the sort is one radix sort over all runs);
  order       bin order of the sort key: 0 row-major cells, 1 Morton (Z-order) cells;
  binShift    bins of 2^binShift x 2^binShift cells (fewer key bits, fewer radix passes);
  graphSteps  steps per HIP graph replay (one replay per sort interval here).
tune_scan times each candidate over whole sort intervals of the live scan (these are real
steps of the simulation) and returns the fastest, optionally caching the choice per scan
shape and device in a JSON file.
"""
import json
import os
import time

import torch

DEFAULT_CANDIDATES = [dict(sortEvery=se, order=o, binShift=b)
                      for se in (10, 25, 50) for o in (0, 1) for b in (0, 1)]


def scan_key(scan):
    """Cache key: device, dtype, run count, capacity and grid size."""
    name = torch.cuda.get_device_name(0) if torch.cuda.is_available() else 'cpu'
    return '%s|%s|R%d|cap%d|np%d' % (name, str(scan.dtype).replace('torch.', ''), scan.R, scan.capacity, scan.np)


def _time_candidate(scan, c, intervals):
    se = int(c['sortEvery'])
    if scan._graph is None or scan._graphSteps != se:
        scan.capture(se)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(intervals):
        scan.sort(c['order'], c['binShift'])
        scan.replay()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / (intervals * se)


def tune_scan(scan, candidates=None, intervals=1, rounds=2, cacheFile=None, log=None):
    """Picks the fastest candidate (a dict of sortEvery, order, binShift) for this scan.
    The candidates are timed round-robin, `rounds` times each, in forward then reverse
    order, and ranked by their mean: a state that drifts steadily (the cloud grows, the
    particles heat up) then weighs on every candidate alike.  Returns
    (best, results) with results a list of (candidate, seconds per step).  With cacheFile, a
    stored choice for the same scan_key is returned without timing, and a new choice is
    stored."""
    key = scan_key(scan)
    if cacheFile and os.path.exists(cacheFile):
        try:
            cache = json.load(open(cacheFile))
        except (OSError, ValueError):
            cache = {}
        if key in cache:
            return dict(cache[key]), []
    cands = list(candidates or DEFAULT_CANDIDATES)
    rounds = max(1, int(rounds))
    tot = [0.] * len(cands)
    for rd in range(rounds):
        order = range(len(cands)) if rd % 2 == 0 else reversed(range(len(cands)))
        for i in order:
            t = _time_candidate(scan, cands[i], intervals)
            tot[i] += t
            if log:
                log('autotune %s: %.1f us/step' % (cands[i], 1e6 * t))
    results = [(dict(c), t / rounds) for c, t in zip(cands, tot)]
    best = min(results, key=lambda ct: ct[1])[0]
    best['graphSteps'] = best['sortEvery']
    if cacheFile:
        cache = {}
        if os.path.exists(cacheFile):
            try:
                cache = json.load(open(cacheFile))
            except (OSError, ValueError):
                cache = {}
        cache[key] = best
        d = os.path.dirname(os.path.abspath(cacheFile))
        os.makedirs(d, exist_ok=True)
        json.dump(cache, open(cacheFile, 'w'), indent=1)
    return best, results
