"""Autotuner for the batched scan's bin, interval and deposit tile sizes.

The step's own kernels have fixed shapes; what changes its speed is how well the particles
stay ordered: the push gathers E and deposits rho through per-wave LDS tiles that cover the
bounding box of a wave's cells, so it is fastest right after a cell sort and slows as the
particles drift.  The knobs are
  sortEvery   steps between two cell sorts (the sort is one radix sort over all runs);
  order       bin order of the sort key: 0 row-major cells, 1 Morton (Z-order) cells;
  binShift    bins of 2^binShift x 2^binShift cells (fewer key bits, fewer radix passes);
  graphSteps  steps per HIP graph replay (one replay per sort interval here);
  span        lanes over which the push sums runs of equal cells before its LDS tile (2-64);
  waveTiles   per-wave LDS tiles instead of one tile per push block of 256 particles.
tune_scan times each candidate over whole sort intervals of the live scan (these are real
steps of the simulation) and returns the fastest, optionally caching the choice per scan
shape and device in a JSON file.  The interval and bin knobs are tuned first (the default
tiles), then the tile knobs with the best of those.
"""
import json
import os
import time

import torch

DEFAULT_CANDIDATES = [dict(sortEvery=se, order=o, binShift=b)
                      for se in (10, 25, 50) for o in (0, 1) for b in (0, 1)]
TILE_CANDIDATES = [dict(span=sp, waveTiles=wt) for sp in (4, 16, 64) for wt in (False, True)]


def scan_key(scan):
    """Cache key: device, dtype, run count, capacity and grid size."""
    name = torch.cuda.get_device_name(0) if torch.cuda.is_available() else 'cpu'
    return '%s|%s|R%d|cap%d|np%d' % (name, str(scan.dtype).replace('torch.', ''), scan.R, scan.capacity, scan.np)


def _time_candidate(scan, c, intervals):
    se = int(c['sortEvery'])
    tiles = (int(c.get('span', 0)), bool(c.get('waveTiles', False)))
    if (int(scan._args.depSpan), bool(scan._args.depTile)) != tiles:
        scan.set_deposit(*tiles)
    if scan._graph is None or scan._graphSteps != se:
        scan.capture(se)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(intervals):
        scan.sort(c['order'], c['binShift'])
        scan.replay()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / (intervals * se)


def _round_robin(scan, cands, intervals, rounds, log):
    tot = [0.] * len(cands)
    for rd in range(rounds):
        order = range(len(cands)) if rd % 2 == 0 else reversed(range(len(cands)))
        for i in order:
            t = _time_candidate(scan, cands[i], intervals)
            tot[i] += t
            if log:
                log('autotune %s: %.1f us/step' % (cands[i], 1e6 * t))
    return [(dict(c), t / rounds) for c, t in zip(cands, tot)]


def tune_scan(scan, candidates=None, intervals=1, rounds=2, cacheFile=None, log=None, tiles=None):
    """Picks the fastest candidate (a dict of sortEvery, order, binShift, span, waveTiles) for
    this scan, in two stages: the candidates (sort interval and bins, default tiles), then the
    tile candidates `tiles` (default TILE_CANDIDATES; [] skips the stage) with the best of the
    first stage.  Within a stage the candidates are timed round-robin, `rounds` times each, in
    forward then reverse order, and ranked by their mean: a state that drifts steadily (the
    cloud grows, the particles heat up) then weighs on every candidate alike.  Returns (best,
    results) with results a list of (candidate, seconds per step) of both stages; the scan is
    left set to the best tiles.  With cacheFile, a stored choice for the same scan_key is
    returned (and its tiles set) without timing, and a new choice is stored."""
    key = scan_key(scan)
    if cacheFile and os.path.exists(cacheFile):
        try:
            cache = json.load(open(cacheFile))
        except (OSError, ValueError):
            cache = {}
        if key in cache:
            best = dict(cache[key])
            scan.set_deposit(int(best.get('span', 0)), bool(best.get('waveTiles', False)))
            return best, []
    cands = [dict(c, span=0, waveTiles=False) if 'span' not in c else dict(c)
             for c in (candidates or DEFAULT_CANDIDATES)]
    rounds = max(1, int(rounds))
    results = _round_robin(scan, cands, intervals, rounds, log)
    best = min(results, key=lambda ct: ct[1])[0]
    tcands = list(TILE_CANDIDATES if tiles is None else tiles)
    if tcands:
        base = {k: best[k] for k in ('sortEvery', 'order', 'binShift')}
        tres = _round_robin(scan, [dict(base, **t) for t in tcands], intervals, rounds, log)
        results += tres
        best = min(tres, key=lambda ct: ct[1])[0]
    best = dict(best)
    best['graphSteps'] = best['sortEvery']
    scan.set_deposit(int(best.get('span', 0)), bool(best.get('waveTiles', False)))
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
