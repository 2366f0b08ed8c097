"""Tuner of the batched launch (BatchGraph): tries launch parameters on the caller's
own inputs, keeps the fastest, and caches the choice per device and batch shape.

    from pylians3_amd import tune
    best = tune.tune_batch(fields, nreal, dims, BoxSize, axis)   # {'chunks_per_real': ..}
    g = BatchGraph(fields, nreal, dims, BoxSize, axis, tuning=best)

Parameters (none changes any bit of a result: the tile and bin sums are exact
integers, whatever their grouping):
  chunks_per_real  sort chunks per realisation: the counting sort's granularity
                   (fewer, longer chunks: fewer tile-count flushes; more: more
                   workgroups in flight for small batches)
  bin_rows         rows of the half-complex field per binning work item (the
                   items' share of LDS-atomic work against the partials they write)
  deposit_order    1: tiles deposited largest first, 0: in key order

The cache is a JSON file, $PYLIANS3_AMD_TUNE_CACHE or ~/.cache/pylians3_amd/tune.json,
keyed by device name, nreal, dims, axis, the fields' MAS and weights, and the
particle capacity rounded to a power of two.  Each candidate is a captured graph
timed over `reps` replays (median, HIP events around the replay on the host clock
after a synchronisation).
"""
import itertools
import json
import os
import time

CANDIDATES = {
    'chunks_per_real': (4, 16, 64),
    'bin_rows': (0, 16, 64),
    'deposit_order': (-1, 0),
}


def cache_path():
    p = os.environ.get('PYLIANS3_AMD_TUNE_CACHE')
    if p:
        return p
    return os.path.join(os.path.expanduser('~'), '.cache', 'pylians3_amd', 'tune.json')


def _key(fields, nreal, dims, axis):
    import torch
    dev = torch.cuda.get_device_name(0)
    cap = max(int(f['pos'].shape[0]) for f in fields)
    cap2 = 1 << max(0, (cap - 1).bit_length())
    desc = ','.join('%s%s' % (f['mas'], 'W' if f.get('W') is not None else '') for f in fields)
    return '%s|nreal=%d|dims=%d|axis=%d|fields=%s|cap=%d' % (dev, nreal, dims, axis, desc, cap2)


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path, table):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + '.tmp%d' % os.getpid()
    with open(tmp, 'w') as f:
        json.dump(table, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def time_graph(g, reps=3):
    """Median seconds of `reps` replays of a BatchGraph (after one untimed)."""
    import numpy as np
    g.run()
    t = []
    for _ in range(reps):
        t0 = time.perf_counter()
        g.run()
        t.append(time.perf_counter() - t0)
    return float(np.median(t))


def tune_batch(fields, nreal, dims, BoxSize, axis=0, candidates=None, reps=3, use_cache=True,
               log=None):
    """Best tuning dict for this batch shape (from the cache when present).  Also
    returns, under 'timings', the seconds measured for every candidate when tuned."""
    from .batch import BatchGraph
    path = cache_path()
    key = _key(fields, nreal, dims, axis)
    table = _load(path) if use_cache else {}
    if use_cache and key in table:
        best = dict(table[key])
        best['cached'] = True
        return best
    cand = dict(CANDIDATES)
    cand.update(candidates or {})
    names = sorted(cand)
    timings = []
    best, best_t = None, None
    for values in itertools.product(*(cand[n] for n in names)):
        tun = dict(zip(names, values))
        g = BatchGraph(fields, nreal, dims, BoxSize, axis, tuning=tun)
        try:
            t = time_graph(g, reps)
        finally:
            g.close()
        timings.append((tun, t))
        if log:
            log('tune %s: %.2f ms' % (tun, 1e3 * t))
        if best_t is None or t < best_t:
            best, best_t = tun, t
    out = dict(best)
    out['seconds'] = best_t
    if use_cache:
        table[key] = out
        _store(path, table)
    out = dict(out)
    out['cached'] = False
    out['timings'] = [dict(t, seconds=s) for t, s in timings]
    return out
