"""Autotuner of the launch parameters of a batch.

Parameters: lanes per atom of the pair kernel (16/32/64), atoms per lane of
the structure-factor pass (2/4/8) and of the force pass (1/2/4), and the k
vectors per row block of the structure-factor and force passes.  Each
parameter affects a known set of kernels; coordinate descent times only
those kernels (HIP events on the launch stream, minimum of a few runs) with
the other parameters fixed.  Results are cached in a JSON file keyed by
architecture, precision, potential mix and size bucket.
"""
import json
import math
import os

import torch

from . import _lib

__all__ = ['tune', 'cache_key']

CANDIDATES = [
    ('pair_lanes', (16, 32, 64), ('pair',)),
    ('apl', (2, 4, 8), ('ewald_sk', 'ewald_kreduce')),
    ('skb', None, ('ewald_sk', 'ewald_kreduce')),
    ('fapl', (1, 2, 4), ('ewald_force',)),
    ('fkb', None, ('ewald_force',)),
]

DEFAULT_CACHE = os.path.join(os.path.expanduser('~'), '.cache', 'yaff_amd', 'autotune.json')


def cache_key(engine):
    try:
        arch = torch.cuda.get_device_properties(0).gcnArchName.split(':')[0]
    except Exception:
        arch = 'gfx950'
    kinds = sorted(set(tuple(sorted(p.kind for p in s.pots)) + (('ewald',) if s.ewald else ())
                       for s in engine.systems))
    natom = sum(s.natom for s in engine.systems)
    return '%s|%s|%s|n%d|s%d' % (arch, 'fp64' if engine.precision == _lib.FP64 else 'mixed', kinds,
                                 int(round(math.log2(max(natom, 1)))), int(round(math.log2(max(engine.nsys, 1)))))


def _time_kernels(engine, names, reps=5):
    """Minimum over reps of the summed duration of the named kernels in one
    force evaluation, after one untimed evaluation (a kernel's first launch
    loads its code); the list is built once, outside the timing.  Timing
    noise only adds time, so the minimum is the steadier estimate."""
    engine.pair(gpos=True, accumulate=False)
    engine.ewald_reci(gpos=True, accumulate=True)
    times = []
    for _ in range(reps):
        torch.cuda.synchronize()
        _lib.profile_enable(True)
        engine.pair(gpos=True, accumulate=False)
        engine.ewald_reci(gpos=True, accumulate=True)
        prof = _lib.profile_read()
        _lib.profile_enable(False)
        times.append(sum(ms for n, ms in prof if n in names))
    return min(times)


def _row_block_sizes(k0):
    """Row-block sizes (k per block) tried around the current k0: every
    size of a ladder (steps of 8 k up to 256, then 16, 32, 64, ...) from
    k0/8 to max(8 k0, 512).  The number of blocks (whole rows each) changes
    in steps, and a launch that ends just past a full round of resident
    blocks idles most of the chip in its last round, so the best size sits
    in a narrow window (the force pass of the 6912-atom rocksalt: 31 us at
    200-216 k per row block, 33-34 us at 192 or 224, from a default of 64)."""
    lo, hi = k0/8.0, max(8*k0, 512)
    out, v, step = [], 16, 8
    while v <= hi:
        if v >= lo:
            out.append(v)
        if v >= 256 and v % (32*step) == 0:
            step *= 2
        v += step
    return out or [k0]


def tune(engine, cache=DEFAULT_CACHE, verbose=False):
    """Coordinate descent over CANDIDATES; applies and returns the best
    parameters (loaded from the cache when the key is known)."""
    key = cache_key(engine)
    table = {}
    if cache and os.path.exists(cache):
        try:
            with open(cache) as f:
                table = json.load(f)
        except (OSError, ValueError):
            table = {}
    if key in table:
        engine.set_tuning(**table[key])
        return dict(table[key])
    engine.ensure_nlist()
    best = {}
    for name, values, kernels in CANDIDATES:
        if name != 'pair_lanes' and not engine.has_ewald:
            continue
        if values is None:
            values = _row_block_sizes(getattr(engine, name))
        results = []
        for v in values:
            engine.set_tuning(**{name: v})
            results.append((_time_kernels(engine, kernels), v))
            if verbose:
                print('autotune %s=%s: %.1f us' % (name, v, results[-1][0]*1e3), flush=True)
        # the three fastest again with more runs: a single noisy minimum over
        # many candidates would otherwise decide
        finals = []
        for t, v in sorted(results)[:3]:
            engine.set_tuning(**{name: v})
            finals.append((min(t, _time_kernels(engine, kernels, reps=11)), v))
        t, v = min(finals)
        engine.set_tuning(**{name: v})
        best[name] = v
    if cache:
        table[key] = best
        try:
            os.makedirs(os.path.dirname(cache), exist_ok=True)
            with open(cache, 'w') as f:
                json.dump(table, f, indent=1, sort_keys=True)
        except OSError:
            pass
    return best
