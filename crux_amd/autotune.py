"""Measured choice of the traversal's launch-shape knobs.

The traversal kernels keep the results a parent update has not consumed yet
in a per-lane LDS stack of 2, 4 or 8 slots.  More slots keep more results
out of HBM; fewer let more blocks share a CU.  Which wins depends on the tree
shape, the state and component counts and the pattern count, so the tuner
times full traversals of the caller's own likelihood at each depth (with the
device's own event timer around the traversal launch) and leaves the fastest
set on the likelihood's context (the setting is per context: it applies to
every likelihood evaluated there).

    from crux_amd.autotune import tune
    best, times = tune(lik)          # times: {depth: traversal ms}

Results are remembered per problem shape (states, component class, pattern
count, taxa), so tuning a second likelihood of the same shape costs nothing.
"""

from __future__ import annotations

import numpy as np

from . import _native as nat

DEPTHS = (2, 4, 8)

_cache: dict = {}


def _shape(lik):
    aln = lik.alignment
    ncomp = sum(len(m.comps) for m in lik.models)
    return (lik.dim, int(ncomp), int(aln.nchars), int(aln.ntaxa))


def time_traversal(lik, root=None, reps=8, warmup=2):
    """Median traversal-kernel time (ms) of a full evaluation of lik."""
    ctx = lik._ctx
    for _ in range(warmup):
        lik.flush()
        lik.lnL(root)
    ctx.set_timing(True)
    try:
        ts = []
        for _ in range(reps):
            lik.flush()
            lik.lnL(root)
            ts.append(float(ctx.last_timing()[1]))
    finally:
        ctx.set_timing(False)
    return float(np.median(ts))


def tune(lik, root=None, depths=DEPTHS, reps=8, use_cache=True):
    """Time full traversals of lik at each stack depth, keep the fastest.

    Returns (best_depth, {depth: median traversal ms}).  The likelihood's
    cache is flushed by the timing runs (it is rebuilt by the next lnL).
    """
    ctx = lik._ctx
    key = _shape(lik)
    if use_cache and key in _cache:
        best, times = _cache[key]
        ctx.set_tuning(nat.CRUX_TUNE_STACK_DEPTH, best)
        return best, dict(times)
    times = {}
    for d in depths:
        ctx.set_tuning(nat.CRUX_TUNE_STACK_DEPTH, int(d))
        times[int(d)] = time_traversal(lik, root, reps=reps)
    best = min(times, key=lambda d: (times[d], d))
    ctx.set_tuning(nat.CRUX_TUNE_STACK_DEPTH, best)
    _cache[key] = (best, dict(times))
    return best, times


def clear():
    _cache.clear()
