"""Measured choice of the traversal's launch shape.

Three knobs set how a traversal class (state count K, component class CM)
is launched, none of which changes a result (every choice computes the same
lnL bit for bit):

* stack depth (2, 4 or 8 LDS slots per lane for results a parent update has
  not consumed yet): more slots keep more results out of HBM, fewer let more
  blocks share a CU;
* 4 states: updates per staged chunk (short, middle, long): the P images and
  control words of a chunk are staged in LDS together, and the block's waves
  wait for their stores once per chunk;
* 20/22 states: waves per SIMD the kernel's registers are held to (1 or 2).

Where the context's split setting runs the class's plans over two launches
(subtree blocks, then the updates above them), the tuner also times the
best shape in one launch and keeps that for the class if it is faster.

Which shape wins depends on the tree, the state and component counts and the
pattern count, so the tuner times full traversals of the caller's own
likelihood at every shape (with the device's own event timer around the
traversal launch) and keeps the fastest as the setting of that likelihood's
class on its context; other classes keep theirs.

    from crux_amd.autotune import tune
    best, times = tune(lik)     # best: (depth, chunk, waves) key values,
                                # times: {key values: traversal ms}

Results are remembered per problem shape (states, components, pattern count,
taxa), so tuning a second likelihood of the same shape costs nothing.
"""

from __future__ import annotations

import numpy as np

from . import _native as nat

DEPTHS = (2, 4, 8)
CHUNKS = (1, 2, 3)      # short, middle, long (CRUX_TUNE_CHUNK values)
WAVES = (1, 2)          # CRUX_TUNE_WAVES values
KEYS = (nat.CRUX_TUNE_STACK_DEPTH, nat.CRUX_TUNE_CHUNK, nat.CRUX_TUNE_WAVES)

_cache: dict = {}


def _ncomp(lik):
    return int(sum(len(m.comps) for m in lik.models))


def launch_of(lik):
    """(stack depth, updates per chunk, waves per SIMD, CM) lik's evaluation
    launches with now."""
    return lik._ctx.launch_shape(lik.dim, _ncomp(lik), int(lik.alignment.nchars))


def _shape(lik):
    aln = lik.alignment
    return (lik.dim, _ncomp(lik), int(aln.nchars), int(aln.ntaxa))


def grid(lik):
    """The (depth, chunk, waves) key values the tuner tries for lik's class
    (0: the key does not apply)."""
    if lik.dim == 4:
        return [(d, c, 0) for d in DEPTHS for c in CHUNKS]
    return [(d, 0, w) for d in DEPTHS for w in WAVES]


def set_class_shape(lik, shape):
    """Set (depth, chunk, waves) for lik's class on its context (None: clear,
    the context's values apply, split setting included)."""
    ctx = lik._ctx
    for key, v in zip(KEYS, shape if shape is not None else (-1, -1, -1)):
        ctx.set_class_tuning(lik.dim, _ncomp(lik), key, v)
    if shape is None:
        set_class_split(lik, None)


def set_class_split(lik, value):
    """Split setting of lik's class (CRUX_TUNE_SPLIT; 0: one launch, None:
    the context's)."""
    lik._ctx.set_class_tuning(lik.dim, _ncomp(lik), nat.CRUX_TUNE_SPLIT, -1 if value is None else value)


def class_split(lik):
    """The class's split setting (None: the context's applies)."""
    v = lik._ctx.get_class_tuning(lik.dim, _ncomp(lik), nat.CRUX_TUNE_SPLIT)
    return None if v < 0 else v


def class_shape(lik):
    ctx = lik._ctx
    return tuple(ctx.get_class_tuning(lik.dim, _ncomp(lik), key) for key in KEYS)


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


def tune(lik, root=None, shapes=None, reps=8, use_cache=True):
    """Time full traversals of lik at every launch shape of its class, keep
    the fastest for the class.

    Returns (best, {shape: median traversal ms}), shapes as (depth, chunk,
    waves) key values.  Choices that resolve to the same launch (a depth
    capped by what fits, waves fixed at depth 8) are timed once.  The
    likelihood's cache is flushed by the timing runs (it is rebuilt by the
    next lnL).
    """
    key = _shape(lik)
    if use_cache and key in _cache:
        best, times, split = _cache[key]
        set_class_shape(lik, best)
        set_class_split(lik, split)
        return best, dict(times)
    set_class_split(lik, None)
    times = {}
    seen = set()
    for sh in shapes if shapes is not None else grid(lik):
        sh = tuple(int(x) for x in sh)
        set_class_shape(lik, sh)
        launch = launch_of(lik)
        if launch in seen:
            continue
        seen.add(launch)
        times[sh] = time_traversal(lik, root, reps=reps)
    best = min(times, key=lambda s: (times[s], s))
    set_class_shape(lik, best)
    # Split plans (two traversal launches for the one class): one launch at
    # the best shape, kept if faster.
    split = None
    ctx = lik._ctx
    lik.flush()
    lik.lnL(root)
    if ctx.last_launches()[0] > 1:
        set_class_split(lik, 0)
        if time_traversal(lik, root, reps=reps) < times[best]:
            split = 0
    set_class_split(lik, split)
    _cache[key] = (best, dict(times), split)
    return best, times


def clear():
    _cache.clear()
