"""Launch-shape tuning: every shape (stack depth, 4-state chunk length,
20/22-state waves per SIMD) gives the same lnL and site lnLs bit for bit, in
every component class; class settings override the context's; the tuner
keeps one of the timed shapes for the likelihood's class only and reuses its
result for a likelihood of the same shape."""

import numpy as np
import pytest

from crux_amd import _native as nat
from crux_amd import autotune, synth
from crux_amd.lik import Lik

pytestmark = pytest.mark.gpu

KEYS = autotune.KEYS


def _lik(kind, ncat, ntaxa=40, nsites=1500, seed=90):
    prob = synth.make_problem(ntaxa, nsites, seed + ncat, kind=kind, ncat=ncat)
    lik = Lik(prob.tree, prob.alignment, ncat=ncat)
    synth.configure(lik, prob.model)
    return prob, lik


@pytest.mark.parametrize("kind,ncat", [("dna", 1), ("dna", 2), ("dna", 4), ("dna", 6), ("dna", 12),
                                       ("aa", 1), ("aa", 2), ("aa", 4), ("aa", 6)])
def test_every_shape_is_bit_identical(kind, ncat):
    print("shapes", kind, ncat, flush=True)
    prob, lik = _lik(kind, ncat)
    ctx = lik._ctx
    K = lik.dim
    try:
        ref = lik.lnL()
        ref_sites = np.array(lik.siteLnLs())
        e = list(lik.tree.edges)[7]
        l0 = e.length
        launches, parts = set(), []
        for sh in autotune.grid(lik):
            autotune.set_class_shape(lik, sh)
            launch = autotune.launch_of(lik)
            launches.add(launch)
            lik.flush()
            got = lik.lnL()
            assert got == ref, (sh, launch, got, ref)
            assert np.array_equal(np.array(lik.siteLnLs()), ref_sites), (sh, launch)
            # A partial re-evaluation (cached children) under the same shape.
            e.length = l0 * 1.25
            parts.append(lik.lnL(e.ring.node))
            e.length = l0
        assert len(set(parts)) == 1, parts
        assert len(launches) >= 3, launches
    finally:
        autotune.set_class_shape(lik, None)


def test_class_settings_override_context():
    _, dna = _lik("dna", 4)
    ctx = dna._ctx
    S = 1500
    try:
        base = ctx.launch_shape(4, 4, S)
        ctx.set_tuning(nat.CRUX_TUNE_STACK_DEPTH, 2)
        assert ctx.launch_shape(4, 4, S)[0] == 2
        ctx.set_class_tuning(4, 3, nat.CRUX_TUNE_STACK_DEPTH, 8)   # C 3 -> class CM 4
        assert ctx.get_class_tuning(4, 4, nat.CRUX_TUNE_STACK_DEPTH) == 8
        assert ctx.launch_shape(4, 4, S)[0] == 8
        assert ctx.launch_shape(4, 2, S)[0] == 2      # other class: the context's
        ctx.set_class_tuning(4, 4, nat.CRUX_TUNE_CHUNK, 1)
        assert ctx.launch_shape(4, 4, S)[1] == 4      # short chunk at depth 8, CM 4
        ctx.set_class_tuning(20, 4, nat.CRUX_TUNE_WAVES, 1)
        assert ctx.launch_shape(20, 4, S)[2] == 1
        ctx.set_class_tuning(20, 4, nat.CRUX_TUNE_STACK_DEPTH, 8)
        assert ctx.launch_shape(20, 4, S)[2] == 1
        with pytest.raises(ValueError):
            ctx.set_class_tuning(4, 4, nat.CRUX_TUNE_CHUNK, 4)
        with pytest.raises(ValueError):
            ctx.set_class_tuning(4, 4, nat.CRUX_TUNE_PRECISION, 1)
        with pytest.raises(ValueError):
            ctx.set_class_tuning(4, 40, nat.CRUX_TUNE_CHUNK, 1)
        with pytest.raises(ValueError):
            ctx.set_tuning(nat.CRUX_TUNE_WAVES, 3)
    finally:
        ctx.set_tuning(nat.CRUX_TUNE_STACK_DEPTH, 0)
        for K, C in ((4, 4), (20, 4)):
            for key in KEYS:
                ctx.set_class_tuning(K, C, key, -1)
    assert ctx.launch_shape(4, 4, S) == base


@pytest.mark.parametrize("kind", ["dna", "aa"])
def test_tuner_keeps_a_timed_shape_and_caches(kind):
    print("autotune", kind, flush=True)
    prob, lik = _lik(kind, 4)
    ctx = lik._ctx
    try:
        ref = lik.lnL()
        other_class = (20, 4) if kind == "dna" else (4, 4)
        before = ctx.launch_shape(*other_class, 1500)
        autotune.clear()
        best, times = autotune.tune(lik, reps=2)
        assert best in times and set(times) <= set(autotune.grid(lik))
        assert len(times) >= 2
        assert all(t > 0.0 for t in times.values())
        assert autotune.class_shape(lik) == best
        assert ctx.launch_shape(*other_class, 1500) == before
        assert lik.lnL() == ref
        # Same shape: answered from the cache.
        other = Lik(prob.fresh_tree(), prob.alignment, ncat=4)
        synth.configure(other, prob.model)
        autotune.set_class_shape(other, None)
        best2, times2 = autotune.tune(other)
        assert best2 == best and times2 == times
        assert autotune.class_shape(other) == best
        assert abs(other.lnL() - ref) <= 1e-9 * abs(ref)
    finally:
        autotune.set_class_shape(lik, None)
        autotune.clear()


@pytest.mark.parametrize("kind", ["dna", "aa"])
def test_tuner_decides_the_split(kind):
    """A class whose plans the context splits over two launches: the tuner
    keeps the split or one launch for the class, whichever its timings
    favour, the lnL is the same bit for bit either way, and clearing the
    class's shape clears its split setting."""
    print("autotune split", kind, flush=True)
    prob = synth.make_problem(120, 1500 if kind == "aa" else 3000, 31, kind=kind, ncat=4)
    lik = Lik(prob.tree, prob.alignment, ncat=4)
    synth.configure(lik, prob.model)
    ctx = lik._ctx
    try:
        autotune.clear()
        lik.flush()
        ref = lik.lnL()
        assert ctx.last_launches()[0] == 2
        autotune.tune(lik, reps=2)
        split = autotune.class_split(lik)
        assert split in (None, 0), split
        lik.flush()
        assert lik.lnL() == ref
        assert ctx.last_launches()[0] == (2 if split is None else 1)
        autotune.set_class_shape(lik, None)
        assert autotune.class_split(lik) is None
    finally:
        autotune.set_class_shape(lik, None)
        autotune.clear()
