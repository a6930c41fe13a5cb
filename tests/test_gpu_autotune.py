"""Stack-depth tuner: every depth gives the same lnL bit for bit, the tuner
keeps one of the timed depths and reuses its result for a likelihood of the
same shape."""

import pytest

from crux_amd import _native as nat
from crux_amd import autotune, synth
from crux_amd.lik import Lik

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["dna", "aa"])
def test_depths_agree_and_tuner_caches(kind):
    print("autotune", kind, flush=True)
    prob = synth.make_problem(40, 1500, 90, kind=kind)
    ctx = nat.get_context(0)
    lik = Lik(prob.tree, prob.alignment, ncat=4)
    synth.configure(lik, prob.model)
    try:
        ref = lik.lnL()
        for d in autotune.DEPTHS:
            ctx.set_tuning(nat.CRUX_TUNE_STACK_DEPTH, d)
            assert ctx.get_tuning(nat.CRUX_TUNE_STACK_DEPTH) == d
            lik.flush()
            assert lik.lnL() == ref, d
        autotune.clear()
        best, times = autotune.tune(lik, reps=2)
        assert best in autotune.DEPTHS
        assert set(times) == set(autotune.DEPTHS)
        assert all(t > 0.0 for t in times.values())
        assert ctx.get_tuning(nat.CRUX_TUNE_STACK_DEPTH) == best
        assert lik.lnL() == ref
        # Same shape: answered from the cache, same depth set.
        other = Lik(prob.fresh_tree(), prob.alignment, ncat=4)
        synth.configure(other, prob.model)
        ctx.set_tuning(nat.CRUX_TUNE_STACK_DEPTH, 0)
        best2, times2 = autotune.tune(other)
        assert best2 == best and times2 == times
        assert ctx.get_tuning(nat.CRUX_TUNE_STACK_DEPTH) == best
        assert abs(other.lnL() - ref) <= 1e-9 * abs(ref)
    finally:
        ctx.set_tuning(nat.CRUX_TUNE_STACK_DEPTH, 0)
        autotune.clear()
