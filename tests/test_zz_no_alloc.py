"""crux_lik_lnl_noalloc: the no-allocation variant of crux_lik_lnl.

Its plan, P matrices and per-wave partial sums live in the library's static
device scratch; the CLV slots are the tree's own (allocated by earlier
evaluations).  Checked through the C ABI against crux_lik_lnl on the same
likelihoods, for two calls in a row (the second on a smaller tree), with no
change in the context's allocated bytes, plus a partial re-evaluation and the
refusal of a batch larger than the scratch.
"""

import ctypes
import math
import sys

import numpy as np
import pytest

from crux_amd import _native as nat
from crux_amd import synth
from crux_amd.lik import Lik

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _noalloc(lik, root):
    out = ctypes.c_double()
    nat.check(nat.lib().crux_lik_lnl_noalloc(lik.handle, int(root), ctypes.byref(out)))
    return out.value


def _regular(lik, root):
    out = ctypes.c_double()
    nat.check(nat.lib().crux_lik_lnl(lik.handle, int(root), ctypes.byref(out)))
    return out.value


def _lik(prob, tree):
    lik = Lik(tree, prob.alignment, ncat=4)
    synth.configure(lik, prob.model)
    return lik


def test_noalloc_matches_regular_two_calls():
    ctx = nat.get_context(0)
    # The 60-taxon amino-acid plan is split over two launches (subtrees,
    # then the updates above them) in both entry points.
    for ntaxa, nsites, seed, kind in ((48, 4000, 71, "dna"), (14, 300, 72, "aa"), (60, 700, 74, "aa")):
        say("noalloc", ntaxa, nsites, kind)
        prob = synth.make_problem(ntaxa, nsites, seed, kind=kind)
        lik = _lik(prob, prob.tree)
        rid = lik._ready(None)
        a = _regular(lik, rid)            # allocates the CLV slots it needs
        lik.flush()                       # slots back to the pool, all stale
        before = ctx.bytes
        b = _noalloc(lik, rid)            # recomputes every CLV from scratch
        assert ctx.bytes == before
        say("  regular %.10f noalloc %.10f exact %s" % (a, b, a == b))
        assert math.isfinite(b)
        assert abs(a - b) <= 1e-12 * abs(a)
        # Cached re-evaluation, then a partial one at a changed branch,
        # rooted at a node the regular path has already rooted at (so the
        # tree holds the slots of that direction), against a twin tree.
        assert _noalloc(lik, rid) == b
        k = 3
        e = list(lik.tree.edges)[k]
        _regular(lik, e.ring.node.id)
        e.length *= 1.37
        lik._ready(None)
        before = ctx.bytes
        c = _noalloc(lik, e.ring.node.id)
        assert ctx.bytes == before
        twin = _lik(prob, prob.fresh_tree())
        e2 = list(twin.tree.edges)[k]
        e2.length *= 1.37
        d = twin.lnL(e2.ring.node)
        assert abs(c - d) <= 1e-12 * abs(d), (c, d)


def test_noalloc_refuses_batches_larger_than_the_scratch():
    # 1,024 taxa at 8 components of 20 states: the P images and tip tables
    # alone take about 27 MB, more than the 16 MiB scratch.
    say("noalloc too large")
    prob = synth.make_problem(1024, 60, 73, kind="aa", ncat=8)
    lik = Lik(prob.tree, prob.alignment, ncat=8)
    synth.configure(lik, prob.model)
    rid = lik._ready(None)
    a = _regular(lik, rid)
    lik.flush()
    with pytest.raises(ValueError, match="scratch"):
        _noalloc(lik, rid)
    # The likelihood is still usable through the regular entry point.
    assert abs(_regular(lik, rid) - a) <= 1e-12 * abs(a)
