"""Incremental planning against planning from scratch.

The planner skips subtrees in which nothing changed since it last walked
them (a length edit marks the rings whose subtree holds the edge, a
topology edit those whose subtree holds a re-linked node), checking only
the cache key of the ring at the top of a skipped subtree: the reference's
keys (parent, sibling count, branch length;
pkg/Crux/Tree/Lik.pyx:1500-1532) decide what is recomputed, as before.  Here one likelihood (and its clone, the other polarity) follows a
random sequence of branch-length edits, no-op edits, NNI moves and re-rootings
without ever being flushed, while a twin on an identical tree receives the
same edits and is flushed before every evaluation.  A CLV computed from
equal inputs is equal bit for bit, so every site lnL must be too.
"""

import math
import sys

import numpy as np
import pytest

from crux_amd import synth
from crux_amd.lik import Lik
from crux_amd.mc3 import _swap_subtrees

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _nni(tree_edges, i, pa, pb):
    """NNI across edge i of one twin: ring pa of the edge's first node and
    ring pb of its second swap subtrees (indices into the rings other than
    the edge's own, in ring order)."""
    e = tree_edges[i]
    u, v = e.ring.node, e.ring.other.node
    ra = [r for r in u.ring if r is not e.ring]
    rb = [r for r in v.ring if r is not e.ring.other]
    _swap_subtrees(ra[pa % len(ra)], rb[pb % len(rb)], u, v)


@pytest.mark.parametrize("ntaxa,nsites,ncat,kind,seed", [
    (48, 1500, 4, "dna", 1),
    (30, 800, 2, "dna", 2),
    (40, 1000, 1, "dna", 3),
    (24, 300, 4, "aa", 4),
])
def test_incremental_planning_matches_fresh(ntaxa, nsites, ncat, kind, seed):
    say("incremental", ntaxa, nsites, ncat, kind, seed)
    prob = synth.make_problem(ntaxa, nsites, seed, kind=kind, ncat=ncat)
    t1, t2 = prob.tree, prob.fresh_tree()
    inc = Lik(t1, prob.alignment, ncat=ncat)
    ref = Lik(t2, prob.alignment, ncat=ncat)
    for L in (inc, ref):
        synth.configure(L, prob.model)
    mate = inc.clone()
    e1, e2 = list(t1.edges), list(t2.edges)
    n1, n2 = list(t1.nodes), list(t2.nodes)
    rng = np.random.default_rng(seed)
    kinds = {"len": 0, "noop": 0, "nni": 0}
    for step in range(90):
        act = int(rng.integers(7))
        if act <= 3:
            k = int(rng.integers(1, 4))
            for i in rng.choice(len(e1), size=k, replace=False).tolist():
                v = e1[i].length * math.exp(0.8 * (rng.random() - 0.5))
                e1[i].length = v
                e2[i].length = v
            kinds["len"] += 1
        elif act == 4:
            i = int(rng.integers(len(e1)))
            e1[i].length = e1[i].length
            e2[i].length = e2[i].length
            kinds["noop"] += 1
        elif act == 5:
            inner = [i for i, e in enumerate(e1)
                     if e.ring.node.degree >= 3 and e.ring.other.node.degree >= 3]
            i = inner[int(rng.integers(len(inner)))]
            pa, pb = int(rng.integers(4)), int(rng.integers(4))
            _nni(e1, i, pa, pb)
            _nni(e2, i, pa, pb)
            kinds["nni"] += 1
        # Root: a random node, or an end of an edge just edited.
        ri = int(rng.integers(len(n1)))
        lik = mate if step % 6 == 5 else inc
        a = np.array(lik.siteLnLs(n1[ri]))
        ref.flush()
        b = np.array(ref.siteLnLs(n2[ri]))
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (step, act, ri)
        # The same root again, with nothing changed: only the root reduction.
        if step % 10 == 0:
            a2 = np.array(lik.siteLnLs(n1[ri]))
            assert np.array_equal(a2.view(np.uint64), b.view(np.uint64)), step
    say("edits", kinds)
