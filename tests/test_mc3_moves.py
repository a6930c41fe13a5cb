"""Tree moves of the MC3 driver (host only): an NNI and its undo; the
reference's eTBR, polytomy merge/split and branch-length moves restore the
tree exactly on rejection, and on a flat likelihood sample the topology
prior (uniform over binary topologies for eTBR; c per missing internal edge
for the polytomy jumps with the split's full Hastings term); the rate-class
table and the resolution-class draw of the prior sampler."""

import numpy as np
import pytest

from crux_amd import synth
from crux_amd.mc3 import _swap_subtrees


def _splits(tree):
    """Set of bipartitions (frozensets of taxon names on one side)."""
    taxa = set(tree.taxa)
    out = set()
    for e in tree.edges:
        side = set()
        stack = [e.ring.other]
        while stack:
            r = stack.pop()
            n = r.node
            if n.taxon is not None:
                side.add(n.taxon)
            for s in r.siblings():
                stack.append(s.other)
        small = side if len(side) <= len(taxa) - len(side) else taxa - side
        out.add(frozenset(small))
    return out


def test_nni_and_undo():
    prob = synth.make_problem(12, 50, 5)
    t = prob.fresh_tree()
    rng = np.random.default_rng(0)
    before = _splits(t)
    degrees = {n.id: n.degree for n in t.nodes}
    inner = [e for e in t.edges if e.ring.node.degree >= 3 and e.ring.other.node.degree >= 3]
    changed = 0
    for _ in range(20):
        e = inner[int(rng.integers(len(inner)))]
        u, v = e.ring.node, e.ring.other.node
        a = [r for r in u.ring if r is not e.ring][0]
        b = [r for r in v.ring if r is not e.ring.other][0]
        A, B = a.other.node, b.other.node
        _swap_subtrees(a, b, u, v)
        # The subtree rings stay on their subtrees; the tree stays binary.
        assert a.other.node is A and b.other.node is B
        assert a.node is v and b.node is u
        assert {n.id: n.degree for n in t.nodes} == degrees
        after = _splits(t)
        changed += after != before
        assert len(after) == len(before)
        _swap_subtrees(a, b, v, u)
        assert _splits(t) == before
    assert changed == 20


# ---- the reference's tree moves on a flat likelihood (host only) ----------

import math
from collections import Counter

from crux_amd import mc3
from crux_amd.alignment import Dna


class _Aln:
    charType = Dna

    def __init__(self, ntaxa):
        self.ntaxa = ntaxa


class _FlatLik:
    """Stands in for a likelihood in tree-move tests: the moves only touch
    the tree, and the test settles every proposal with the same lnL."""

    dim = 4

    def __init__(self, tree):
        self.tree = tree
        self.alignment = _Aln(tree.ntaxa)

    def nmodels(self):
        return 1


def _chain(tree, props, seed, **kw):
    params = mc3.Mc3Params(props={k: props.get(k, 0.0) for k in mc3.PROPS}, **kw)
    c = mc3.Chain(_FlatLik(tree), 0, 0, seed, params, mix="reference")
    c.lnL = 0.0
    return c


def _shape(tree):
    """Topology as its set of splits plus every edge's ends, lengths and the
    node each ring sits on: equal after an undo exactly when the tree and its
    ring polarity (which the CLV caches depend on) are restored."""
    ends = sorted((e.id, e.ring.node.id, e.ring.other.node.id, e.length) for e in tree.edges)
    return _splits(tree), ends, tree.base.id


def _check_tree(tree, ntaxa):
    taxa = [n.taxon for n in tree.nodes if n.taxon is not None]
    assert sorted(taxa) == sorted(set(taxa)) and len(taxa) == ntaxa
    for n in tree.nodes:
        assert n.degree == 1 if n.taxon is not None else n.degree >= 3
    # A tree: edges = nodes - 1, every ring's node on its cycle.
    assert len(tree.edges) == len(tree.nodes) - 1
    for e in tree.edges:
        for r in (e.ring, e.ring.other):
            assert any(x is r for x in r.node.ring)


def _yule(ntaxa, seed):
    rng = np.random.default_rng(seed)
    return synth.yule_tree(ntaxa, rng)[1]


@pytest.mark.parametrize("kind", ["etbr", "polytomyJump", "brlen"])
def test_reference_tree_moves_undo(kind):
    t = _yule(14, 3)
    c = _chain(t, {kind: 1.0, "polytomyJump": 1.0}, 11)
    seen = Counter()
    for i in range(300):
        before = _shape(t)
        lik, root = c.propose()
        assert lik is c.lik and root is not None
        _check_tree(t, 14)
        kind_now = c._pend.kind
        seen[kind_now] += 1
        if i % 3 == 0:
            # Accept (moves the base; polytomies accumulate).
            c.settle(1e300)
        else:
            c._pend.lnPrior = -math.inf
            c.settle(0.0)
            assert _shape(t) == before, (i, kind_now)
        _check_tree(t, 14)
    assert seen[kind] > 50 and seen["polytomyJump"] > 50
    # Merges and splits both happened: some polytomy was made.
    assert c.accepts["polytomyJump"] > 0


def _topology(tree):
    """The tree's non-trivial splits."""
    return frozenset(s for s in _splits(tree) if len(s) > 1)


def _run_flat(tree, props, steps, seed, **kw):
    c = _chain(tree, props, seed, **kw)
    counts = Counter()
    for _ in range(steps):
        c.propose()
        c.settle(0.0)
        counts[_topology(tree)] += 1
    return counts


def test_flat_likelihood_etbr_samples_binary_topologies_uniformly():
    # Five taxa, binary trees only: 15 topologies, each 1/15 under a flat
    # likelihood if eTBR's Hastings terms are right.
    t = _yule(5, 1)
    counts = _run_flat(t, {"etbr": 1.0, "brlen": 1.0}, 60000, 5)
    assert len(counts) == 15
    f = np.array(list(counts.values())) / sum(counts.values())
    assert np.all(np.abs(f - 1 / 15) < 0.25 / 15), sorted(f)


@pytest.mark.parametrize("c", [1.0, 2.0])
def test_flat_likelihood_polytomy_jumps_sample_the_topology_prior(c):
    # Five taxa: 15 binary topologies (2 internal edges), 10 with one internal
    # edge, the star tree.  The prior weighs a tree by c per missing internal
    # edge; with the new edge length's proposal density in the split's
    # Hastings term, the chain samples it exactly.
    t = _yule(5, 2)
    counts = _run_flat(t, {"polytomyJump": 1.0, "brlen": 1.0}, 80000, 7,
                       polytomyJumpPrior=c, splitJacobian=True)
    assert len(counts) == 26
    by = Counter()
    for top, n in counts.items():
        by[len(top)] += n
    tot = sum(by.values())
    z = 15 + 10 * c + c * c
    want = {2: 15 / z, 1: 10 * c / z, 0: c * c / z}
    for k, w in want.items():
        assert abs(by[k] / tot - w) < 0.12 * w, (k, by[k] / tot, w)


def test_rate_class_table():
    # Groupings of the 6 GTR rates: Stirling numbers S(6, k), in the
    # reference table's order (Mc3/__init__.pyx:100-240).
    t = mc3.rate_classes(6)
    assert [len(x) for x in t] == [1, 31, 90, 65, 15, 1]
    assert t[0] == [[0, 0, 0, 0, 0, 0]] and t[5] == [[0, 1, 2, 3, 4, 5]]
    assert t[1][:4] == [[0, 0, 0, 0, 0, 1], [0, 0, 0, 0, 1, 0], [0, 0, 0, 0, 1, 1],
                        [0, 0, 0, 1, 0, 0]]
    assert t[2][:3] == [[0, 0, 0, 0, 1, 2], [0, 0, 0, 1, 0, 2], [0, 0, 0, 1, 1, 2]]
    for k, rows in enumerate(t):
        assert all(max(r) == k for r in rows)
        assert rows == sorted(rows)


def test_resolution_draw():
    # P(k) proportional to prior^-(k-1): prior 1 is uniform, prior 2 halves
    # each step.
    us = (np.arange(20000) + 0.5) / 20000
    for prior, n in ((1.0, 6), (2.0, 3)):
        ks = Counter(mc3._resolution(u, prior, n) for u in us)
        w = np.array([prior ** -(k - 1) for k in range(1, n + 1)])
        w /= w.sum()
        for k in range(1, n + 1):
            assert abs(ks[k] / len(us) - w[k - 1]) < 1e-3
