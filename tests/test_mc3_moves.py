"""Tree moves of the MC3 driver (host only): an NNI and its undo."""

import numpy as np

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
