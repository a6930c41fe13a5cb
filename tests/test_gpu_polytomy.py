"""Multifurcating trees on the device: nodes of degree 4-8 and roots of
degree > 3 (Tree.collapse of zero-length internal edges,
pkg/Crux/Tree/__init__.pyx:475-517; the reference plans deg - 1 steps per
node with ntrail, pkg/Crux/Tree/Lik.pyx:1512-1532, 1569-1596).

A fused update takes at most three children, so a node with more is chained:
the first update writes the node's slot, each later one multiplies the slot
back in as an identity child; a root with more than three children
accumulates into a scratch slot.  Checked against the FP64 oracle (itself
pinned on polytomies by brute force) and against the invariant P(0) = I:
the binary tree with those internal edges at length 0 has the same lnL.
"""

import math
import sys

import numpy as np
import pytest

from crux_amd import _native as nat
from crux_amd import synth
from crux_amd.lik import Lik
from oracle.oracle import OracleLik

pytestmark = pytest.mark.gpu

REL = 1e-6


def close(a, b, rel=REL):
    return abs(a - b) <= rel * abs(b) + 1e-9


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _internal_edges(tree):
    return [i for i, e in enumerate(tree.edges)
            if e.ring.node.degree > 1 and e.ring.other.node.degree > 1]


def _cluster(tree, rng, k):
    """Indices (in tree.edges order) of k internal edges forming a connected
    cluster around a random internal node: collapsing them makes one node of
    degree 3 + k."""
    edges = list(tree.edges)
    index = {id(e): i for i, e in enumerate(edges)}
    inner = [n for n in tree.nodes if n.degree > 1]
    start = inner[int(rng.integers(len(inner)))]
    picked, frontier, seen = [], [start], {id(start)}
    while frontier and len(picked) < k:
        n = frontier.pop(0)
        for r in n.ring:
            m = r.other.node
            if m.degree > 1 and id(m) not in seen:
                seen.add(id(m))
                picked.append(index[id(r.edge)])
                frontier.append(m)
                if len(picked) == k:
                    break
    assert len(picked) == k
    return picked


def _polytomy_problem(ntaxa, nsites, seed, kind, ncat, cluster=5, scattered=4):
    """(device Lik on the collapsed tree, oracle on the same collapsed tree,
    device Lik on the binary tree with the zeroed edges, problem)."""
    prob = synth.make_problem(ntaxa, nsites, seed, kind=kind, ncat=ncat)
    rng = np.random.default_rng(seed + 1000)
    t_bin = prob.fresh_tree()
    zero = set(_cluster(t_bin, rng, cluster))
    rest = [i for i in _internal_edges(t_bin) if i not in zero]
    zero |= set(rng.choice(rest, size=scattered, replace=False).tolist())
    trees = [prob.tree, prob.fresh_tree(), t_bin]
    for t in trees:
        edges = list(t.edges)
        for i in zero:
            edges[i].length = 0.0
    for t in trees[:2]:
        assert t.collapse() == len(zero)
    dev = Lik(trees[0], prob.alignment, ncat=ncat)
    ora = OracleLik(trees[1], prob.alignment, ncat=ncat, nthreads=4)
    binl = Lik(trees[2], prob.alignment, ncat=ncat)
    for L in (dev, ora, binl):
        synth.configure(L, prob.model)
    return dev, ora, binl, prob


def _check_polytomy(dev, ora, binl):
    degs = sorted(n.degree for n in dev.tree.nodes)
    assert degs[-1] >= 8, degs
    b = ora.lnL()
    a = dev.lnL()
    say("poly", degs[-3:], a, b)
    assert close(a, b), (a, b)
    assert close(binl.lnL(), b)
    sd, so = np.array(dev.siteLnLs()), np.array(ora.siteLnLs())
    assert np.allclose(sd, so, rtol=1e-6, atol=1e-5)
    nodes_d = list(dev.tree.nodes)
    nodes_o = list(ora.tree.nodes)
    # Every node as root, the polytomies first: the root of degree 8 chains
    # its children through the scratch slot; later roots re-read cached CLVs.
    order = sorted(range(len(nodes_d)), key=lambda i: -nodes_d[i].degree)
    for i in order:
        assert close(dev.lnL(nodes_d[i]), b), (i, nodes_d[i].degree)
    # A branch next to the largest polytomy changes: partial re-evaluation
    # rooted there, against the oracle on the same change.
    big = order[0]
    r_d = nodes_d[big].ring
    r_o = nodes_o[big].ring
    for step in range(3):
        r_d.edge.length = r_d.edge.length * 1.4 + 0.01
        r_o.edge.length = r_o.edge.length * 1.4 + 0.01
        a2 = dev.lnL(nodes_d[big])
        b2 = ora.lnL(nodes_o[big])
        assert close(a2, b2), (step, a2, b2)
        a3 = dev.lnL(r_d.other.node)
        assert close(a3, b2), (step, a3, b2)
        r_d, r_o = r_d.next, r_o.next


@pytest.mark.parametrize("ncat", [1, 2, 4, 8])
def test_polytomy_dna(ncat):
    say("poly dna", ncat)
    dev, ora, binl, _ = _polytomy_problem(60, 800, 900 + ncat, "dna", ncat)
    _check_polytomy(dev, ora, binl)


@pytest.mark.parametrize("ncat,kind", [(4, "aa"), (1, "aa"), (2, "aa")])
def test_polytomy_amino_acids(ncat, kind):
    say("poly", kind, ncat)
    dev, ora, binl, _ = _polytomy_problem(30, 500, 950 + ncat, kind, ncat)
    _check_polytomy(dev, ora, binl)


def test_polytomy_plain_mode_and_clone():
    # One launch per node update (identity children read back from HBM),
    # and the clone polarity on a multifurcating tree.
    say("poly plain")
    dev, ora, binl, prob = _polytomy_problem(40, 600, 977, "dna", 4)
    b = ora.lnL()
    ctx = nat.get_context(0)
    ctx.set_mode(1)
    try:
        dev.flush()
        assert close(dev.lnL(), b)
        big = max(dev.tree.nodes, key=lambda n: n.degree)
        assert close(dev.lnL(big), b)
    finally:
        ctx.set_mode(0)
    c = dev.clone()
    c.setAlpha(0, 1.9)
    ora.setAlpha(0, 1.9)
    big = max(dev.tree.nodes, key=lambda n: n.degree)
    big_o = max(ora.tree.nodes, key=lambda n: n.degree)
    assert close(c.lnL(big), ora.lnL(big_o))
    ora.setAlpha(0, prob.model.alpha)
    assert close(dev.lnL(big), b)


def test_star_root_degree_six_and_leaf_root():
    # Six taxa joined at one node: a root of degree 6 (three chained updates
    # through the scratch slot), and every leaf as root.
    say("star")
    from crux_amd.alignment import Alignment
    from crux_amd.tree import Tree
    rng = np.random.default_rng(5)
    seqs = {"T%d" % i: "".join("ACGT"[j] for j in rng.integers(4, size=200)) for i in range(6)}
    aln = Alignment("".join(">%s\n%s\n" % kv for kv in seqs.items()))
    aln.compact()
    nw = "(" + ",".join("T%d:%.3f" % (i, 0.05 + 0.07 * i) for i in range(6)) + ");"
    for ncat in (1, 4):
        dev = Lik(Tree(nw, rooted=False), aln, ncat=ncat)
        ora = OracleLik(Tree(nw, rooted=False), aln, ncat=ncat)
        for L in (dev, ora):
            L.setRclass(0, [0, 1, 2, 3, 4, 5], [1.0, 2.5, 0.7, 1.2, 3.1, 1.0])
            L.setFreq(0, 0, 0.4)
            if ncat > 1:
                L.setAlpha(0, 0.8)
        b = ora.lnL()
        for n in dev.tree.nodes:
            assert close(dev.lnL(n), b), (ncat, n.degree)
        assert dev.tree.base.degree == 6
