"""The short-plan 4-state kernel (partial re-evaluations and MC3 proposals of
at most 24 node updates, two or four components) against the general
traversal kernel: the same site lnLs bit for bit on the same plans, and the
FP64 oracle within the 4-state tolerance.

The plans cover what the short kernel takes from each source: tip rows and
CLVs cached by earlier evaluations (loaded two updates ahead), the previous
update's result (registers), the LDS stack (several changed branches in one
evaluation), the plan's own results re-read from HBM when the stack
overflows, identity children (multifurcations) and roots of degree > 3.
The caller pattern is the reference's proposal loop: one or a few branch
lengths change, then lnL at a node next to the change
(pkg/Crux/Mc3/Chain.pyx:455, pkg/Crux/Tree/Lik.pyx:1395-1601).
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


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _moves(lik, rng, n, multi):
    """n proposals: (edge indices, length factors, root node index)."""
    edges = list(lik.tree.edges)
    nodes = list(lik.tree.nodes)
    out = []
    for it in range(n):
        k = int(rng.integers(2, 7)) if multi and it % 4 == 3 else 1
        idx = rng.choice(len(edges), size=k, replace=False).tolist()
        fac = np.exp(0.9 * (rng.random(k) - 0.5)).tolist()
        if it % 3:
            root = nodes.index(edges[idx[0]].ring.node)
        else:
            root = int(rng.integers(len(nodes)))
        out.append((idx, fac, root))
    return out


def _run(lik, moves, mode, oracle=None):
    """Site lnLs of the base evaluation and after each move under short-plan
    mode `mode`, and the launches that ran the short-plan kernel; the tree
    is restored after."""
    ctx = nat.get_context(0)
    ctx.set_tuning(nat.CRUX_TUNE_SHORT_PLANS, mode)
    edges = list(lik.tree.edges)
    nodes = list(lik.tree.nodes)
    saved = [e.length for e in edges]
    lik.flush()
    res = [np.array(lik.siteLnLs())]
    nshort = 0
    try:
        for idx, fac, root in moves:
            for i, f in zip(idx, fac):
                edges[i].length = edges[i].length * f
            res.append(np.array(lik.siteLnLs(nodes[root])))
            nshort += ctx.last_launches()[1]
    finally:
        for e, l in zip(edges, saved):
            e.length = l
        ctx.set_tuning(nat.CRUX_TUNE_SHORT_PLANS, 1)
    return res, nshort


def _oracle_check(prob, ncat, moves, res, tree=None):
    """The oracle on the same moves (every third one), within tolerance."""
    ora = OracleLik(tree if tree is not None else prob.fresh_tree(), prob.alignment, ncat=ncat,
                    nthreads=4)
    synth.configure(ora, prob.model)
    edges = list(ora.tree.edges)
    nodes = list(ora.tree.nodes)
    b = np.array(ora.siteLnLs())
    assert np.allclose(res[0], b, rtol=REL, atol=1e-5)
    for k, (idx, fac, root) in enumerate(moves):
        for i, f in zip(idx, fac):
            edges[i].length = edges[i].length * f
        if k % 3 == 0:
            o = np.array(ora.siteLnLs(nodes[root]))
            assert np.allclose(res[k + 1], o, rtol=REL, atol=1e-5), k
            assert math.isclose(res[k + 1].sum(), o.sum(), rel_tol=REL)


@pytest.mark.parametrize("ntaxa,nsites,ncat,seed", [
    (24, 700, 2, 1),     # two components
    (60, 3000, 4, 2),    # four components
    (40, 2500, 3, 3),    # three components: class of four, one idle
    (150, 9000, 4, 4),   # longer paths, more than 64 patterns per tile row
])
def test_short_plans_match_general_kernel(ntaxa, nsites, ncat, seed):
    say("short-plans", ntaxa, nsites, ncat, seed)
    prob = synth.make_problem(ntaxa, nsites, seed, ncat=ncat)
    lik = Lik(prob.tree, prob.alignment, ncat=ncat)
    synth.configure(lik, prob.model)
    moves = _moves(lik, np.random.default_rng(seed), 24, multi=True)
    on, nshort = _run(lik, moves, 1)
    off, noff = _run(lik, moves, 0)
    assert noff == 0
    # Most proposals are short plans (a multi-branch change or a distant
    # root can exceed 24 updates and take the general kernel).
    assert nshort >= len(moves) // 2, nshort
    for k, (a, b) in enumerate(zip(on, off)):
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), k
    _oracle_check(prob, ncat, moves, on)


def test_short_plans_polytomies_and_stack_overflow():
    """A tree with nodes of degree 6-8 (identity children, a root of degree
    8) and multi-branch changes whose plans need more than the short
    kernel's two stack slots (own results re-read from HBM)."""
    say("short-plans polytomy")
    ncat = 4
    prob = synth.make_problem(70, 2000, 5, ncat=ncat)
    rng = np.random.default_rng(55)
    trees = [prob.tree, prob.fresh_tree()]
    edges0 = list(trees[0].edges)
    inner = [i for i, e in enumerate(edges0)
             if e.ring.node.degree > 1 and e.ring.other.node.degree > 1]
    zero = rng.choice(inner, size=14, replace=False).tolist()
    for t in trees:
        edges = list(t.edges)
        for i in zero:
            edges[i].length = 0.0
        assert t.collapse() == len(zero)
    degs = sorted(n.degree for n in trees[0].nodes)
    say("degrees", degs[-4:])
    assert degs[-1] >= 5, degs
    lik = Lik(trees[0], prob.alignment, ncat=ncat)
    synth.configure(lik, prob.model)
    moves = _moves(lik, rng, 30, multi=True)
    nodes = list(lik.tree.nodes)
    big = max(range(len(nodes)), key=lambda i: nodes[i].degree)
    # Rooted at the largest polytomy, with branches around it changed.
    ring = list(nodes[big].ring)
    moves += [([list(lik.tree.edges).index(r.edge) for r in ring[:3]], [1.3, 0.8, 1.1], big)]
    on, nshort = _run(lik, moves, 1)
    off, _ = _run(lik, moves, 0)
    assert nshort >= len(moves) // 2, nshort
    for k, (a, b) in enumerate(zip(on, off)):
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), k
    _oracle_check(prob, ncat, moves, on, tree=trees[1])
