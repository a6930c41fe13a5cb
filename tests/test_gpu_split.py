"""Plans split over two launches against one launch.

A class launched with few site tiles (per SIMD for 20/22 states, per CU
for 4 states) runs each long plan as its largest subtrees of about T
updates (one block per subtree and tile, first launch) and the updates
above them (second launch, reading the subtrees' results from their
slots).  Every CLV is the same product of the
same inputs either way, so every site lnL must be equal bit for bit; the
oracle checks the split results within the FP32-storage tolerance.

The cases cover one, two, four and eight components (idle lanes), 22
states, polytomies (identity children, a root of degree > 3), stack depths
that overflow (results re-read from HBM) or never do, more tips than one
staged chunk (chunks staged where the upper plan reaches them), partial
re-evaluations after branch changes (the reference's proposal loop,
pkg/Crux/Mc3/Chain.pyx:455, pkg/Crux/Tree/Lik.pyx:1395-1601) and several
trees in one batch.
"""

import sys

import numpy as np
import pytest

from crux_amd import _native as nat
from crux_amd import synth
from crux_amd.alignment import Protein
from crux_amd.lik import Lik
from oracle.oracle import OracleLik

pytestmark = pytest.mark.gpu

REL = 1e-6


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _walk(lik, split, steps, depth=0):
    """Site lnLs of a full evaluation and of `steps` (edge, factor, root)
    re-evaluations at split setting `split`, and the traversal launches of
    the full evaluation; the tree is restored after."""
    ctx = nat.get_context(0)
    ctx.set_tuning(nat.CRUX_TUNE_SPLIT, split)
    ctx.set_tuning(nat.CRUX_TUNE_STACK_DEPTH, depth)
    edges = list(lik.tree.edges)
    nodes = list(lik.tree.nodes)
    saved = [e.length for e in edges]
    try:
        lik.flush()
        res = [np.array(lik.siteLnLs())]
        launches = ctx.last_launches()[0]
        for i, f, r in steps:
            edges[i].length = edges[i].length * f
            res.append(np.array(lik.siteLnLs(nodes[r])))
        # A full evaluation at another root after the edits.
        lik.flush()
        res.append(np.array(lik.siteLnLs(nodes[steps[0][2]])))
    finally:
        for e, l in zip(edges, saved):
            e.length = l
        ctx.set_tuning(nat.CRUX_TUNE_SPLIT, 3)
        ctx.set_tuning(nat.CRUX_TUNE_STACK_DEPTH, 0)
    return res, launches


def _steps(lik, rng, n):
    ne, nn = len(list(lik.tree.edges)), len(list(lik.tree.nodes))
    return [(int(rng.integers(ne)), float(np.exp(rng.random() - 0.5)), int(rng.integers(nn)))
            for _ in range(n)]


def _same(a, b):
    assert len(a) == len(b)
    for k, (x, y) in enumerate(zip(a, b)):
        assert np.array_equal(x.view(np.uint64), y.view(np.uint64)), k


def _oracle(prob, ncat, res, tree=None):
    ora = OracleLik(tree if tree is not None else prob.fresh_tree(), prob.alignment, ncat=ncat,
                    nthreads=4)
    synth.configure(ora, prob.model)
    o = np.array(ora.siteLnLs())
    assert np.allclose(res[0], o, rtol=REL, atol=1e-5)


def _check(prob, ncat, seed, depths=(0,), tree=None, splits=(3, 1, 16)):
    lik = Lik(prob.tree, prob.alignment, ncat=ncat)
    synth.configure(lik, prob.model)
    steps = _steps(lik, np.random.default_rng(seed), 6)
    for depth in depths:
        off, n_off = _walk(lik, 0, steps, depth)
        assert n_off == 1, n_off
        for split in splits:
            on, n_on = _walk(lik, split, steps, depth)
            say("  depth", depth, "split", split, "launches", n_on)
            assert n_on == 2, n_on
            _same(on, off)
    _oracle(prob, ncat, on, tree)


@pytest.mark.parametrize("ntaxa,nsites,ncat,seed,kind", [
    (120, 1500, 4, 1, "aa"),    # four components, tips in two staged chunks
    (90, 900, 1, 2, "aa"),      # one component
    (70, 800, 2, 3, "aa"),      # two components
    (60, 700, 3, 4, "aa"),      # class of four, one idle component
    (50, 500, 6, 5, "aa"),      # class of eight, two idle components
    (150, 3000, 4, 11, "dna"),  # 4 states: lane-halves kernel, four components
    (100, 2000, 1, 12, "dna"),  # one-component 4-state kernel
    (80, 1500, 2, 13, "dna"),   # two components
    (60, 1200, 6, 14, "dna"),   # class of eight, two idle components
])
def test_split_matches_one_launch(ntaxa, nsites, ncat, seed, kind):
    say("split", ntaxa, nsites, ncat, seed, kind)
    prob = synth.make_problem(ntaxa, nsites, seed, kind=kind, ncat=ncat)
    _check(prob, ncat, seed, depths=(0, 2, 8) if ncat == 4 else (0,))


def test_split_amino_acid_workload():
    """256 taxa x 20k sites, four components: 1,177 site tiles, subtrees of
    about 100 (3 blocks per SIMD) and 40 updates (8 per SIMD)."""
    say("split 256 x 20000")
    prob = synth.make_problem(256, 20000, 4, kind="aa", ncat=4)
    _check(prob, 4, 9, splits=(3, 8))


def test_split_protein_22_states():
    say("split protein")
    rng = np.random.default_rng(7)
    rooted, tree, nw = synth.yule_tree(80, rng, 0.15, newick=True)
    m = synth.Model(rng.lognormal(0, 1, 231), rng.dirichlet(np.ones(22)), 0.7, 4, Protein)
    aln = synth.simulate(rooted, m, 700, rng)
    aln.compact()
    prob = synth.Problem(tree, rooted, aln, m, 700, nw)
    _check(prob, 4, 7)


@pytest.mark.parametrize("kind", ["aa", "dna"])
def test_split_polytomies(kind):
    """Nodes of degree 5 and more (identity children) and a re-rooting at the
    largest one."""
    say("split polytomies", kind)
    prob = synth.make_problem(110, 900 if kind == "aa" else 2500, 8, kind=kind, ncat=4)
    rng = np.random.default_rng(8)
    trees = [prob.tree, prob.fresh_tree()]
    edges0 = list(trees[0].edges)
    inner = [i for i, e in enumerate(edges0)
             if e.ring.node.degree > 1 and e.ring.other.node.degree > 1]
    zero = rng.choice(inner, size=20, replace=False).tolist()
    for t in trees:
        edges = list(t.edges)
        for i in zero:
            edges[i].length = 0.0
        assert t.collapse() == len(zero)
    assert max(n.degree for n in trees[0].nodes) >= 5
    _check(prob, 4, 8, depths=(0, 2), tree=trees[1])


def test_split_batch_of_trees():
    """Several trees in one batch: amino-acid trees of two component classes
    (their subtrees in each class's first launch, their upper updates in a
    second launch per class after every first one), a tree too small to
    split (whole, in the first launch) and a DNA tree (split likewise)."""
    say("split batch")
    spec = [(64, 600, "aa", 4), (100, 900, "aa", 4), (40, 300, "aa", 4), (130, 1200, "aa", 4),
            (12, 200, "aa", 4), (80, 900, "aa", 1), (50, 2000, "dna", 4)]
    probs = [synth.make_problem(n, s, 20 + i, kind=k, ncat=c) for i, (n, s, k, c) in enumerate(spec)]
    liks = []
    for p, (_, _, _, c) in zip(probs, spec):
        L = Lik(p.tree, p.alignment, ncat=c)
        synth.configure(L, p.model)
        liks.append(L)
    ctx = nat.get_context(0)
    out = {}
    try:
        for split in (0, 3):
            ctx.set_tuning(nat.CRUX_TUNE_SPLIT, split)
            for L in liks:
                L.flush()
            roots = [L._ready(None) for L in liks]
            vals = nat.batch_wait(ctx, nat.batch_submit(ctx, liks, roots))
            out[split] = (np.array(vals), ctx.last_launches()[0],
                          [np.array(L.siteLnLs()) for L in liks])
    finally:
        ctx.set_tuning(nat.CRUX_TUNE_SPLIT, 3)
    # One launch per class unsplit; one more per class split.
    assert out[0][1] == 3 and out[3][1] == 6, (out[0][1], out[3][1])
    assert np.array_equal(out[0][0].view(np.uint64), out[3][0].view(np.uint64))
    for a, b in zip(out[0][2], out[3][2]):
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    for p, (_, _, _, c), v in zip(probs, spec, out[3][0]):
        ora = OracleLik(p.fresh_tree(), p.alignment, ncat=c, nthreads=4)
        synth.configure(ora, p.model)
        assert abs(v - ora.lnL()) <= REL * abs(v)
