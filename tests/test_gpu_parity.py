"""Device likelihood vs the FP64 oracle and the reference's golden values.

Tolerance: the device path stores CLVs in FP32 with power-of-two site
scaling and sums site lnLs in compensated FP64, so it is held to
|dlnL| <= 1e-6 * |lnL| (+1e-9 absolute for lnL near 0) against the FP64
oracle, and to the printed precision of the reference's golden outputs.
Site lnLs are held to 1e-5 absolute + 1e-6 relative.
"""

import math
import sys

import numpy as np
import pytest

from crux_amd.alignment import Alignment, AminoAcid20, Protein
from crux_amd.lik import Lik
from crux_amd import synth
from crux_amd import _native as nat
from oracle.oracle import OracleLik
from tests.golden_util import run_case, tol

pytestmark = pytest.mark.gpu

REL = 1e-6


def close(a, b, rel=REL, abs_=1e-9):
    if math.isinf(a) or math.isinf(b):
        return a == b
    return abs(a - b) <= rel * abs(b) + abs_


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_golden_cases(golden):
    fails = []
    for case in golden:
        say("golden", case["name"])
        approx = case.get("approx_tree", False)
        for ev, lnl, sites in run_case(case, Lik):
            if "lnL_below" in ev:
                if not lnl < ev["lnL_below"]:
                    fails.append((case["name"], lnl))
                continue
            if "sig" in ev:
                t = 0.5 * 10.0 ** (math.floor(math.log10(abs(ev["lnL"]))) - ev["sig"] + 1)
                t = max(t, REL * abs(ev["lnL"]))
            else:
                t = tol(ev, approx=approx) + REL * abs(ev["lnL"])
            if abs(lnl - ev["lnL"]) > t:
                fails.append((case["name"], lnl, ev["lnL"], t))
            if sites is not None:
                ts = tol(ev, "site_dec", approx) + 1e-6
                assert len(sites) == len(ev["sites"])
                for a, b in zip(sites, ev["sites"]):
                    if abs(a - b) > ts:
                        fails.append((case["name"], "site", a, b))
    assert not fails, fails


def _pair(prob, **kw):
    """A device Lik and an oracle Lik on independent copies of the problem."""
    t_dev = prob.tree
    t_ora = prob.fresh_tree() if prob.newick else prob.tree.dup()
    dev = Lik(t_dev, prob.alignment, **kw)
    ora = OracleLik(t_ora, prob.alignment, nthreads=4, **kw)
    synth.configure(dev, prob.model)
    synth.configure(ora, prob.model)
    return dev, ora


@pytest.mark.parametrize("ntaxa,nsites,seed,kind", [
    (5, 50, 11, "dna"), (16, 1000, 1, "dna"), (40, 3000, 12, "dna"), (64, 700, 13, "dna"),
    (12, 400, 14, "aa"), (30, 1500, 15, "aa"),
])
def test_random_trees_vs_oracle(ntaxa, nsites, seed, kind):
    say("random", ntaxa, nsites, seed, kind)
    prob = synth.make_problem(ntaxa, nsites, seed, kind=kind)
    dev, ora = _pair(prob, ncat=4)
    a, b = dev.lnL(), ora.lnL()
    assert close(a, b), (a, b)
    sd, so = np.array(dev.siteLnLs()), np.array(ora.siteLnLs())
    assert np.allclose(sd, so, rtol=1e-6, atol=1e-5)
    # Every node as root, in a different order than the oracle visits them.
    nodes_d = list(dev.tree.nodes)
    nodes_o = list(ora.tree.nodes)
    idx = list(range(len(nodes_d)))[::-3]
    for i in idx:
        a = dev.lnL(nodes_d[i])
        assert close(a, b), (i, a, b)


@pytest.mark.parametrize("ncat,kind", [(1, "dna"), (2, "dna"), (3, "dna"), (6, "dna"), (12, "dna"),
                                       (1, "aa"), (2, "aa"), (3, "aa"), (6, "aa"), (1, "protein"),
                                       (2, "protein"), (6, "protein")])
def test_state_and_component_classes(ncat, kind):
    # Every component class of the 4-state kernels (1, 2, 4 with an idle
    # component, 8 with two idle, 16 with four) and of the 20/22-state kernel
    # (1, 2, 4 with an idle component, 8), with every other node as root: the
    # roots after the first re-evaluate from cached CLVs read back from HBM.
    say("classes", ncat, kind)
    rng = np.random.default_rng(60 + ncat)
    if kind in ("dna", "aa"):
        n = 70 if kind == "dna" else 18
        prob = synth.make_problem(n, 600, 60 + ncat, kind=kind, ncat=ncat)
    else:
        # From Newick, so that both trees list their edges in one order.
        rooted, tree, nw = synth.yule_tree(14, rng, 0.15, newick=True)
        m = synth.Model(rng.lognormal(0, 1, 231), rng.dirichlet(np.ones(22)), 0.6, ncat, Protein)
        aln = synth.simulate(rooted, m, 500, rng)
        aln.compact()
        prob = synth.Problem(tree, rooted, aln, m, 500, nw)
    dev, ora = _pair(prob, ncat=ncat)
    a, b = dev.lnL(), ora.lnL()
    assert close(a, b), (a, b)
    sd, so = np.array(dev.siteLnLs()), np.array(ora.siteLnLs())
    assert np.allclose(sd, so, rtol=1e-6, atol=1e-5)
    for node in list(dev.tree.nodes)[::2]:
        assert close(dev.lnL(node), b), node
    # One branch changed: the path to it is recomputed, the rest read back.
    e_d, e_o = list(dev.tree.edges)[7], list(ora.tree.edges)[7]
    e_d.length *= 1.7
    e_o.length *= 1.7
    a, b = dev.lnL(e_d.ring.node), ora.lnL(e_o.ring.node)
    assert close(a, b), (a, b)


def test_aa_many_tips():
    # More tips than one staged chunk of tip codes (64): later chunks are
    # staged where the plan reaches them.
    say("aa many tips")
    prob = synth.make_problem(150, 400, 23, kind="aa")
    dev, ora = _pair(prob, ncat=4)
    a, b = dev.lnL(), ora.lnL()
    assert close(a, b), (a, b)
    sd, so = np.array(dev.siteLnLs()), np.array(ora.siteLnLs())
    assert np.allclose(sd, so, rtol=1e-6, atol=1e-5)


def test_aa_partial_reevaluation():
    # Branch-length changes on a 20-state tree: each evaluation recomputes the
    # path to the changed edge, the siblings' CLVs come from HBM.
    prob = synth.make_problem(40, 900, 22, kind="aa")
    dev, ora = _pair(prob, ncat=4)
    dev.lnL()
    rng = np.random.default_rng(6)
    edges_d = list(dev.tree.edges)
    edges_o = list(ora.tree.edges)
    for step in range(10):
        k = int(rng.integers(len(edges_d)))
        f = math.exp(2 * math.log(1.6) * (rng.random() - 0.5))
        edges_d[k].length = edges_d[k].length * f
        edges_o[k].length = edges_o[k].length * f
        a = dev.lnL(edges_d[k].ring.node)
        b = ora.lnL(edges_o[k].ring.node)
        assert close(a, b), (step, a, b)


def test_compute_dtypes_full_size_c1():
    # BASELINE config 1 shape: 16 taxa, 1000 sites, GTR+G4 alpha=0.5.
    prob = synth.make_problem(16, 1000, 1)
    dev, ora = _pair(prob, ncat=4)
    a, b = dev.lnL(), ora.lnL()
    say("C1", a, b, abs(a - b) / abs(b))
    assert close(a, b)


def test_partial_reevaluation_matches_fresh():
    prob = synth.make_problem(48, 2000, 21)
    dev, ora = _pair(prob, ncat=4)
    dev.lnL()
    rng = np.random.default_rng(5)
    edges_d = list(dev.tree.edges)
    edges_o = list(ora.tree.edges)
    for step in range(12):
        k = int(rng.integers(len(edges_d)))
        f = math.exp(2 * math.log(1.6) * (rng.random() - 0.5))
        edges_d[k].length = edges_d[k].length * f
        edges_o[k].length = edges_o[k].length * f
        node_d = edges_d[k].ring.node
        node_o = edges_o[k].ring.node
        a = dev.lnL(node_d)
        b = ora.lnL(node_o)
        assert close(a, b), (step, a, b)
        if rng.random() < 0.5:
            dev.tree.base = node_d      # accept: keep the new root
            ora.tree.base = node_o
    # Cached call returns the same value without work.
    a1 = dev.lnL()
    a2 = dev.lnL()
    assert a1 == a2


def test_mixture_invar_median_and_clone():
    prob = synth.make_problem(24, 1200, 31)
    kw = dict(nmodels=2, ncat=4, catMedian=True, invar=True)
    dev, ora = _pair(prob, **kw)
    for L in (dev, ora):
        synth.configure(L, prob.model, 1)
        L.setRmult(1, 2.5)
        L.setWeight(0, 0.7)
        L.setWeight(1, 0.3)
        L.setWVar(0, 0.8)
        L.setWInvar(0, 0.2)
        L.setAlpha(1, 1.7)
    a, b = dev.lnL(), ora.lnL()
    assert close(a, b), (a, b)
    # The clone shares the tree through the other polarity; changing it must
    # not disturb the original's cache.
    c = dev.clone()
    c.setAlpha(0, 3.0)
    ora.setAlpha(0, 3.0)
    b2 = ora.lnL()
    assert close(c.lnL(), b2)
    assert dev.lnL() == a
    ora.setAlpha(0, prob.model.alpha)


def test_protein_22_states():
    rng = np.random.default_rng(3)
    rooted, tree = synth.yule_tree(10, rng, 0.2)
    m = synth.Model(rng.lognormal(0, 1, 231), rng.dirichlet(np.ones(22)), 0.8, 4, Protein)
    aln = synth.simulate(rooted, m, 300, rng)
    aln.compact()
    prob = synth.Problem(tree, rooted, aln, m, 300, "")
    dev, ora = _pair(prob, ncat=4)
    a, b = dev.lnL(), ora.lnL()
    assert close(a, b), (a, b)


def test_batch_and_plain_modes_agree():
    probs = [synth.make_problem(n, s, 40 + i, kind=k) for i, (n, s, k) in
             enumerate([(20, 900, "dna"), (33, 2500, "dna"), (14, 300, "aa"), (50, 4000, "dna")])]
    liks = []
    expect = []
    for p in probs:
        dev, ora = _pair(p, ncat=4)
        liks.append(dev)
        expect.append(ora.lnL())
    ctx = nat.get_context(0)
    for L in liks:
        L._ready(None)
    got = nat.batch_lnl(ctx, liks, [L.tree.base.id for L in liks])
    for g, e in zip(got, expect):
        assert close(g, e), (g, e)
    # Plain mode: one launch per node update.
    for L in liks:
        L.invalidate = True
        L._push_model()
    ctx.set_mode(1)
    try:
        for L, e in zip(liks, expect):
            L.invalidate = True
            L._model_dirty = True
            assert close(L.lnL(), e)
    finally:
        ctx.set_mode(0)


@pytest.mark.parametrize("nsites,ncat", [(12, 4), (31, 3), (33, 4), (64, 2), (70, 6), (130, 4)])
def test_pattern_counts_around_tiles(nsites, ncat):
    # Pattern counts below, at and just past the 32-site wave and 64-site
    # tile boundaries (a block's second wave may have no site at all), with
    # idle components in the last lane half; every other node as root
    # re-reads cached CLVs from HBM.
    say("tiles", nsites, ncat)
    prob = synth.make_problem(9, nsites, 500 + nsites, ncat=ncat, compact=False)
    dev, ora = _pair(prob, ncat=ncat)
    b = ora.lnL()
    assert close(dev.lnL(), b)
    sd, so = np.array(dev.siteLnLs()), np.array(ora.siteLnLs())
    assert np.allclose(sd, so, rtol=1e-6, atol=1e-5)
    for node in list(dev.tree.nodes)[::2]:
        assert close(dev.lnL(node), b), node


def test_large_batch_matches_single_evaluations():
    # A batch whose plan (operations, control words, P jobs) spans several
    # hundred KiB: the P-matrix launch copies it to the device over several
    # passes of its copy blocks.  Every lik's batched lnL equals its own
    # single evaluation exactly (same kernels, sums in tile order).
    say("large batch")
    probs = [synth.make_problem(60 + (i % 5) * 7, 700 + 37 * i, 300 + i) for i in range(40)]
    liks = []
    for p in probs:
        L = Lik(p.tree, p.alignment, ncat=4)
        synth.configure(L, p.model)
        liks.append(L)
    ctx = nat.get_context(0)
    for L in liks:
        L._ready(None)
    got = nat.batch_lnl(ctx, liks, [L.tree.base.id for L in liks])
    for L, g in zip(liks, got):
        L.flush()
        assert L.lnL() == g
    # And against the FP64 oracle for a few of them.
    for i in (0, 17, 39):
        ora = OracleLik(probs[i].fresh_tree(), probs[i].alignment, ncat=4, nthreads=4)
        synth.configure(ora, probs[i].model)
        assert close(got[i], ora.lnL())


def test_zero_length_conflict_is_neg_inf_or_tiny():
    aln = Alignment(">one\nACGT\n>two\nACGT\n>three\nAAAA\n")
    from crux_amd.tree import Tree
    t = Tree("(one:0.0,two:0.0,three:0.0);", rooted=False)
    L = Lik(t, aln)
    for n in t.nodes:
        assert L.lnL(n) < -50.0


def test_errors():
    from crux_amd.tree import Tree
    aln = Alignment(">a\nACGT\n>b\nACGA\n>c\nAAGT\n")
    with pytest.raises(ValueError):
        Lik(Tree("(a:1,b:1,c:1);"), aln)          # rooted
    t = Tree("(a:1,b:1,c:1);", rooted=False)
    L = Lik(t, aln)
    t.edges[0].length = -0.5
    with pytest.raises(ValueError):
        L.lnL()
    t.edges[0].length = 0.5
    assert math.isfinite(L.lnL())
    t2 = Tree("(a:1,b:1,x:1);", rooted=False)
    L2 = Lik(t2, aln)
    with pytest.raises(ValueError):
        L2.lnL()
    with pytest.raises(ValueError):
        L.setWInvar(0, 0.5)


def test_c2_scale_parity():
    # BASELINE config 2 scale: 128 taxa, 100k sites (~60k patterns), GTR+G4.
    prob = synth.make_problem(128, 100000, 2, mean_len=0.025)
    dev, ora = _pair(prob, ncat=4)
    ora.nthreads = 8
    a, b = dev.lnL(), ora.lnL()
    say("C2", prob.alignment.nchars, a, b, abs(a - b) / abs(b))
    assert close(a, b), (a, b)
    # Single-branch partial re-evaluation at full size.
    e_d = list(dev.tree.edges)[37]
    e_o = list(ora.tree.edges)[37]
    e_d.length *= 1.3
    e_o.length *= 1.3
    a, b = dev.lnL(e_d.ring.node), ora.lnL(e_o.ring.node)
    assert close(a, b), (a, b)
