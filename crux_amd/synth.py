"""Deterministic synthetic workloads for tests and benchmarks.

* Trees: pure-birth (Yule) rooted binary shapes -- start from two lineages and
  split a uniformly chosen tip until there are N -- with branch lengths drawn
  Exp(mean), then derooted into the unrooted form the likelihood takes (the
  root edge merge goes through float32 as in the reference's deroot).
* DNA models: GTR exchangeabilities ~ U(0.5, 2) in the order AC AG AT CG CT GT,
  frequencies ~ Dirichlet(5,5,5,5), discrete Gamma(alpha) with 4 categories.
* Amino-acid models: 190 exchangeabilities ~ LogNormal(0, 1), frequencies ~
  Dirichlet(1 x 20), 20-state alphabet.
* Alignments: sites simulated down the rooted tree (each site drawn a Gamma
  category uniformly), then compressed to unique patterns with integer
  weights.

Everything is seeded with numpy's PCG64 generator.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import _native as nat
from .alignment import Alignment, AminoAcid20, Dna
from .tree import Edge, Node, Tree

__all__ = ["Model", "yule_newick", "yule_tree", "dna_model", "aa_model", "simulate",
           "make_problem", "Problem"]


@dataclass
class Model:
    rtri: np.ndarray
    pi: np.ndarray
    alpha: float = math.inf
    ncat: int = 1
    charType: type = Dna

    @property
    def dim(self):
        return len(self.pi)


def dna_model(rng, alpha=0.5, ncat=4):
    return Model(rng.uniform(0.5, 2.0, 6), rng.dirichlet(np.full(4, 5.0)), alpha, ncat, Dna)


def aa_model(rng, alpha=0.5, ncat=4):
    return Model(rng.lognormal(0.0, 1.0, 190), rng.dirichlet(np.ones(20)), alpha, ncat,
                 AminoAcid20)


def yule_newick(ntaxa, rng, mean_len=0.1, prefix="T"):
    """Newick string of a Yule tree with Exp(mean_len) branch lengths."""
    assert ntaxa >= 2
    # children[i] = [] for tips; node 0 is the root.
    children = {0: [1, 2]}
    tips = [1, 2]
    nxt = 3
    while len(tips) < ntaxa:
        k = int(rng.integers(len(tips)))
        t = tips[k]
        children[t] = [nxt, nxt + 1]
        tips[k] = nxt
        tips.append(nxt + 1)
        nxt += 2
    lens = rng.exponential(mean_len, nxt)
    names = {}
    order = list(tips)
    perm = rng.permutation(len(order))
    for i, t in enumerate(order):
        names[t] = "%s%d" % (prefix, perm[i])

    # Iterative post-order to avoid deep recursion on caterpillar shapes.
    out = {}
    stack = [(0, False)]
    while stack:
        n, done = stack.pop()
        if n not in children:
            out[n] = names[n]
        elif done:
            out[n] = "(" + ",".join("%s:%.17g" % (out[c], lens[c]) for c in children[n]) + ")"
        else:
            stack.append((n, True))
            stack.extend((c, False) for c in children[n])
    return out[0] + ";"


def yule_tree(ntaxa, rng, mean_len=0.1, newick=False):
    """(rooted Tree, unrooted Tree) with the same shape and lengths (and the
    Newick string when ``newick``)."""
    nw = yule_newick(ntaxa, rng, mean_len)
    if newick:
        return Tree(nw), Tree(nw, rooted=False), nw
    return Tree(nw), Tree(nw, rooted=False)


def simulate(rooted, model, nsites, rng, taxa=None):
    """Alignment simulated down a rooted tree under ``model``.

    Each site draws a Gamma category; along an edge the child state of a
    site is the first state whose cumulative P(t) row (of the site's
    category and parent state) reaches a uniform draw.  The inversion starts
    from a guide table of 4K cells per row (the first state whose cumulative
    probability reaches the cell's lower end) and steps forward from there,
    about one step per site, so an edge costs O(sites) rather than
    O(sites x states)."""
    V, Vi, w, pin, qn = nat.q_decomp(model.rtri, model.pi)
    if model.ncat > 1 and not math.isinf(model.alpha):
        rates = nat.gamma_rates(model.alpha, model.ncat, False)
    else:
        rates = np.ones(1)
    cat = rng.integers(len(rates), size=nsites)
    states = np.frombuffer("".join(model.charType.get().pcodes()).encode(), dtype=np.uint8)
    K = len(model.pi)
    C = len(rates)
    M = 4 * K
    grid = np.arange(M) / M
    rowbase = cat * K
    seqs = {}
    base = rooted.base
    root_seq = rng.choice(K, size=nsites, p=pin)
    stack = [(r.other, root_seq) for r in base.ring] if base.ring is not None else []
    if base.taxon is not None:
        seqs[base.taxon] = root_seq
    while stack:
        ring, pseq = stack.pop()
        cdf = np.empty((C * K, K))
        for ci, r in enumerate(rates):
            P = nat.pt(V, Vi, w, ring.edge.length * r * qn)
            c = np.maximum.accumulate(np.cumsum(P, axis=1), axis=1)
            c[:, -1] = 1.0
            cdf[ci * K:(ci + 1) * K] = c
        guide = (cdf[:, None, :] < grid[None, :, None]).sum(axis=2)
        flat = cdf.ravel()
        row = rowbase + pseq
        u = rng.random(nsites)
        seq = guide[row, (u * M).astype(np.int64)]
        base_k = row * K
        # Step forward only the sites that are not there yet.
        act = np.nonzero(flat[base_k + seq] < u)[0]
        while act.size:
            seq[act] += 1
            act = act[flat[base_k[act] + seq[act]] < u[act]]
        if ring.node.taxon is not None:
            seqs[ring.node.taxon] = seq
        stack.extend((r.other, seq) for r in ring.siblings())
    names = taxa if taxa is not None else sorted(seqs, key=lambda s: (len(s), s))
    rows = np.stack([states[seqs[n]] for n in names])
    return Alignment(charType=model.charType, taxa=names, rows=rows)


@dataclass
class Problem:
    tree: Tree            # unrooted
    rooted: Tree
    alignment: Alignment  # compressed patterns
    model: Model
    nsites_raw: int
    newick: str = ""
    extra: dict = field(default_factory=dict)

    def fresh_tree(self):
        """Another unrooted copy built the same way as ``tree`` (same node and
        edge ids and ring order)."""
        return Tree(self.newick, rooted=False)


def make_problem(ntaxa, nsites, seed, kind="dna", alpha=0.5, ncat=4, mean_len=0.1,
                 compact=True):
    rng = np.random.default_rng(seed)
    rooted, unrooted, nw = yule_tree(ntaxa, rng, mean_len, newick=True)
    model = dna_model(rng, alpha, ncat) if kind == "dna" else aa_model(rng, alpha, ncat)
    aln = simulate(rooted, model, nsites, rng)
    if compact:
        aln.compact()
    return Problem(unrooted, rooted, aln, model, nsites, nw)


def configure(lik, model, m=0):
    """Set a Lik's model ``m`` to ``model`` (exchangeabilities, frequencies,
    Gamma shape)."""
    nr = len(model.rtri)
    lik.setRclass(m, list(range(nr)), [float(x) for x in model.rtri])
    for i, f in enumerate(model.pi):
        lik.setFreq(m, i, float(f))
    if model.ncat > 1 and not math.isinf(model.alpha):
        lik.setAlpha(m, model.alpha)
