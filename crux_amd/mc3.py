"""Batched Metropolis-coupled MCMC driver over the device likelihood.

The caller the hot path serves: MC3 chains, each proposing a move per step
on its own tree, with heated replicas that swap heats.  The reference
evaluates one proposal per lnL call (one Lik per chain, Chain.pyx:1816-1862);
here every chain's proposal of a step goes into ONE device pass
(``crux_batch_lnl``): all dirty node updates of all trees run in a single
traversal launch, so the per-launch and per-call costs are paid once per step
instead of once per chain.  ``Mc3Batch.run`` also overlaps the host with the
device: two halves of the runs alternate, one half's pass running while the
host settles and re-proposes the other.

Two proposal mixes:

* ``mix="brlen-nni"`` (the batched-MC3 workload): branch-length multipliers
  (Chain.pyx:438-475) and nearest-neighbour interchanges across a random
  internal edge (branch lengths unchanged, symmetric, Hastings term 0);
* ``mix="reference"``: the reference's advance0 (Chain.pyx:1816-1862) over
  its fourteen proposals with the reference's default weights
  (Mc3/__init__.pyx:276-289), disabled as initProps does
  (Mc3/__init__.pyx:904-947): parameter moves evaluated on the likelihood's
  clone (mixture weight, frequency, rate multiplier, relative rate, inverse
  Gamma shape, +I weights, Chain.pyx:72-437), branch lengths, extending TBR
  (Chain.pyx:477-819), and the reversible jumps (rate-class merge/split,
  polytomy merge/split, +G and +I on/off, equal/estimated frequencies,
  mixture add/remove, Chain.pyx:821-1815).  Starting points are drawn from
  the priors as randomLik / randomDnaQ do (Mc3/__init__.pyx:1457-1640).
  The mixture prior is truncated at the components a device likelihood
  holds (16 at 4 states, 8 at 20/22): an addition past that is rejected.

Acceptance follows the reference throughout: p = exp(((lnL1 - lnL) +
lnPrior) heat + lnHastings), accepted when p >= u; heats 1 / (1 + i 0.05)
for replica i (Chain.pyx:53); one heat swap attempt per run and step
between two replicas drawn from the run's swap generator, accepted with
exp((L_b - L_a) h_a + (L_a - L_b) h_b) (Chain.pyx:1871-1911).  Random
streams are numpy's, not the reference's SFMT.

    python -m crux_amd.mc3 [--runs 32] [--coupled 4] [--steps 50] [--small]
"""

from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import _native as nat
from . import synth
from .alignment import Dna
from .lik import Lik
from .tree import Edge, Node

# Proposal kinds in the reference's order (Chain.pxd:5-19).
PROPS = ("weight", "freq", "rmult", "rate", "rateShapeInv", "invar", "brlen", "etbr",
         "rateJump", "polytomyJump", "rateShapeInvJump", "invarJump", "freqJump", "mixtureJump")
# Their default relative weights (Mc3/__init__.pyx:276-289).
DEFAULT_PROPS = (30.0, 10.0, 10.0, 30.0, 10.0, 5.0, 20.0, 80.0, 10.0, 40.0, 10.0, 5.0, 10.0, 5.0)

BRLEN_LAMBDA = 2.0 * math.log(1.6)
BRLEN_PRIOR = 10.0
HEAT_DELTA = 0.05


class Mc3Params:
    """Tuning and prior parameters with the reference's defaults
    (Mc3/__init__.pyx:240-289); any of them can be given as a keyword."""

    def __init__(self, **kw):
        lam = 2.0 * math.log(1.6)
        self.nmodels = 1
        self.ncat = 4
        self.catMedian = False
        self.invar = False
        self.heatDelta = HEAT_DELTA
        self.weightLambda = lam
        self.freqLambda = lam
        self.rmultLambda = lam
        self.rateLambda = lam
        self.rateShapeInvLambda = lam
        self.invarLambda = lam
        self.brlenLambda = lam
        self.etbrPExt = 0.8
        self.rateShapeInvPrior = 1.0
        self.invarPrior = 0.5
        self.brlenPrior = BRLEN_PRIOR
        self.rateJumpPrior = 1.0
        self.polytomyJumpPrior = 1.0
        self.rateShapeInvJumpPrior = 1.0
        self.invarJumpPrior = 1.0
        self.freqJumpPrior = 1.0
        self.mixtureJumpPrior = 1.0 / 3.0
        # The reference's polytomy split leaves the proposal density of the
        # new edge's length out of its Hastings term (it computes the
        # Jacobian term and does not add it, Chain.pyx:1287-1291), while the
        # merge includes it; True adds it, which makes the two moves each
        # other's inverse (the topology prior is then sampled exactly).
        self.splitJacobian = False
        self.props = dict(zip(PROPS, DEFAULT_PROPS))
        for k, v in kw.items():
            if k == "props":
                self.props.update(v)
            elif hasattr(self, k):
                setattr(self, k, v)
            else:
                raise TypeError("unknown parameter %r" % k)

    def pdf(self, alignment):
        """Proposal probabilities for an alignment, with the proposals the
        reference's initProps disables set to 0 (Mc3/__init__.pyx:904-947)."""
        props = dict(self.props)
        nstates = alignment.charType.get().nstates
        if alignment.charType is not Dna:
            props["mixtureJump"] = 0.0
        if self.nmodels < 2 and props["mixtureJump"] == 0.0:
            props["weight"] = 0.0
            props["rmult"] = 0.0
        if nstates <= 2:
            props["rateJump"] = 0.0
        if alignment.ntaxa < 3:
            props["etbr"] = 0.0
        if alignment.ntaxa < 4:
            props["polytomyJump"] = 0.0
        if self.ncat < 2:
            props["rateShapeInv"] = 0.0
            props["rateShapeInvJump"] = 0.0
        if not self.invar:
            props["invar"] = 0.0
            props["invarJump"] = 0.0
        total = sum(props.values())
        if total == 0.0:
            raise ValueError("No proposals are enabled")
        return {k: props[k] / total for k in PROPS}


_rclass_tables: dict = {}


def rate_classes(rlen):
    """Every grouping of ``rlen`` relative rates into classes, as rclass
    lists (restricted growth strings), by number of classes: table[k - 1]
    holds the groupings into k classes in lexicographic order (the
    reference's DNA table, Mc3/__init__.pyx:100-240, is rlen 6)."""
    if rlen not in _rclass_tables:
        out = [[] for _ in range(rlen)]

        def grow(prefix, top):
            if len(prefix) == rlen:
                out[top].append(list(prefix))
                return
            for c in range(top + 2):
                grow(prefix + [c], max(top, c))

        grow([0], 0)
        _rclass_tables[rlen] = out
    return _rclass_tables[rlen]


def _exp_draw(rng):
    """-log(1 - u): an Exp(1) variate as the reference draws it."""
    return -math.log(1.0 - rng.random())


def _resolution(u, prior, n):
    """Index k in 1..n drawn with P(k) proportional to prior^-(k - 1), by the
    reference's cumulative walk (Mc3/__init__.pyx:1481-1521)."""
    norm = 0.0
    factor = 1.0
    for _ in range(n):
        norm += 1.0 / factor
        factor *= prior
    cum = 0.0
    factor = 1.0
    for k in range(1, n):
        cum += (1.0 / factor) / norm
        if cum >= u:
            return k
        factor *= prior
    return n


def random_q(lik, m, params, pdf, rng):
    """Draw model m's parameters from their priors (randomDnaQ,
    Mc3/__init__.pyx:1457-1592)."""
    nstates = lik.dim
    rlen = nstates * (nstates - 1) // 2
    if lik.nmodels() > 1:
        if pdf["weight"] > 0.0:
            lik.setWeight(m, _exp_draw(rng))
        if pdf["rmult"] > 0.0:
            lik.setRmult(m, _exp_draw(rng))
    if pdf["freq"] > 0.0:
        if pdf["freqJump"] > 0.0:
            estimate = _resolution(rng.random(), params.freqJumpPrior, 2) == 2
        else:
            estimate = True
        if estimate:
            for i in range(nstates):
                lik.setFreq(m, i, _exp_draw(rng))
    if pdf["rate"] > 0.0:
        if pdf["rateJump"] > 0.0 and rlen <= 6:
            nrates = min(rlen, _resolution(rng.random(), params.rateJumpPrior, rlen))
            table = rate_classes(rlen)[nrates - 1]
            rclass = table[int(rng.integers(len(table)))]
        else:
            # The reference's grouping table covers DNA only; other state
            # counts start with every rate in its own class.
            nrates = rlen
            rclass = list(range(rlen))
        lik.setRclass(m, rclass)
        for i in range(nrates):
            lik.setRate(m, i, _exp_draw(rng))
    if pdf["rateShapeInv"] > 0.0 and params.ncat > 1:
        if pdf["rateShapeInvJump"] > 0.0:
            gamma = _resolution(rng.random(), params.rateShapeInvJumpPrior, 2) == 2
        else:
            gamma = True
        if gamma:
            lik.setAlpha(m, _exp_draw(rng) * params.rateShapeInvPrior)
    if pdf["invar"] > 0.0 and params.invar:
        if pdf["invarJump"] > 0.0:
            plus_i = _resolution(rng.random(), params.invarJumpPrior, 2) == 2
        else:
            plus_i = True
        if plus_i:
            lik.setWVar(m, _exp_draw(rng) * (1.0 - params.invarPrior))
            lik.setWInvar(m, _exp_draw(rng) * params.invarPrior)


def _model_comps(params):
    """Components one model of the sampler adds (+I counts as one)."""
    return params.ncat + (1 if params.invar else 0)


def random_lik(tree, alignment, params, rng, context=None):
    """A Lik on ``tree`` (unrooted) with branch lengths and model parameters
    drawn from their priors (randomLik, Mc3/__init__.pyx:1594-1640)."""
    pdf = params.pdf(alignment)
    for e in tree.edges:
        e.length = _exp_draw(rng) / params.brlenPrior
    nmodels = params.nmodels
    if pdf["mixtureJump"] > 0.0:
        # Geometric, truncated at the components a likelihood may hold.
        nmodels = 1
        while rng.random() > params.mixtureJumpPrior and \
                (nmodels + 1) * _model_comps(params) <= nat.MAX_COMPONENTS[alignment.charType.get().nstates]:
            nmodels += 1
    lik = Lik(tree, alignment, nmodels, params.ncat, params.catMedian, params.invar,
              context=context)
    for m in range(nmodels):
        random_q(lik, m, params, pdf, rng)
    return lik


def _index(rng, n):
    """A uniform index in [0, n) from one uniform draw (numpy's scalar
    Generator.integers costs several times as much per call)."""
    return min(int(rng.random() * n), n - 1)


def _accept_prob(x):
    """exp(x) for an acceptance test, NaN (a failed evaluation) as 0."""
    if x != x:
        return 0.0
    return math.inf if x > 700.0 else math.exp(x)


def _move(e, old, new):
    """Move the end of edge e at node ``old`` to node ``new``, keeping which
    of its rings is e.ring (so cached CLVs keep their polarity), as the
    reference does at every re-attachment."""
    if e.ring.node is old:
        other = e.ring.other.node
        e.detach()
        e.attach(new, other)
    else:
        other = e.ring.node
        e.detach()
        e.attach(other, new)


class _Pending:
    """A proposal waiting for its lnL."""

    __slots__ = ("kind", "lik", "root", "lnPrior", "lnProp", "accept", "reject")

    def __init__(self, kind, lik, root, lnPrior, lnProp, accept=None, reject=None):
        self.kind = kind
        self.lik = lik
        self.root = root
        self.lnPrior = lnPrior
        self.lnProp = lnProp
        self.accept = accept
        self.reject = reject


class Chain:
    """One heated replica: its own tree and Lik (plus the Lik's clone for
    parameter moves) on a shared alignment."""

    def __init__(self, lik, run, index, seed, params=None, mix="brlen-nni"):
        self.run = run
        self.index = index
        self.params = params if params is not None else Mc3Params()
        self.heat = 1.0 / (1.0 + index * self.params.heatDelta)
        self.rng = np.random.default_rng(seed)
        self.lik = lik
        self.tree = lik.tree
        self.lnL = None
        self.mix = mix
        if mix == "reference":
            self.pdf = self.params.pdf(lik.alignment)
            cum, self.cdf = 0.0, []
            for k in PROPS:
                cum += self.pdf[k]
                self.cdf.append((cum, k))
        elif mix != "brlen-nni":
            raise ValueError("unknown proposal mix %r" % mix)
        kinds = PROPS + ("nni", "swap")
        self.accepts = {k: 0 for k in kinds}
        self.rejects = {k: 0 for k in kinds}
        self._pend = None
        # Branch-length and NNI moves keep the set of edges and which of them
        # are internal, so under that mix both lists are taken once instead
        # of re-walking the tree every step.
        self._fixed_edges = None
        self._inner = None
        if mix == "brlen-nni":
            self._fixed_edges = list(self.tree.edges)
            self._inner = [e for e in self._fixed_edges
                           if e.ring.node.degree >= 3 and e.ring.other.node.degree >= 3]

    @classmethod
    def for_problem(cls, prob, run, index, seed, ncat=4, context=None):
        """A replica of the batched-MC3 workload: the problem's generating
        model and tree, branch-length and NNI moves."""
        tree = prob.fresh_tree()
        lik = Lik(tree, prob.alignment, ncat=ncat, context=context)
        synth.configure(lik, prob.model)
        return cls(lik, run, index, seed)

    # -- the step -----------------------------------------------------------
    def propose(self, p_nni=0.2):
        """Apply a proposal; return (likelihood, root node) to evaluate it at
        (root None: the tree's base)."""
        if self.mix == "brlen-nni":
            pend = None
            if self.rng.random() < p_nni:
                pend = self._p_nni()
            if pend is None:
                pend = self._p_brlen()
        else:
            # advance0: draw a kind until one applies (Chain.pyx:1824-1861).
            pend = None
            while pend is None:
                u = self.rng.random()
                kind = next((k for c, k in self.cdf if u < c), self.cdf[-1][1])
                pend = getattr(self, "_p_" + kind)()
        self._pend = pend
        return pend.lik, pend.root

    def settle(self, lnL1):
        """Accept or reject the pending proposal given its lnL."""
        pd = self._pend
        self._pend = None
        u = self.rng.random()
        p = _accept_prob(((lnL1 - self.lnL) + pd.lnPrior) * self.heat + pd.lnProp)
        if p >= u:
            self.lnL = lnL1
            self.lik = pd.lik
            if pd.accept is not None:
                pd.accept()
            self.accepts[pd.kind] += 1
        else:
            if pd.reject is not None:
                pd.reject()
            self.rejects[pd.kind] += 1

    # -- helpers ------------------------------------------------------------
    def _draw_lnm(self, lam):
        return lam * (self.rng.random() - 0.5)

    def _pick_model(self, ok):
        """Uniform choice among the models for which ok(m) holds (None if
        none): the reference's mChoice mapping."""
        ms = [m for m in range(self.lik.nmodels()) if ok(m)]
        if not ms:
            return None
        return ms[int(self.rng.integers(len(ms)))]

    def _freqs_equal(self, lik, m):
        f = lik.getFreq(m, 0)
        return all(lik.getFreq(m, i) == f for i in range(1, lik.dim))

    def _n_freqs_estim(self, lik):
        return sum(1 for m in range(lik.nmodels()) if not self._freqs_equal(lik, m))

    @staticmethod
    def _n_rates_estim(lik):
        return sum(1 for m in range(lik.nmodels()) if lik.getNrates(m) != 1)

    @staticmethod
    def _n_gamma(lik):
        return sum(1 for m in range(lik.nmodels()) if lik.getAlpha(m) != math.inf)

    @staticmethod
    def _n_invar(lik):
        return sum(1 for m in range(lik.nmodels()) if lik.getWInvar(m) != 0.0)

    # -- parameter moves on the clone (Chain.pyx:72-437) ---------------------
    def _p_weight(self):
        if self.lik.nmodels() == 1:
            return None
        lik1 = self.lik.clone()
        m = int(self.rng.integers(lik1.nmodels()))
        lnM = self._draw_lnm(self.params.weightLambda)
        w0 = lik1.getWeight(m)
        w1 = w0 * math.exp(lnM)
        lik1.setWeight(m, w1)
        return _Pending("weight", lik1, None, -(w1 - w0), lnM)

    def _p_freq(self):
        if self.pdf["freqJump"] != 0.0:
            m = self._pick_model(lambda i: not self._freqs_equal(self.lik, i))
            if m is None:
                return None
        else:
            m = int(self.rng.integers(self.lik.nmodels()))
        lik1 = self.lik.clone()
        f = int(self.rng.integers(lik1.dim))
        lnM = self._draw_lnm(self.params.freqLambda)
        f0 = lik1.getFreq(m, f)
        f1 = f0 * math.exp(lnM)
        lik1.setFreq(m, f, f1)
        return _Pending("freq", lik1, None, -(f1 - f0), lnM)

    def _p_rmult(self):
        if self.lik.nmodels() == 1:
            return None
        lik1 = self.lik.clone()
        m = int(self.rng.integers(lik1.nmodels()))
        lnM = self._draw_lnm(self.params.rmultLambda)
        r0 = lik1.getRmult(m)
        r1 = r0 * math.exp(lnM)
        lik1.setRmult(m, r1)
        return _Pending("rmult", lik1, None, -(r1 - r0), lnM)

    def _p_rate(self):
        m = self._pick_model(lambda i: self.lik.getNrates(i) != 1)
        if m is None:
            return None
        lik1 = self.lik.clone()
        r = int(self.rng.integers(lik1.getNrates(m)))
        lnM = self._draw_lnm(self.params.rateLambda)
        r0 = lik1.getRate(m, r)
        r1 = r0 * math.exp(lnM)
        lik1.setRate(m, r, r1)
        return _Pending("rate", lik1, None, -(r1 - r0), lnM)

    def _p_rateShapeInv(self):
        m = self._pick_model(lambda i: self.lik.getAlpha(i) != math.inf)
        if m is None:
            return None
        lik1 = self.lik.clone()
        lnM = self._draw_lnm(self.params.rateShapeInvLambda)
        a0inv = 1.0 / lik1.getAlpha(m)
        a1inv = a0inv * math.exp(lnM)
        lik1.setAlpha(m, 1.0 / a1inv)
        lnPrior = -self.params.rateShapeInvPrior * (a1inv - a0inv)
        return _Pending("rateShapeInv", lik1, None, lnPrior, lnM)

    def _p_invar(self):
        n = self._n_invar(self.lik)
        if n == 0:
            return None
        # The reference draws one of the n +I models but maps the draw over
        # every model (its filter, wInvar != INFINITY, holds for all of them:
        # Chain.pyx:388-393), so model index = the draw.
        m = int(self.rng.integers(n))
        lik1 = self.lik.clone()
        which = int(self.rng.integers(2))
        lnM = self._draw_lnm(self.params.invarLambda)
        if which == 0:
            w0 = lik1.getWVar(m)
            w1 = w0 * math.exp(lnM)
            lik1.setWVar(m, w1)
            lnPrior = -(1.0 - self.params.invarPrior) * (w1 - w0)
        else:
            w0 = lik1.getWInvar(m)
            w1 = w0 * math.exp(lnM)
            lik1.setWInvar(m, w1)
            lnPrior = -self.params.invarPrior * (w1 - w0)
        return _Pending("invar", lik1, None, lnPrior, lnM)

    # -- tree moves ---------------------------------------------------------
    def _edges(self):
        return self._fixed_edges if self._fixed_edges is not None else self.tree.edges

    def _p_brlen(self):
        """Chain.pyx:438-475."""
        edges = self._edges()
        edge = edges[_index(self.rng, len(edges))]
        lnM = self._draw_lnm(self.params.brlenLambda)
        v0 = edge.length
        v1 = v0 * math.exp(lnM)
        edge.length = v1
        node = edge.ring.node

        def accept():
            # Re-base so the next unrooted call reuses the cache
            # (Chain.pyx:466-469).
            self.tree.setBase(node)

        def reject():
            edge.length = v0

        return _Pending("brlen", self.lik, node, -self.params.brlenPrior * (v1 - v0), lnM,
                        accept, reject)

    def _p_nni(self):
        """Nearest-neighbour interchange across a random internal edge."""
        inner = self._inner if self._inner is not None else \
            [e for e in self._edges() if e.ring.node.degree >= 3 and e.ring.other.node.degree >= 3]
        if not inner:
            return None
        e = inner[_index(self.rng, len(inner))]
        u, v = e.ring.node, e.ring.other.node
        ra = [r for r in u.ring if r is not e.ring]
        rb = [r for r in v.ring if r is not e.ring.other]
        a = ra[_index(self.rng, len(ra))]
        b = rb[_index(self.rng, len(rb))]
        _swap_subtrees(a, b, u, v)
        # a now hangs on v and b on u: swap back on rejection.
        return _Pending("nni", self.lik, u, 0.0, 0.0, None,
                        lambda: _swap_subtrees(a, b, v, u))

    def _extend(self, ring0):
        """One end of an eTBR move (Chain.pyx:521-600 for X, 602-681 for Y):
        from node n0 = ring0.node, step along a random edge eX away from
        ring0's edge, then keep extending past each node with probability
        etbrPExt; move n0's other edges to n1 (the node behind eX) and
        re-attach eX and the edge eR reached last so that n0 sits on eR.
        Returns the state needed for the Hastings term and the undo, or None
        when n0 is a leaf (the constrained start)."""
        n0 = ring0.node
        if n0.degree <= 1:
            return None
        rng = self.rng
        pext = self.params.etbrPExt
        r = ring0.next
        for _ in range(int(rng.integers(n0.degree - 1))):
            r = r.next
        eX = r.edge
        rfrom = r.other
        n1 = nR0 = rfrom.node
        nR1 = n0
        nR0_uncon = nR0.degree > 1
        while nR0_uncon:
            if pext < rng.random():
                break
            r = rfrom.next
            for _ in range(int(rng.integers(nR0.degree - 1))):
                r = r.next
            rfrom = r.other
            nR1 = nR0
            nR0 = rfrom.node
            nR0_uncon = nR0.degree > 1
        st = {"n0": n0, "n1": n1, "eX": eX, "nR0": nR0, "nR1": nR1, "nR0_uncon": nR0_uncon,
              "sibs": [], "eR": None}
        if nR0 is not n1:
            sibs = []
            r = ring0.next
            while r is not ring0:
                rnext = r.next
                e = r.edge
                if e is not eX:
                    sibs.append(e)
                    _move(e, n0, n1)
                r = rnext
            _move(eX, n1, nR1)
            eR = rfrom.edge
            _move(eR, nR1, n0)
            st["sibs"] = sibs
            st["eR"] = eR
        return st

    @staticmethod
    def _unextend(st):
        """Undo _extend (Chain.pyx:780-846)."""
        if st["nR0"] is st["n1"]:
            return
        n0, n1, nR1 = st["n0"], st["n1"], st["nR1"]
        _move(st["eR"], n0, nR1)
        _move(st["eX"], nR1, n1)
        for e in st["sibs"]:
            _move(e, n1, n0)

    def _p_etbr(self):
        """Extending TBR (Chain.pyx:477-819): both ends of a random edge eA
        are extended and re-attached, and eA, eX and eY get branch-length
        multipliers."""
        assert self.tree.ntaxa >= 3
        edges = self._edges()
        eA = edges[int(self.rng.integers(len(edges)))]
        sx = self._extend(eA.ring)
        sy = self._extend(eA.ring.other)
        nX0 = eA.ring.node
        P = self.params
        lnProp = 0.0
        lnPrior = 0.0
        undo = []
        for e in [eA] + [s["eX"] for s in (sx, sy) if s is not None]:
            lnM = self._draw_lnm(P.brlenLambda)
            lnProp += lnM
            v0 = e.length
            v1 = v0 * math.exp(lnM)
            e.length = v1
            lnPrior += -P.brlenPrior * (v1 - v0)
            undo.append((e, v0))
        # Hastings terms of the extension (Chain.pyx:754-767): only the
        # constrained/unconstrained mixed cases contribute.
        log1m = math.log(1.0 - P.etbrPExt)
        for s, other in ((sx, sy), (sy, sx)):
            if s is not None and s["nR0"] is not s["n1"]:
                if other is not None:
                    if not s["nR0_uncon"]:
                        lnProp += log1m
                elif s["nR0_uncon"]:
                    lnProp -= log1m

        def accept():
            self.tree.setBase(nX0)

        def reject():
            for e, v0 in undo:
                e.length = v0
            for s in (sx, sy):
                if s is not None:
                    self._unextend(s)

        return _Pending("etbr", self.lik, nX0, lnPrior, lnProp, accept, reject)

    # -- reversible jumps ---------------------------------------------------
    def _p_rateJump(self):
        """Rate-class merge or split (Chain.pyx:821-1074)."""
        assert self.lik.dim > 2
        m = int(self.rng.integers(self.lik.nmodels()))
        rclass = list(self.lik.getRclass(m))
        nrates = self.lik.getNrates(m)
        if nrates == 1:
            merge = False
        elif nrates == len(rclass):
            merge = True
        else:
            merge = bool(self.rng.integers(2))
        return self._rate_merge(m, rclass, nrates) if merge else \
            self._rate_split(m, rclass, nrates)

    def _rate_merge(self, m, rclass, nrates):
        rng = self.rng
        lik1 = self.lik.clone()
        a = int(rng.integers(nrates))
        b = int(rng.integers(nrates - 1))
        if b >= a:
            b += 1
        if a > b:
            a, b = b, a
        na = sum(1 for r in rclass if r == a)
        nb = sum(1 for r in rclass if r == b)
        rateA, rateB = lik1.getRate(m, a), lik1.getRate(m, b)
        rateAB = (na * rateA + nb * rateB) / (na + nb)
        rclass = [a if r == b else (r - 1 if r > b else r) for r in rclass]
        rates = [lik1.getRate(m, r) for r in range(nrates)]
        rates[a] = rateAB
        rates.pop(b)
        lik1.setRclass(m, rclass, rates)
        pr = self.pdf["rateJump"]
        if nrates == 2:
            p_split = pr
            if self._n_rates_estim(lik1) == 0:
                p_split /= (1.0 - self.pdf["rate"])
            p_merge = pr / 2.0
        elif nrates == len(rclass):
            p_split, p_merge = pr / 2.0, pr
        else:
            p_split = p_merge = pr / 2.0
        lnPrior = math.log(self.params.rateJumpPrior) + math.log(p_split / p_merge)
        ns = [0] * nrates
        rev_split = 0
        for r in rclass:
            if ns[r] == 1:
                rev_split += 1
            ns[r] += 1
        lnProp = math.log((len(rclass) * nrates) /
                          (rev_split * (2.0 ** (na + nb) - 2.0) * rateAB * (na + nb)))
        return _Pending("rateJump", lik1, None, lnPrior, lnProp)

    def _rate_split(self, m, rclass, nrates):
        rng = self.rng
        lik1 = self.lik.clone()
        ns = [0] * nrates
        for r in rclass:
            ns[r] += 1
        splittable = [i for i in range(nrates) if ns[i] > 1]
        r0 = splittable[int(rng.integers(len(splittable)))]
        n0 = ns[r0]
        rate0 = lik1.getRate(m, r0)
        ra, rb = r0, nrates
        na = nb = 0
        a0 = int(rng.integers(n0))
        b0 = int(rng.integers(n0 - 1))
        if b0 >= a0:
            b0 += 1
        for i, r in enumerate(rclass):
            if r != r0:
                continue
            k = na + nb
            in_a = True if k == a0 else False if k == b0 else int(rng.integers(2)) == 0
            if in_a:
                rclass[i] = ra
                na += 1
            else:
                rclass[i] = rb
                nb += 1
        u = (n0 * rng.random() - na) * rate0
        rateA = rate0 + u / na
        rateB = rate0 - u / nb
        rates0 = [lik1.getRate(m, i) for i in range(nrates)] + [-1.0]
        rates0[ra] = rateA
        rates0[rb] = rateB
        # Renumber the classes in order of first appearance.
        reorder = {}
        for r in rclass:
            if r not in reorder:
                reorder[r] = len(reorder)
        rclass = [reorder[r] for r in rclass]
        rates1 = [0.0] * (nrates + 1)
        for i in range(nrates + 1):
            rates1[reorder[i]] = rates0[i]
        lik1.setRclass(m, rclass, rates1)
        pr = self.pdf["rateJump"]
        if nrates == len(rclass) - 1:
            p_merge, p_split = pr, pr / 2.0
        elif nrates == 1:
            p_merge, p_split = pr / 2.0, pr
            if self._n_rates_estim(lik1) == 1:
                p_split /= (1.0 - self.pdf["rate"])
        else:
            p_merge = p_split = pr / 2.0
        lnPrior = -math.log(self.params.rateJumpPrior) + math.log(p_merge / p_split)
        lnProp = math.log((len(splittable) * (2.0 ** n0 - 2.0) * rate0 * n0) /
                          (len(rclass) * (nrates + 1)))
        return _Pending("rateJump", lik1, None, lnPrior, lnProp)

    def _p_polytomyJump(self):
        """Polytomy merge or split (Chain.pyx:1076-1345)."""
        tree = self.tree
        assert tree.ntaxa > 3
        nedges, ntaxa = len(tree.edges), tree.ntaxa
        if nedges == ntaxa:
            merge = False
        elif nedges == 2 * ntaxa - 3:
            merge = True
        else:
            merge = bool(self.rng.integers(2))
        return self._polytomy_merge(nedges, ntaxa) if merge else \
            self._polytomy_split(nedges, ntaxa)

    def _polytomy_merge(self, nedges, ntaxa):
        tree, P = self.tree, self.params
        pick = int(self.rng.integers(nedges - ntaxa))
        j = 0
        for edge in list(tree.edges):
            nodeA, nodeB = edge.ring.node, edge.ring.other.node
            degA, degB = nodeA.degree, nodeB.degree
            if degA > 1 and degB > 1:
                if j == pick:
                    break
                j += 1
        edge.detach()
        sibsB = []
        r = nodeB.ring
        for _ in range(degB - 1):
            rnext = r.next
            e = r.edge
            sibsB.append(e)
            _move(e, nodeB, nodeA)
            r = rnext
        base = tree.base
        if base is nodeB:
            tree.setBase(nodeA)
        lnPrior = math.log(P.polytomyJumpPrior) - math.log(P.brlenPrior) - \
            (-P.brlenPrior * edge.length)
        if nedges == 2 * ntaxa - 3 and nedges - 1 != ntaxa:
            lnGam = math.log(0.5)
        elif nedges != 2 * ntaxa - 3 and nedges - 1 == ntaxa:
            lnGam = math.log(2.0)
        else:
            lnGam = 0.0
        np1 = sum(1 for n in tree.nodes if n.degree > 3)
        lnJacob = math.log(P.brlenPrior) + (-P.brlenPrior * edge.length)
        d = degA + degB - 2
        lnProp = lnGam + math.log(nedges - ntaxa) - math.log(np1) - \
            math.log(2.0 ** (d - 1.0) - d - 1.0) + lnJacob

        def reject():
            for e in sibsB:
                _move(e, nodeA, nodeB)
            edge.attach(nodeA, nodeB)
            if base is not tree.base:
                tree.setBase(base)

        return _Pending("polytomyJump", self.lik, nodeA, lnPrior, lnProp, None, reject)

    def _polytomy_split(self, nedges, ntaxa):
        tree, P, rng = self.tree, self.params, self.rng
        polys = [n for n in tree.nodes if n.degree > 3]
        assert polys
        nodeA = polys[int(rng.integers(len(polys)))]
        nodeB = Node(tree)
        edge = Edge(tree)
        edge.length = _exp_draw(rng) / P.brlenPrior
        deg0 = nodeA.degree
        # Four distinct siblings pre-assigned, two to each side.
        a0, aa0, b0, bb0 = (int(x) for x in rng.choice(deg0, size=4, replace=False))
        degA = degB = 0
        r = nodeA.ring
        for i in range(deg0):
            rnext = r.next
            if i in (a0, aa0):
                in_a = True
            elif i in (b0, bb0):
                in_a = False
            else:
                in_a = int(rng.integers(2)) == 0
            if in_a:
                degA += 1
            else:
                _move(r.edge, nodeA, nodeB)
                degB += 1
            r = rnext
        assert degA and degB and degA + degB == deg0
        edge.attach(nodeA, nodeB)
        lnPrior = -math.log(P.polytomyJumpPrior) + math.log(P.brlenPrior) + \
            (-P.brlenPrior * edge.length)
        if nedges == ntaxa and nedges + 1 != 2 * ntaxa - 3:
            lnGam = math.log(0.5)
        elif nedges != ntaxa and nedges + 1 == 2 * ntaxa - 3:
            lnGam = math.log(2.0)
        else:
            lnGam = 0.0
        lnProp = lnGam + math.log(len(polys)) + \
            math.log(2.0 ** (deg0 - 1) - deg0 - 1.0) - math.log(nedges - ntaxa + 1)
        if P.splitJacobian:
            lnProp += -math.log(P.brlenPrior) - (-P.brlenPrior * edge.length)

        def reject():
            edge.detach()
            r = nodeB.ring
            for _ in range(degB):
                rnext = r.next
                _move(r.edge, nodeB, nodeA)
                r = rnext

        return _Pending("polytomyJump", self.lik, nodeA, lnPrior, lnProp, None, reject)

    def _p_rateShapeInvJump(self):
        """+G off or on (Chain.pyx:1347-1445)."""
        P = self.params
        assert P.ncat > 1
        m = int(self.rng.integers(self.lik.nmodels()))
        alpha = self.lik.getAlpha(m)
        lik1 = self.lik.clone()
        if alpha != math.inf:
            w0 = 1.0 / alpha
            lik1.setAlpha(m, math.inf)
            lnPrior = math.log(P.rateShapeInvJumpPrior) - math.log(P.rateShapeInvPrior) - \
                (-P.rateShapeInvPrior * w0)
            lnProp = math.log(P.rateShapeInvPrior) + (-P.rateShapeInvPrior * w0)
            if self._n_gamma(lik1) == 0:
                lnProp += -math.log(1.0 - self.pdf["rateShapeInv"])
        else:
            w1 = _exp_draw(self.rng) / P.rateShapeInvPrior
            lik1.setAlpha(m, 1.0 / w1)
            lnPrior = -math.log(P.rateShapeInvJumpPrior) + math.log(P.rateShapeInvPrior) + \
                (-P.rateShapeInvPrior * w1)
            lnProp = -math.log(P.rateShapeInvPrior) - (-P.rateShapeInvPrior * w1)
            if self._n_gamma(lik1) == 1:
                lnProp += math.log(1.0 - self.pdf["rateShapeInv"])
        return _Pending("rateShapeInvJump", lik1, None, lnPrior, lnProp)

    def _p_invarJump(self):
        """+I off or on (Chain.pyx:1447-1549)."""
        P = self.params
        m = int(self.rng.integers(self.lik.nmodels()))
        wInvar = self.lik.getWInvar(m)
        lik1 = self.lik.clone()
        if wInvar != 0.0:
            wVar = self.lik.getWVar(m)
            lnPrior = math.log(P.invarJumpPrior)
            lnHast = 0.0
            lnPrior += -math.log(1.0 - P.invarPrior) + (1.0 - P.invarPrior) * wVar
            lnHast += math.log(1.0 - P.invarPrior) - P.invarPrior * wVar
            lik1.setWVar(m, 1.0)
            lnPrior += math.log(P.invarPrior) + wInvar
            lnHast += -math.log(P.invarPrior) - wInvar
            lik1.setWInvar(m, 0.0)
            if self._n_invar(lik1) == 0:
                lnHast += -math.log(1.0 - self.pdf["invar"])
        else:
            lnPrior = -math.log(P.invarJumpPrior)
            lnHast = 0.0
            wVar = _exp_draw(self.rng) * (1.0 - P.invarPrior)
            lnPrior += -math.log(1.0 - P.invarPrior) - wVar
            lnHast += math.log(1.0 - P.invarPrior) + wVar
            lik1.setWVar(m, wVar)
            wI = _exp_draw(self.rng) * P.invarPrior
            lnPrior += math.log(P.invarPrior) - wI
            lnHast += -math.log(P.invarPrior) + wI
            lik1.setWInvar(m, wI)
            if self._n_invar(lik1) == 1:
                lnHast += math.log(1.0 - self.pdf["invar"])
        return _Pending("invarJump", lik1, None, lnPrior, lnHast)

    def _p_freqJump(self):
        """Equal or estimated frequencies (Chain.pyx:1551-1641)."""
        P = self.params
        m = int(self.rng.integers(self.lik.nmodels()))
        lik1 = self.lik.clone()
        if self._freqs_equal(self.lik, m):
            lnPrior = -math.log(P.freqJumpPrior)
            lnHast = 0.0
            for i in range(lik1.dim):
                f = _exp_draw(self.rng)
                lnPrior += -f
                lnHast += f
                lik1.setFreq(m, i, f)
            if self._n_freqs_estim(lik1) == 1:
                lnHast += math.log(1.0 - self.pdf["freq"])
        else:
            lnPrior = math.log(P.freqJumpPrior)
            lnHast = 0.0
            for i in range(lik1.dim):
                f = lik1.getFreq(m, i)
                lnPrior += f
                lnHast += -f
                lik1.setFreq(m, i, 1.0 / lik1.dim)
            if self._n_freqs_estim(lik1) == 0:
                lnHast += -math.log(1.0 - self.pdf["freq"])
        return _Pending("freqJump", lik1, None, lnPrior, lnHast)

    def _q_terms(self, lik, m, sign):
        """Prior and Hastings contributions of model m's parameters for a
        mixture jump (Chain.pyx:1672-1695): +sign to the prior and -sign to
        the Hastings term for each Exp(1) parameter, and the Gamma shape's
        density."""
        P = self.params
        nf = 1 if self._freqs_equal(lik, m) else lik.dim
        vals = [lik.getWeight(m), lik.getRmult(m)]
        if nf > 1:
            vals += [lik.getFreq(m, i) for i in range(nf)]
        nr = lik.getNrates(m)
        if nr > 1:
            vals += [lik.getRate(m, i) for i in range(nr)]
        lnPrior = sign * sum(vals)
        lnHast = -sign * sum(vals)
        alpha = lik.getAlpha(m)
        if alpha != math.inf:
            om = 1.0 / alpha
            lnPrior += sign * (math.log(P.rateShapeInvPrior) + om * P.rateShapeInvPrior)
            lnHast -= sign * (math.log(P.rateShapeInvPrior) + om * P.rateShapeInvPrior)
        return lnPrior, lnHast

    def _p_mixtureJump(self):
        """Mixture model removal or addition (Chain.pyx:1643-1814)."""
        nmodels = self.lik.nmodels()
        remove = False if nmodels == 1 else bool(self.rng.integers(2))
        P, pdf = self.params, self.pdf
        lik0 = self.lik
        if not remove and len(lik0.comps) + _model_comps(P) > nat.MAX_COMPONENTS[lik0.dim]:
            # The mixture prior is truncated at the components a likelihood
            # may hold: an addition past it has prior 0 and is rejected (the
            # unchanged likelihood is evaluated, its cache answers).
            return _Pending("mixtureJump", lik0, None, -math.inf, 0.0)
        lik1 = lik0.clone()
        if remove:
            m = int(self.rng.integers(nmodels))
            lnPrior, lnHast = self._q_terms(lik1, m, +1.0)
            lik1.delModel(m)
            lnPrior += -math.log(1.0 - P.mixtureJumpPrior)
            if nmodels == 2:
                lnHast += math.log(2.0)
                lnHast += math.log(1.0 - pdf["weight"] - pdf["rmult"])
        else:
            m = lik1.addModel(1.0, P.ncat, P.catMedian, P.invar)
            random_q(lik1, m, P, pdf, self.rng)
            lnPrior, lnHast = self._q_terms(lik1, m, -1.0)
            lnPrior += math.log(1.0 - P.mixtureJumpPrior)
            if nmodels == 1:
                lnHast += math.log(0.5)
                lnHast += math.log(1.0 - pdf["weight"] - pdf["rmult"])
        # Proposals a jump enables or disables (counts 0 <-> 1).
        for count, kind in ((self._n_freqs_estim, "freq"), (self._n_rates_estim, "rate"),
                            (self._n_gamma, "rateShapeInv")):
            before, after = count(lik0), count(lik1)
            if remove and before == 1 and after == 0:
                lnHast += -math.log(1.0 - pdf[kind])
            elif not remove and before == 0 and after == 1:
                lnHast += math.log(1.0 - pdf[kind])
        return _Pending("mixtureJump", lik1, None, lnPrior, lnHast)


def _swap_subtrees(a, b, u, v):
    """Exchange the subtree behind ring a (at node u) with the one behind ring
    b (at node v).  The rings on the subtree side stay where they are, so the
    cached CLVs of both subtrees stay valid."""
    ea, eb = a.edge, b.edge
    A, B = a.other.node, b.other.node
    ea.detach()
    eb.detach()
    # Edge.attach(x, y) puts edge.ring at x; keep a.other at A, b.other at B.
    if a is ea.ring:
        ea.attach(v, A)
    else:
        ea.attach(A, v)
    if b is eb.ring:
        eb.attach(u, B)
    else:
        eb.attach(B, u)


class Mc3Batch:
    """``nruns`` independent runs of ``ncoupled`` heated replicas each.

    ``mix="brlen-nni"``: each replica starts from its problem's generating
    tree and model; ``mix="reference"``: from branch lengths and model
    parameters drawn from the priors of ``params`` (randomLik) on the
    problem's tree."""

    def __init__(self, problems, ncoupled=4, seed=0, ncat=4, context=None, mix="brlen-nni",
                 params=None):
        self.ctx = context if context is not None else nat.get_context(0)
        self.chains = []
        self.runs = []
        self.mix = mix
        if mix == "reference" and params is None:
            params = Mc3Params(ncat=ncat)
        for r, prob in enumerate(problems):
            run = []
            for i in range(ncoupled):
                cseed = seed * 100003 + r * 97 + i
                if mix == "reference":
                    rng = np.random.default_rng(cseed + 0x5eed)
                    lik = random_lik(prob.fresh_tree(), prob.alignment, params, rng,
                                     context=self.ctx)
                    run.append(Chain(lik, r, i, cseed, params, mix="reference"))
                else:
                    run.append(Chain.for_problem(prob, r, i, cseed, ncat=ncat, context=self.ctx))
            self.runs.append(run)
            self.chains.extend(run)
        # Heat swaps draw from one generator per run, so the order in which
        # runs are stepped does not change any run's trajectory.
        self.swap_rngs = [np.random.default_rng((seed + 1) * 7919 + r) for r in range(len(problems))]
        self.evals = 0

    def initialize(self):
        out = self.evaluate([(c.lik, c.tree.base) for c in self.chains], batched=True)
        for c, v in zip(self.chains, out):
            c.lnL = float(v)

    def evaluate(self, pairs, batched=True):
        """lnL of every (likelihood, root) pair (root None: the tree's
        base): one fused device pass, or one call per evaluation."""
        if batched:
            liks = [lik for lik, _ in pairs]
            roots = [lik._ready(r) for lik, r in pairs]
            out = nat.batch_lnl(self.ctx, liks, roots)
        else:
            out = np.array([lik.lnL(r) for lik, r in pairs])
        self.evals += len(pairs)
        return out

    def step(self, batched=True, p_nni=0.2):
        pairs = [c.propose(p_nni) for c in self.chains]
        lnls = self.evaluate(pairs, batched=batched)
        for c, v in zip(self.chains, lnls):
            c.settle(float(v))
        self._swaps(range(len(self.runs)))

    def run(self, nsteps, p_nni=0.2):
        """``nsteps`` steps of every chain with the host and the device
        overlapped: the runs are split into two halves whose fused device
        passes alternate (``crux_batch_submit``), so while one half's batch
        runs, the host settles the other half's results and plans its next
        proposals.  Every chain sees exactly the sequence of ``step``."""
        nr = len(self.runs)
        halves = [list(range(0, (nr + 1) // 2)), list(range((nr + 1) // 2, nr))]
        halves = [h for h in halves if h]
        chains = [[c for r in h for c in self.runs[r]] for h in halves]
        pending = [None] * len(halves)

        def submit(g):
            pairs = [c.propose(p_nni) for c in chains[g]]
            liks = [lik for lik, _ in pairs]
            roots = [lik._ready(r) for lik, r in pairs]
            pending[g] = nat.batch_submit(self.ctx, liks, roots)

        for g in range(len(halves)):
            submit(g)
        for k in range(nsteps):
            for g in range(len(halves)):
                lnls = nat.batch_wait(self.ctx, pending[g])
                for c, v in zip(chains[g], lnls):
                    c.settle(float(v))
                self._swaps(halves[g])
                self.evals += len(chains[g])
                if k + 1 < nsteps:
                    submit(g)

    def _swaps(self, runs):
        """One heat-swap attempt per run (Chain.pyx:1871-1911): replicas a
        and b drawn from the run's swap generator, accepted with
        exp((L_b - L_a) h_a + (L_a - L_b) h_b) >= u."""
        for r in runs:
            run = self.runs[r]
            if len(run) < 2:
                continue
            rng = self.swap_rngs[r]
            a = _index(rng, len(run))
            b = _index(rng, len(run) - 1)
            if b >= a:
                b += 1
            u = rng.random()
            ca, cb = run[a], run[b]
            p = _accept_prob((cb.lnL - ca.lnL) * ca.heat + (ca.lnL - cb.lnL) * cb.heat)
            if p >= u:
                ca.heat, cb.heat = cb.heat, ca.heat
                ca.accepts["swap"] += 1
            else:
                ca.rejects["swap"] += 1


def make_problems(nruns, seed=3, small=False):
    """Alignments of the batched-MC3 workload: 64-512 taxa and 1k-50k sites,
    log-uniform, GTR+G4 (one per run; its replicas share it)."""
    rng = np.random.default_rng(seed)
    lo_t, hi_t, lo_s, hi_s = (16, 64, 200, 2000) if small else (64, 512, 1000, 50000)
    probs = []
    for r in range(nruns):
        nt = int(round(math.exp(rng.uniform(math.log(lo_t), math.log(hi_t)))))
        ns = int(round(math.exp(rng.uniform(math.log(lo_s), math.log(hi_s)))))
        probs.append(synth.make_problem(nt, ns, seed * 1000 + r))
    return probs


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--runs", type=int, default=32)
    ap.add_argument("--coupled", type=int, default=4)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--plain-steps", type=int, default=5)
    ap.add_argument("--small", action="store_true")
    ap.add_argument("--mix", default="brlen-nni", choices=("brlen-nni", "reference"))
    args = ap.parse_args(argv)
    t0 = time.time()
    probs = make_problems(args.runs, small=args.small)
    print("problems: %d runs, %.1fs" % (len(probs), time.time() - t0), file=sys.stderr, flush=True)
    mc = Mc3Batch(probs, ncoupled=args.coupled, mix=args.mix)
    mc.initialize()
    print("initialized %d chains" % len(mc.chains), file=sys.stderr, flush=True)
    # Batched: every chain's proposal of a step in one device pass.
    mc.step(batched=True)
    t = time.perf_counter()
    e0 = mc.evals
    for _ in range(args.steps):
        mc.step(batched=True)
    serial_eps = (mc.evals - e0) / (time.perf_counter() - t)
    print("batched, one pass per step: %.1f evals/s" % serial_eps, file=sys.stderr, flush=True)
    # The same with two halves of the runs alternating on the device.
    t = time.perf_counter()
    e0 = mc.evals
    mc.run(args.steps)
    batched_eps = (mc.evals - e0) / (time.perf_counter() - t)
    print("batched, host and device overlapped: %.1f evals/s" % batched_eps, file=sys.stderr,
          flush=True)
    # One call per evaluation, one launch per node update.
    mc.ctx.set_mode(1)
    try:
        t = time.perf_counter()
        e0 = mc.evals
        for _ in range(args.plain_steps):
            mc.step(batched=False)
        plain_eps = (mc.evals - e0) / (time.perf_counter() - t)
    finally:
        mc.ctx.set_mode(0)
    print("plain: %.1f evals/s" % plain_eps, file=sys.stderr, flush=True)
    kinds = [k for k in mc.chains[0].accepts
             if any(c.accepts[k] + c.rejects[k] for c in mc.chains)]
    acc = {k: sum(c.accepts[k] for c in mc.chains) for k in kinds}
    rej = {k: sum(c.rejects[k] for c in mc.chains) for k in kinds}
    print(json.dumps({"workload": "batched MC3: %d runs x %d heated replicas, GTR+G4, "
                                  "%s moves" % (args.runs, args.coupled, args.mix),
                      "chains": len(mc.chains), "batched_evals_per_s": round(batched_eps, 1),
                      "batched_one_pass_per_step_evals_per_s": round(serial_eps, 1),
                      "plain_evals_per_s": round(plain_eps, 1),
                      "speedup": round(batched_eps / plain_eps, 2),
                      "accepts": acc, "rejects": rej}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
