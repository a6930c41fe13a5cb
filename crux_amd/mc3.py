"""Batched Metropolis-coupled MCMC driver over the device likelihood.

The caller the hot path serves: MC3 chains, each proposing a branch-length
or NNI move on its own tree per step, with heated replicas that swap states.
The reference evaluates one proposal per lnL call (Chain.pyx:438-475, one
Lik per chain); here every chain's proposal of a step goes into ONE device
pass (``crux_batch_lnl``): all dirty node updates of all trees run in a
single traversal launch, so the per-launch and per-call costs are paid once
per step instead of once per chain.  ``Mc3Batch.run`` also overlaps the host
with the device: two halves of the runs alternate, one half's pass running
while the host settles and re-proposes the other.

Proposal and acceptance arithmetic follow the reference:
  * branch length: multiplier m = exp(lambda (u - 1/2)), lambda = 2 ln 1.6,
    exponential prior with rate 10, Hastings term ln m, evaluated rooted at
    the changed edge (Chain.pyx:438-475, Mc3/__init__.pyx:264-269);
  * acceptance p = exp(((lnL1 - lnL) + lnPrior) heat + lnProp), accept when
    p >= u (Chain.pyx:104-105);
  * heats 1 / (1 + i 0.05) for replica i (Chain.pyx:53, Mc3/__init__.pyx:252);
  * a swap between two replicas of a run is accepted with
    exp((h_i - h_j)(post_j - post_i)), post = lnL + lnPrior.
The topology move is a nearest-neighbour interchange across a random
internal edge (the reference's eTBR, Chain.pyx:486-760, with the
extension limited to one edge: symmetric, Hastings term 0), branch lengths
unchanged.

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
from .lik import Lik

BRLEN_LAMBDA = 2.0 * math.log(1.6)
BRLEN_PRIOR = 10.0
HEAT_DELTA = 0.05


class Chain:
    """One heated replica: its own tree and Lik on a shared alignment."""

    def __init__(self, prob, run, index, seed, ncat=4, context=None):
        self.run = run
        self.index = index
        self.heat = 1.0 / (1.0 + index * HEAT_DELTA)
        self.rng = np.random.default_rng(seed)
        self.tree = prob.fresh_tree()
        self.lik = Lik(self.tree, prob.alignment, ncat=ncat, context=context)
        synth.configure(self.lik, prob.model)
        self.lnL = None
        self.lnPrior = -BRLEN_PRIOR * sum(e.length for e in self.tree.edges)
        # Moves keep the set of edges and which of them are internal (an NNI
        # swaps one end of two edges between two internal nodes), so both
        # lists are taken once instead of re-walking the tree every step.
        self._edges = list(self.tree.edges)
        self._inner = [e for e in self._edges
                       if e.ring.node.degree >= 3 and e.ring.other.node.degree >= 3]
        self.accepts = {"brlen": 0, "nni": 0, "swap": 0}
        self.rejects = {"brlen": 0, "nni": 0, "swap": 0}
        self._undo = None

    # -- proposals ----------------------------------------------------------
    def propose(self, p_nni=0.2):
        """Apply a proposal; return the node to evaluate the proposal at."""
        if self.rng.random() < p_nni:
            root = self._propose_nni()
            if root is not None:
                return root
        return self._propose_brlen()

    def _propose_brlen(self):
        edges = self._edges
        edge = edges[int(self.rng.integers(len(edges)))]
        lnM = BRLEN_LAMBDA * (self.rng.random() - 0.5)
        v0 = edge.length
        v1 = v0 * math.exp(lnM)
        edge.length = v1
        self._undo = ("brlen", edge, v0, -BRLEN_PRIOR * (v1 - v0), lnM)
        return edge.ring.node

    def _propose_nni(self):
        inner = self._inner
        if not inner:
            return None
        e = inner[int(self.rng.integers(len(inner)))]
        u, v = e.ring.node, e.ring.other.node
        ra = [r for r in u.ring if r is not e.ring]
        rb = [r for r in v.ring if r is not e.ring.other]
        a = ra[int(self.rng.integers(len(ra)))]
        b = rb[int(self.rng.integers(len(rb)))]
        _swap_subtrees(a, b, u, v)
        self._undo = ("nni", a, b, u, v)
        return u

    def settle(self, lnL1):
        """Accept or reject the pending proposal given its lnL."""
        kind = self._undo[0]
        if kind == "brlen":
            _, edge, v0, dprior, lnProp = self._undo
        else:
            dprior, lnProp = 0.0, 0.0
        u = self.rng.random()
        with np.errstate(over="ignore"):
            p = math.exp(min(700.0, ((lnL1 - self.lnL) + dprior) * self.heat + lnProp))
        if p >= u:
            self.lnL = lnL1
            self.lnPrior += dprior
            self.accepts[kind] += 1
            if kind == "brlen":
                # Re-base so the next unrooted call reuses the cache
                # (Chain.pyx:466-469).
                self.tree.setBase(edge.ring.node)
        else:
            if kind == "brlen":
                edge.length = v0
            else:
                _, a, b, uu, vv = self._undo
                # a now hangs on vv and b on uu: swap back.
                _swap_subtrees(a, b, vv, uu)
            self.rejects[kind] += 1
        self._undo = None


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
    """``nruns`` independent runs of ``ncoupled`` heated replicas each."""

    def __init__(self, problems, ncoupled=4, seed=0, ncat=4, context=None):
        self.ctx = context if context is not None else nat.get_context(0)
        self.chains = []
        self.runs = []
        for r, prob in enumerate(problems):
            run = [Chain(prob, r, i, seed * 100003 + r * 97 + i, ncat=ncat, context=self.ctx)
                   for i in range(ncoupled)]
            self.runs.append(run)
            self.chains.extend(run)
        # Heat swaps draw from one generator per run, so the order in which
        # runs are stepped does not change any run's trajectory.
        self.swap_rngs = [np.random.default_rng((seed + 1) * 7919 + r) for r in range(len(problems))]
        self.evals = 0

    def initialize(self):
        out = self.evaluate([(c, c.tree.base) for c in self.chains], batched=True)
        for c, v in zip(self.chains, out):
            c.lnL = v

    def evaluate(self, pairs, batched=True):
        """lnL of every (chain, root) pair: one fused device pass, or one
        call per chain."""
        if batched:
            liks = [c.lik for c, _ in pairs]
            roots = [c.lik._ready(r) for c, r in pairs]
            out = nat.batch_lnl(self.ctx, liks, roots)
        else:
            out = np.array([c.lik.lnL(r) for c, r in pairs])
        self.evals += len(pairs)
        return out

    def step(self, batched=True, p_nni=0.2):
        pairs = [(c, c.propose(p_nni)) for c in self.chains]
        lnls = self.evaluate(pairs, batched=batched)
        for (c, _), v in zip(pairs, lnls):
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
            pairs = [(c, c.propose(p_nni)) for c in chains[g]]
            liks = [c.lik for c, _ in pairs]
            roots = [c.lik._ready(r) for c, r in pairs]
            pending[g] = (nat.batch_submit(self.ctx, liks, roots), pairs)

        for g in range(len(halves)):
            submit(g)
        for k in range(nsteps):
            for g in range(len(halves)):
                ticket, pairs = pending[g]
                lnls = nat.batch_wait(self.ctx, ticket)
                for (c, _), v in zip(pairs, lnls):
                    c.settle(float(v))
                self._swaps(halves[g])
                self.evals += len(pairs)
                if k + 1 < nsteps:
                    submit(g)

    def _swaps(self, runs):
        for r in runs:
            run = self.runs[r]
            if len(run) < 2:
                continue
            rng = self.swap_rngs[r]
            i, j = rng.choice(len(run), size=2, replace=False)
            a, b = run[i], run[j]
            pa, pb = a.lnL + a.lnPrior, b.lnL + b.lnPrior
            lnp = (a.heat - b.heat) * (pb - pa)
            if lnp >= 0.0 or rng.random() < math.exp(lnp):
                a.heat, b.heat = b.heat, a.heat
                a.accepts["swap"] += 1
            else:
                a.rejects["swap"] += 1


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
    args = ap.parse_args(argv)
    t0 = time.time()
    probs = make_problems(args.runs, small=args.small)
    print("problems: %d runs, %.1fs" % (len(probs), time.time() - t0), file=sys.stderr, flush=True)
    mc = Mc3Batch(probs, ncoupled=args.coupled)
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
    acc = {k: sum(c.accepts[k] for c in mc.chains) for k in ("brlen", "nni", "swap")}
    rej = {k: sum(c.rejects[k] for c in mc.chains) for k in ("brlen", "nni", "swap")}
    print(json.dumps({"workload": "batched MC3: %d runs x %d heated replicas, GTR+G4, "
                                  "branch-length + NNI moves" % (args.runs, args.coupled),
                      "chains": len(mc.chains), "batched_evals_per_s": round(batched_eps, 1),
                      "batched_one_pass_per_step_evals_per_s": round(serial_eps, 1),
                      "plain_evals_per_s": round(plain_eps, 1),
                      "speedup": round(batched_eps / plain_eps, 2),
                      "accepts": acc, "rejects": rej}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
