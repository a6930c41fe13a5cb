"""Continuous-batching likelihood server over one device context.

Requests arrive over time (Poisson), each naming a resident tree (a
session): 70% single-branch re-evaluations (one edge length changed by an
MC3-style multiplier, lnL rooted at that edge) and 30% full-tree lnLs (cache
flushed, every CLV recomputed).  Trees, alignments, CLV caches and model
eigensystems stay on the device between requests.

The loop keeps up to two fused batches in flight (``crux_batch_submit`` /
``crux_batch_wait``): while the device runs one batch, every request that
has arrived is admitted into the next one, as long as its session is not
already in that batch and its tree fits the HBM budget (sessions not in
flight are evicted least-recently-used to make room; a request whose tree
cannot fit waits).  Each batch is planned and uploaded while the previous
one runs, and batches run back to back on the device in submission order,
so the requests of one session are served in arrival order.  Reported:
throughput and p50/p99 latency (arrival to result) per arrival rate.

    python -m crux_amd.serve [--sessions 16] [--requests 2000] [--rates 2000,8000,...]
"""

from __future__ import annotations

import argparse
import collections
import json
import math
import sys
import time

import numpy as np

from . import _native as nat
from . import synth
from .lik import Lik

BRLEN_LAMBDA = 2.0 * math.log(1.6)


class Session:
    """A resident tree: problem, device Lik, and its request bookkeeping."""

    def __init__(self, sid, prob, ncat=4):
        self.sid = sid
        self.prob = prob
        self.ncat = ncat
        self.lik = None
        self.tree = None
        self.edges = None
        self.last_used = 0.0
        self.in_flight = 0

    def footprint(self, C=4):
        """Device bytes of a fully cached tree (both CLV directions of every
        internal edge, FP32, plus exponents, tips and site buffers)."""
        aln = self.prob.alignment
        S, K, N = aln.nchars, 4 if self.prob.kind == "dna" else 20, aln.ntaxa
        slots = 2 * (2 * N - 3)
        return slots * S * (C * K * 4 + 4) + N * S * 2 + S * 16

    @property
    def resident(self):
        return self.lik is not None

    def load(self, ctx):
        # The host tree outlives evictions: it carries the session's state
        # (branch lengths edited by earlier requests).
        if self.tree is None:
            self.tree = self.prob.fresh_tree()
            self.edges = list(self.tree.edges)
        self.lik = Lik(self.tree, self.prob.alignment, ncat=self.ncat, context=ctx)
        synth.configure(self.lik, self.prob.model)

    def evict(self):
        """Free the device state (CLV cache, tree mirror, model)."""
        # The lik first (it refers to the device tree), then the tree.
        if self.lik._handle:
            nat.lib().crux_lik_destroy(self.lik._handle)
            self.lik._handle = None
        self.lik._dev.close()
        self.lik = None


class Request:
    __slots__ = ("rid", "sid", "kind", "t_arrive", "t_done", "lnl", "edge", "mult")

    def __init__(self, rid, sid, kind, t_arrive, edge, mult):
        self.rid = rid
        self.sid = sid
        self.kind = kind
        self.t_arrive = t_arrive
        self.edge = edge          # edge index (single-branch requests)
        self.mult = mult          # branch-length multiplier
        self.t_done = None
        self.lnl = None


def make_sessions(n, seed=5, taxa=(32, 1024), sites=(500, 200000), aa_ratio=0.5):
    """Trees of log-uniform size, DNA and 20-state amino-acid (GTR+G4)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        nt = int(round(math.exp(rng.uniform(math.log(taxa[0]), math.log(taxa[1])))))
        ns = int(round(math.exp(rng.uniform(math.log(sites[0]), math.log(sites[1])))))
        kind = "aa" if rng.random() < aa_ratio else "dna"
        p = synth.make_problem(nt, ns, seed * 1000 + i, kind=kind)
        p.kind = kind
        out.append(Session(i, p))
    return out


def make_requests(sessions, n, rate, seed=6, p_branch=0.7):
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for rid in range(n):
        t += rng.exponential(1.0 / rate)
        s = sessions[int(rng.integers(len(sessions)))]
        kind = "branch" if rng.random() < p_branch else "full"
        ne = 2 * s.prob.alignment.ntaxa - 3
        reqs.append(Request(rid, s.sid, kind, t, int(rng.integers(ne)),
                            math.exp(BRLEN_LAMBDA * (rng.random() - 0.5))))
    return reqs


def _apply(sess, req):
    """Mutate the session for the request; return the node to root at."""
    if req.kind == "branch":
        e = sess.edges[req.edge % len(sess.edges)]
        e.length = e.length * req.mult
        return e.ring.node
    sess.lik.flush()
    return sess.tree.base


class Server:
    def __init__(self, sessions, ctx=None, hbm_budget=None, max_batch=256):
        self.ctx = ctx if ctx is not None else nat.get_context(0)
        self.sessions = {s.sid: s for s in sessions}
        self.budget = hbm_budget
        self.max_batch = max_batch
        self.batches = 0
        self.batched_requests = 0
        self.evictions = 0

    def _fits(self, sess):
        """Admission by HBM budget: make room for a tree that is not resident
        by evicting idle sessions, least recently used first."""
        if self.budget is None or sess.resident:
            return True
        # Each resident tree is charged its fully cached footprint, so a
        # tree's CLV cache can grow to all its directed edges within budget.
        need = sess.footprint()
        while sum(s.footprint() for s in self.sessions.values() if s.resident) + need > self.budget:
            idle = [s for s in self.sessions.values() if s.resident and s.in_flight == 0]
            if not idle:
                # Nothing left to evict: wait for in-flight work, unless the
                # device holds no tree at all (a tree above the budget then
                # runs alone rather than never).
                return not any(s.resident for s in self.sessions.values())
            min(idle, key=lambda s: s.last_used).evict()
            self.evictions += 1
        return True

    def run(self, requests, clock=None):
        """Serve requests (sorted by arrival time, seconds from start) in
        real time.  Returns the requests with t_done and lnl filled in."""
        clock = clock or time.perf_counter
        t0 = clock()
        arrivals = collections.deque(requests)
        queue = collections.deque()
        inflight = collections.deque()   # (ticket, [requests])
        done = 0
        n = len(requests)
        while done < n:
            now = clock() - t0
            while arrivals and arrivals[0].t_arrive <= now:
                queue.append(arrivals.popleft())
            if queue and len(inflight) < 2:
                batch, seen, rest = [], set(), collections.deque()
                while queue and len(batch) < self.max_batch:
                    r = queue.popleft()
                    sess = self.sessions[r.sid]
                    if r.sid in seen or not self._fits(sess):
                        rest.append(r)
                        continue
                    seen.add(r.sid)
                    if not sess.resident:
                        sess.load(self.ctx)
                    # Busy from now on: a later admission may not evict it.
                    sess.in_flight += 1
                    batch.append(r)
                queue.extendleft(reversed(rest))
                if batch:
                    liks, roots = [], []
                    for r in batch:
                        sess = self.sessions[r.sid]
                        node = _apply(sess, r)
                        liks.append(sess.lik)
                        roots.append(sess.lik._ready(node))
                        sess.last_used = now
                    inflight.append((nat.batch_submit(self.ctx, liks, roots), batch))
                    self.batches += 1
                    self.batched_requests += len(batch)
                    continue
            if inflight:
                ticket, batch = inflight.popleft()
                out = nat.batch_wait(self.ctx, ticket)
                t = clock() - t0
                for r, v in zip(batch, out):
                    r.lnl = float(v)
                    r.t_done = t
                    self.sessions[r.sid].in_flight -= 1
                done += len(batch)
                continue
            if arrivals:
                dt = arrivals[0].t_arrive - (clock() - t0)
                if dt > 0:
                    time.sleep(min(dt, 0.001))
        return requests


def replay_alone(sessions, requests, ctx=None):
    """The same requests one at a time, in arrival order, one lnL call each
    on fresh sessions: the reference for the mixing-equivalence check."""
    ctx = ctx if ctx is not None else nat.get_context(0)
    fresh = {s.sid: Session(s.sid, s.prob, s.ncat) for s in sessions}
    out = []
    for r in requests:
        sess = fresh[r.sid]
        if not sess.resident:
            sess.load(ctx)
        node = _apply(sess, r)
        out.append(sess.lik.lnL(node))
    for s in fresh.values():
        if s.resident:
            s.evict()
    return out


def summarize(reqs, wall):
    lat = np.array([r.t_done - r.t_arrive for r in reqs]) * 1e3
    return {"requests": len(reqs), "throughput_rps": round(len(reqs) / wall, 1),
            "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "mean_ms": round(float(lat.mean()), 3)}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--sessions", type=int, default=16)
    ap.add_argument("--requests", type=int, default=1500)
    ap.add_argument("--rates", default="500,2000,8000,32000")
    ap.add_argument("--taxa", default="32,1024")
    ap.add_argument("--sites", default="500,200000")
    ap.add_argument("--budget-gib", type=float, default=0.0,
                    help="HBM budget for resident trees (0: none)")
    args = ap.parse_args(argv)
    taxa = tuple(int(x) for x in args.taxa.split(","))
    sites = tuple(int(x) for x in args.sites.split(","))
    t0 = time.time()
    sessions = make_sessions(args.sessions, taxa=taxa, sites=sites)
    print("sessions: %d (%.1fs)" % (len(sessions), time.time() - t0), file=sys.stderr, flush=True)
    ctx = nat.get_context(0)
    budget = int(args.budget_gib * (1 << 30)) or None
    srv = Server(sessions, ctx, hbm_budget=budget)
    # Warm: every session resident with its cache built.
    warm = make_requests(sessions, 4 * len(sessions), 1e9, seed=1)
    srv.run(warm)
    lines = []
    for rate in [float(x) for x in args.rates.split(",")]:
        reqs = make_requests(sessions, args.requests, rate, seed=int(rate) % 1000 + 7)
        b0, r0 = srv.batches, srv.batched_requests
        t = time.perf_counter()
        srv.run(reqs)
        wall = time.perf_counter() - t
        s = summarize(reqs, wall)
        s["offered_rps"] = rate
        s["mean_batch"] = round((srv.batched_requests - r0) / max(1, srv.batches - b0), 1)
        print("rate %g: %s" % (rate, s), file=sys.stderr, flush=True)
        lines.append(s)
    print(json.dumps({"workload": "served likelihood: %d resident trees (%s taxa, %s sites, "
                                  "DNA/AA 1:1, GTR+G4), 70%% single-branch / 30%% full"
                                  % (len(sessions), args.taxa, args.sites),
                      "sweep": lines}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
