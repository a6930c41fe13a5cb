"""Served likelihood: a Poisson stream of requests against resident trees,
driven through the native continuous-batching server (crux_server_*).

Requests arrive over time (Poisson), each naming a resident tree (a
session): 70% single-branch re-evaluations (one edge length changed by an
MC3-style multiplier, lnL rooted at that edge) and 30% full-tree lnLs (cache
flushed, every CLV recomputed).  Trees, alignments, CLV caches and model
eigensystems stay on the device between requests.

Batching, ordering and admission are the native server's: a server thread
forms a batch from every request that has arrived while the device runs the
previous one (up to two batches in flight), up to two requests of a tree per
batch (one per CLV polarity: each session serves through a likelihood and
its clone), each planned against the tree as the requests before it leave
it; the context's HBM budget is enforced by releasing the CLV pools of idle
trees, least recently used first.  This module submits each request at its
arrival time and reports throughput and p50/p99 latency (arrival to result)
per arrival rate.

    python -m crux_amd.serve [--sessions 16] [--requests 2000] [--rates 2000,8000,...]
"""

from __future__ import annotations

import argparse
import ctypes
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
    """A resident tree: problem, a device Lik and its clone (the two CLV
    polarities the server may use)."""

    def __init__(self, sid, prob, ncat=4):
        self.sid = sid
        self.prob = prob
        self.ncat = ncat
        self.lik = None
        self.mate = None
        self.tree = None
        self.edges = None

    def footprint(self, C=4):
        """Device bytes of a fully cached tree in one polarity (CLVs of every
        directed internal edge, FP32, plus exponents, tips and site buffers)."""
        aln = self.prob.alignment
        S, K, N = aln.nchars, 4 if self.prob.kind == "dna" else 20, aln.ntaxa
        slots = 2 * (2 * N - 3)
        return slots * S * (C * K * 4 + 4) + N * S * 2 + S * 16

    @property
    def resident(self):
        return self.lik is not None

    def load(self, ctx):
        # The host tree outlives the device objects: it carries the session's
        # state (branch lengths edited by earlier requests).
        if self.tree is None:
            self.tree = self.prob.fresh_tree()
            self.edges = list(self.tree.edges)
        self.lik = Lik(self.tree, self.prob.alignment, ncat=self.ncat, context=ctx)
        synth.configure(self.lik, self.prob.model)
        self.mate = self.lik.clone()
        # Models and topology on the device before the server takes over.
        self.lik._ready(None)
        self.mate._ready(None)

    def evict(self):
        """Free the device state (CLV caches, tree mirror, models)."""
        for L in (self.mate, self.lik):
            if L is not None and L._handle:
                nat.lib().crux_lik_destroy(L._handle)
                L._handle = None
        self.lik._dev.close()
        self.lik = None
        self.mate = None


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
    """Apply the request to the host tree; return (root node, edited edge ids,
    their new lengths, flush)."""
    if req.kind == "branch":
        e = sess.edges[req.edge % len(sess.edges)]
        e.length = e.length * req.mult
        return e.ring.node, [e.id], [e.length], False
    return sess.tree.base, [], [], True


class Server:
    """The native continuous-batching server on one context."""

    def __init__(self, sessions, ctx=None, max_batch=256):
        self.ctx = ctx if ctx is not None else nat.get_context(0)
        self.sessions = {s.sid: s for s in sessions}
        h = nat._vp()
        nat.check(nat.lib().crux_server_create(self.ctx.handle, int(max_batch), ctypes.byref(h)))
        self.handle = h
        self._added = set()

    def _ensure(self, sess):
        if not sess.resident:
            sess.load(self.ctx)
        if sess.sid not in self._added:
            for L in (sess.lik, sess.mate):
                nat.check(nat.lib().crux_server_add_lik(self.handle, L.handle))
            self._added.add(sess.sid)

    def stats(self):
        """(batches, requests batched, trees released for the HBM budget,
        failed requests)."""
        st = (ctypes.c_uint64 * 4)()
        nat.check(nat.lib().crux_server_stats(self.handle, st))
        return tuple(int(x) for x in st)

    @property
    def batches(self):
        return self.stats()[0]

    @property
    def batched_requests(self):
        return self.stats()[1]

    @property
    def evictions(self):
        return self.stats()[2]

    def run(self, requests):
        """Submit each request at its arrival time (seconds from now), then
        collect every result.  Returns the requests with lnl and t_done
        (seconds from the start) filled in."""
        lib = nat.lib()
        for s in {r.sid for r in requests}:
            self._ensure(self.sessions[s])
        ids = []
        t0 = time.perf_counter_ns()
        out = ctypes.c_uint64()
        for r in requests:
            while (time.perf_counter_ns() - t0) * 1e-9 < r.t_arrive:
                pass
            sess = self.sessions[r.sid]
            node, edges, lens, flush = _apply(sess, r)
            n = len(edges)
            nat.check(lib.crux_server_submit(
                self.handle, sess.lik.handle, node.id, n, (ctypes.c_int32 * max(1, n))(*edges),
                (ctypes.c_double * max(1, n))(*lens), 1 if flush else 0, ctypes.byref(out)))
            ids.append(out.value)
        v = ctypes.c_double()
        td = ctypes.c_int64()
        for r, i in zip(requests, ids):
            nat.check(lib.crux_server_wait(self.handle, i, ctypes.byref(v), ctypes.byref(td)))
            r.lnl = v.value
            r.t_done = (td.value - t0) * 1e-9
        return requests

    def close(self):
        """Stop the server (after its requests) and free the sessions."""
        if self.handle:
            for sid in list(self._added):
                sess = self.sessions[sid]
                if sess.resident:
                    for L in (sess.lik, sess.mate):
                        nat.lib().crux_server_remove_lik(self.handle, L.handle)
            nat.lib().crux_server_destroy(self.handle)
            self.handle = None
            self._added.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


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
        node, _, _, flush = _apply(sess, r)
        if flush:
            sess.lik.flush()
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
                    help="HBM budget of the serving context (0: none)")
    args = ap.parse_args(argv)
    taxa = tuple(int(x) for x in args.taxa.split(","))
    sites = tuple(int(x) for x in args.sites.split(","))
    t0 = time.time()
    sessions = make_sessions(args.sessions, taxa=taxa, sites=sites)
    print("sessions: %d (%.1fs)" % (len(sessions), time.time() - t0), file=sys.stderr, flush=True)
    budget = int(args.budget_gib * (1 << 30))
    ctx = nat.Context(0, budget) if budget else nat.get_context(0)
    srv = Server(sessions, ctx)
    # Warm: every session resident with its cache built.
    srv.run(make_requests(sessions, 4 * len(sessions), 1e9, seed=1))
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
    srv.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
