"""Serving loop on the device: every request's result, served in the mix
(continuous batching, two batches in flight, branch-length edits and cache
flushes interleaved across sessions), equals the same request served alone
(one lnL call at a time, in arrival order, on fresh sessions), and the
submit/wait surface returns each ticket's results once."""

import math

import numpy as np
import pytest

from crux_amd import _native as nat
from crux_amd import serve

pytestmark = pytest.mark.gpu


def test_mixed_equals_alone():
    print("serve: 6 sessions, 300 requests", flush=True)
    sessions = serve.make_sessions(6, seed=9, taxa=(8, 40), sites=(100, 3000))
    reqs = serve.make_requests(sessions, 300, rate=1e6, seed=3)
    srv = serve.Server(sessions, max_batch=5)
    srv.run(reqs)
    assert all(r.t_done is not None for r in reqs)
    assert srv.batches > 20
    alone = serve.replay_alone(sessions, reqs)
    for r, a in zip(reqs, alone):
        assert abs(r.lnl - a) <= 1e-6 * abs(a) + 1e-9, (r.rid, r.sid, r.kind, r.lnl, a)


def test_budget_evicts_and_waits():
    sessions = serve.make_sessions(4, seed=11, taxa=(16, 16), sites=(600, 600), aa_ratio=0.0)
    one = max(s.footprint() for s in sessions)
    # A context of its own, so the budget counts only this server's trees.
    ctx = nat.Context(0)
    # Room for about one tree at a time: sessions are evicted and reloaded
    # (their host trees keep the edits), and results do not change.
    srv = serve.Server(sessions, ctx=ctx, hbm_budget=int(1.2 * one), max_batch=4)
    reqs = serve.make_requests(sessions, 80, rate=1e6, seed=4)
    srv.run(reqs)
    assert srv.evictions > 0
    for sess in sessions:
        if sess.resident:
            sess.evict()
    alone = serve.replay_alone(sessions, reqs)
    for r, a in zip(reqs, alone):
        assert abs(r.lnl - a) <= 1e-6 * abs(a) + 1e-9


def test_ticket_collected_once():
    sessions = serve.make_sessions(2, seed=12, taxa=(8, 12), sites=(100, 200))
    ctx = nat.get_context(0)
    for s in sessions:
        s.load(ctx)
    liks = [s.lik for s in sessions]
    roots = [L._ready(None) for L in liks]
    t1 = nat.batch_submit(ctx, liks, roots)
    t2 = nat.batch_submit(ctx, liks, roots)
    t3 = nat.batch_submit(ctx, liks, roots)   # completes the oldest first
    r3 = nat.batch_wait(ctx, t3)
    r1 = nat.batch_wait(ctx, t1)
    r2 = nat.batch_wait(ctx, t2)
    assert np.allclose(r1, r2, rtol=0, atol=0) and np.allclose(r1, r3, rtol=0, atol=0)
    with pytest.raises(ValueError):
        nat.batch_wait(ctx, t1)
