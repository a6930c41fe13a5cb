"""Serving loop on the device: every request's result, served in the mix
(continuous batching, two batches in flight, branch-length edits and cache
flushes interleaved across sessions), equals the same request served alone
(one lnL call at a time, in arrival order, on fresh sessions), and the
submit/wait surface returns each ticket's results once."""

import ctypes
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
    srv = serve.Server(sessions, max_batch=12)
    srv.run(reqs)
    assert all(r.t_done is not None for r in reqs)
    batches, batched, _, failed = srv.stats()
    print("batches", batches, "mean batch", batched / batches, flush=True)
    assert batched == 300 and failed == 0
    srv.close()
    alone = serve.replay_alone(sessions, reqs)
    for r, a in zip(reqs, alone):
        assert abs(r.lnl - a) <= 1e-6 * abs(a) + 1e-9, (r.rid, r.sid, r.kind, r.lnl, a)


def test_two_requests_of_a_tree_in_one_batch():
    # Two requests of one tree submitted together run in one batch through
    # the two polarities, each seeing exactly the edits before it.
    sessions = serve.make_sessions(1, seed=21, taxa=(30, 30), sites=(2000, 2000), aa_ratio=0.0)
    srv = serve.Server(sessions, max_batch=8)
    warm = serve.make_requests(sessions, 4, 1e9, seed=2)
    srv.run(warm)
    reqs = serve.make_requests(sessions, 40, 1e9, seed=5, p_branch=1.0)
    for r in reqs:
        r.t_arrive = 0.0
    b0, q0, _, _ = srv.stats()
    srv.run(reqs)
    b1, q1, _, failed = srv.stats()
    print("batches", b1 - b0, "requests", q1 - q0, flush=True)
    assert failed == 0 and q1 - q0 == 40
    assert b1 - b0 < 40            # some batch held two requests of the tree
    srv.close()
    alone = serve.replay_alone(sessions, warm + reqs)
    for r, a in zip(warm + reqs, alone):
        assert abs(r.lnl - a) <= 1e-6 * abs(a) + 1e-9, (r.rid, r.lnl, a)


def test_budget_releases_idle_trees():
    # The CLV bytes of one tree, measured on a context without a budget.
    probe_sessions = serve.make_sessions(4, seed=11, taxa=(16, 16), sites=(600, 600), aa_ratio=0.0)
    probe = nat.Context(0)
    for s in probe_sessions:
        s.load(probe)
    base = probe.bytes
    psrv = serve.Server(probe_sessions, ctx=probe)
    # One tree through both polarities, rooted all over: its full CLV pools.
    psrv.run(serve.make_requests(probe_sessions[:1], 60, 1e9, seed=4))
    one = probe.bytes - base
    psrv.close()
    for s in probe_sessions:
        s.evict()
    probe.close()
    assert one > 0
    # Room for about one and a half trees' CLVs: trees are released and
    # recomputed, results do not change.
    sessions = serve.make_sessions(4, seed=11, taxa=(16, 16), sites=(600, 600), aa_ratio=0.0)
    ctx = nat.Context(0, base + int(1.5 * one))
    srv = serve.Server(sessions, ctx=ctx, max_batch=4)
    reqs = serve.make_requests(sessions, 80, rate=1e6, seed=4)
    srv.run(reqs)
    assert srv.evictions > 0
    srv.close()
    for s in sessions:
        s.evict()
    alone = serve.replay_alone(sessions, reqs)
    for r, a in zip(reqs, alone):
        assert abs(r.lnl - a) <= 1e-6 * abs(a) + 1e-9


def test_server_errors_and_poll():
    sessions = serve.make_sessions(1, seed=31, taxa=(10, 10), sites=(300, 300), aa_ratio=0.0)
    srv = serve.Server(sessions)
    srv.run(serve.make_requests(sessions, 2, 1e9, seed=3))
    sess = sessions[0]
    lib = nat.lib()
    rid = ctypes.c_uint64()
    e = sess.edges[3]
    nat.check(lib.crux_server_submit(srv.handle, sess.lik.handle, e.ring.node.id, 1,
                                     (ctypes.c_int32 * 1)(e.id), (ctypes.c_double * 1)(-1.0), 0,
                                     ctypes.byref(rid)))
    v, td = ctypes.c_double(), ctypes.c_int64()
    with pytest.raises(ValueError, match="Negative"):
        nat.check(lib.crux_server_wait(srv.handle, rid.value, ctypes.byref(v), ctypes.byref(td)))
    # A valid request after the failed one; poll until done.
    nat.check(lib.crux_server_submit(srv.handle, sess.lik.handle, e.ring.node.id, 1,
                                     (ctypes.c_int32 * 1)(e.id), (ctypes.c_double * 1)(0.2), 0,
                                     ctypes.byref(rid)))
    done = ctypes.c_int()
    for _ in range(200000):
        nat.check(lib.crux_server_poll(srv.handle, rid.value, ctypes.byref(done)))
        if done.value:
            break
    assert done.value
    nat.check(lib.crux_server_wait(srv.handle, rid.value, ctypes.byref(v), ctypes.byref(td)))
    with pytest.raises(ValueError):
        nat.check(lib.crux_server_poll(srv.handle, rid.value, ctypes.byref(done)))
    e.length = 0.2
    assert abs(v.value - sess.lik.lnL(e.ring.node)) <= 1e-9 * abs(v.value)
    assert srv.stats()[3] == 1
    srv.close()


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
