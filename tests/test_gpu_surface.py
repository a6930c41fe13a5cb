"""The rest of the Lik call surface on the device: pickling (Lik.pyx:326-386),
simulate() (Lik.pyx:662-833) and the non-blocking batch poll.  dup() is
covered by the Lik_i golden cases (test/Lik_i.cx.exp) in test_gpu_parity."""

import math
import pickle
import sys
import time

import numpy as np
import pytest

from crux_amd import _native as nat
from crux_amd import synth
from crux_amd.lik import Lik
from oracle.oracle import OracleLik

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def close(a, b, rel=1e-6):
    return abs(a - b) <= rel * abs(b) + 1e-9


def _mixture_lik(prob):
    lik = Lik(prob.tree, prob.alignment, nmodels=2, ncat=4, invar=True)
    synth.configure(lik, prob.model, 0)
    synth.configure(lik, prob.model, 1)
    lik.setRmult(1, 1.8)
    lik.setWeight(1, 0.4)
    lik.setWVar(0, 0.75)
    lik.setWInvar(0, 0.25)
    lik.setAlpha(1, 2.2)
    return lik


def test_pickle_round_trip():
    say("pickle")
    prob = synth.make_problem(20, 700, 81)
    lik = _mixture_lik(prob)
    a = lik.lnL()
    data = pickle.dumps(lik)
    back = pickle.loads(data)
    assert back.nmodels() == 2
    assert back.getRmult(1) == 1.8 and back.getWInvar(0) == pytest.approx(0.25)
    assert back.getAlpha(1) == 2.2 and back.getInvar(0) and back.getInvar(1)
    b = back.lnL()
    assert abs(a - b) <= 1e-12 * abs(a), (a, b)
    # The copy is independent: changing it leaves the original alone.
    list(back.tree.edges)[2].length *= 2.0
    assert back.lnL() != b
    assert lik.lnL() == a


def test_simulate_then_evaluate():
    say("simulate")
    prob = synth.make_problem(12, 400, 82)
    lik = Lik(prob.tree, prob.alignment, ncat=4)
    synth.configure(lik, prob.model)
    sim = lik.simulate(5000, rng=np.random.default_rng(3))
    aln = sim.alignment
    assert aln.nchars == 5000 and aln.ntaxa == prob.alignment.ntaxa
    # Simulated states follow the model's frequencies (root and stationarity).
    counts = np.array([(aln.rows == ord(c)).sum() for c in "ACGT"], float)
    pin = prob.model.pi / prob.model.pi.sum()
    assert np.allclose(counts / counts.sum(), pin, atol=0.02)
    # The simulated likelihood evaluates on the device as in the oracle.
    ora = OracleLik(sim.tree.dup(), aln, ncat=4)
    synth.configure(ora, prob.model)
    assert close(sim.lnL(), ora.lnL())
    # The generating branch lengths fit the simulated data better than
    # lengths scaled by 3.
    base = sim.lnL()
    for e in sim.tree.edges:
        e.length *= 3.0
    assert sim.lnL() < base


def test_batch_poll():
    say("poll")
    probs = [synth.make_problem(30 + 5 * i, 2000 + 500 * i, 83 + i) for i in range(3)]
    liks = []
    for p in probs:
        L = Lik(p.tree, p.alignment, ncat=4)
        synth.configure(L, p.model)
        L._ready(None)
        liks.append(L)
    ctx = nat.get_context(0)
    roots = [L.tree.base.id for L in liks]
    expect = nat.batch_lnl(ctx, liks, roots)
    for L in liks:
        L.flush()
    tk = nat.batch_submit(ctx, liks, roots)
    t0 = time.time()
    while not nat.batch_poll(ctx, tk):
        assert time.time() - t0 < 30.0
        time.sleep(0.0005)
    assert nat.batch_poll(ctx, tk)          # still ready until collected
    got = nat.batch_wait(ctx, tk)
    assert np.array_equal(got, expect)
    with pytest.raises(ValueError):
        nat.batch_poll(ctx, tk)             # collected once


def test_tree_destroyed_before_its_likelihoods():
    # A tree may go before likelihoods created on it, including a clone that
    # never evaluated (it holds no polarity): they are detached, and
    # destroying them afterwards frees only their handles.
    say("tree before liks")
    prob = synth.make_problem(10, 200, 84)
    lik = Lik(prob.tree, prob.alignment, ncat=2)
    synth.configure(lik, prob.model)
    mate = lik.clone()
    lik.lnL()
    lib = nat.lib()
    lik._dev.close()
    for L in (mate, lik):
        assert lib.crux_lik_destroy(L._handle) == 0
        L._handle = None
