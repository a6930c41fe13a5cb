"""Batched MC3 driver on the device: every chain's tracked lnL after a mix of
batched and one-call-per-eval steps (branch-length and NNI moves, accepts,
rejects and heat swaps) equals the FP64 oracle's lnL of the chain's current
tree, and a fused pass over all chains equals one call per chain."""

import math

import numpy as np
import pytest

from crux_amd import _native as nat
from crux_amd import mc3, synth
from oracle.oracle import OracleLik

pytestmark = pytest.mark.gpu


def test_mc3_batched_matches_oracle():
    print("mc3: 3 runs x 2 replicas, small problems", flush=True)
    probs = [synth.make_problem(n, s, 70 + i) for i, (n, s) in
             enumerate([(16, 400), (27, 900), (40, 300)])]
    mc = mc3.Mc3Batch(probs, ncoupled=2, seed=4)
    mc.initialize()
    for _ in range(6):
        mc.step(batched=True, p_nni=0.4)
    mc.ctx.set_mode(1)
    try:
        for _ in range(2):
            mc.step(batched=False, p_nni=0.4)
    finally:
        mc.ctx.set_mode(0)
    for _ in range(4):
        mc.step(batched=True, p_nni=0.4)
    acc = sum(c.accepts["brlen"] + c.accepts["nni"] for c in mc.chains)
    rej = sum(c.rejects["brlen"] + c.rejects["nni"] for c in mc.chains)
    assert acc > 0 and rej > 0
    for c in mc.chains:
        t = c.tree.dup()
        ora = OracleLik(t, c.lik.alignment, ncat=4, nthreads=2)
        synth.configure(ora, probs[c.run].model)
        ref = ora.lnL()
        assert abs(c.lnL - ref) <= 1e-6 * abs(ref), (c.run, c.index, c.lnL, ref)
    # One fused pass over every chain vs one call per chain.
    pairs = [(c, c.tree.base) for c in mc.chains]
    fused = mc.evaluate(pairs, batched=True)
    single = mc.evaluate(pairs, batched=False)
    for a, b in zip(fused, single):
        assert abs(a - b) <= 1e-12 * abs(b)


def test_mc3_overlapped_run_matches_steps():
    # The host/device-overlapped driver gives every chain exactly the
    # trajectory of the one-pass-per-step driver.
    print("mc3: overlapped vs stepped, 4 runs x 2 replicas", flush=True)
    probs = [synth.make_problem(n, s, 80 + i) for i, (n, s) in
             enumerate([(14, 300), (22, 500), (9, 200), (30, 400)])]
    a = mc3.Mc3Batch(probs, ncoupled=2, seed=9)
    b = mc3.Mc3Batch(probs, ncoupled=2, seed=9)
    a.initialize()
    b.initialize()
    for _ in range(8):
        a.step(batched=True, p_nni=0.3)
    b.run(8, p_nni=0.3)
    assert a.evals == b.evals
    for ca, cb in zip(a.chains, b.chains):
        assert ca.lnL == cb.lnL, (ca.run, ca.index, ca.lnL, cb.lnL)
        assert ca.heat == cb.heat
        assert ca.accepts == cb.accepts and ca.rejects == cb.rejects
