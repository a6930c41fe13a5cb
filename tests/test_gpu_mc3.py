"""Batched MC3 driver on the device: every chain's tracked lnL after a mix of
batched and one-call-per-eval steps (branch-length and NNI moves, accepts,
rejects and heat swaps) equals the FP64 oracle's lnL of the chain's current
tree, and a fused pass over all chains equals one call per chain.  The same
under the reference's full proposal mix (parameter moves on the clone, eTBR,
polytomy, rate-class, +G, +I, frequency and mixture jumps), the oracle
rebuilt from each chain's current parameters."""

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
    pairs = [(c.lik, c.tree.base) for c in mc.chains]
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


def oracle_of(lik, nthreads=2):
    """An FP64 oracle likelihood with lik's current tree and parameters."""
    ora = OracleLik(lik.tree.dup(), lik.alignment, nmodels=0, nthreads=nthreads)
    for i, m in enumerate(lik.models):
        ora.addModel(lik.getWeight(i), m.ncat, m.catMedian, m.invar)
        ora.setRmult(i, lik.getRmult(i))
        ora.setRclass(i, lik.getRclass(i),
                      [float(lik.getRate(i, j)) for j in range(lik.getNrates(i))])
        for j in range(lik.dim):
            ora.setFreq(i, j, lik.getFreq(i, j))
        if lik.getAlpha(i) != math.inf:
            ora.setAlpha(i, lik.getAlpha(i))
        if m.invar:
            ora.setWVar(i, lik.getWVar(i))
            ora.setWInvar(i, lik.getWInvar(i))
    return ora


def test_reference_mix_matches_oracle():
    print("mc3: reference proposal mix, 3 runs x 2 replicas", flush=True)
    probs = [synth.make_problem(n, s, 90 + i) for i, (n, s) in
             enumerate([(12, 300), (20, 500), (9, 200)])]
    params = mc3.Mc3Params(ncat=4, invar=True)
    mc = mc3.Mc3Batch(probs, ncoupled=2, seed=5, mix="reference", params=params)
    mc.initialize()
    poly = False
    for k in range(60):
        mc.step(batched=True)
        poly |= any(n.degree > 3 for c in mc.chains for n in c.tree.nodes)
        if k == 30:
            mc.ctx.set_mode(1)
            try:
                mc.step(batched=False)
            finally:
                mc.ctx.set_mode(0)
    mc.run(20)
    tried = {k: sum(c.accepts[k] + c.rejects[k] for c in mc.chains) for k in mc3.PROPS}
    print("  proposals", tried, flush=True)
    for k in ("freq", "rate", "rateShapeInv", "invar", "brlen", "etbr", "rateJump",
              "polytomyJump", "rateShapeInvJump", "invarJump", "freqJump", "mixtureJump"):
        assert tried[k] > 0, k
    assert sum(c.accepts["polytomyJump"] for c in mc.chains) > 0
    assert poly
    for c in mc.chains:
        ref = oracle_of(c.lik).lnL()
        assert abs(c.lnL - ref) <= 1e-6 * abs(ref), (c.run, c.index, c.lnL, ref)
