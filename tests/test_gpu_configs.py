"""The batched, amino-acid and served configurations at their full sizes.

* amino acids, 256 taxa x 20,000 sites, 20 states +G4: full traversal, site
  lnLs and partial re-evaluations against the FP64 oracle;
* batched MC3, 32 runs x 4 heated replicas on 64-512 taxa x 1k-50k sites
  (log-uniform): fused batched steps, overlapped and one pass per step, then
  a sample of chains' tracked lnL against the FP64 oracle on their current
  trees;
* served likelihood, 16 resident trees of 32-1,024 taxa x 500-200k sites,
  DNA and amino acids: every result of a Poisson mix of branch-length and
  full-tree requests equals the same request served alone, and the final
  state of three sessions (the largest tree, the longest DNA alignment and
  a small tree) matches the FP64 oracle.
"""

import os
import sys

import numpy as np
import pytest

from crux_amd import mc3, serve, synth
from crux_amd.lik import Lik
from oracle.oracle import OracleLik

pytestmark = pytest.mark.gpu

REL = 1e-6
THREADS = min(16, len(os.sched_getaffinity(0)))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def close(a, b, rel=REL):
    return abs(a - b) <= rel * abs(b)


def test_amino_acids_256_taxa_20k_sites():
    say("aa 256 x 20000")
    prob = synth.make_problem(256, 20000, 4, kind="aa")
    dev = Lik(prob.tree, prob.alignment, ncat=4)
    ora = OracleLik(prob.fresh_tree(), prob.alignment, ncat=4, nthreads=THREADS)
    synth.configure(dev, prob.model)
    synth.configure(ora, prob.model)
    a, b = dev.lnL(), ora.lnL()
    say("  lnL", a, b, abs(a - b) / abs(b))
    assert close(a, b)
    sd, so = np.array(dev.siteLnLs()), np.array(ora.siteLnLs())
    assert np.allclose(sd, so, rtol=1e-6, atol=1e-4)
    ed, eo = list(dev.tree.edges), list(ora.tree.edges)
    for k, f in ((17, 1.3), (200, 0.7), (401, 1.6)):
        ed[k].length *= f
        eo[k].length *= f
        a, b = dev.lnL(ed[k].ring.node), ora.lnL(eo[k].ring.node)
        assert close(a, b), (k, a, b)


def test_batched_mc3_32_runs_x_4_replicas():
    say("mc3 32 x 4")
    probs = mc3.make_problems(32)
    sizes = [(p.alignment.ntaxa, p.alignment.nchars) for p in probs]
    assert min(s[0] for s in sizes) >= 64 and max(s[0] for s in sizes) <= 512
    mc = mc3.Mc3Batch(probs, ncoupled=4, seed=2)
    assert len(mc.chains) == 128
    mc.initialize()
    for _ in range(3):
        mc.step(batched=True)
    mc.run(4)
    mc.step(batched=False)
    acc = sum(c.accepts["brlen"] + c.accepts["nni"] for c in mc.chains)
    rej = sum(c.rejects["brlen"] + c.rejects["nni"] for c in mc.chains)
    assert acc > 0 and rej > 0
    for c in mc.chains[::16]:
        ora = OracleLik(c.tree.dup(), c.lik.alignment, ncat=4, nthreads=THREADS)
        synth.configure(ora, probs[c.run].model)
        ref = ora.lnL()
        say("  chain", c.run, c.index, c.lnL, ref)
        assert close(c.lnL, ref), (c.run, c.index, c.lnL, ref)


def test_served_mix_16_sessions_full_size():
    say("serve 16 sessions")
    sessions = serve.make_sessions(16)
    taxa = [s.prob.alignment.ntaxa for s in sessions]
    sites = [s.prob.nsites_raw for s in sessions]
    assert max(taxa) > 512 and max(sites) > 100000
    reqs = serve.make_requests(sessions, 160, rate=20000.0, seed=13)
    srv = serve.Server(sessions)
    srv.run(reqs)
    assert all(r.t_done is not None for r in reqs)
    assert srv.batched_requests / srv.batches > 1.0
    srv.close()
    alone = serve.replay_alone(sessions, reqs)
    for r, a in zip(reqs, alone):
        assert abs(r.lnl - a) <= REL * abs(a), (r.rid, r.sid, r.kind, r.lnl, a)
    # The last result of a session is its final tree's lnL.
    last = {}
    for r in reqs:
        last[r.sid] = r
    by = {s.sid: s for s in sessions}
    served = [sid for sid in last]
    big = max(served, key=lambda i: by[i].prob.alignment.ntaxa)
    dna = [i for i in served if by[i].prob.kind == "dna"]
    longest = max(dna, key=lambda i: by[i].prob.alignment.nchars) if dna else big
    small = min(served, key=lambda i: by[i].prob.alignment.ntaxa * by[i].prob.alignment.nchars)
    for sid in {big, longest, small}:
        s = by[sid]
        ora = OracleLik(s.tree.dup(), s.prob.alignment, ncat=4, nthreads=THREADS)
        synth.configure(ora, s.prob.model)
        ref = ora.lnL()
        say("  session", sid, s.prob.kind, s.prob.alignment.ntaxa, s.prob.alignment.nchars,
            last[sid].lnl, ref)
        assert close(last[sid].lnl, ref), (sid, last[sid].lnl, ref)
    for s in sessions:
        if s.resident:
            s.evict()
