"""The measurement hooks the benchmarks rely on: the byte accounting behind
the roofline figure, the device timing and the host timing.

The roofline's algorithmic bytes of a full 4-state traversal are, per the
byte model, S*C*K*4 + 4S per parent write (CLV plus exponents), S per tip
row and 4S + 8S at the root (weights in, site lnLs out).  An unrooted
binary tree of N taxa evaluated at an internal node has N - 2 updates, one
of them the root, so a full traversal moves

    (N - 3) * (S*C*16 + 4S) + N*S + 12*S

bytes when every child comes from registers or the LDS stack.  The test
counts N and S independently of the engine and compares.
"""

import sys

import pytest

from crux_amd import _native as nat
from crux_amd import synth
from crux_amd.lik import Lik

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("ntaxa,nsites,ncat", [(20, 500, 4), (37, 900, 2)])
def test_full_traversal_bytes(ntaxa, nsites, ncat):
    say("traffic", ntaxa, nsites, ncat)
    prob = synth.make_problem(ntaxa, nsites, 40 + ntaxa, ncat=ncat)
    lik = Lik(prob.tree, prob.alignment, ncat=ncat)
    synth.configure(lik, prob.model)
    root = prob.tree.base
    ctx = lik._ctx
    # Eight stack slots hold every pending result of these small trees (the
    # Sethi-Ullman depth of N leaves is about log2 N), so no child is read
    # back from HBM.
    old = ctx.get_tuning(nat.CRUX_TUNE_STACK_DEPTH)
    ctx.set_tuning(nat.CRUX_TUNE_STACK_DEPTH, 8)
    try:
        lik.flush()
        lik.lnL(root)
        moved, model_bytes, nodes = ctx.last_traffic()
    finally:
        ctx.set_tuning(nat.CRUX_TUNE_STACK_DEPTH, old)
    S = prob.alignment.nchars
    N = prob.alignment.ntaxa
    C = ncat
    assert nodes == N - 3
    assert moved == (N - 3) * (S * C * 16 + 4 * S) + N * S + 12 * S
    # The all-children-from-HBM model adds one CLV read per internal child.
    assert model_bytes >= moved


def test_device_timing_fields():
    say("timing")
    prob = synth.make_problem(24, 2000, 9)
    lik = Lik(prob.tree, prob.alignment, ncat=4)
    synth.configure(lik, prob.model)
    ctx = lik._ctx
    ctx.set_timing(True)
    try:
        lik.flush()
        lik.lnL()
        pm, trav, gap, total = (float(x) for x in ctx.last_timing())
    finally:
        ctx.set_timing(False)
    # P matrices, traversal, the gap between them and the span of both.
    assert pm > 0.0 and trav > 0.0 and gap >= 0.0
    assert total >= trav and total >= pm
    assert abs(total - (pm + gap + trav)) <= 1e-3 + 0.01 * total


def test_host_timing_fields():
    # Host microseconds of the last batch, recorded whether or not device
    # timing is on: planning, building, launch calls, wait + final sums.
    say("host timing")
    prob = synth.make_problem(24, 2000, 9)
    lik = Lik(prob.tree, prob.alignment, ncat=4)
    synth.configure(lik, prob.model)
    ctx = lik._ctx
    lik.flush()
    lik.lnL()
    plan, build, launch, wait = (float(x) for x in ctx.last_host_timing())
    assert plan > 0.0 and build > 0.0 and launch > 0.0 and wait > 0.0
    assert plan + build + launch + wait < 1e6
    # A second batch overwrites the figures.
    e = list(lik.tree.edges)[3]
    e.length *= 1.1
    lik.lnL(e.ring.node)
    assert float(ctx.last_host_timing()[0]) > 0.0
