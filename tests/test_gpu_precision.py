"""Input precision of the 20/22-state products (CRUX_TUNE_PRECISION).

FP32 is the default; the 16-bit forms feed the matrix core bfloat16 or
float16 inputs with FP32 accumulation, or float16 high and low parts (three
products).  Each is held to the error bound stated here against the FP64
oracle, with the relative errors measured on this GPU at 256 taxa x 20,000
sites (4 categories) in brackets: FP32 and the split float16 to 1e-6 (1.9e-8,
7.3e-8), float16 to 2e-5 (4.8e-6), bfloat16 to 5e-4 (9.6e-5).  4-state
likelihoods do not depend on the setting.
"""

import sys

import numpy as np
import pytest

from crux_amd import _native as nat
from crux_amd import synth
from crux_amd.lik import Lik
from oracle.oracle import OracleLik

pytestmark = pytest.mark.gpu

BOUND = {nat.CRUX_PREC_F32: 1e-6, nat.CRUX_PREC_F16X3: 1e-6, nat.CRUX_PREC_F16: 2e-5,
         nat.CRUX_PREC_BF16: 5e-4}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture
def ctx():
    c = nat.get_context(0)
    yield c
    c.set_tuning(nat.CRUX_TUNE_PRECISION, nat.CRUX_PREC_F32)


@pytest.mark.parametrize("ncat,kind", [(4, "aa"), (1, "aa"), (2, "aa"), (6, "aa")])
def test_precisions_against_oracle(ctx, ncat, kind):
    prob = synth.make_problem(40, 2000, 600 + ncat, kind=kind, ncat=ncat)
    dev = Lik(prob.tree, prob.alignment, ncat=ncat)
    ora = OracleLik(prob.fresh_tree(), prob.alignment, ncat=ncat, nthreads=4)
    synth.configure(dev, prob.model)
    synth.configure(ora, prob.model)
    ref = ora.lnL()
    f32 = dev.lnL()
    e = list(dev.tree.edges)[9]
    e2 = list(ora.tree.edges)[9]
    e.length *= 1.4
    e2.length *= 1.4
    ref2 = ora.lnL(e2.ring.node)
    e.length /= 1.4
    for prec, bound in BOUND.items():
        ctx.set_tuning(nat.CRUX_TUNE_PRECISION, prec)
        assert ctx.get_tuning(nat.CRUX_TUNE_PRECISION) == prec
        a = dev.lnL()
        rel = abs(a - ref) / abs(ref)
        say("prec", prec, "ncat", ncat, "rel %.2e" % rel)
        assert rel <= bound, (prec, rel)
        # Partial re-evaluation at the precision: cached CLVs of this
        # precision plus the recomputed path.
        e.length *= 1.4
        b = dev.lnL(e.ring.node)
        assert abs(b - ref2) <= bound * abs(ref2), (prec, b, ref2)
        e.length /= 1.4
        dev.lnL()
    # Back at FP32 the cached 16-bit CLVs are recomputed: the FP32 value
    # returns exactly.
    ctx.set_tuning(nat.CRUX_TUNE_PRECISION, nat.CRUX_PREC_F32)
    assert dev.lnL() == f32


def test_protein_split_precision(ctx):
    from crux_amd.alignment import Protein
    rng = np.random.default_rng(17)
    rooted, tree, nw = synth.yule_tree(16, rng, 0.15, newick=True)
    m = synth.Model(rng.lognormal(0, 1, 231), rng.dirichlet(np.ones(22)), 0.7, 4, Protein)
    aln = synth.simulate(rooted, m, 700, rng)
    aln.compact()
    prob = synth.Problem(tree, rooted, aln, m, 700, nw)
    dev = Lik(prob.tree, aln, ncat=4)
    ora = OracleLik(prob.fresh_tree(), aln, ncat=4)
    synth.configure(dev, m)
    synth.configure(ora, m)
    ref = ora.lnL()
    for prec in (nat.CRUX_PREC_F16X3, nat.CRUX_PREC_F16):
        ctx.set_tuning(nat.CRUX_TUNE_PRECISION, prec)
        assert abs(dev.lnL() - ref) <= BOUND[prec] * abs(ref)


def test_dna_ignores_precision(ctx):
    prob = synth.make_problem(30, 1500, 610)
    dev = Lik(prob.tree, prob.alignment, ncat=4)
    synth.configure(dev, prob.model)
    a = dev.lnL()
    for prec in BOUND:
        ctx.set_tuning(nat.CRUX_TUNE_PRECISION, prec)
        dev.flush()
        assert dev.lnL() == a
    with pytest.raises(ValueError):
        ctx.set_tuning(nat.CRUX_TUNE_PRECISION, 7)
