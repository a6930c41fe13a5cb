"""Amino-acid (20-state) full-traversal benchmark.

Workload: 256-taxon Yule tree, 20,000 simulated sites compressed to
patterns, random reversible 20x20 exchangeabilities ~LogNormal(0,1), freqs
~Dirichlet(1 x 20), +G4 (alpha 0.5).  Reports full-traversal lnL evals/s,
the traversal kernel time (HIP events), its roofline -- algorithmic HBM
bytes (CLV and exponent writes, tip codes, weights in, site lnLs out) per
launch against 8 TB/s, and the matrix-core FP32 work (20x20 P times CLV per
internal child edge, component and pattern) against 157.3 TFLOP/s; at about
10 flop/B the kernel sits below the ~20 flop/B ridge, so HBM bounds it --
|dlnL| against the FP64 oracle and the oracle's CPU throughput on a bounded
sample.

    python -m crux_amd.bench_aa [--steps 20] [--no-cpu]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import _native as nat
from . import synth
from .lik import Lik

FP32_MATRIX_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args(argv)
    t0 = time.time()
    prob = synth.make_problem(256, 20000, 4, kind="aa")
    S = prob.alignment.nchars
    print("problem: 256 taxa, %d patterns, %.1fs" % (S, time.time() - t0), file=sys.stderr,
          flush=True)
    ctx = nat.get_context(0)
    lik = Lik(prob.tree, prob.alignment, ncat=4)
    synth.configure(lik, prob.model)
    root = prob.tree.base

    def full():
        lik.flush()
        return lik.lnL(root)

    lnl = full()
    for _ in range(args.warmup):
        full()
    t = time.perf_counter()
    for _ in range(args.steps):
        full()
    eps = args.steps / (time.perf_counter() - t)
    # Traversal kernel duration from events the runtime stamps at the
    # kernel's start and end, averaged over as many evaluations as timed.
    ctx.set_timing(True)
    tr = []
    for _ in range(args.steps):
        full()
        tr.append(ctx.last_timing()[1])
    ctx.set_timing(False)
    trav = float(np.mean(tr))
    nbytes, _, nodes = ctx.last_traffic()
    # Matrix-core FP32 work: per internal child edge and component, a 20x20
    # matrix times a 20-vector per pattern (2 flops per multiply-add); tip
    # children are table rows (no product).
    ninternal = prob.alignment.ntaxa - 3
    flops = 2.0 * 20 * 20 * S * 4 * ninternal
    tflops = flops / (trav * 1e-3) / 1e12
    gbs = nbytes / (trav * 1e-3) / 1e9
    print("full: %.1f evals/s; traverse %.3f ms; %.1f GB/s (%.1f%% of HBM); %.1f TFLOP/s" % (
        eps, trav, gbs, 100 * gbs / HBM_PEAK_GBS, tflops), file=sys.stderr, flush=True)
    cpu = None
    dl = None
    if not args.no_cpu:
        from oracle.oracle import OracleLik
        ncpu = min(os.cpu_count() or 1, 16)
        ora = OracleLik(prob.fresh_tree(), prob.alignment, ncat=4, nthreads=ncpu)
        synth.configure(ora, prob.model)
        t = time.perf_counter()
        ref = ora.lnL()
        n = 1
        while time.perf_counter() - t < args.cpu_seconds and n < 20:
            ora.lnL()
            n += 1
        cpu_eps = n / (time.perf_counter() - t)
        dl = {"abs": abs(lnl - ref), "rel": abs(lnl - ref) / abs(ref)}
        cpu = {"value": round(cpu_eps, 3), "unit": "lnL evals/s", "cores": ncpu, "kind": "port",
               "sample": "%d full evaluations by the FP64 C restatement (%d threads)" % (n, ncpu)}
        print("oracle: %.3f evals/s, |dlnL| %.3e rel %.3e" % (cpu_eps, dl["abs"], dl["rel"]),
              file=sys.stderr, flush=True)
    print(json.dumps({"workload": "amino acid: 256 taxa, 20,000 sites -> %d patterns, "
                                  "LogNormal exchangeabilities, Dirichlet(1) freqs, +G4" % S,
                      "evals_per_s": round(eps, 2), "traverse_ms": round(trav, 4),
                      "node_updates": nodes,
                      "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                                   "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                                   "bytes": nbytes},
                      "matrix_core": {"tflops": round(tflops, 2),
                                      "peak_tflops_fp32": FP32_MATRIX_PEAK_TFLOPS,
                                      "frac": round(tflops / FP32_MATRIX_PEAK_TFLOPS, 4)},
                      "lnL": lnl,
                      "dlnL": dl, "cpu_baseline": cpu}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
