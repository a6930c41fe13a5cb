#!/usr/bin/env python3
"""Benchmark of the tree-likelihood hot path on one MI355X.

Workload (default): 128-taxon Yule tree, 100,000 simulated DNA sites
compressed to ~60k patterns with integer weights, GTR+Gamma4 (alpha 0.5),
FP32 CLVs with power-of-two site scaling, compensated FP64 site sums.  A
step is one full-traversal lnL evaluation (every CLV recomputed) with all
data resident in HBM.  Also reported: single-branch partial re-evaluation
throughput (re-root at a changed edge, as an MC3 branch-length move does),
the plain one-launch-per-node path, |dlnL| against the FP64 oracle, the
roofline of the traversal kernel and the oracle's CPU throughput.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def _cpu_model():
    """The host CPU's model name (first processor in /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--partial-steps", type=int, default=400)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--stack-depth", type=int, default=None,
                    help="4-state kernel LDS stack depth (2, 4, 8; default: tuned)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (the reference "
                         "kernel for scaling HBM byte counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the likelihood on a single "
              "device (see DESIGN.md for how the path shards); --gpus %d is not "
              "supported" % args.gpus, file=sys.stderr)
        return 2
    if args.copy_calibration:
        import torch
        n = args.copy_calibration << 20
        a = torch.empty(n, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": n, "copies": 5}))
        return 0

    from crux_amd import _native as nat
    from crux_amd import synth
    from crux_amd.lik import Lik

    t0 = time.time()
    prob = synth.make_problem(128, 100000, 2, mean_len=0.025)
    S = prob.alignment.nchars
    log("problem: 128 taxa, %d patterns, %.1fs" % (S, time.time() - t0))

    ctx = nat.get_context(0)
    if args.stack_depth is not None:
        ctx.set_tuning(nat.CRUX_TUNE_STACK_DEPTH, args.stack_depth)
    lik = Lik(prob.tree, prob.alignment, ncat=4)
    synth.configure(lik, prob.model)
    root = prob.tree.base
    lnl0 = lik.lnL(root)

    # --- full traversals (fused) --------------------------------------------
    def full():
        lik.flush()
        return lik.lnL(root)

    for _ in range(args.warmup):
        full()
    ctx.set_timing(True)
    full()
    tm = ctx.last_timing()
    traffic, formula_bytes, nodes = ctx.last_traffic()
    ctx.set_timing(False)
    t = time.perf_counter()
    for _ in range(args.steps):
        full()
    dt = time.perf_counter() - t
    full_eps = args.steps / dt
    log("full: %.1f evals/s, %.3f ms/eval; kernel ms P=%.3f trav=%.3f" %
        (full_eps, 1e3 * dt / args.steps, tm[0], tm[1]))

    # Traversal kernel duration: HIP events around the traversal launch on
    # the context's stream, averaged over as many evaluations as the timed
    # region runs (a second pass: the events' own calls stay out of the
    # evals/s figure).
    ctx.set_timing(True)
    tr = []
    for _ in range(args.steps):
        full()
        tr.append(ctx.last_timing()[1])
    ctx.set_timing(False)
    trav_ms = float(np.mean(tr))
    achieved = traffic / (trav_ms * 1e-3) / 1e9

    # --- plain path: one launch per node update ---------------------------
    ctx.set_mode(1)
    for _ in range(3):
        full()
    t = time.perf_counter()
    n_plain = max(10, args.steps // 4)
    for _ in range(n_plain):
        full()
    plain_eps = n_plain / (time.perf_counter() - t)
    ctx.set_mode(0)
    log("plain: %.1f evals/s" % plain_eps)

    # --- single-branch partial re-evaluation -----------------------------
    rng = np.random.default_rng(7)
    edges = list(prob.tree.edges)
    lam = 2.0 * math.log(1.6)
    lik.lnL(root)

    def partial():
        e = edges[int(rng.integers(len(edges)))]
        e.length = e.length * math.exp(lam * (rng.random() - 0.5))
        v = lik.lnL(e.ring.node)
        prob.tree.base = e.ring.node
        return v

    part_nodes = []
    for _ in range(20):
        partial()
        part_nodes.append(ctx.last_traffic()[2])
    t = time.perf_counter()
    for _ in range(args.partial_steps):
        partial()
    part_eps = args.partial_steps / (time.perf_counter() - t)
    log("partial: %.1f evals/s" % part_eps)

    # --- FP64 oracle: parity and CPU baseline ---------------------------------
    cpu = cpu1 = cpup = None
    dl_abs = dl_rel = None
    if not args.no_cpu:
        from oracle.oracle import OracleLik
        ncpu = min(os.cpu_count() or 1, 16)
        prob2 = synth.make_problem(128, 100000, 2, mean_len=0.025)
        ora = OracleLik(prob2.tree, prob2.alignment, ncat=4, nthreads=ncpu)
        synth.configure(ora, prob2.model)
        lik2 = Lik(prob2.fresh_tree(), prob2.alignment, ncat=4)
        synth.configure(lik2, prob2.model)
        dev_l = lik2.lnL()
        t = time.perf_counter()
        ref_l = ora.lnL()
        n = 1
        while time.perf_counter() - t < args.cpu_seconds and n < 50:
            ora.lnL()
            n += 1
        cpu_eps = n / (time.perf_counter() - t)
        dl_abs = abs(dev_l - ref_l)
        dl_rel = dl_abs / abs(ref_l)
        log("oracle: lnL %.6f dev %.6f |d| %.3e rel %.3e; %.2f evals/s on %d threads" %
            (ref_l, dev_l, dl_abs, dl_rel, cpu_eps, ncpu))
        cpu = {"value": round(cpu_eps, 3), "unit": "lnL evals/s", "cores": ncpu,
               "kind": "port", "cpu_model": _cpu_model(),
               "sample": "%d full-traversal evaluations of the same 128-taxon, %d-pattern "
                         "GTR+G4 workload by the FP64 C restatement of the reference "
                         "engine (oracle/cxlik_oracle.c) with %d stripe threads" % (n, S, ncpu)}
        # The same restatement on one thread (a few evaluations).
        ora1 = OracleLik(prob2.fresh_tree(), prob2.alignment, ncat=4, nthreads=1)
        synth.configure(ora1, prob2.model)
        t = time.perf_counter()
        ora1.lnL()
        n1 = 1
        while time.perf_counter() - t < args.cpu_seconds / 4 and n1 < 5:
            ora1.lnL()
            n1 += 1
        cpu1 = {"value": round(n1 / (time.perf_counter() - t), 3), "unit": "lnL evals/s",
                "cores": 1, "kind": "port", "sample": "%d evaluations, one thread" % n1}
        log("oracle, one thread: %.2f evals/s" % cpu1["value"])
        # Partial re-evaluation on the CPU: the same restatement executing as
        # many CLV updates (plus the root) as the device's partial
        # re-evaluations above needed on average.
        k = int(round(float(np.mean(part_nodes))))
        tp = ora.time_partial(k, reps=20, warmup=3)
        cpup = {"value": round(1.0 / tp, 2), "unit": "lnL evals/s", "cores": ncpu,
                "kind": "port",
                "sample": "median of 20 executions (after 3) of the last %d CLV updates and "
                          "the root of the same workload's plan on %d stripe threads: the "
                          "work of a partial re-evaluation (the device's averaged %.1f node "
                          "updates)" % (k, ncpu, float(np.mean(part_nodes)))}
        log("oracle, partial: %.1f evals/s" % cpup["value"])

    # Write bandwidth the HIP runtime's fill kernel reaches on this GPU
    # (hipMemsetD32Async over 512 MiB, timed with HIP events on the null
    # stream): the practical ceiling of an almost write-only kernel, next to
    # the published peak.
    write_peak = None
    try:
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        nbytes = 512 << 20
        buf, e0, e1 = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        ok = hip.hipMalloc(ctypes.byref(buf), ctypes.c_size_t(nbytes)) == 0
        ok = ok and hip.hipEventCreate(ctypes.byref(e0)) == 0 and hip.hipEventCreate(ctypes.byref(e1)) == 0
        if ok:
            fill = lambda: hip.hipMemsetD32Async(buf, ctypes.c_int(0x3F800000),
                                                 ctypes.c_size_t(nbytes // 4), None)
            for _ in range(3):
                fill()
            hip.hipEventRecord(e0, None)
            for _ in range(10):
                fill()
            hip.hipEventRecord(e1, None)
            hip.hipEventSynchronize(e1)
            ms = ctypes.c_float()
            hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
            write_peak = nbytes * 10 / (ms.value * 1e-3) / 1e9
            hip.hipEventDestroy(e0)
            hip.hipEventDestroy(e1)
            hip.hipFree(buf)
            log("fill kernel: %.1f GB/s" % write_peak)
    except OSError as exc:
        log("fill kernel not measured: %s" % exc)

    # HBM bytes per traversal launch as rocprofv3's FETCH_SIZE + WRITE_SIZE
    # counters read them for this workload (scaled by the copy-kernel
    # calibration), recorded next to the profile summary they come from.
    traffic_hbm = None
    prof = os.path.join(ROOT, "profiles", "k_trav4h_dna_c2.json")
    if os.path.exists(prof):
        with open(prof) as f:
            rec = json.load(f)
        if rec["workload"] == {"taxa": 128, "patterns": S, "categories": 4}:
            traffic_hbm = int(rec["traffic_bytes"])

    line = {
        "metric": "lnL evals/s (full traversal, 128 taxa x ~60k patterns GTR+G4)",
        "value": round(full_eps, 2),
        "unit": "lnL evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 / full_eps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 CLV / f64 P,log-sum",
        "data": "synthetic (Yule tree, GTR+G4 simulated sites, seeded)",
        "config": {"workload": "single large DNA likelihood: 128 taxa, 100,000 sites -> "
                               "%d patterns, GTR+G4 alpha=0.5, full traversal" % S,
                   "patterns": S, "taxa": 128, "categories": 4},
        "partial_evals_per_s": round(part_eps, 2),
        "plain_gpu_evals_per_s": round(plain_eps, 2),
        "lnL": lnl0,
        "dlnL_abs": dl_abs,
        "dlnL_rel": dl_rel,
        "kernel_ms": {"pmat": round(float(tm[0]), 4), "traverse": round(trav_ms, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_hbm, "algorithmic_bytes": traffic, "node_updates": nodes,
                     "model_bytes": formula_bytes,
                     "model_GBps": round(formula_bytes / (trav_ms * 1e-3) / 1e9, 1),
                     "peak_kind": "published MI355X HBM3E bandwidth (spec)",
                     "measured_write_GBps": round(write_peak, 1) if write_peak else None,
                     "frac_of_measured_write": round(achieved / write_peak, 4) if write_peak else None},
        "cpu_baseline": cpu,
        "cpu_baseline_1thread": cpu1,
        "cpu_baseline_partial": cpup,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
