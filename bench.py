#!/usr/bin/env python3
"""Benchmark of the tree-likelihood hot path on one MI355X.

Headline workload (``value``): the single large DNA likelihood -- 128-taxon
Yule tree, 100,000 simulated sites compressed to ~60k patterns with integer
weights, GTR+Gamma4 (alpha 0.5), FP32 CLVs with power-of-two site scaling,
compensated FP64 site sums.  A step is one full-traversal lnL evaluation
(every CLV recomputed) with all data resident in HBM.

Also reported in the same JSON line:
  * single-branch partial re-evaluation (re-root at a changed edge, as an
    MC3 branch-length move does) and the plain one-launch-per-node path;
  * the traversal kernel's roofline (algorithmic bytes / kernel time);
  * amino acids, 256 taxa x 20,000 sites, 20 states +G4: evals/s, HBM and
    matrix-core fractions, |dlnL| against the FP64 oracle;
  * batched MC3, 32 runs x 4 heated replicas on 64-512 taxa x 1k-50k sites:
    fused batched passes (host/device overlapped and one pass per step),
    one fused call per evaluation, and one launch per node update;
  * served likelihood: Poisson arrivals of 70% single-branch / 30% full-tree
    requests on 16 resident trees of 32-1,024 taxa x 500-200k sites, DNA and
    amino acids: throughput and p50/p99 latency per arrival rate;
  * the FP64 CPU baseline (the oracle's C restatement of the reference's
    engine) on the host cores of the same machine: median of 20 evaluations
    after 3 warm-ups, at one thread and at every available core.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--skip c3,c4,c5,cpu]
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
FP32_MATRIX_PEAK_TFLOPS = 157.3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    """The host CPU as /proc/cpuinfo names its first processor: model name,
    family / model / stepping, clock and cache size (a virtualised host may
    report only a short model name)."""
    info = {}
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if not ln.strip():
                    break
                k, _, v = ln.partition(":")
                info[k.strip()] = v.strip()
    except OSError:
        return None
    name = info.get("model name")
    extra = ["%s %s" % (k, info[k]) for k in ("cpu family", "model", "stepping", "cpu MHz",
                                              "cache size") if k in info]
    return "%s (%s)" % (name, ", ".join(extra)) if extra else name


def available_cores():
    """Cores this process may use: its CPU affinity, capped by a cgroup CPU
    quota when one is set."""
    n = len(os.sched_getaffinity(0))
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            quota, period = f.read().split()[:2]
        if quota != "max":
            n = min(n, max(1, int(math.ceil(int(quota) / int(period)))))
    except (OSError, ValueError):
        pass
    return n


# ---------------------------------------------------------------------------
# Single large DNA likelihood (the headline).

def tuner_sweep(lik, root):
    """Traversal time at every launch shape of lik's class (the autotuner's
    grid), next to the untuned default the other figures are measured at.
    Leaves the class untuned."""
    from crux_amd import autotune

    ctx = lik._ctx
    default = autotune.launch_of(lik)
    best, times = autotune.tune(lik, root, reps=10, use_cache=False)
    tuned = autotune.launch_of(lik)
    autotune.set_class_shape(lik, None)
    # The default is one of the grid's launches: its time is that launch's.
    by, dflt_ms = {}, None
    for sh, ms in times.items():
        autotune.set_class_shape(lik, sh)
        launch = autotune.launch_of(lik)
        by["depth %d chunk %d waves %d" % launch[:3]] = round(ms, 4)
        if launch == default:
            dflt_ms = ms
    autotune.set_class_shape(lik, None)
    return {"default_launch": {"depth": default[0], "chunk": default[1], "waves": default[2]},
            "tuned_launch": {"depth": tuned[0], "chunk": tuned[1], "waves": tuned[2]},
            "default_ms": None if dflt_ms is None else round(dflt_ms, 4),
            "tuned_ms": round(times[best], 4), "traverse_ms_by_shape": by}


def bench_dna(args, out):
    from crux_amd import _native as nat
    from crux_amd import synth
    from crux_amd.lik import Lik

    t0 = time.time()
    prob = synth.make_problem(128, 100000, 2, mean_len=0.025)
    S = prob.alignment.nchars
    log("dna: 128 taxa, %d patterns, %.1fs" % (S, time.time() - t0))
    ctx = nat.get_context(0)
    if args.stack_depth is not None:
        ctx.set_tuning(nat.CRUX_TUNE_STACK_DEPTH, args.stack_depth)
    lik = Lik(prob.tree, prob.alignment, ncat=4)
    synth.configure(lik, prob.model)
    root = prob.tree.base
    lnl0 = lik.lnL(root)

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
    log("dna full: %.1f evals/s, %.3f ms/eval" % (full_eps, 1e3 * dt / args.steps))

    # The same full evaluations with two in flight (crux_batch_submit /
    # crux_batch_wait): each is planned and uploaded while the previous one
    # runs on the device.  Reported beside `value`, which is the synchronous
    # one-call-per-evaluation rate of the reference's lnL.
    def submit_full():
        lik.flush()
        return nat.batch_submit(ctx, [lik], [lik._ready(root)])

    pipe_eps = None
    try:
        pending = submit_full()
        t = time.perf_counter()
        for _ in range(args.steps):
            nxt = submit_full()
            lnl_p = float(nat.batch_wait(ctx, pending)[0])
            pending = nxt
        lnl_p = float(nat.batch_wait(ctx, pending)[0])
        dt = time.perf_counter() - t
        pipe_eps = args.steps / dt
        log("dna full, two in flight: %.1f evals/s (lnL %s the synchronous one)"
            % (pipe_eps, "equals" if lnl_p == lnl0 else "DIFFERS FROM"))
    except Exception as e:  # the headline figures do not depend on it
        log("dna full, two in flight: failed (%s)" % e)

    # Traversal kernel duration: HIP events the runtime stamps at the
    # kernel's start and end on the context's stream, averaged over as many
    # evaluations as the timed region ran (a second pass, so the events stay
    # out of the evals/s figure).
    ctx.set_timing(True)
    tr, pm, gap = [], [], []
    for _ in range(max(args.steps, 20)):
        full()
        ms = ctx.last_timing()
        pm.append(ms[0])
        tr.append(ms[1])
        gap.append(ms[2])
    ctx.set_timing(False)
    trav_ms = float(np.mean(tr))
    achieved = traffic / (trav_ms * 1e-3) / 1e9
    # Host side of an evaluation (median over 50, device timing off):
    # planning, building the batch, launch calls, completion wait + sums.
    hs = []
    for _ in range(50):
        full()
        hs.append(ctx.last_host_timing().copy())
    host_us = [round(float(x), 2) for x in np.median(np.array(hs), axis=0)]
    log("dna host us: plan %.1f build %.1f launch %.1f wait+sum %.1f" % tuple(host_us))
    log("dna traversal %.4f ms, P-matrix %.4f ms, gap %.4f ms: %.1f GB/s" %
        (trav_ms, float(np.mean(pm)), float(np.mean(gap)), achieved))

    tune = tuner_sweep(lik, root) if args.stack_depth is None else None
    if tune:
        log("dna shapes: default %.4f ms, tuned %.4f ms %s" %
            (tune["default_ms"], tune["tuned_ms"], tune["tuned_launch"]))

    # Plain path: one launch per node update.
    ctx.set_mode(1)
    try:
        for _ in range(3):
            full()
        n_plain = max(10, args.steps // 2)
        t = time.perf_counter()
        for _ in range(n_plain):
            full()
        plain_eps = n_plain / (time.perf_counter() - t)
    finally:
        ctx.set_mode(0)
    log("dna plain: %.1f evals/s" % plain_eps)

    # Single-branch partial re-evaluation, re-rooted at the changed edge.
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
    log("dna partial: %.1f evals/s (%.1f node updates)" % (part_eps, float(np.mean(part_nodes))))

    # The same partial re-evaluations with two in flight: the next edit is
    # planned while the previous evaluation runs (requests that do not wait
    # on each other's results, as in serving; an MC3 chain's next proposal
    # does wait on its acceptance).
    def submit_partial():
        e = edges[int(rng.integers(len(edges)))]
        e.length = e.length * math.exp(lam * (rng.random() - 0.5))
        tk = nat.batch_submit(ctx, [lik], [lik._ready(e.ring.node)])
        prob.tree.base = e.ring.node
        return tk

    part_pipe_eps = None
    try:
        pending = submit_partial()
        t = time.perf_counter()
        for _ in range(args.partial_steps):
            nxt = submit_partial()
            nat.batch_wait(ctx, pending)
            pending = nxt
        nat.batch_wait(ctx, pending)
        part_pipe_eps = args.partial_steps / (time.perf_counter() - t)
        log("dna partial, two in flight: %.1f evals/s" % part_pipe_eps)
    except Exception as e:  # the headline figures do not depend on it
        log("dna partial, two in flight: failed (%s)" % e)

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

    out.update({
        "value": round(full_eps, 2),
        "ms_per_step": round(1e3 / full_eps, 4),
        "config": {"workload": "single large DNA likelihood: 128 taxa, 100,000 sites -> "
                               "%d patterns, GTR+G4 alpha=0.5, full traversal" % S,
                   "patterns": S, "taxa": 128, "categories": 4},
        "full_evals_per_s_two_in_flight": None if pipe_eps is None else round(pipe_eps, 2),
        "partial_evals_per_s": round(part_eps, 2),
        "partial_evals_per_s_two_in_flight": None if part_pipe_eps is None else round(part_pipe_eps, 2),
        "partial_node_updates": round(float(np.mean(part_nodes)), 2),
        "plain_gpu_evals_per_s": round(plain_eps, 2),
        "speedup_vs_plain_gpu": round(full_eps / plain_eps, 2),
        "lnL": lnl0,
        "autotune": tune,
        "kernel_ms": {"pmat": round(float(np.mean(pm)), 4), "traverse": round(trav_ms, 4),
                      "launch_gap": round(float(np.mean(gap)), 4)},
        "host_us": dict(zip(("plan", "build", "launch", "wait_and_sum"), host_us)),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_hbm, "algorithmic_bytes": traffic, "node_updates": nodes,
                     "model_bytes": formula_bytes,
                     "model_GBps": round(formula_bytes / (trav_ms * 1e-3) / 1e9, 1),
                     "peak_kind": "published MI355X HBM3E bandwidth (spec)"},
    })
    return prob, int(round(float(np.mean(part_nodes))))


def measured_write_peak(out):
    """Write bandwidth of the HIP runtime's fill kernel on this GPU
    (hipMemsetD32Async over 512 MiB, HIP events on the null stream): the
    practical ceiling of an almost write-only kernel."""
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError as exc:
        log("fill kernel not measured: %s" % exc)
        return
    nbytes = 512 << 20
    buf, e0, e1 = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    ok = hip.hipMalloc(ctypes.byref(buf), ctypes.c_size_t(nbytes)) == 0
    ok = ok and hip.hipEventCreate(ctypes.byref(e0)) == 0 and hip.hipEventCreate(ctypes.byref(e1)) == 0
    if not ok:
        return
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
    peak = nbytes * 10 / (ms.value * 1e-3) / 1e9
    hip.hipEventDestroy(e0)
    hip.hipEventDestroy(e1)
    hip.hipFree(buf)
    log("fill kernel: %.1f GB/s" % peak)
    rf = out["roofline"]
    rf["measured_write_GBps"] = round(peak, 1)
    rf["frac_of_measured_write"] = round(rf["achieved"] / peak, 4)


# ---------------------------------------------------------------------------
# FP64 CPU baseline and parity.

def cpu_baseline(args, out, prob2, part_k, aa_prob):
    """The oracle's FP64 C restatement of CxLikExecute on this host, median
    of 20 after 3 warm-ups (CL buffers reused as the reference reuses its
    CL objects), at one thread and at every available core; also the FP64
    lnL each device result is compared with."""
    from crux_amd import synth
    from oracle.oracle import OracleLik

    ncpu = available_cores()
    model = cpu_model()

    def oracle(prob, nthreads):
        o = OracleLik(prob.fresh_tree(), prob.alignment, ncat=4, nthreads=nthreads)
        synth.configure(o, prob.model)
        return o

    res = {}
    # Headline workload, all cores and one thread.
    o = oracle(prob2, ncpu)
    ref = o.lnL()
    tn = o.time_full(reps=20, warmup=3)
    o1 = oracle(prob2, 1)
    t1 = o1.time_full(reps=20, warmup=3)
    tp = o.time_partial(part_k, reps=20, warmup=3)
    log("cpu dna: %.3f s (%d threads), %.3f s (1 thread), partial %.4f s" % (tn, ncpu, t1, tp))
    res["dna_ref_lnL"] = ref
    out["cpu_baseline"] = {
        "value": round(1.0 / tn, 3), "unit": "lnL evals/s", "cores": ncpu, "kind": "port",
        "cpu_model": model,
        "sample": "median of 20 full-traversal evaluations (after 3 warm-ups) of the same "
                  "128-taxon, %d-pattern GTR+G4 workload by the FP64 C restatement of the "
                  "reference engine (oracle/cxlik_oracle.c) with %d stripe threads"
                  % (prob2.alignment.nchars, ncpu)}
    out["cpu_baseline_1thread"] = {
        "value": round(1.0 / t1, 3), "unit": "lnL evals/s", "cores": 1, "kind": "port",
        "sample": "median of 20 full evaluations after 3 warm-ups, one thread"}
    out["cpu_baseline_partial"] = {
        "value": round(1.0 / tp, 2), "unit": "lnL evals/s", "cores": ncpu, "kind": "port",
        "sample": "median of 20 executions (after 3) of the last %d CLV updates and the root "
                  "of the workload's plan: the work of a partial re-evaluation (the device's "
                  "mean node-update count)" % part_k}
    # The reference test size (16 taxa, 1,000 sites).
    p1 = synth.make_problem(16, 1000, 1)
    c1n = oracle(p1, ncpu).time_full(reps=20, warmup=3)
    c1s = oracle(p1, 1).time_full(reps=20, warmup=3)
    out["cpu_baseline_c1"] = {
        "workload": "16 taxa, 1,000 sites -> %d patterns, GTR+G4" % p1.alignment.nchars,
        "evals_per_s_all_cores": round(1.0 / c1n, 1), "evals_per_s_1thread": round(1.0 / c1s, 1),
        "cores": ncpu, "lnL": oracle(p1, 1).lnL()}
    # Amino acids.
    if aa_prob is not None:
        oa = oracle(aa_prob, ncpu)
        res["aa_ref_lnL"] = oa.lnL()
        ta = oa.time_full(reps=20, warmup=3)
        ta1 = oracle(aa_prob, 1).time_full(reps=3, warmup=1)
        res["aa_cpu"] = {"value": round(1.0 / ta, 3), "unit": "lnL evals/s", "cores": ncpu,
                         "kind": "port", "value_1thread": round(1.0 / ta1, 3),
                         "sample": "median of 20 after 3 warm-ups (%d threads); one thread: "
                                   "median of 3 after 1" % ncpu}
        log("cpu aa: %.3f s (%d threads), %.3f s (1 thread)" % (ta, ncpu, ta1))
    return res


# ---------------------------------------------------------------------------
# Amino acids, 20 states.

def bench_aa(args):
    from crux_amd import _native as nat
    from crux_amd import synth
    from crux_amd.lik import Lik

    t0 = time.time()
    prob = synth.make_problem(256, 20000, 4, kind="aa")
    S = prob.alignment.nchars
    log("aa: 256 taxa, %d patterns, %.1fs" % (S, time.time() - t0))
    ctx = nat.get_context(0)
    lik = Lik(prob.tree, prob.alignment, ncat=4)
    synth.configure(lik, prob.model)
    root = prob.tree.base

    def full():
        lik.flush()
        return lik.lnL(root)

    lnl = full()
    for _ in range(3):
        full()
    n = max(10, args.steps)
    t = time.perf_counter()
    for _ in range(n):
        full()
    eps = n / (time.perf_counter() - t)
    ctx.set_timing(True)
    tr = []
    for _ in range(n):
        full()
        tr.append(ctx.last_timing()[1])
    ctx.set_timing(False)
    trav = float(np.mean(tr))
    nbytes, _, nodes = ctx.last_traffic()
    # Matrix-core FP32 work: per internal child edge and component, a 20x20
    # matrix times a 20-vector per pattern (tip children are table rows).
    flops = 2.0 * 20 * 20 * S * 4 * (prob.alignment.ntaxa - 3)
    tflops = flops / (trav * 1e-3) / 1e12
    gbs = nbytes / (trav * 1e-3) / 1e9
    log("aa full: %.1f evals/s; traversal %.4f ms; %.1f GB/s; %.1f TFLOP/s" % (eps, trav, gbs, tflops))
    tune = tuner_sweep(lik, root)
    log("aa shapes: default %.4f ms, tuned %.4f ms %s" %
        (tune["default_ms"], tune["tuned_ms"], tune["tuned_launch"]))
    # Opt-in 16-bit inputs to the matrix core (FP32 accumulation; CLVs
    # stay FP32): traversal time and lnL of each, compared with the FP64
    # oracle below.
    precs = {}
    for name, prec in (("bf16", nat.CRUX_PREC_BF16), ("f16", nat.CRUX_PREC_F16),
                       ("f16x3", nat.CRUX_PREC_F16X3)):
        ctx.set_tuning(nat.CRUX_TUNE_PRECISION, prec)
        v = full()
        for _ in range(3):
            full()
        ctx.set_timing(True)
        tp = []
        for _ in range(n):
            full()
            tp.append(ctx.last_timing()[1])
        ctx.set_timing(False)
        ms = float(np.mean(tp))
        precs[name] = {"lnL": v, "traverse_ms": round(ms, 4),
                       "hbm_frac": round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        log("aa %s: traversal %.4f ms, lnL %.6f" % (name, ms, v))
    ctx.set_tuning(nat.CRUX_TUNE_PRECISION, nat.CRUX_PREC_F32)
    return prob, {
        "workload": "amino acids: 256 taxa, 20,000 sites -> %d patterns, LogNormal "
                    "exchangeabilities, Dirichlet(1) freqs, +G4 alpha=0.5, full traversal" % S,
        "evals_per_s": round(eps, 2), "traverse_ms": round(trav, 4), "node_updates": nodes,
        "lnL": lnl,
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes": nbytes},
        "matrix_core": {"tflops": round(tflops, 2), "peak_tflops_fp32": FP32_MATRIX_PEAK_TFLOPS,
                        "frac": round(tflops / FP32_MATRIX_PEAK_TFLOPS, 4)},
        "input_precisions": precs,
        "autotune": tune,
    }


# ---------------------------------------------------------------------------
# Batched MC3.

def bench_mc3(args):
    from crux_amd import mc3
    from crux_amd import synth
    from oracle.oracle import OracleLik

    t0 = time.time()
    probs = mc3.make_problems(32)
    log("mc3: 32 problems, %.1fs" % (time.time() - t0))
    mc = mc3.Mc3Batch(probs, ncoupled=4)
    mc.initialize()
    mc.step(batched=True)

    def rate(fn, steps):
        e0 = mc.evals
        t = time.perf_counter()
        fn(steps)
        return (mc.evals - e0) / (time.perf_counter() - t)

    def stepped(k, batched=True):
        for _ in range(k):
            mc.step(batched=batched)

    one_pass = rate(stepped, args.mc3_steps)
    overlapped = rate(mc.run, args.mc3_steps)
    fused_single = rate(lambda k: stepped(k, batched=False), 2)
    mc.ctx.set_mode(1)
    try:
        plain = rate(lambda k: stepped(k, batched=False), 2)
    finally:
        mc.ctx.set_mode(0)
    # Tracked lnL of a sample of chains against the FP64 oracle on their
    # current trees.
    worst = 0.0
    for c in mc.chains[::37]:
        o = OracleLik(c.tree.dup(), c.lik.alignment, ncat=4, nthreads=available_cores())
        synth.configure(o, probs[c.run].model)
        ref = o.lnL()
        worst = max(worst, abs(c.lnL - ref) / abs(ref))
    log("mc3: overlapped %.0f, one pass per step %.0f, fused one call per eval %.0f, "
        "plain %.0f evals/s; worst |dlnL|/|lnL| %.2e" % (overlapped, one_pass, fused_single,
                                                          plain, worst))
    return {"workload": "batched MC3: 32 runs x 4 heated replicas (128 chains), 64-512 taxa x "
                        "1k-50k sites log-uniform, GTR+G4, branch-length + NNI moves",
            "batched_evals_per_s": round(overlapped, 1),
            "batched_one_pass_per_step_evals_per_s": round(one_pass, 1),
            "fused_one_call_per_eval_evals_per_s": round(fused_single, 1),
            "plain_evals_per_s": round(plain, 1),
            "speedup_vs_plain": round(overlapped / plain, 2),
            "speedup_vs_fused_unbatched": round(overlapped / fused_single, 2),
            "dlnL_rel_max_sample": worst}


# ---------------------------------------------------------------------------
# Served likelihood.

def bench_serve(args):
    from crux_amd import _native as nat
    from crux_amd import serve

    t0 = time.time()
    sessions = serve.make_sessions(16)
    log("serve: 16 sessions, %.1fs" % (time.time() - t0))
    ctx = nat.get_context(0)
    srv = serve.Server(sessions, ctx)
    srv.run(serve.make_requests(sessions, 64, 1e9, seed=1))
    sweep = []
    for r in args.serve_rates:
        n = int(min(args.serve_requests, max(200, 2.0 * r)))
        reqs = serve.make_requests(sessions, n, r, seed=int(r) % 1000 + 7)
        b0, q0 = srv.batches, srv.batched_requests
        t = time.perf_counter()
        srv.run(reqs)
        s = serve.summarize(reqs, time.perf_counter() - t)
        s["offered_rps"] = r
        s["mean_batch"] = round((srv.batched_requests - q0) / max(1, srv.batches - b0), 1)
        log("serve rate %g: %s" % (r, s))
        sweep.append(s)
    srv.close()
    return {"workload": "served likelihood: Poisson arrivals, 70% single-branch / 30% full, "
                        "16 resident trees of 32-1,024 taxa x 500-200k sites, DNA and 20-state "
                        "amino acids 1:1, GTR+G4",
            "sweep": sweep}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--partial-steps", type=int, default=400)
    ap.add_argument("--mc3-steps", type=int, default=10)
    ap.add_argument("--serve-rates", default="500,2000,8000,100000")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--skip", default="", help="comma list of c3,c4,c5,cpu to leave out")
    ap.add_argument("--stack-depth", type=int, default=None,
                    help="4-state kernel LDS stack depth (2, 4, 8; default: tuned)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (the reference "
                         "kernel for scaling HBM byte counters) and exit")
    args = ap.parse_args()
    args.serve_rates = [float(x) for x in args.serve_rates.split(",") if x]
    skip = set(x for x in args.skip.split(",") if x)
    if args.gpus != 1:
        print("bench.py measures one GPU: the likelihood runs on a single device; "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
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

    line = {
        "metric": "lnL evals/s (full traversal, 128 taxa x ~60k patterns GTR+G4)",
        "value": None,
        "unit": "lnL evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 CLV / f64 P,log-sum",
        "data": "synthetic (Yule trees, GTR+G4 / 20-state sites simulated, seeded)",
    }
    prob, part_k = bench_dna(args, line)
    measured_write_peak(line)
    aa_prob, aa = (None, None) if "c4" in skip else bench_aa(args)
    mc3r = None if "c3" in skip else bench_mc3(args)
    srv = None if "c5" in skip else bench_serve(args)
    refs = {}
    if "cpu" not in skip:
        from crux_amd import synth
        prob2 = synth.make_problem(128, 100000, 2, mean_len=0.025)
        refs = cpu_baseline(args, line, prob2, part_k, aa_prob)
        ref = refs["dna_ref_lnL"]
        line["dlnL_abs"] = abs(line["lnL"] - ref)
        line["dlnL_rel"] = line["dlnL_abs"] / abs(ref)
        line["speedup_vs_cpu"] = round(line["value"] / line["cpu_baseline"]["value"], 1)
        line["partial_speedup_vs_cpu"] = round(
            line["partial_evals_per_s"] / line["cpu_baseline_partial"]["value"], 1)
        log("dna |dlnL| %.3e rel %.3e" % (line["dlnL_abs"], line["dlnL_rel"]))
    if aa is not None:
        if "aa_ref_lnL" in refs:
            aa["dlnL_abs"] = abs(aa["lnL"] - refs["aa_ref_lnL"])
            aa["dlnL_rel"] = aa["dlnL_abs"] / abs(refs["aa_ref_lnL"])
            aa["cpu_baseline"] = refs["aa_cpu"]
            aa["speedup_vs_cpu"] = round(aa["evals_per_s"] / refs["aa_cpu"]["value"], 1)
            for v in aa["input_precisions"].values():
                v["dlnL_rel"] = abs(v["lnL"] - refs["aa_ref_lnL"]) / abs(refs["aa_ref_lnL"])
        line["amino_acids"] = aa
    if mc3r is not None:
        line["mc3"] = mc3r
    if srv is not None:
        line["serve"] = srv
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
