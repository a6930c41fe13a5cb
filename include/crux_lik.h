/*
 * crux_lik.h -- C ABI of the MI355X tree-likelihood engine.
 *
 * The reference's likelihood sits behind a Python class (Lik,
 * pkg/Crux/Tree/Lik.pxd:23-98) over a C core whose ABI is three functions on
 * caller-owned host FP64 buffers (lib/Crux/Tree/CxLik.h:279-285):
 *
 *     bool CxLikQDecomp(int n, double *RTri, double *PiDiag,
 *                       double *PiDiagNorm, double *qNorm,
 *                       double *qEigVecCube, double *qEigVals);
 *     void CxLikPt(int n, double *P, double *qEigVecCube, double *qEigVals,
 *                  double v);
 *     void CxLikExecute(CxtLik *lik);
 *
 * Host FP64 CLVs cannot stay device-resident, so this ABI replaces that
 * struct-passing core with opaque handles whose state lives in HBM: a tree
 * (alignment patterns + topology + a CLV pool with two polarity slots per
 * directed edge) and a lik (one mixture model + one polarity over a tree).
 * The Python Lik class (crux_amd/lik.py) keeps the reference's call surface
 * on top of it.  All entry points return an int status (CRUX_OK == 0); a
 * -inf lnL is a result, not an error.  A handle is used by one thread at a
 * time; calls on different trees of one context may come from different
 * threads.
 */
#ifndef CRUX_LIK_H
#define CRUX_LIK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    CRUX_OK = 0,
    CRUX_EINVAL = 1,    /* bad argument or shape */
    CRUX_ENOMEM = 2,    /* device allocation failed or HBM budget exceeded */
    CRUX_ENEGLEN = 3,   /* negative branch length (Lik.pyx:1423-1424) */
    CRUX_EDECOMP = 4,   /* eigendecomposition failed (Lik.pyx:472-475) */
    CRUX_EDEVICE = 5,   /* HIP runtime error */
    CRUX_ETAXON = 6,    /* leaf without a usable taxon (Lik.pyx:1449-1464) */
    CRUX_ESTATE = 7     /* call out of order (no model, no topology, ...) */
};

typedef struct crux_ctx crux_ctx;
typedef struct crux_tree crux_tree;
typedef struct crux_lik crux_lik;

/* Text of a status code, and the detail message of the calling thread's
 * last failure. */
const char *crux_strerror(int status);
const char *crux_last_error(void);

/* ---- FP64 model setup on the host ------------------------------------- */

/* Replaces CxLikQDecomp (CxLik.c:177-250).  rtri holds the n(n-1)/2
 * exchangeabilities in the reference's upper-triangle order (CxLik.c:36-53,
 * AC AG AT CG CT GT for DNA); pi the n (unnormalised) frequencies.  Outputs:
 * pi_norm[n], *q_norm (1 / mean rate of the unscaled Q, +inf for Q == 0),
 * evec[n*n] (V, row-major, columns are right eigenvectors), evec_inv[n*n]
 * (V^-1) and evals[n], with Q = V diag(evals) V^-1. */
int crux_q_decomp(int n, const double *rtri, const double *pi, double *pi_norm,
                  double *q_norm, double *evec, double *evec_inv, double *evals);

/* Replaces CxLikPt (CxLik.c:258-292): P = max(0, V exp(evals*t) V^-1),
 * row-major, row = parent state. */
int crux_pt(int n, const double *evec, const double *evec_inv,
            const double *evals, double t, double *P);

/* Gamma rate categories, replacing the CxMath use in Lik.setAlpha
 * (Lik.pyx:1196-1257; CxMath.pxd:125-368): cmult[ncat] with mean 1 from
 * category means (median == 0) or medians (median != 0). */
int crux_gamma_rates(double alpha, int ncat, int median, double *cmult);

/* ---- device context --------------------------------------------------- */

/* One context per device.  hbm_budget caps the bytes the context allocates
 * (0 = no cap).  The threaded CPU pool of the reference (CxLik.c:849-968)
 * has no counterpart: sites map to GPU lanes. */
int crux_ctx_create(int device, uint64_t hbm_budget, crux_ctx **out);
int crux_ctx_destroy(crux_ctx *ctx);
/* Bytes currently allocated by the context. */
uint64_t crux_ctx_bytes(const crux_ctx *ctx);

/* ---- trees -------------------------------------------------------------- */

/* Alignment data of a tree: nsites patterns (integer weights) over ntaxa
 * rows.  codes[ntaxa*nsites] indexes code_vecs[ncodes*dim], the 0/1 leaf
 * vector of each distinct state set (gap already mapped to "any",
 * Lik.pyx:1466-1475).  Replaces the leaf CL construction of
 * Lik._planRecurse and charFreqs (Lik.pyx:313-314,1446-1475). */
int crux_tree_create(crux_ctx *ctx, int dim, int nsites, int ntaxa,
                     const uint8_t *codes, int ncodes, const double *code_vecs,
                     const uint32_t *weights, crux_tree **out);
int crux_tree_destroy(crux_tree *tree);

/* Topology as ring arrays (a ring is one end of an edge; ring r belongs to
 * edge r/2 and its other end is r^1): ring_node[2*nedges], ring_next[2*nedges]
 * (next ring around the same node, -1 if unused), node_ring[nnodes] (first
 * ring of the node, -1 if none), node_taxon[nnodes] (alignment row, -1 for
 * an internal node, -2 for a label missing from the alignment) and
 * edge_len[nedges].  Ring and node ids must stay stable across calls so
 * cached CLVs survive; cache validity is re-derived per lnL call exactly as
 * Lik._planRecurse does (Lik.pyx:1500-1532).  The CLV slots of rings that
 * are detached (ring_node -1) are released. */
int crux_tree_set_topology(crux_tree *tree, int nnodes, int nedges,
                           const int32_t *ring_node, const int32_t *ring_next,
                           const int32_t *node_ring, const int32_t *node_taxon,
                           const double *edge_len);
/* Update some branch lengths (Edge.length writes). */
int crux_tree_set_lengths(crux_tree *tree, int n, const int32_t *edges,
                          const double *lens);
/* Partial topology update after Edge.attach/detach calls (the ring list
 * edits of pkg/Crux/Tree/__init__.pyx Edge.attach/detach): new ring_node and
 * ring_next of nrings rings, new node_ring of nnodes nodes, in the encoding
 * of crux_tree_set_topology; node and ring counts are unchanged.  The ring
 * cycles of every node touched are checked as crux_tree_set_topology does.
 * Cached CLVs are kept; their cache keys decide what is recomputed, as after
 * a full topology upload.  The CLV slots of rings left detached are
 * released (the reference drops a merged edge's cached CLs the same way,
 * pkg/Crux/Mc3/Chain.pyx:1171-1174). */
int crux_tree_update_rings(crux_tree *tree, int nrings, const int32_t *rings,
                           const int32_t *ring_node, const int32_t *ring_next,
                           int nnodes, const int32_t *nodes, const int32_t *node_ring);

/* ---- likelihoods ---------------------------------------------------------- */

/* A mixture model evaluated on a tree with CLV slots of one polarity
 * (Lik.clone() makes the mate with the other polarity, Lik.pyx:835-883). */
int crux_lik_create(crux_tree *tree, int polarity, crux_lik **out);
int crux_lik_destroy(crux_lik *lik);

/* Active mixture components (weightScaled != 0; zero-weight components are
 * skipped everywhere, CxLik.c:337,357,604,618).  For component c of model
 * comp_model[c]: weight comp_weight[c] (weightScaled) and time factor
 * comp_rate[c] = cmult * rmult * wNorm (CxLik.c:344-345).  Per model m:
 * evec/evec_inv/evals as from crux_q_decomp and pi_norm.  invalidate != 0
 * drops every cached CLV of this polarity (Lik.prep/_plan invalidate flag,
 * Lik.pyx:1639,1664,1278). */
int crux_lik_set_model(crux_lik *lik, int nmodels, const double *evec,
                       const double *evec_inv, const double *evals,
                       const double *pi_norm, int ncomp,
                       const int32_t *comp_model, const double *comp_weight,
                       const double *comp_rate, int invalidate);

/* lnL with root_node as the root (Lik.lnL, Lik.pyx:1687-1737): plans the
 * dirty CLVs, runs them on the device, sums the pattern-weighted site lnLs
 * with compensated FP64 accumulation.  Synchronous. */
int crux_lik_lnl(crux_lik *lik, int root_node, double *lnl);

/* No counterpart in the reference: crux_lik_lnl without any device or
 * pinned-memory allocation, for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * evaluation's plan, P matrices and per-wave partial sums go into a static
 * scratch of 16 MiB per device, held by the library until the process ends;
 * the CLV slots are the tree's own.  Fails with CRUX_EINVAL when the
 * evaluation needs more scratch than that, and with CRUX_ENOMEM when the
 * tree's CLV pool lacks a slot the plan needs or is laid out for another
 * component count (an evaluation with crux_lik_lnl at that root sets it up).
 * Every call on a device shares its scratch, which keeps the last call's
 * contents (nothing in it needs to start at zero); the library serializes
 * such calls within a process, since overlapping ones would race.
 * Synchronous. */
int crux_lik_lnl_noalloc(crux_lik *lik, int root_node, double *lnl);

/* Pattern-weighted site lnLs of the last evaluation at root_node
 * (Lik.siteLnLs, Lik.pyx:1739-1763); evaluates first when dirty. */
int crux_lik_site_lnl(crux_lik *lik, int root_node, double *out, int n);

/* Drop this polarity's cached CLVs (Lik.flush, Lik.pyx:1765-1787). */
int crux_lik_flush(crux_lik *lik);

/* ---- batches ------------------------------------------------------------ */

/* Evaluate n likelihoods in one fused device pass (one P-matrix launch, one
 * traversal launch per state-count/component-count class; the waves' partial
 * sums are finished on the host).  Each lik must be on a different
 * tree-polarity.  out[i] receives
 * lnL of liks[i] rooted at roots[i].  Synchronous.  Replaces, for a batch,
 * the per-Lik CxLikExecute calls of Lik.lnL (CxLik.h:285, Lik.pyx:1699). */
int crux_batch_lnl(crux_ctx *ctx, int n, crux_lik *const *liks,
                   const int32_t *roots, double *out);

/* The same, split: crux_batch_submit plans the batch and enqueues its device
 * work, returning at once with a ticket; crux_batch_wait(ticket) blocks until
 * that batch is done and copies its n results.  Two batches can be in flight
 * per context: the host plans and uploads the next while the device runs the
 * current one (a third submit first completes the oldest).  Batches run in
 * submission order, so a tree may appear in consecutive in-flight batches.
 * Every ticket is collected once.  No counterpart in the reference: this is
 * the asynchronous submit/poll surface of a serving loop. */
int crux_batch_submit(crux_ctx *ctx, int n, crux_lik *const *liks,
                      const int32_t *roots, uint64_t *ticket);
int crux_batch_wait(crux_ctx *ctx, uint64_t ticket, int n, double *out);

/* Non-blocking completion test of a submitted batch: *done = 1 once its
 * results are ready (crux_batch_wait then returns them at once), 0 while the
 * device still runs it.  Replaces, for a serving loop, the blocking wait of
 * CxLikExecute (CxLik.c:914-968) with a poll. */
int crux_batch_poll(crux_ctx *ctx, uint64_t ticket, int *done);

/* ---- serving loop -------------------------------------------------------- */

/* Continuous batching of likelihood requests on one context (no counterpart
 * in the reference, whose callers evaluate one proposal per synchronous
 * Lik.lnL call, pkg/Crux/Mc3/Chain.pyx:455).  A server thread forms a batch
 * from every request that has arrived while the device runs the previous
 * one (up to two batches in flight) and runs it as one fused device pass.
 * Requests of one tree run in submission order; a batch takes up to two
 * requests of a tree, one per polarity of the likelihoods added for it (a
 * likelihood and its clone, which must hold the same model), each planned
 * against the tree as the requests before it leave it.  Admission follows
 * the context's HBM budget: when a request needs memory the budget does not
 * allow, the CLV pools of the least recently used tree that is neither in
 * flight nor in the batch being formed are released (recomputed when next
 * used) and planning is retried; a request that still
 * does not fit waits for the batches in flight, and fails with CRUX_ENOMEM
 * only when nothing else holds memory.  While a likelihood is added, its
 * tree is edited and evaluated by the server only. */
typedef struct crux_server crux_server;
int crux_server_create(crux_ctx *ctx, int max_batch, crux_server **out);
/* Finishes every submitted request, then stops the server thread. */
int crux_server_destroy(crux_server *srv);
int crux_server_add_lik(crux_server *srv, crux_lik *lik);
/* Waits for the tree's requests to finish, then forgets the likelihood. */
int crux_server_remove_lik(crux_server *srv, crux_lik *lik);
/* Request: set the lengths of nedges edges (edges[i] gets lens[i]), drop
 * the likelihood's cached CLVs when flush != 0 (a full-tree evaluation),
 * then evaluate lnL rooted at root_node (the request may be served by the
 * tree's other added likelihood).  *req names it for poll and wait. */
int crux_server_submit(crux_server *srv, crux_lik *lik, int root_node, int nedges,
                       const int32_t *edges, const double *lens, int flush, uint64_t *req);
/* *done = 1 once the request has finished (never blocks). */
int crux_server_poll(crux_server *srv, uint64_t req, int *done);
/* Blocks until the request has finished; returns its status and, when it
 * succeeded, its lnL and completion time (CLOCK_MONOTONIC, ns).  Each
 * request is collected once. */
int crux_server_wait(crux_server *srv, uint64_t req, double *lnl, int64_t *t_done_ns);
/* [0] batches, [1] requests batched, [2] trees whose CLVs were released for
 * the HBM budget, [3] failed requests. */
int crux_server_stats(crux_server *srv, uint64_t *stats4);

/* Execution switches of the context (bench and tests):
 *   mode 0: fused (default) -- one traversal launch per batch;
 *   mode 1: plain -- one launch per node update plus one for the root,
 *           the "one call per eval, no fusion" comparison path. */
int crux_ctx_set_mode(crux_ctx *ctx, int mode);

/* Launch-shape tuning of the context (set by the autotuner,
 * crux_amd/autotune.py).  No counterpart in the reference, whose stripe
 * width is fixed (CxLik.h:176-273).  None of the shape keys changes a
 * result: every choice computes the same lnL bit for bit.
 *   CRUX_TUNE_STACK_DEPTH: LDS stack slots per lane of the traversal
 *   kernels (2, 4 or 8, capped by what fits; 0 = default: 4 for 20/22
 *   states, for 4 states see below).  Fewer slots let more blocks share a
 *   CU; more slots keep more pending results out of HBM.
 *   CRUX_TUNE_CHUNK: updates per staged chunk of the 4-state kernels, whose
 *   P images and control words are staged in LDS a chunk at a time and whose
 *   waves wait for their stores at each chunk's end: 0 = default, 1 short,
 *   2 middle, 3 long (per component class 1/2/4/8/16: 4-6-8, 7-14-16,
 *   6-11-16 (4-8-12 at stack depth 8), 2-4-6, 1-2-3 updates).
 *   Untuned 4-state depth and chunk follow the launch's blocks per CU: a
 *   block walks the whole plan for its 64 sites, so a launch with more
 *   blocks than fit on the chip at once runs a second round of whole block
 *   lifetimes, and a shape with less LDS per block may fit them all.
 *   CRUX_TUNE_WAVES: waves per SIMD the 20/22-state kernel's registers are
 *   held to, 1 (up to 512 registers per lane) or 2 (up to 256); 0 = default
 *   (2 for 20 states with up to 4 components, else 1).
 *   CRUX_TUNE_PRECISION: input precision of the 20/22-state P * CLV
 *   products on the matrix core, accumulated in FP32 (CLVs are stored in
 *   FP32 either way): CRUX_PREC_F32 (default), CRUX_PREC_BF16, CRUX_PREC_F16,
 *   or CRUX_PREC_F16X3 (float16 high and low parts, three products).  The
 *   16-bit forms run at stack depth 4 and the default waves.  Likelihoods
 *   evaluated at another precision recompute their CLVs.
 *   CRUX_TUNE_SHORT_PLANS: 1 (default) runs 4-state classes whose plans are
 *   all short (at most 24 updates, two or four components: partial
 *   re-evaluations, MC3 proposals) on the short-plan kernel, which stages
 *   the whole plan at once and loads tip rows and cached children two
 *   updates ahead; 0 runs them on the general kernel.  Same results, bit
 *   for bit, either way.
 *   CRUX_TUNE_SPLIT: blocks a class launch aims at, per SIMD for 20/22
 *   states and per two CUs for 4 states (1 to 16, default 3; 0 never
 *   splits).  A class launched with fewer site tiles than that (or fewer
 *   than twice that many, with at least 200 updates of work per block
 *   aimed at) runs each plan of 32 updates or more, outside the short-plan
 *   kernel, as its largest subtrees of at most (updates x tiles) / (blocks
 *   aimed at) updates and half the plan, one block per subtree and tile in
 *   a first launch, and the updates above them in a second launch that
 *   reads the subtrees' results from their slots.  Same results, bit for
 *   bit, either way.  Also a class key (crux_ctx_set_class_tuning), which
 *   crux_amd.autotune sets to 0 where one launch measures faster.
 * Settings apply to batches submitted after the call; do not change them
 * while another thread submits on the context. */
enum { CRUX_PREC_F32 = 0, CRUX_PREC_BF16 = 1, CRUX_PREC_F16 = 2, CRUX_PREC_F16X3 = 3 };
enum {
    CRUX_TUNE_STACK_DEPTH = 1,
    CRUX_TUNE_PRECISION = 2,
    CRUX_TUNE_CHUNK = 3,
    CRUX_TUNE_WAVES = 4,
    CRUX_TUNE_SHORT_PLANS = 5,
    CRUX_TUNE_SPLIT = 6
};
int crux_ctx_set_tuning(crux_ctx *ctx, int key, int value);
int crux_ctx_get_tuning(const crux_ctx *ctx, int key, int *value);

/* The same shape keys and CRUX_TUNE_SPLIT (not precision or short plans)
 * for one traversal class: K states and C components (the class of the
 * smallest instantiated component count >= C).  value -1 clears the class's
 * setting (the context's value applies).  A batch mixing classes launches
 * each with its own shape. */
int crux_ctx_set_class_tuning(crux_ctx *ctx, int K, int C, int key, int value);
int crux_ctx_get_class_tuning(const crux_ctx *ctx, int K, int C, int key, int *value);

/* The launch shape a batch of one evaluation of S patterns (S after
 * compaction) would use now for K states and C components: shape4 = {stack
 * depth, updates per chunk (4 states; else 0), waves per SIMD (20/22 states;
 * else 0), component class CM}.  Untuned 4-state keys depend on the
 * launch's blocks per CU (a batch's class launch covers every evaluation of
 * the class in it). */
int crux_ctx_launch_shape(const crux_ctx *ctx, int K, int C, int S, int *shape4);

/* Timing of the last evaluation on the device stream, in milliseconds, from
 * events the runtime stamps when each kernel starts and ends: [0] P-matrix
 * kernel, [1] traversal kernel(s), [2] the gap between the two (launch
 * latency), [3] total from the P-matrix kernel's start to the traversal's
 * end (the final sum runs on the host).
 * Only filled when timing was enabled with crux_ctx_set_timing(ctx, 1). */
int crux_ctx_set_timing(crux_ctx *ctx, int on);
int crux_ctx_last_timing(const crux_ctx *ctx, double *ms4);

/* Host time of the last batch, in microseconds, always recorded: [0]
 * planning its likelihoods (cache walk and plan emission), [1] building the
 * batch (stack schedule, control words, P jobs, entries), [2] the launch
 * calls, [3] waiting for its completion and the final sums.  No counterpart
 * in the reference. */
int crux_ctx_last_host_timing(const crux_ctx *ctx, double *us4);

/* Algorithmic HBM traffic of the last evaluation in bytes: `bytes` counts
 * what the kernels must move (CLV and exponent writes, CLVs of children cached
 * by earlier evaluations, tip codes, weights, site lnLs); `formula_bytes`
 * counts every child CLV as read from HBM (the per-node-update byte model
 * B = sum(child reads) + parent write).  `nodes` is the number of node
 * updates. */
int crux_ctx_last_traffic(const crux_ctx *ctx, uint64_t *bytes, uint64_t *formula_bytes,
                          int *nodes);

/* Traversal kernel launches of the last batch (one per state/component
 * class with work; one per node update in one-launch-per-update mode), and
 * how many of them ran the short-plan kernel (4 states, two or four
 * components, every plan of the class at most 24 node updates; see
 * CRUX_TUNE_SHORT_PLANS).  A measurement hook with no counterpart in the
 * reference. */
int crux_ctx_last_launches(const crux_ctx *ctx, int *launches, int *short_launches);

#ifdef __cplusplus
}
#endif

#endif /* CRUX_LIK_H */
