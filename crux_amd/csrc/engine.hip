// Host engine of libcrux_amd: device context, trees with device-resident CLV
// pools, the likelihood planner and batched execution.
//
// Planner semantics follow Lik._plan/_planRecurse (pkg/Crux/Tree/Lik.pyx:
// 1427-1601): every directed edge ("ring") has, per polarity, a cached CLV
// with a validity flag and a cache key (parent ring, number of siblings,
// branch length).  A key mismatch invalidates the parent; a recomputed CLV
// invalidates its parent in turn; the root CL is recomputed whenever any of
// its children changed.  Unlike the reference, all child steps of one parent
// become ONE fused node update (children multiplied in registers), and the
// plan is executed for all sites by a single device launch.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "crux_lik.h"
#include "device.h"
#include "internal.h"

namespace crux {

int traverse_cmax(int K, int C);
// Optional events (null: none) are stamped when the kernel starts and ends.
int launch_traverse(int K, int CM, const Entry *entries, const int32_t *block_entry,
                    int blk0, int nblocks, const TravShape &sh, hipStream_t stream, hipEvent_t t0,
                    hipEvent_t t1);
void launch_pmat(const PJob *jobs, int njobs, size_t shmem, const void *src, void *dst,
                 size_t nbytes, hipStream_t stream, hipEvent_t t0, hipEvent_t t1);

static thread_local char g_errbuf[512] = "";

int set_error(int status, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_errbuf, sizeof g_errbuf, fmt, ap);
    va_end(ap);
    return status;
}

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t err_ = (expr);                                                      \
        if (err_ != hipSuccess)                                                        \
            return set_error(err_ == hipErrorOutOfMemory ? CRUX_ENOMEM : CRUX_EDEVICE, \
                             "%s: %s", #expr, hipGetErrorString(err_));                \
    } while (0)

static size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

}  // namespace crux

using namespace crux;

// ---------------------------------------------------------------------------
// Context.

struct crux_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t budget = 0;
    uint64_t used = 0;             // bytes allocated (under amu)
    int mode = 0;
    int ncu = 256;                 // compute units of the device
    int depth4 = 0;                // traversal LDS stack depth (0 = default)
    int chunk4 = 0;                // 4-state chunk choice (0 = default, 1-3)
    int waves20 = 0;               // 20/22-state waves per SIMD (0 = default)
    int prec = kPrecF32;           // input precision of 20/22-state products
    bool short_plans = true;       // 4-state short plans on the short-plan kernel
    int split20 = 3;               // 20/22-state split: blocks per SIMD aimed at (0: never)
    // Per-class settings of the three shape keys (-1: the context's value),
    // indexed by class_index(K, CM).
    int ctune[kNumClasses][4];   // depth, chunk, waves, split (-1: the context's)
    float *d_zero = nullptr;       // kZeroBytes of zeros (20/22-state kernel)
    bool timing = false;
    double last_ms[4] = {0, 0, 0, 0};
    // Host time of the last batch (us): planning, building, launch calls,
    // completion wait + final sums; plan_us accumulates the planning of the
    // likelihoods of the batch being formed.
    double host_us[4] = {0, 0, 0, 0};
    double plan_us = 0.0;
    uint64_t last_bytes = 0;      // HBM bytes the last evaluation had to move
    uint64_t last_formula = 0;    // same with every child CLV read from HBM
    int last_nodes = 0;
    int last_launches = 0;        // traversal launches of the last batch
    int last_short = 0;           // ... of which ran the short-plan kernel
    // Two staging sets, so the host plans and uploads one batch while the
    // device runs the previous one (crux_batch_submit / crux_batch_wait).
    // Batches run in submission order on the one stream, which is what makes
    // a tree's cache bookkeeping (done when a batch is planned) valid for the
    // next batch before the previous one has finished.
    struct Staging {
        // The plan: written by the host into coherent pinned memory
        // (h_plan; the device reads it at dh_plan) and copied to device
        // memory (d_plan) by the P-matrix launch, which reads its own jobs
        // straight from h_plan.
        char *h_plan = nullptr;
        char *dh_plan = nullptr;
        char *d_plan = nullptr;
        size_t plan_cap = 0;
        float *d_pmat = nullptr;       // P blocks and tip tables
        size_t pmat_cap = 0;           // floats
        // Per-wave (sum, compensation) pairs, [sum ntiles][2], in coherent
        // pinned host memory: the traversal writes them there and the host
        // finishes each evaluation's sum once the batch is done.
        double *h_partials = nullptr;
        double *d_partials = nullptr;  // the device's address of h_partials
        size_t partials_cap = 0;       // doubles

        hipEvent_t done = nullptr;
        // Timing events of this slot's batch: the P-matrix kernel's start
        // and end, the first traversal's start and the last one's end.
        hipEvent_t ev[4] = {};
        bool pending = false;
        bool timed = false;
        uint64_t ticket = 0;
        std::vector<crux_lik *> liks;                  // the batch, in call order
        struct Done {
            crux_lik *L;
            size_t poff;     // first partial pair of the evaluation
            size_t ntiles;
        };
        std::vector<Done> items;                       // evaluated in this batch
    } stg[2];
    int cur = 0;                   // slot the next batch uses
    // crux_lik_lnl_noalloc: a staging set over the device's static scratch
    // and host memory kept here (no device or pinned allocation per call).
    Staging fixed_stg;
    void *scratch_ptr = nullptr;   // this device's copy of `scratch`
    std::vector<char> fixed_plan;
    std::vector<double> fixed_part;
    bool no_alloc = false;         // set while a no-allocation call plans
    uint64_t next_ticket = 1;
    std::unordered_map<uint64_t, std::vector<double>> results;  // completed, not collected
    int ntrees = 0;                // trees alive on this context (under amu)
    bool dying = false;            // destroy requested while trees were alive
    std::mutex mu;                 // one host thread in the context at a time
    // The allocation counter and tree count: trees of one context may be
    // created, set up and destroyed on different threads.
    std::mutex amu;
};

// Launch shape of a traversal class (K, CM) as a batch runs it, `tiles`
// blocks in the class's launch: the class's tuned values, else the
// context's, resolved to instantiated ones; untuned keys take the default
// for the launch's tiles per CU.  The 16-bit-input 20/22-state kernels are
// instantiated at stack depth 4 and the default waves only.  The planner
// assigns stack slots for the same depth.
static TravShape trav_shape(const crux_ctx *ctx, int K, int CM, long tiles)
{
    const int ci = class_index(K, CM);
    auto pick = [&](int k, int dflt) {
        const int v = ci >= 0 ? ctx->ctune[ci][k] : -1;
        return v >= 0 ? v : dflt;
    };
    TravShape sh;
    sh.prec = K == 4 ? (int)kPrecF32 : ctx->prec;
    const bool wide = K == 4 || sh.prec == kPrecF32;
    const int dt = pick(0, ctx->depth4), ct = pick(1, ctx->chunk4);
    if (K == 4) {
        int dd, dc;
        default_shape4(CM, tiles, ctx->ncu, &dd, &dc);
        sh.depth = stack_depth4(CM, dt, dd);
        sh.chunk = ct > 0 ? ct - 1 : dc;
    } else {
        sh.depth = wide ? stack_depth20(dt) : 4;
        sh.chunk = 0;
    }
    sh.waves = trav20_waves(kpad(K), CM, sh.depth, wide ? pick(2, ctx->waves20) : 0);
    return sh;
}

static int dev_alloc(crux_ctx *ctx, void **p, size_t bytes)
{
    if (ctx->no_alloc)
        return set_error(CRUX_ENOMEM,
                         "a no-allocation evaluation needs %zu more bytes of CLV slots; the "
                         "tree's CLV pool must already hold every slot the plan uses "
                         "(evaluate it once with crux_lik_lnl at this root)", bytes);
    {
        std::lock_guard<std::mutex> g(ctx->amu);
        if (ctx->budget && ctx->used + bytes > ctx->budget)
            return set_error(CRUX_ENOMEM, "HBM budget exceeded: %llu + %llu > %llu bytes",
                             (unsigned long long)ctx->used, (unsigned long long)bytes,
                             (unsigned long long)ctx->budget);
        ctx->used += bytes;    // reserved before the allocation
    }
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess)
        e = hipMalloc(p, bytes ? bytes : 16);
    if (e != hipSuccess) {
        *p = nullptr;
        std::lock_guard<std::mutex> g(ctx->amu);
        ctx->used -= bytes;
        return set_error(e == hipErrorOutOfMemory ? CRUX_ENOMEM : CRUX_EDEVICE,
                         "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    }
    return CRUX_OK;
}

static void dev_free(crux_ctx *ctx, void *p, size_t bytes)
{
    if (p) {
        (void)hipFree(p);
        std::lock_guard<std::mutex> g(ctx->amu);
        ctx->used -= bytes;
    }
}

template <typename T>
static int grow(crux_ctx *ctx, T **p, size_t *cap, size_t need)
{
    if (need <= *cap)
        return CRUX_OK;
    size_t ncap = std::max(need, *cap * 2);
    dev_free(ctx, *p, *cap * sizeof(T));
    *p = nullptr;
    *cap = 0;
    int rc = dev_alloc(ctx, (void **)p, ncap * sizeof(T));
    if (rc)
        return rc;
    *cap = ncap;
    return CRUX_OK;
}

// ---------------------------------------------------------------------------
// Trees and likelihoods.

namespace {

struct RingCache {
    int32_t slot = -1;
    bool valid = false;
    int32_t parent = INT32_MIN;   // ring id of the parent CL, kRootKey for the root CL
    int32_t nsibs = -1;
    double elen = NAN;
};

constexpr int32_t kRootKey = -1;

struct Arena {
    int C = 0;
    uint64_t gen = 0;        // bumped whenever the slots are released wholesale
    size_t slot_bytes = 0;   // CLV floats + int32 exponents, each 256-aligned
    size_t clv_bytes = 0;
    int chunk_slots = 0;
    std::vector<char *> chunks;
    std::vector<int32_t> free_slots;
    int32_t nslots = 0;

    float *clv(int32_t slot) const
    {
        return (float *)(chunks[slot / chunk_slots] + (size_t)(slot % chunk_slots) * slot_bytes);
    }
    int32_t *exps(int32_t slot) const
    {
        return (int32_t *)((char *)clv(slot) + clv_bytes);
    }
};

struct PlanChild {
    int kind;
    int32_t ring;     // internal/tip: child ring
    int32_t taxon;    // tip: alignment row
    int32_t slot;     // internal/ident: CLV slot
    double len;
};

// The children of one fused update (at most kOpMaxChild), held inline.
struct ChildList {
    PlanChild v[kOpMaxChild];
    int n = 0;
    void push_back(const PlanChild &c) { v[n++] = c; }
    size_t size() const { return (size_t)n; }
    PlanChild &operator[](size_t i) { return v[i]; }
    const PlanChild &operator[](size_t i) const { return v[i]; }
    PlanChild *begin() { return v; }
    PlanChild *end() { return v + n; }
    const PlanChild *begin() const { return v; }
    const PlanChild *end() const { return v + n; }
};

struct PlanOp {
    int32_t parent_slot;   // -1: root reduce
    ChildList ch;
    int32_t from_stack = 0;            // bit j: child j comes from the LDS stack
    int32_t fwd = 0;                   // bit j: child j is the previous update's result
    int32_t own = 0;                   // bit j: child j is this plan's result, re-read from HBM
    int8_t pos[kOpMaxChild + 1] = {-1, -1, -1, -1};
    int32_t tip_load = -1;             // tip chunk staged before this op (entry-relative)
    int32_t tip_first = 0;             // index of the op's first tip in the evaluation's list
    int32_t tip_count = 0;
};

}  // namespace

struct crux_tree {
    crux_ctx *ctx = nullptr;
    int K = 0, S = 0, ntaxa = 0, ncodes = 0;
    double *d_codevecs = nullptr;
    uint8_t *d_masks = nullptr;     // padded tip rows: 4-bit state mask (K == 4) or code
    uint32_t *d_weights = nullptr;
    bool has_topology = false;
    int nnodes = 0, nedges = 0;
    std::vector<int32_t> ring_node, ring_next, node_ring, node_taxon, degree;
    std::vector<double> edge_len;
    std::vector<RingCache> cache[2];
    // Per polarity and ring: nothing behind the ring (its node's side of the
    // tree away from its edge) changed since that polarity's planner last
    // walked it and left every CL there valid and keyed.  Closed under
    // subtrees: a clean ring's whole subtree is clean.  The planner skips a
    // clean subtree with a valid CL at its top (checking only that ring's
    // own key); a length edit clears the rings whose subtree holds the edge,
    // a topology edit clears the rings whose subtree holds a re-linked node.
    std::vector<uint8_t> clean[2];
    Arena arena[2];
    crux_lik *owner[2] = {nullptr, nullptr};
    std::vector<crux_lik *> liks;   // every likelihood created on the tree

    size_t mask_stride() const { return align_up((size_t)S, kTile4); }
    // One power-of-two exponent per site, shared by all components.
    size_t exp_count(int C) const { (void)C; return (size_t)S; }
};

struct crux_lik {
    crux_tree *tree = nullptr;
    int pol = 0;
    int C = 0;
    bool have_model = false;
    bool invalidate = true;
    bool root_valid = false;
    int prec = kPrecF32;           // precision the cached CLVs were computed at (K != 4)
    int32_t root_tmp_slot = -1;    // scratch CL for roots of degree > kOpMaxChild
    uint64_t root_tmp_gen = 0;     // arena generation root_tmp_slot belongs to
    double *d_model = nullptr;     // [C][model_stride(K)]
    size_t model_bytes = 0;
    double *d_piw = nullptr;       // [C][K]
    double *d_site = nullptr;      // [S]
    double last_lnl = 0.0;
    bool have_lnl = false;
    // plan of the current execution
    std::vector<PlanOp> ops;
};

static void arena_reset(crux_tree *t, int pol)
{
    Arena &a = t->arena[pol];
    for (char *c : a.chunks)
        dev_free(t->ctx, c, (size_t)a.chunk_slots * a.slot_bytes);
    a.chunks.clear();
    a.free_slots.clear();
    a.nslots = 0;
    a.C = 0;
    a.gen++;
    for (RingCache &rc : t->cache[pol]) {
        rc.slot = -1;
        rc.valid = false;
    }
}

static int arena_get(crux_tree *t, int pol, int32_t *slot)
{
    Arena &a = t->arena[pol];
    if (a.free_slots.empty()) {
        if (a.chunk_slots == 0) {
            // About 64 MiB per allocation, at least one slot.
            a.chunk_slots = (int)std::max<size_t>(1, ((size_t)64 << 20) / a.slot_bytes);
            a.chunk_slots = std::min(a.chunk_slots, std::max(1, t->ntaxa));
        }
        char *p = nullptr;
        int rc = dev_alloc(t->ctx, (void **)&p, (size_t)a.chunk_slots * a.slot_bytes);
        if (rc)
            return rc;
        a.chunks.push_back(p);
        for (int i = a.chunk_slots - 1; i >= 0; i--)
            a.free_slots.push_back(a.nslots + i);
        a.nslots += a.chunk_slots;
    }
    *slot = a.free_slots.back();
    a.free_slots.pop_back();
    return CRUX_OK;
}

static void arena_put(crux_tree *t, int pol, int32_t slot)
{
    if (slot >= 0)
        t->arena[pol].free_slots.push_back(slot);
}

// A ring detached from the tree when its topology is sent: its cached CLVs
// (both polarities) are dead, since a ring that comes back is re-keyed
// anyway, so the slots return to the free lists.  Moves that detach and
// re-attach an edge between two evaluations keep its ring attached at every
// sync and its caches.
static void release_ring(crux_tree *t, int32_t r)
{
    for (int p = 0; p < 2; p++)
        if (r < (int32_t)t->cache[p].size()) {
            arena_put(t, p, t->cache[p][r].slot);
            t->cache[p][r] = RingCache{};
        }
}

// A length edit of edge e: clear the rings whose subtree holds e, i.e. on
// each side of e every ring of every node reached except the one the walk
// came in by.  Subtrees grow away from e and clean rings are closed under
// subtrees, so the walk stops at a ring that is already unclean.
static void mark_edit(crux_tree *t, int32_t e)
{
    thread_local std::vector<int32_t> todo;
    todo.clear();
    todo.push_back(2 * e);
    todo.push_back(2 * e + 1);
    while (!todo.empty()) {
        const int32_t in = todo.back();
        todo.pop_back();
        if (t->ring_node[in] < 0)
            continue;
        for (int32_t r = t->ring_next[in]; r != in; r = t->ring_next[r]) {
            if (!t->clean[0][r] && !t->clean[1][r])
                continue;
            t->clean[0][r] = 0;
            t->clean[1][r] = 0;
            todo.push_back(r ^ 1);
        }
    }
}

// A topology edit that re-linked rings at the given nodes: clear every ring
// of those nodes and every ring whose subtree holds one of them (walking
// outward from each node as mark_edit does from an edge).  Stopping at a ring
// already unclean stays right across the edit: between that ring and one
// farther out with no edited node in between, the tree is as it was, so the
// farther ring was unclean already; with an edited node in between, the walk
// from that node reaches it.
static void mark_nodes(crux_tree *t, const std::vector<int32_t> &nodes)
{
    thread_local std::vector<int32_t> todo;
    todo.clear();
    for (int32_t n : nodes) {
        const int32_t h = t->node_ring[n];
        if (h < 0)
            continue;
        int32_t r = h;
        do {
            t->clean[0][r] = 0;
            t->clean[1][r] = 0;
            todo.push_back(r ^ 1);
            r = t->ring_next[r];
        } while (r != h);
    }
    while (!todo.empty()) {
        const int32_t in = todo.back();
        todo.pop_back();
        if (t->ring_node[in] < 0)
            continue;
        for (int32_t r = t->ring_next[in]; r != in; r = t->ring_next[r]) {
            if (!t->clean[0][r] && !t->clean[1][r])
                continue;
            t->clean[0][r] = 0;
            t->clean[1][r] = 0;
            todo.push_back(r ^ 1);
        }
    }
}

static int arena_layout(crux_tree *t, int pol, int C)
{
    Arena &a = t->arena[pol];
    if (a.C == C)
        return CRUX_OK;
    arena_reset(t, pol);
    a.C = C;
    if (t->K == 4) {
        // One plane [S][4] for one component; tiles of 32 sites holding all
        // components of the class otherwise.
        const int cm = traverse_cmax(4, C);
        a.clv_bytes = cm <= 1 ? align_up((size_t)t->S * 4 * sizeof(float), 256)
                              : align_up((size_t)(t->S + 31) / 32 * 32 * cm * 4 * sizeof(float), 256);
    } else {
        // Whole tiles of 64 (site, component) lanes of the component class.
        const int cm = traverse_cmax(t->K, C), g = sites20(cm > 0 ? cm : 1);
        a.clv_bytes = align_up((size_t)(t->S + g - 1) / g * 64 * kpad(t->K) * sizeof(float), 256);
    }
    a.slot_bytes = a.clv_bytes + align_up(t->exp_count(C) * sizeof(int32_t), 256);
    a.chunk_slots = 0;
    // The kernels address a slot with 32-bit byte offsets.
    if (a.slot_bytes > (size_t)INT32_MAX) {
        a.C = 0;
        return set_error(CRUX_EINVAL, "CLV slot of %zu bytes: patterns x components too large",
                         a.slot_bytes);
    }
    return CRUX_OK;
}

// ---------------------------------------------------------------------------
// Planner (Lik._plan / _planRecurse restated over ring arrays).

namespace {

struct Planner {
    crux_lik *L;
    crux_tree *t;
    int pol;
    bool inval;
    std::vector<PlanOp> &ops;
    int rc = CRUX_OK;
    std::vector<PlanChild> chbuf;   // children of the node being emitted

    bool &valid_of(int32_t key)
    {
        return key == kRootKey ? L->root_valid : t->cache[pol][key].valid;
    }

    void flush(int32_t ring)
    {
        // Drop a CL this plan will not use (Lik.pyx:1492-1496).
        RingCache &c = t->cache[pol][ring];
        if (t->degree[t->ring_node[ring]] > 1 && c.slot >= 0) {
            arena_put(t, pol, c.slot);
            c.slot = -1;
        }
        c.valid = false;
        c.parent = INT32_MIN;
    }

    PlanChild child_of(int32_t ring, double len)
    {
        PlanChild pc{};
        pc.ring = ring;
        pc.len = len;
        int node = t->ring_node[ring];
        if (t->degree[node] == 1) {
            pc.kind = kChildTip;
            pc.taxon = t->node_taxon[node];
            pc.slot = -1;
        } else {
            pc.kind = kChildInternal;
            pc.taxon = -1;
            pc.slot = t->cache[pol][ring].slot;
        }
        return pc;
    }

    // Emit the fused node update(s) for target slot (or the root reduce when
    // target == -1) over the given children.
    void emit(int32_t target, std::vector<PlanChild> &ch)
    {
        size_t i = 0;
        int32_t acc_slot = target;
        if (target < 0 && ch.size() > (size_t)kOpMaxChild) {
            if (L->root_tmp_slot < 0) {
                rc = arena_get(t, pol, &L->root_tmp_slot);
                if (rc)
                    return;
                L->root_tmp_gen = t->arena[pol].gen;
            }
            acc_slot = L->root_tmp_slot;
        }
        bool first = true;
        while (i < ch.size()) {
            PlanOp op;
            size_t room = kOpMaxChild;
            if (!first) {
                PlanChild id{};
                id.kind = kChildIdent;
                id.ring = -1;
                id.taxon = -1;
                id.slot = acc_slot;
                id.len = 0.0;
                op.ch.push_back(id);
                room -= 1;
            }
            while (room > 0 && i < ch.size()) {
                op.ch.push_back(ch[i++]);
                room--;
            }
            bool last = (i == ch.size());
            op.parent_slot = last ? target : acc_slot;
            ops.push_back(std::move(op));
            first = false;
        }
    }

    void recurse(int32_t ring, int32_t parent_key, int nsibs, double elen)
    {
        if (rc)
            return;
        const int node = t->ring_node[ring];
        const int deg = t->degree[node];
        RingCache &me = t->cache[pol][ring];
        if (!inval && t->clean[pol][ring] && (deg == 1 || (me.valid && me.slot >= 0))) {
            // Nothing behind this ring changed since its subtree was last
            // walked: the walk would find every key there equal and every CL
            // valid.  Only this ring's own key can differ.
            if (me.parent != parent_key || me.nsibs != nsibs || me.elen != elen) {
                valid_of(parent_key) = false;
                me.parent = parent_key;
                me.nsibs = nsibs;
                me.elen = elen;
            }
            return;
        }
        if (deg == 1) {
            if (t->node_taxon[node] == -1) {
                rc = set_error(CRUX_ETAXON, "Leaf node missing taxon");
                return;
            }
            if (t->node_taxon[node] < 0 || t->node_taxon[node] >= t->ntaxa) {
                rc = set_error(CRUX_ETAXON, "Taxon missing from alignment's taxa map");
                return;
            }
        } else if (me.slot < 0) {
            rc = arena_get(t, pol, &me.slot);
            if (rc)
                return;
            me.valid = false;
        }
        for (int32_t r = t->ring_next[ring]; r != ring; r = t->ring_next[r]) {
            if (inval)
                flush(r);
            recurse(r ^ 1, ring, deg, t->edge_len[r >> 1]);
            if (rc)
                return;
        }
        if (inval || me.parent != parent_key || me.nsibs != nsibs || me.elen != elen) {
            valid_of(parent_key) = false;
            me.parent = parent_key;
            me.nsibs = nsibs;
            me.elen = elen;
        }
        if (deg != 1 && (inval || !me.valid)) {
            std::vector<PlanChild> &ch = chbuf;
            ch.clear();
            for (int32_t r = t->ring_next[ring]; r != ring; r = t->ring_next[r]) {
                double len = t->edge_len[r >> 1];
                if (len < 0.0) {
                    rc = set_error(CRUX_ENEGLEN, "Negative branch length");
                    return;
                }
                ch.push_back(child_of(r ^ 1, len));
            }
            emit(me.slot, ch);
            me.valid = true;
            valid_of(parent_key) = false;
        }
        if (!rc)
            t->clean[pol][ring] = 1;
    }

    void plan(int root)
    {
        const int32_t ring = t->node_ring[root];
        const int deg = t->degree[root];
        if (ring < 0) {
            rc = set_error(CRUX_EINVAL, "root node %d has no edges", root);
            return;
        }
        if (deg == 1) {
            // Leaf root: its neighbour over the edge, then the leaf itself
            // over a zero-length branch (Lik.pyx:1551-1568).
            recurse(ring ^ 1, kRootKey, 2, t->edge_len[ring >> 1]);
            recurse(ring, kRootKey, 2, 0.0);
            if (rc)
                return;
            if (inval || !L->root_valid) {
                double len = t->edge_len[ring >> 1];
                if (len < 0.0) {
                    rc = set_error(CRUX_ENEGLEN, "Negative branch length");
                    return;
                }
                std::vector<PlanChild> &ch = chbuf;
                ch.clear();
                ch.push_back(child_of(ring ^ 1, len));
                ch.push_back(child_of(ring, 0.0));
                emit(-1, ch);
                L->root_valid = true;
            }
        } else {
            int32_t r = ring;
            do {
                if (inval)
                    flush(r);
                recurse(r ^ 1, kRootKey, deg, t->edge_len[r >> 1]);
                if (rc)
                    return;
                r = t->ring_next[r];
            } while (r != ring);
            if (inval || !L->root_valid) {
                std::vector<PlanChild> &ch = chbuf;
                ch.clear();
                r = ring;
                do {
                    double len = t->edge_len[r >> 1];
                    if (len < 0.0) {
                        rc = set_error(CRUX_ENEGLEN, "Negative branch length");
                        return;
                    }
                    ch.push_back(child_of(r ^ 1, len));
                    r = t->ring_next[r];
                } while (r != ring);
                emit(-1, ch);
                L->root_valid = true;
            }
        }
    }
};

}  // namespace

// Shape of an update for the traversals' specialised bodies: 1 + 4 s0 + s1
// for two children from sources s0, s1 (ChildSrc) under a parent with no
// identity child, else 0.
static int op_shape(const Op &op, const PlanOp &pop)
{
    // Results of this plan that did not fit the stack are re-read from HBM
    // where they are used (the generic walk waits for every store first);
    // the shapes' HBM children are CLVs cached by earlier plans, which the
    // 4-state kernel loads one update ahead.
    if (op.nchild != 2 || pop.parent_slot < 0 || pop.own)
        return 0;
    int src[2];
    for (int j = 0; j < 2; j++) {
        const int kind = (op.kinds >> (2 * j)) & 3;
        if (kind == kChildIdent)
            return 0;
        src[j] = kind == kChildTip ? kSrcTip
                 : ((pop.fwd >> j) & 1) ? kSrcFwd
                 : ((pop.from_stack >> j) & 1) ? kSrcStk
                                              : kSrcHbm;
    }
    if (src[0] == kSrcFwd && src[1] == kSrcFwd)
        return 0;
    return 1 + 4 * src[0] + src[1];
}

static double us_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
}

static int plan_lik_(crux_lik *L, int root);

// Plans one likelihood; its host time adds to the batch being formed.
static int plan_lik(crux_lik *L, int root)
{
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = plan_lik_(L, root);
    L->tree->ctx->plan_us += us_since(t0);
    return rc;
}

static int plan_lik_(crux_lik *L, int root)
{
    crux_tree *t = L->tree;
    L->ops.clear();
    if (!t->has_topology)
        return set_error(CRUX_ESTATE, "tree has no topology");
    if (!L->have_model)
        return set_error(CRUX_ESTATE, "no model set");
    if (root < 0 || root >= t->nnodes || t->node_ring[root] < 0)
        return set_error(CRUX_EINVAL, "bad root node %d", root);
    // Another lik of this polarity owned the CLVs: everything is stale.
    bool inval = L->invalidate || t->owner[L->pol] != L;
    if (t->arena[L->pol].C != L->C) {
        if (int rc = arena_layout(t, L->pol, L->C))
            return rc;
        inval = true;
    }
    // CLVs cached at another input precision are stale.
    if (t->K != 4 && L->prec != t->ctx->prec) {
        inval = true;
        L->prec = t->ctx->prec;
    }
    // A scratch root slot from before the arena was last released (by this
    // lik or another of the polarity) is no longer this lik's.
    if (L->root_tmp_slot >= 0 && L->root_tmp_gen != t->arena[L->pol].gen)
        L->root_tmp_slot = -1;
    t->owner[L->pol] = L;
    Planner p{L, t, L->pol, inval, L->ops};
    if (inval)
        L->root_valid = false;
    p.plan(root);
    if (p.rc) {
        // Leave everything of this polarity invalid; the next call replans.
        L->invalidate = true;
        L->ops.clear();
        return p.rc;
    }
    L->invalidate = false;
    return CRUX_OK;
}

// ---------------------------------------------------------------------------
// Execution of a batch of planned likelihoods.

// Re-sequence a plan for the 4-state kernel's per-lane LDS stack and assign
// stack slots.  The node updates of a plan form a tree (each update's result
// is consumed by exactly one later update, the root update by none).  It is
// re-emitted in post-order with the child subtree that needs the deepest
// stack first (Sethi-Ullman order), which bounds the depth by about log2 of
// the number of updates.  Results that would sit deeper than `depth` slots
// are not pushed; their consumer reads them back from HBM, where every
// non-root update is stored anyway.
static void schedule_stack(std::vector<PlanOp> &ops, int depth)
{
    const int n = (int)ops.size();
    if (n == 0)
        return;
    // Scratch reused across calls (the planner runs under the context lock;
    // thread_local keeps contexts on different threads apart).
    thread_local std::vector<int> writer;                  // slot -> latest op writing it
    thread_local std::vector<std::array<int, kOpMaxChild>> producer;
    thread_local std::vector<int> consumer, need, order, push_slot, kidn;
    thread_local std::vector<std::array<int, kOpMaxChild>> kids;
    thread_local std::vector<std::pair<int, int>> st;
    thread_local std::vector<char> forwarded;
    thread_local std::vector<PlanOp> out;
    // producer[i][j]: op that produced child j of op i in this plan, or -1.
    int32_t maxslot = -1;
    for (int i = 0; i < n; i++) {
        maxslot = std::max(maxslot, ops[i].parent_slot);
        for (const PlanChild &c : ops[i].ch)
            maxslot = std::max(maxslot, c.slot);
    }
    writer.assign((size_t)maxslot + 1, -1);
    producer.resize(n);
    consumer.assign(n, -1);
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < kOpMaxChild; j++)
            producer[i][j] = -1;
        for (size_t j = 0; j < ops[i].ch.size(); j++) {
            const PlanChild &c = ops[i].ch[j];
            if (c.kind == kChildTip || c.slot < 0)
                continue;
            const int w = writer[c.slot];
            if (w >= 0 && consumer[w] < 0) {
                producer[i][j] = w;
                consumer[w] = i;
            }
        }
        if (ops[i].parent_slot >= 0)
            writer[ops[i].parent_slot] = i;
    }
    // Stack need per op (children are earlier in the original order); the
    // producing children sorted deepest first.
    need.assign(n, 1);
    kids.resize(n);
    kidn.assign(n, 0);
    for (int i = 0; i < n; i++) {
        int &k = kidn[i];
        for (int j = 0; j < kOpMaxChild; j++)
            if (producer[i][j] >= 0)
                kids[i][k++] = producer[i][j];
        std::stable_sort(kids[i].begin(), kids[i].begin() + k,
                         [&](int a, int b) { return need[a] > need[b]; });
        int d = 1;
        for (int q = 0; q < k; q++)
            d = std::max(d, q + need[kids[i][q]]);
        need[i] = d;
    }
    // Emit post-order from each root (ops without a consumer), deepest first.
    order.clear();
    for (int r = 0; r < n; r++) {
        if (consumer[r] >= 0)
            continue;
        st.clear();
        st.push_back({r, 0});
        while (!st.empty()) {
            auto &top = st.back();
            if (top.second < kidn[top.first]) {
                const int k = kids[top.first][top.second++];
                st.push_back({k, 0});
            } else {
                order.push_back(top.first);
                st.pop_back();
            }
        }
    }
    // Assign slots by simulating the stack.  A result whose consumer is the
    // very next update stays in the kernel's registers (forwarded) and takes
    // no slot; in post-order that is the last child of every update with an
    // internal child, so the stack only holds the earlier siblings.
    push_slot.assign(n, -1);
    forwarded.assign(n, 0);
    out.clear();
    int sp = 0;
    for (size_t k = 0; k < order.size(); k++) {
        const int i = order[k];
        PlanOp op = ops[i];
        op.from_stack = 0;
        op.fwd = 0;
        op.own = 0;
        for (int j = 0; j < kOpMaxChild; j++)
            op.pos[j] = -1;
        int popped = 0;
        for (size_t j = 0; j < op.ch.size(); j++) {
            int p = producer[i][j];
            if (p >= 0 && forwarded[p]) {
                op.fwd |= 1 << j;
            } else if (p >= 0 && push_slot[p] >= 0) {
                op.from_stack |= 1 << j;
                op.pos[j] = (int8_t)push_slot[p];
                popped++;
            } else if (p >= 0) {
                op.own |= 1 << j;
            }
        }
        sp -= popped;
        op.pos[kOpMaxChild] = -1;
        if (consumer[i] >= 0 && k + 1 < order.size() && order[k + 1] == consumer[i]) {
            forwarded[i] = 1;
        } else if (consumer[i] >= 0 && sp < depth) {
            push_slot[i] = sp;
            op.pos[kOpMaxChild] = (int8_t)sp;
            sp++;
        }
        out.push_back(op);
    }
    ops.swap(out);
}

// Number the tip children of a 4-state plan in execution order and decide
// where the kernel stages their state masks into LDS: the chunk of up to
// kTipChunk rows starting at tip 0 is staged when the launch starts, and a
// new chunk starts at the first tip of the first update whose tips would not
// fit in the current one.  A tip child's pos byte is its row in its chunk.
// When each update is its own launch, the update's tips form the launch's own
// list (tip_first, tip_count), staged at its start.  Returns the taxon row of
// every tip child in order.
static std::vector<int32_t> assign_tips(std::vector<PlanOp> &ops, bool per_launch)
{
    std::vector<int32_t> taxa;
    int base = 0;
    for (PlanOp &op : ops) {
        op.tip_load = -1;
        op.tip_first = (int32_t)taxa.size();
        op.tip_count = 0;
        for (const PlanChild &c : op.ch)
            op.tip_count += c.kind == kChildTip;
        if (op.tip_count == 0)
            continue;
        const int k = (int)taxa.size();
        if (!per_launch && k + op.tip_count - base > kTipChunk) {
            op.tip_load = k;
            base = k;
        }
        for (size_t j = 0; j < op.ch.size(); j++) {
            if (op.ch[j].kind != kChildTip)
                continue;
            op.pos[j] = (int8_t)(per_launch ? (int)taxa.size() - k : (int)taxa.size() - base);
            taxa.push_back(op.ch[j].taxon);
        }
    }
    return taxa;
}

// A 20/22-state plan run over two launches.  A launch with few blocks per
// SIMD leaves most SIMDs running one wave that waits on its own loads, and
// the SIMDs that got a second block set the launch's time; walking each site
// tile's subtrees as blocks of their own gives the first launch several
// shorter blocks per tile, started as SIMDs free up, and the second launch
// the updates above them.
struct PlanSplit {
    std::vector<std::pair<int32_t, int32_t>> units;   // first and last update of each subtree
    std::vector<int32_t> top;                         // the updates above them, in plan order
    std::vector<uint8_t> from_unit;                   // per update: bit j, child j is a subtree's result
    std::vector<int32_t> chunk;                       // per update: tip chunk staged when it runs
};

// Splits a plan into its largest subtrees of at most T updates (each the
// range of the plan that ends at its root: a plan lists a subtree's updates
// together, which is checked) and the updates above them, which read the
// subtrees' results from their slots.  Returns false, leaving sp empty, where
// that does not pay or does not apply: fewer than two subtrees, more than a
// third of the updates above them, or a result read by anything but one
// later update's internal child.
static bool split_plan(const std::vector<PlanOp> &ops, int T, PlanSplit &sp)
{
    sp = PlanSplit{};
    const int n = (int)ops.size();
    std::vector<int32_t> size(n, 1), cons(n, -1), cons_kind(n, -1);
    std::vector<std::array<int32_t, kOpMaxChild>> prod(n);
    int32_t nslot = 0;
    for (const PlanOp &op : ops) {
        nslot = std::max(nslot, op.parent_slot + 1);
        for (const PlanChild &c : op.ch)
            if (c.kind != kChildTip)
                nslot = std::max(nslot, c.slot + 1);
    }
    std::vector<int32_t> last(nslot, -1);   // slot -> the update that last wrote it
    for (int i = 0; i < n; i++) {
        prod[i].fill(-1);
        for (size_t j = 0; j < ops[i].ch.size(); j++) {
            const PlanChild &c = ops[i].ch[j];
            if (c.kind == kChildTip || c.slot < 0 || last[c.slot] < 0)
                continue;   // a tip, or cached by an earlier plan
            const int p = last[c.slot];
            if (cons[p] >= 0)
                return false;
            cons[p] = i;
            cons_kind[p] = c.kind;
            prod[i][j] = p;
            size[i] += size[p];
            last[c.slot] = -1;
        }
        if (ops[i].parent_slot >= 0)
            last[ops[i].parent_slot] = i;
    }
    std::vector<int32_t> unit_of(n, -1);
    std::vector<int32_t> stack;
    for (int i = 0; i < n; i++) {
        if (size[i] > T || cons[i] < 0 || size[cons[i]] <= T || cons_kind[i] != kChildInternal ||
            ops[i].parent_slot < 0)
            continue;
        // Its subtree: exactly the size[i] updates ending at i.
        int lo = i, cnt = 0;
        stack.assign(1, i);
        while (!stack.empty()) {
            const int k = stack.back();
            stack.pop_back();
            cnt++;
            lo = std::min(lo, k);
            for (int j = 0; j < kOpMaxChild; j++)
                if (prod[k][j] >= 0)
                    stack.push_back(prod[k][j]);
        }
        if (cnt != size[i] || lo != i - size[i] + 1)
            return false;
        for (int k = lo; k <= i; k++)
            unit_of[k] = (int32_t)sp.units.size();
        sp.units.push_back({lo, i});
    }
    for (int i = 0; i < n; i++)
        if (unit_of[i] < 0)
            sp.top.push_back(i);
    if (sp.units.size() < 2 || 3 * sp.top.size() > (size_t)n) {
        sp = PlanSplit{};
        return false;
    }
    sp.from_unit.assign(n, 0);
    for (int i : sp.top)
        for (int j = 0; j < kOpMaxChild; j++)
            if (prod[i][j] >= 0 && unit_of[prod[i][j]] >= 0)
                sp.from_unit[i] |= (uint8_t)(1 << j);
    sp.chunk.assign(n, 0);
    int32_t base = 0;
    for (int i = 0; i < n; i++) {
        if (ops[i].tip_load >= 0)
            base = ops[i].tip_load;
        sp.chunk[i] = base;
    }
    return true;
}

using Staging = crux_ctx::Staging;

// A gfx950 buffer resource descriptor (V#) for raw access to nbytes at p:
// base address (48 bits), stride 0, no swizzle, range nbytes, and the dword-3
// flags the kernels' own descriptors use.  Accesses at or past the range read
// zero and write nothing.
static void put_vsharp(int32_t *w, const void *p, uint32_t nbytes)
{
    const uint64_t a = (uint64_t)(uintptr_t)p;
    w[0] = (int32_t)(uint32_t)a;
    w[1] = (int32_t)((uint32_t)(a >> 32) & 0xffffu);
    w[2] = (int32_t)nbytes;
    w[3] = (int32_t)kVsharpFlags;
}

// Neumaier addition of (x, xc) into (s, c); an infinite running sum drops
// the compensation so that -inf stays -inf (the kernels' wave sums do the
// same).
static void nsum_merge_host(double &s, double &c, double x, double xc)
{
    const double t = s + x;
    if (std::isfinite(t)) {
        if (std::fabs(s) >= std::fabs(x))
            c += (s - t) + x;
        else
            c += (x - t) + s;
    } else {
        c = 0.0;
    }
    s = t;
    if (std::isfinite(s))
        c += xc;
}

// Finish the batch of staging set S: wait for its event, hand the results to
// the likelihoods, keep them under the batch's ticket for crux_batch_wait.
static int finish(crux_ctx *ctx, Staging &S)
{
    if (!S.pending)
        return CRUX_OK;
    const auto t0 = std::chrono::steady_clock::now();
    const hipError_t e = hipEventSynchronize(S.done);
    S.pending = false;
    if (e != hipSuccess) {
        // The batch's results are lost; its likelihoods replan everything.
        for (auto &it : S.items)
            it.L->invalidate = true;
        S.liks.clear();
        S.items.clear();
        return set_error(CRUX_EDEVICE, "batch %llu: %s", (unsigned long long)S.ticket,
                         hipGetErrorString(e));
    }
    if (S.timed) {
        // This slot's own events, all complete once `done` is.
        float a = 0, b = 0, c = 0, d = 0;
        if (hipEventElapsedTime(&a, S.ev[0], S.ev[1]) == hipSuccess &&
            hipEventElapsedTime(&b, S.ev[2], S.ev[3]) == hipSuccess &&
            hipEventElapsedTime(&c, S.ev[1], S.ev[2]) == hipSuccess &&
            hipEventElapsedTime(&d, S.ev[0], S.ev[3]) == hipSuccess) {
            ctx->last_ms[0] = a;
            ctx->last_ms[1] = b;
            ctx->last_ms[2] = c;
            ctx->last_ms[3] = d;
        }
    }
    // Each evaluation's lnL: compensated (Neumaier) sum of its waves'
    // partial pairs in tile order, so the result is the same run to run.
    for (auto &it : S.items) {
        double sum = 0.0, cmp = 0.0;
        for (size_t k = 0; k < it.ntiles; k++)
            nsum_merge_host(sum, cmp, S.h_partials[2 * (it.poff + k)],
                            S.h_partials[2 * (it.poff + k) + 1]);
        it.L->last_lnl = sum + cmp;
        it.L->have_lnl = true;
    }
    std::vector<double> r(S.liks.size());
    for (size_t i = 0; i < S.liks.size(); i++)
        r[i] = S.liks[i]->last_lnl;
    ctx->results[S.ticket] = std::move(r);
    S.liks.clear();
    S.items.clear();
    ctx->host_us[3] = us_since(t0);
    return CRUX_OK;
}

// Finish the batch in staging slot k, and first any older batch in the other
// slot (results are handed to the likelihoods in submission order).
static int complete_slot(crux_ctx *ctx, int k)
{
    Staging &S = ctx->stg[k];
    if (!S.pending)
        return CRUX_OK;
    Staging &O = ctx->stg[k ^ 1];
    if (O.pending && O.ticket < S.ticket) {
        int rc = complete_slot(ctx, k ^ 1);
        if (rc)
            return rc;
    }
    return finish(ctx, S);
}

static int drain(crux_ctx *ctx)
{
    while (ctx->stg[0].pending || ctx->stg[1].pending) {
        const int k = ctx->stg[0].pending &&
                              (!ctx->stg[1].pending || ctx->stg[0].ticket < ctx->stg[1].ticket)
                          ? 0
                          : 1;
        int rc = complete_slot(ctx, k);
        if (rc)
            return rc;
    }
    return CRUX_OK;
}

// Plan and enqueue one batch of planned likelihoods into staging set S; the
// results are collected with collect() under the returned ticket.  With
// `fixed`, S's buffers are the device's static scratch (d_plan, d_pmat,
// d_partials) and host memory the context keeps (h_plan, h_partials), laid
// out here per batch; nothing is allocated, and a batch that does not fit
// the scratch is refused.
static int submit_to(crux_ctx *ctx, Staging &S, bool fixed, std::vector<crux_lik *> &liks,
                     uint64_t *ticket)
{
    const auto t_build = std::chrono::steady_clock::now();
    ctx->host_us[0] = ctx->plan_us;
    ctx->plan_us = 0.0;
    // Classes by (K, CM) for the traversal launches.
    struct Item {
        crux_lik *L;
        int idx;    // output index
        int CM;
        std::vector<int32_t> tips;   // 4-state: taxon row of each tip child
        PlanSplit split;             // 20/22 states: the plan over two launches
    };
    std::vector<Item> items;
    for (size_t i = 0; i < liks.size(); i++) {
        crux_lik *L = liks[i];
        if (L->ops.empty())
            continue;
        int cm = traverse_cmax(L->tree->K, L->C);
        if (cm < 0)
            return set_error(CRUX_EINVAL, "unsupported state/component count K=%d C=%d",
                             L->tree->K, L->C);
        items.push_back({L, (int)i, cm});
    }
    ctx->last_bytes = 0;
    ctx->last_formula = 0;
    ctx->last_nodes = 0;
    ctx->last_launches = 0;
    ctx->last_short = 0;
    if (items.empty()) {
        // Nothing to run: the values are the last ones, once any pending
        // batch has delivered them.
        int rc0 = drain(ctx);
        if (rc0)
            return rc0;
        std::vector<double> r(liks.size());
        for (size_t i = 0; i < liks.size(); i++)
            r[i] = liks[i]->last_lnl;
        *ticket = ctx->next_ticket++;
        ctx->results[*ticket] = std::move(r);
        return CRUX_OK;
    }
    std::stable_sort(items.begin(), items.end(), [](const Item &a, const Item &b) {
        if (a.L->tree->K != b.L->tree->K)
            return a.L->tree->K < b.L->tree->K;
        return a.CM < b.CM;
    });

    const bool plain = ctx->mode == 1;
    // Launch shape of each class, from its blocks in this batch.
    long cls_tiles[kNumClasses] = {};
    for (Item &it : items)
        cls_tiles[class_index(it.L->tree->K, it.CM)] +=
            (it.L->tree->S + tile_sites(it.L->tree->K, it.CM) - 1) / tile_sites(it.L->tree->K, it.CM);
    TravShape cls_shape[kNumClasses];
    // A 4-state class whose plans are all short runs the short-plan kernel.
    bool cls_short[kNumClasses];
    for (int ci = 0; ci < kNumClasses; ci++)
        cls_short[ci] = true;
    for (Item &it : items) {
        const int ci = class_index(it.L->tree->K, it.CM);
        if (it.L->tree->K != 4 || (it.CM != 2 && it.CM != 4) || it.L->ops.size() > (size_t)kPartMaxOps)
            cls_short[ci] = false;
    }
    for (Item &it : items) {
        const int ci = class_index(it.L->tree->K, it.CM);
        cls_shape[ci] = trav_shape(ctx, it.L->tree->K, it.CM, cls_tiles[ci]);
        if (cls_short[ci] && ctx->mode != 1 && ctx->short_plans) {
            cls_shape[ci].part = true;
            cls_shape[ci].depth = kPartDepth;
        }
    }
    // A 20/22-state class launched with fewer site tiles than split20 per
    // SIMD splits its long plans into subtrees of about the number of updates
    // that makes that many blocks (at most half a plan); so does one with
    // fewer than twice that many tiles and at least 200 x split20 updates of
    // work per SIMD (enough for the balance the subtree blocks bring to
    // outweigh the second launch).
    int cls_split[kNumClasses] = {};
    if (!plain) {
        double wu[kNumClasses] = {};
        int cls_k[kNumClasses] = {};
        for (Item &it : items) {
            const crux_tree *t = it.L->tree;
            const int ci = class_index(t->K, it.CM);
            cls_k[ci] = t->K;
            if (t->K != 4 || !cls_shape[ci].part)
                wu[ci] += (double)((t->S + tile_sites(t->K, it.CM) - 1) / tile_sites(t->K, it.CM)) *
                          (double)it.L->ops.size();
        }
        // 20/22 states: one-wave blocks against SIMDs; 4 states: blocks of
        // one 64-site tile against half the CUs (a CU holds several of them
        // at once, and near one block per CU splitting no longer pays),
        // or any number of tiles once the launch is long enough for the
        // balance to outweigh the second launch (400 updates per block
        // aimed at: 154k at 256 CUs, where a 128-taxon plan over 947 tiles
        // has 119k and runs 12% slower split).
        for (int ci = 0; ci < kNumClasses; ci++) {
            // The class's own setting, else the context's.
            const int sv = ctx->ctune[ci][3] >= 0 ? ctx->ctune[ci][3] : ctx->split20;
            if (sv <= 0)
                continue;
            const bool k4 = cls_k[ci] == 4;
            const double want = (double)sv * (k4 ? 0.5 * ctx->ncu : 4.0 * ctx->ncu);
            const double tl = (double)cls_tiles[ci];
            if (wu[ci] > 0 && (tl < want || (k4 ? wu[ci] >= 400 * want
                                                 : tl < 2 * want && wu[ci] >= 200 * want)))
                cls_split[ci] = std::max(16, (int)std::ceil(wu[ci] / want));
        }
    }
    // Sizes.
    size_t nops = 0, njobs = 0, pm_floats = 0, tiles = 0, nplain = 0, ntips_total = 0;
    size_t ctl_words = 0, nsplit_ent = 0, nsplit_blk = 0;
    size_t shmem = 0;
    for (Item &it : items) {
        crux_lik *L = it.L;
        crux_tree *t = L->tree;
        const int K = t->K, C = L->C;
        // The traversal kernels hand results to parents through their LDS
        // stack; one-launch-per-update mode cannot (LDS does not outlive a
        // launch).
        if (!plain)
            schedule_stack(L->ops, cls_shape[class_index(K, it.CM)].depth);
        // The 20/22-state and short-plan traversals run the plan's last
        // update apart from the others: only that one may lack a parent (it
        // is the root).
        if (K != 4 || cls_shape[class_index(K, it.CM)].part)
            for (size_t o = 0; o + 1 < L->ops.size(); o++)
                if (L->ops[o].parent_slot < 0)
                    return set_error(CRUX_EINVAL, "plan has a root update before its end");
        it.tips = assign_tips(L->ops, plain);
        ntips_total += it.tips.size();
        // Subtrees of at most half the plan, so that there are at least two.
        const int T = std::min(cls_split[class_index(K, it.CM)], (int)(L->ops.size() / 2));
        it.split = PlanSplit{};
        if (T >= 16 && split_plan(L->ops, T, it.split)) {
            nsplit_ent += it.split.units.size();
            nsplit_blk += it.split.units.size() * ((t->S + tile_sites(K, it.CM) - 1) / tile_sites(K, it.CM));
            ctl_words += it.split.top.size() * (K != 4 ? kCtlWords20 : kCtlWords);
            if (K == 4) {
                // The upper updates' own Op copies and P images (their
                // children's P matrices computed a second time).
                nops += it.split.top.size();
                pm_floats += align_up(it.split.top.size() * p4_floats(it.CM), 64);
                for (int i : it.split.top)
                    for (const PlanChild &c : L->ops[i].ch)
                        njobs += c.kind != kChildIdent;
            }
        }
        nops += L->ops.size();
        ctl_words += L->ops.size() * (K != 4 ? kCtlWords20
                                      : cls_shape[class_index(K, it.CM)].part ? kCtlWordsPart
                                                                              : kCtlWords);
        if (plain)
            nplain += L->ops.size();
        if (K == 4)
            pm_floats += align_up(L->ops.size() * p4_floats(it.CM), 64);
        for (PlanOp &op : L->ops)
            for (PlanChild &c : op.ch) {
                if (c.kind == kChildIdent)
                    continue;
                njobs++;
                if (K == 4) {
                    shmem = std::max(shmem, (size_t)C * K * sizeof(double));
                } else {
                    pm_floats += align_up(c.kind == kChildTip ? (size_t)t->ncodes * it.CM * kpad(K)
                                                              : (size_t)img_words(kpad(K), it.CM, ctx->prec),
                                          64);
                    shmem = std::max(shmem, ((size_t)C * K + (size_t)C * K * K) * sizeof(double));
                }
            }
        tiles += (t->S + tile_sites(K, it.CM) - 1) / tile_sites(K, it.CM);
    }
    const size_t nent = items.size() + nplain + nsplit_ent;
    const size_t off_ent = 0;
    const size_t off_ops = align_up(off_ent + nent * sizeof(Entry), 256);
    const size_t off_jobs = align_up(off_ops + nops * sizeof(Op), 256);
    const size_t off_blk = align_up(off_jobs + njobs * sizeof(PJob), 256);
    const size_t nblk_total = plain ? tiles * 0 : tiles;
    size_t blk_count = 0;
    for (Item &it : items) {
        const size_t ts = tile_sites(it.L->tree->K, it.CM);
        size_t nt = (it.L->tree->S + ts - 1) / ts;
        blk_count += plain ? nt * (it.L->ops.size() + 1) : nt;
    }
    blk_count += nsplit_blk;
    (void)nblk_total;
    const size_t off_tips = align_up(off_blk + blk_count * sizeof(int32_t), 256);
    const size_t off_ctl = align_up(off_tips + ntips_total * sizeof(void *), 256);
    const size_t plan_bytes = align_up(off_ctl + ctl_words * sizeof(int32_t), 256);
    if (shmem > 64 * 1024)
        return set_error(CRUX_EINVAL, "P-matrix job too large for LDS");

    if (fixed) {
        // [plan | P matrices | partials] in the static scratch; the plan and
        // partials have host copies in the context's own host memory.
        const size_t off_pm = align_up(plan_bytes, 256);
        const size_t off_part = align_up(off_pm + pm_floats * sizeof(float), 256);
        const size_t need = off_part + 2 * tiles * sizeof(double);
        if (need > kScratchBytes)
            return set_error(CRUX_EINVAL,
                             "batch needs %zu bytes of scratch; the static scratch holds %zu "
                             "(use crux_lik_lnl)", need, (size_t)kScratchBytes);
        char *base = (char *)ctx->scratch_ptr;
        if (ctx->fixed_plan.size() < plan_bytes)
            ctx->fixed_plan.resize(plan_bytes);
        if (ctx->fixed_part.size() < 2 * tiles)
            ctx->fixed_part.resize(2 * tiles);
        S.h_plan = ctx->fixed_plan.data();
        S.dh_plan = nullptr;
        S.d_plan = base;
        S.plan_cap = plan_bytes;
        S.d_pmat = (float *)(base + off_pm);
        S.pmat_cap = pm_floats;
        S.h_partials = ctx->fixed_part.data();
        S.d_partials = (double *)(base + off_part);
        S.partials_cap = 2 * tiles;
    }
    if (!fixed && plan_bytes > S.plan_cap) {
        size_t ncap = std::max(plan_bytes, S.plan_cap * 2);
        if (S.h_plan)
            (void)hipHostFree(S.h_plan);
        dev_free(ctx, S.d_plan, S.plan_cap);
        S.h_plan = nullptr;
        S.dh_plan = nullptr;
        S.d_plan = nullptr;
        S.plan_cap = 0;
        HIP_TRY(hipHostMalloc((void **)&S.h_plan, ncap, hipHostMallocCoherent | hipHostMallocMapped));
        HIP_TRY(hipHostGetDevicePointer((void **)&S.dh_plan, S.h_plan, 0));
        int rc = dev_alloc(ctx, (void **)&S.d_plan, ncap);
        if (rc)
            return rc;
        S.plan_cap = ncap;
    }
    int rc = fixed ? CRUX_OK : grow(ctx, &S.d_pmat, &S.pmat_cap, std::max<size_t>(pm_floats, 64));
    if (rc)
        return rc;
    if (!fixed && 2 * tiles > S.partials_cap) {
        const size_t ncap = std::max<size_t>(2 * tiles, 2 * S.partials_cap);
        if (S.h_partials)
            (void)hipHostFree(S.h_partials);
        S.h_partials = nullptr;
        S.d_partials = nullptr;
        S.partials_cap = 0;
        HIP_TRY(hipHostMalloc((void **)&S.h_partials, ncap * sizeof(double),
                              hipHostMallocCoherent | hipHostMallocMapped));
        HIP_TRY(hipHostGetDevicePointer((void **)&S.d_partials, S.h_partials, 0));
        S.partials_cap = ncap;
    }
    char *H = S.h_plan, *D = S.d_plan;
    Entry *hent = (Entry *)(H + off_ent);
    Op *hops = (Op *)(H + off_ops);
    PJob *hjobs = (PJob *)(H + off_jobs);
    int32_t *hblk = (int32_t *)(H + off_blk);
    const Entry *dent = (const Entry *)(D + off_ent);
    const Op *dops = (const Op *)(D + off_ops);
    // The P-matrix kernel reads its jobs from the pinned plan itself (from
    // the device copy in the scratch when fixed).
    const PJob *djobs = (const PJob *)((fixed ? D : S.dh_plan) + off_jobs);
    const int32_t *dblk = (const int32_t *)(D + off_blk);

    size_t io = 0, ij = 0, ipm = 0, itile = 0, iblk = 0, ient = 0, itip = 0, ictl = 0;
    uint64_t traffic = 0, formula = 0;
    int nodes = 0;
    // Fused launches: groups of consecutive items of one (K, CM) class.
    struct Launch {
        int K, CM;
        size_t blk0, nblk;
    };
    std::vector<Launch> launches;
    std::vector<size_t> item_poff(items.size()), item_ntl(items.size());
    // Blocks of one class for the launch being formed (a new launch where
    // the class changes).
    auto add_blocks = [&](int K, int CM, int32_t ent, size_t n) {
        for (size_t b = 0; b < n; b++)
            hblk[iblk++] = ent;
        if (launches.empty() || launches.back().K != K || launches.back().CM != CM)
            launches.push_back({K, CM, iblk - n, 0});
        launches.back().nblk += n;
    };
    // Entries of split plans' upper updates, launched after all others.
    struct Deferred {
        Entry e;
        int K, CM;
        size_t ntl;
    };
    std::vector<Deferred> deferred;
    std::vector<size_t> job_first;

    for (size_t ii = 0; ii < items.size(); ii++) {
        crux_lik *L = items[ii].L;
        crux_tree *t = L->tree;
        const int K = t->K, C = L->C;
        const Arena &A = t->arena[L->pol];
        const int CM = items[ii].CM, KP = kpad(K);
        const bool part = K == 4 && cls_shape[class_index(K, CM)].part;
        const int nw = K != 4 ? kCtlWords20 : part ? kCtlWordsPart : kCtlWords;
        const size_t ntl = (t->S + tile_sites(K, CM) - 1) / tile_sites(K, CM);
        const size_t op0 = io, ctl0 = ictl;
        const uint64_t clv_b = (uint64_t)t->S * C * KP * 4, exp_b = (uint64_t)t->exp_count(C) * 4;
        // 4-state: the P images of this evaluation, one per update.
        const size_t pdense = ipm;
        if (K == 4)
            ipm += align_up(L->ops.size() * p4_floats(CM), 64);
        std::vector<size_t> &job0 = job_first;   // first P job of each update
        job0.assign(L->ops.size() + 1, 0);
        auto write_ctl = [&](int32_t *cw, const Op &op, const PlanOp &pop, int32_t tip_load) {
            memset(cw, 0, nw * sizeof(int32_t));
            cw[0] = op.nchild | (op.kinds << 8) | (pop.from_stack << 16) | (pop.fwd << 20) |
                    ((pop.parent_slot >= 0 ? 1 : 0) << 24) | (op_shape(op, pop) << 25);
            cw[1] = op.pos;
            cw[2] = tip_load;
            // has_parent, and in bits 8+ the children that are this
            // plan's own results read back from HBM (not loadable ahead).
            cw[3] = (pop.parent_slot >= 0 ? 1 : 0) | (pop.own << kCtlOwnShift);
            const uint64_t pp = (uint64_t)(uintptr_t)(pop.parent_slot >= 0 ? A.clv(pop.parent_slot) : nullptr);
            const uint64_t pe = (uint64_t)(uintptr_t)(pop.parent_slot >= 0 ? A.exps(pop.parent_slot) : nullptr);
            cw[4] = (int32_t)(uint32_t)pp;
            cw[5] = (int32_t)(uint32_t)(pp >> 32);
            cw[6] = (int32_t)(uint32_t)pe;
            cw[7] = (int32_t)(uint32_t)(pe >> 32);
            if (part) {
                // Every child's slot or tip row, and which of them the
                // short-plan kernel loads ahead: tip rows, and CLVs
                // cached by an earlier plan (not forwarded, not on the
                // stack, not this plan's own results).  A child read
                // where it is used takes its slot from these words too.
                const int skip = pop.from_stack | pop.fwd | pop.own;
                for (int j = 0; j < op.nchild; j++) {
                    const int kind = (op.kinds >> (2 * j)) & 3;
                    uint64_t pc = 0;
                    if (kind == kChildTip) {
                        pc = (uint64_t)(uintptr_t)op.ch[j].codes;
                        cw[3] |= 1 << (kCtlTipShift + j);
                    } else {
                        pc = (uint64_t)(uintptr_t)op.ch[j].clv;
                        if (kind == kChildInternal && !((skip >> j) & 1))
                            cw[3] |= 1 << (kCtlAheadShift + j);
                    }
                    cw[kCtlAhead + 2 * j] = (int32_t)(uint32_t)pc;
                    cw[kCtlAhead + 2 * j + 1] = (int32_t)(uint32_t)(pc >> 32);
                }
            }
            if (K == 4)
                return;
            // 20/22 states: buffer descriptors the kernel loads into
            // scalar registers as they are (see Ctl20 in device.h): the
            // parent's slot, each child's P image or tip table and each
            // child's slot; the zero block where there is none (an empty
            // range for slots, so every load of one reads zero).
            const uint32_t slot_range = (uint32_t)(A.clv_bytes + (size_t)t->S * sizeof(int32_t));
            put_vsharp(cw + kCtl20Parent, pop.parent_slot >= 0 ? (const void *)A.clv(pop.parent_slot)
                                                               : ctx->d_zero,
                       pop.parent_slot >= 0 ? slot_range : 0u);
            for (int j = 0; j < kOpMaxChild; j++) {
                const OpChild &oc = op.ch[j];
                const bool have = j < op.nchild;
                // The data's own size, so that an offset past it reads
                // nothing.
                const int kind = (op.kinds >> (2 * j)) & 3;
                const uint32_t dat_range =
                    !(have && oc.mat) ? (uint32_t)kZeroBytes
                    : kind == kChildTip ? (uint32_t)((size_t)t->ncodes * CM * KP * sizeof(float))
                                        : (uint32_t)(img_words(KP, CM, ctx->prec) * sizeof(float));
                put_vsharp(cw + kCtl20Data + 4 * j, have && oc.mat ? (const void *)oc.mat : ctx->d_zero,
                           dat_range);
                put_vsharp(cw + kCtl20Slot + 4 * j, have && oc.clv ? (const void *)oc.clv : ctx->d_zero,
                           have && oc.clv ? slot_range : 0u);
            }
        };
        for (PlanOp &pop : L->ops) {
            const size_t olocal = io - op0;
            job0[olocal] = ij;
            Op &op = hops[io++];
            memset(&op, 0, sizeof op);
            op.nchild = (int32_t)pop.ch.size();
            op.from_stack = pop.from_stack;
            uint32_t packed = 0;
            for (int j = 0; j <= kOpMaxChild; j++)
                packed |= (uint32_t)(uint8_t)pop.pos[j] << (8 * j);
            op.pos = (int32_t)packed;
            op.tip_load = pop.tip_load;
            int32_t *cw = (int32_t *)(H + off_ctl) + ictl;
            ictl += nw;
            if (pop.parent_slot >= 0) {
                op.parent = A.clv(pop.parent_slot);
                op.parent_exp = A.exps(pop.parent_slot);
                traffic += clv_b + exp_b;
                formula += clv_b + exp_b;
                nodes++;
            } else {
                traffic += 4ull * t->S + 8ull * t->S;   // weights in, site lnL out
                formula += 4ull * t->S + 8ull * t->S;
            }
            for (size_t j = 0; j < pop.ch.size(); j++) {
                PlanChild &pc = pop.ch[j];
                OpChild &oc = op.ch[j];
                op.kinds |= pc.kind << (2 * j);
                if (pc.kind == kChildTip) {
                    // Padded tip row: state masks (4 states) or code indices.
                    oc.codes = t->d_masks + (size_t)pc.taxon * t->mask_stride();
                    traffic += (uint64_t)t->S;
                    formula += (uint64_t)t->S;
                } else {
                    oc.clv = A.clv(pc.slot);
                    oc.exps = A.exps(pc.slot);
                    formula += clv_b + exp_b;
                    if (!(((pop.from_stack | pop.fwd) >> j) & 1))
                        traffic += clv_b + exp_b;
                }
                if (pc.kind != kChildIdent) {
                    PJob &jb = hjobs[ij++];
                    jb.len = pc.len;
                    jb.model = L->d_model;
                    jb.codevecs = t->d_codevecs;
                    // The 4-state kernel forms tip products from P itself.
                    jb.layout = K == 4 ? kPImage4 : pc.kind == kChildTip ? kPTips : kPImage;
                    jb.C = C;
                    jb.K = K;
                    jb.ncodes = t->ncodes;
                    jb.cm = CM;
                    jb.kp = KP;
                    jb.child = (int32_t)j;
                    jb.prec = K == 4 ? kPrecF32 : ctx->prec;
                    if (K == 4) {
                        jb.out = S.d_pmat + pdense + olocal * p4_floats(CM);
                    } else {
                        jb.out = S.d_pmat + ipm;
                        ipm += align_up(jb.layout == kPTips ? (size_t)t->ncodes * CM * KP
                                                            : (size_t)img_words(KP, CM, ctx->prec),
                                        64);
                    }
                    oc.mat = jb.out;
                }
            }
            write_ctl(cw, op, pop, op.tip_load);
        }
        job0[L->ops.size()] = ij;
        Entry base{};
        base.ops = dops + op0;
        base.ctl = (const int32_t *)(D + off_ctl) + ctl0;
        base.nops = (int32_t)L->ops.size();
        base.S = t->S;
        base.C = C;
        base.K = K;
        base.ntiles = (int32_t)ntl;
        base.piw = L->d_piw;
        base.weights = t->d_weights;
        base.site_out = L->d_site;
        base.partials = S.d_partials + 2 * itile;
        item_poff[ii] = itile;
        item_ntl[ii] = ntl;
        base.ntips = (int32_t)items[ii].tips.size();
        base.exp_off = (int32_t)A.clv_bytes;
        base.tip_rows = nullptr;
        base.pdense = K == 4 ? S.d_pmat + pdense : nullptr;
        if (!items[ii].tips.empty()) {
            const uint8_t **hrows = (const uint8_t **)(H + off_tips) + itip;
            for (size_t k = 0; k < items[ii].tips.size(); k++)
                hrows[k] = t->d_masks + (size_t)items[ii].tips[k] * t->mask_stride();
            base.tip_rows = (const uint8_t *const *)(D + off_tips) + itip;
            itip += items[ii].tips.size();
        }
        itile += ntl;
        const PlanSplit &sp = items[ii].split;
        if (!plain && !sp.units.empty()) {
            // The subtrees run on the plan's own records; the updates above
            // them on records of their own, which read the subtrees' results
            // from their slots and stage the tip chunks the plan had staged
            // for them.
            // Largest subtrees first: the launch's last blocks to start are
            // its shortest.
            std::vector<size_t> order(sp.units.size());
            for (size_t k = 0; k < order.size(); k++)
                order[k] = k;
            std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
                return sp.units[a].second - sp.units[a].first > sp.units[b].second - sp.units[b].first;
            });
            for (size_t k : order) {
                const auto &u = sp.units[k];
                Entry e = base;
                e.ops = base.ops + u.first;
                e.ctl = base.ctl + (size_t)u.first * nw;
                if (K == 4)
                    e.pdense = base.pdense + (size_t)u.first * p4_floats(CM);
                e.nops = u.second - u.first + 1;
                e.tip0 = sp.chunk[u.first];
                if (e.ntips > 0 && e.tip0 >= e.ntips)
                    return set_error(CRUX_EINVAL, "split plan: tip chunk %d outside %d tips", e.tip0, e.ntips);
                e.partials = nullptr;
                e.tile0 = (int32_t)iblk;
                hent[ient] = e;
                add_blocks(K, items[ii].CM, (int32_t)ient, ntl);
                ient++;
            }
            Entry top = base;
            top.ctl = (const int32_t *)(D + off_ctl) + ictl;
            top.nops = (int32_t)sp.top.size();
            // Tip chunks: the one of the first update with a tip child is
            // staged when the entry starts, another only where an update
            // with a tip child needs it; updates above the subtrees with no
            // tip child at all stage nothing (and have no tip list).
            int top_tips = 0;
            int32_t staged = -1;
            for (int i : sp.top)
                if (L->ops[i].tip_count > 0) {
                    top_tips += L->ops[i].tip_count;
                    if (staged < 0)
                        staged = sp.chunk[i];
                }
            top.tip0 = staged < 0 ? 0 : staged;
            if (top_tips == 0) {
                top.ntips = 0;
                top.tip_rows = nullptr;
            }
            for (size_t k = 0; k < sp.top.size(); k++) {
                const int i = sp.top[k];
                PlanOp p2 = L->ops[i];
                const int m = sp.from_unit[i];
                // A subtree's result the plan handed over in registers or on
                // the stack is read from its slot (a read the traffic count,
                // the fused plan's floor, leaves out).
                p2.from_stack &= ~m;
                p2.fwd &= ~m;
                p2.own &= ~m;
                int32_t tl = -1;
                if (L->ops[i].tip_count > 0 && sp.chunk[i] != staged) {
                    tl = sp.chunk[i];
                    staged = tl;
                }
                // A staged chunk lies inside the tip list (the kernels read
                // its rows from there).
                if (tl >= 0 && (top.ntips <= tl || top.tip_rows == nullptr))
                    return set_error(CRUX_EINVAL, "split plan: tip chunk %d outside %d tips", tl, top.ntips);
                write_ctl((int32_t *)(H + off_ctl) + ictl, hops[op0 + i], p2, tl);
                ictl += nw;
            }
            if (K == 4) {
                // The 4-state kernels read the updates' Ops and P images in
                // plan order: the upper updates get contiguous copies, their
                // P images from jobs of their own.
                const size_t o_top = io, pm_top = ipm;
                ipm += align_up(sp.top.size() * p4_floats(CM), 64);
                for (size_t k = 0; k < sp.top.size(); k++) {
                    const int i = sp.top[k];
                    hops[io++] = hops[op0 + i];
                    for (size_t q = job0[i]; q < job0[i + 1]; q++) {
                        PJob &jb = hjobs[ij++];
                        jb = hjobs[q];
                        jb.out = S.d_pmat + pm_top + k * p4_floats(CM);
                    }
                }
                top.ops = dops + o_top;
                top.pdense = S.d_pmat + pm_top;
            }
            deferred.push_back({top, K, items[ii].CM, ntl});
        } else if (!plain) {
            base.tile0 = (int32_t)iblk;
            hent[ient] = base;
            add_blocks(K, items[ii].CM, (int32_t)ient, ntl);
            ient++;
        } else {
            // One launch per node update, each on a one-op entry.
            for (size_t o = 0; o < L->ops.size(); o++) {
                Entry e1 = base;
                e1.ops = dops + op0 + o;
                e1.nops = 1;
                if (base.pdense)
                    e1.pdense = base.pdense + o * p4_floats(CM);
                e1.ctl = base.ctl + o * nw;
                if (base.tip_rows) {
                    e1.tip_rows = base.tip_rows + L->ops[o].tip_first;
                    e1.ntips = L->ops[o].tip_count;
                }
                e1.tile0 = (int32_t)iblk;
                hent[ient] = e1;
                for (size_t b = 0; b < ntl; b++)
                    hblk[iblk++] = (int32_t)ient;
                launches.push_back({K, items[ii].CM, iblk - ntl, ntl});
                ient++;
            }
        }
    }

    for (size_t d = 0; d < deferred.size(); d++) {
        Deferred &df = deferred[d];
        df.e.tile0 = (int32_t)iblk;
        hent[ient] = df.e;
        if (d == 0)
            launches.push_back({df.K, df.CM, iblk, 0});
        add_blocks(df.K, df.CM, (int32_t)ient, df.ntl);
        ient++;
    }

    hipStream_t st = ctx->stream;
    ctx->host_us[1] = us_since(t_build);
    const auto t_launch = std::chrono::steady_clock::now();
    HIP_TRY(hipSetDevice(ctx->device));
    // One launch computes the P matrices and copies the plan from the pinned
    // buffer to device memory (no runtime copy: on this ROCm the launch
    // after a host-to-device copy costs the host several microseconds).
    // Timing: ev[0]/ev[1] stamp that kernel's start and end, ev[2]/ev[3]
    // the first traversal launch's start and the last one's end.
    const bool tm = ctx->timing;
    if (fixed) {
        // A stream-ordered copy into the scratch: the previous call's
        // kernels have read their plan by the time it runs.
        HIP_TRY(hipMemcpyAsync(D, H, plan_bytes, hipMemcpyHostToDevice, st));
        launch_pmat(djobs, (int)ij, shmem, nullptr, nullptr, 0, st, tm ? S.ev[0] : nullptr,
                    tm ? S.ev[1] : nullptr);
    } else {
        launch_pmat(djobs, (int)ij, shmem, S.dh_plan, D, plan_bytes, st, tm ? S.ev[0] : nullptr,
                    tm ? S.ev[1] : nullptr);
    }
    HIP_TRY(hipGetLastError());
    size_t first = launches.size(), last = 0;
    for (size_t i = 0; i < launches.size(); i++)
        if (launches[i].nblk > 0) {
            first = std::min(first, i);
            last = i;
        }
    int nlaunch = 0, nshort = 0;
    for (size_t i = 0; i < launches.size(); i++) {
        const Launch &ln = launches[i];
        if (ln.nblk == 0)
            continue;
        nlaunch++;
        nshort += cls_shape[class_index(ln.K, ln.CM)].part ? 1 : 0;
        if (launch_traverse(ln.K, ln.CM, dent, dblk, (int)ln.blk0, (int)ln.nblk,
                            cls_shape[class_index(ln.K, ln.CM)], st,
                            tm && i == first ? S.ev[2] : nullptr,
                            tm && i == last ? S.ev[3] : nullptr) != 0)
            return set_error(CRUX_EINVAL, "no traversal kernel for K=%d", ln.K);
        HIP_TRY(hipGetLastError());
    }
    if (tm && first == launches.size()) {
        HIP_TRY(hipEventRecord(S.ev[2], st));
        HIP_TRY(hipEventRecord(S.ev[3], st));
    }
    if (fixed)
        HIP_TRY(hipMemcpyAsync(S.h_partials, S.d_partials, 2 * tiles * sizeof(double),
                               hipMemcpyDeviceToHost, st));
    HIP_TRY(hipEventRecord(S.done, st));
    ctx->host_us[2] = us_since(t_launch);
    ctx->last_bytes = traffic;
    ctx->last_formula = formula;
    ctx->last_nodes = nodes;
    ctx->last_launches = nlaunch;
    ctx->last_short = nshort;
    S.pending = true;
    S.timed = ctx->timing;
    S.ticket = ctx->next_ticket++;
    S.liks = liks;
    S.items.clear();
    for (size_t ii = 0; ii < items.size(); ii++)
        S.items.push_back({items[ii].L, item_poff[ii], item_ntl[ii]});
    *ticket = S.ticket;
    return CRUX_OK;
}

// Submit into the next of the two rotating staging slots (the host plans
// one batch while the device runs the other).
static int submit(crux_ctx *ctx, std::vector<crux_lik *> &liks, uint64_t *ticket)
{
    const int slot = ctx->cur;
    if (ctx->stg[slot].pending) {
        int rc0 = complete_slot(ctx, slot);
        if (rc0)
            return rc0;
    }
    int rc = submit_to(ctx, ctx->stg[slot], false, liks, ticket);
    if (rc == CRUX_OK && ctx->stg[slot].pending)
        ctx->cur = slot ^ 1;
    return rc;
}

// Results of a submitted batch, waiting for it if needed.
static int collect(crux_ctx *ctx, uint64_t ticket, double *out, size_t n)
{
    auto f = ctx->results.find(ticket);
    if (f == ctx->results.end()) {
        for (int k = 0; k < 2; k++)
            if (ctx->stg[k].pending && ctx->stg[k].ticket == ticket) {
                int rc = complete_slot(ctx, k);
                if (rc)
                    return rc;
            }
        f = ctx->results.find(ticket);
        if (f == ctx->results.end())
            return set_error(CRUX_EINVAL, "unknown or collected batch ticket %llu",
                             (unsigned long long)ticket);
    }
    if (f->second.size() != n) {
        ctx->results.erase(f);
        return set_error(CRUX_EINVAL, "batch ticket %llu holds %zu results, not %zu",
                         (unsigned long long)ticket, f->second.size(), n);
    }
    std::copy(f->second.begin(), f->second.end(), out);
    ctx->results.erase(f);
    return CRUX_OK;
}

static int execute(crux_ctx *ctx, std::vector<crux_lik *> &liks, double *out)
{
    uint64_t t = 0;
    int rc = submit(ctx, liks, &t);
    if (rc)
        return rc;
    return collect(ctx, t, out, liks.size());
}

// ---------------------------------------------------------------------------
// Static scratch of crux_lik_lnl_noalloc: the batch's plan, P matrices and
// per-wave partial sums, kScratchBytes on every device the library runs on,
// held until the process ends.  Host code uses the address
// hipGetSymbolAddress returns on each device, never the array's name.

static __device__ __align__(256) char scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static void *g_scratch_dev[kMaxDevices];
// Calls on one device share its scratch: they are serialized here.
static std::mutex g_scratch_use[kMaxDevices];

// With the context's device current: this device's address of `scratch`.
static int scratch_setup(crux_ctx *ctx)
{
    if (ctx->device >= kMaxDevices)
        return set_error(CRUX_EINVAL, "device %d: no scratch", ctx->device);
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_dev[ctx->device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(CRUX_EDEVICE, "hipGetSymbolAddress(scratch): %s",
                             hipGetErrorString(e));
        g_scratch_dev[ctx->device] = p;
    }
    ctx->scratch_ptr = g_scratch_dev[ctx->device];
    return CRUX_OK;
}

// ---------------------------------------------------------------------------
// C ABI.

extern "C" {

const char *crux_strerror(int status)
{
    switch (status) {
    case CRUX_OK: return "ok";
    case CRUX_EINVAL: return "invalid argument";
    case CRUX_ENOMEM: return "out of device memory or HBM budget";
    case CRUX_ENEGLEN: return "negative branch length";
    case CRUX_EDECOMP: return "eigendecomposition failed";
    case CRUX_EDEVICE: return "device error";
    case CRUX_ETAXON: return "leaf taxon error";
    case CRUX_ESTATE: return "call out of order";
    default: return "unknown status";
    }
}

const char *crux_last_error(void) { return g_errbuf; }

int crux_ctx_create(int device, uint64_t hbm_budget, crux_ctx **out)
{
    if (!out)
        return set_error(CRUX_EINVAL, "out is NULL");
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n)
        return set_error(CRUX_EINVAL, "device %d of %d", device, n);
    crux_ctx *ctx = new crux_ctx;
    ctx->device = device;
    ctx->budget = hbm_budget;
    for (auto &row : ctx->ctune)
        for (int &v : row)
            v = -1;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess)
        e = hipDeviceGetAttribute(&ctx->ncu, hipDeviceAttributeMultiprocessorCount, device);
    if (e == hipSuccess)
        e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    for (int k = 0; k < 3 && e == hipSuccess; k++) {
        Staging &S = k < 2 ? ctx->stg[k] : ctx->fixed_stg;
        e = hipEventCreateWithFlags(&S.done, hipEventDisableTiming);
        for (int i = 0; i < 4 && e == hipSuccess; i++)
            e = hipEventCreate(&S.ev[i]);
    }
    if (e == hipSuccess)
        e = hipMalloc((void **)&ctx->d_zero, kZeroBytes);
    if (e == hipSuccess)
        e = hipMemset(ctx->d_zero, 0, kZeroBytes);
    if (e != hipSuccess) {
        if (ctx->d_zero)
            (void)hipFree(ctx->d_zero);
        delete ctx;
        return set_error(CRUX_EDEVICE, "context setup: %s", hipGetErrorString(e));
    }
    if (int rc = scratch_setup(ctx)) {
        (void)hipFree(ctx->d_zero);
        delete ctx;
        return rc;
    }
    *out = ctx;
    return CRUX_OK;
}

int crux_ctx_destroy(crux_ctx *ctx)
{
    if (!ctx)
        return CRUX_OK;
    {
        std::lock_guard<std::mutex> g(ctx->amu);
        if (ctx->ntrees > 0) {
            // Trees still refer to the context (and their liks to the
            // trees): the last crux_tree_destroy finishes the job.
            ctx->dying = true;
            return CRUX_OK;
        }
    }
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (Staging &S : ctx->stg) {
        if (S.h_plan)
            (void)hipHostFree(S.h_plan);
        if (S.h_partials)
            (void)hipHostFree(S.h_partials);
        dev_free(ctx, S.d_plan, S.plan_cap);
        dev_free(ctx, S.d_pmat, S.pmat_cap * sizeof(float));
        if (S.done)
            (void)hipEventDestroy(S.done);
        for (int i = 0; i < 4; i++)
            if (S.ev[i])
                (void)hipEventDestroy(S.ev[i]);
    }
    if (ctx->fixed_stg.done)
        (void)hipEventDestroy(ctx->fixed_stg.done);
    for (int i = 0; i < 4; i++)
        if (ctx->fixed_stg.ev[i])
            (void)hipEventDestroy(ctx->fixed_stg.ev[i]);
    (void)hipFree(ctx->d_zero);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return CRUX_OK;
}

uint64_t crux_ctx_bytes(const crux_ctx *ctx)
{
    if (!ctx)
        return 0;
    std::lock_guard<std::mutex> g(const_cast<crux_ctx *>(ctx)->amu);
    return ctx->used;
}

int crux_ctx_set_mode(crux_ctx *ctx, int mode)
{
    if (!ctx || mode < 0 || mode > 1)
        return set_error(CRUX_EINVAL, "bad mode");
    ctx->mode = mode;
    return CRUX_OK;
}

// Range check of a shape key's value (stack depth, chunk choice, waves).
static int check_shape_value(int key, int value)
{
    switch (key) {
    case CRUX_TUNE_STACK_DEPTH:
        if (value < 0 || value > 64)
            return set_error(CRUX_EINVAL, "stack depth %d", value);
        return CRUX_OK;
    case CRUX_TUNE_CHUNK:
        if (value < 0 || value > 3)
            return set_error(CRUX_EINVAL, "chunk choice %d (0 default, 1-3)", value);
        return CRUX_OK;
    case CRUX_TUNE_WAVES:
        if (value < 0 || value > 2)
            return set_error(CRUX_EINVAL, "waves per SIMD %d (0 default, 1 or 2)", value);
        return CRUX_OK;
    case CRUX_TUNE_SPLIT:
        if (value < 0 || value > 16)
            return set_error(CRUX_EINVAL, "split %d (0 to 16)", value);
        return CRUX_OK;
    default:
        return set_error(CRUX_EINVAL, "unknown tuning key %d", key);
    }
}

static int *ctx_shape_field(crux_ctx *ctx, int key)
{
    return key == CRUX_TUNE_STACK_DEPTH ? &ctx->depth4
           : key == CRUX_TUNE_CHUNK     ? &ctx->chunk4
                                        : &ctx->waves20;
}

int crux_ctx_set_tuning(crux_ctx *ctx, int key, int value)
{
    if (!ctx)
        return set_error(CRUX_EINVAL, "ctx is NULL");
    if (key == CRUX_TUNE_PRECISION) {
        if (value < kPrecF32 || value > kPrecF16x3)
            return set_error(CRUX_EINVAL, "precision %d", value);
        ctx->prec = value;
        return CRUX_OK;
    }
    if (key == CRUX_TUNE_SHORT_PLANS) {
        if (value != 0 && value != 1)
            return set_error(CRUX_EINVAL, "short plans %d (0 or 1)", value);
        ctx->short_plans = value != 0;
        return CRUX_OK;
    }
    if (key == CRUX_TUNE_SPLIT) {
        if (value < 0 || value > 16)
            return set_error(CRUX_EINVAL, "split %d (0 to 16)", value);
        ctx->split20 = value;
        return CRUX_OK;
    }
    if (int rc = check_shape_value(key, value))
        return rc;
    *ctx_shape_field(ctx, key) = value;
    return CRUX_OK;
}

int crux_ctx_get_tuning(const crux_ctx *ctx, int key, int *value)
{
    if (!ctx || !value)
        return set_error(CRUX_EINVAL, "NULL argument");
    if (key == CRUX_TUNE_PRECISION) {
        *value = ctx->prec;
        return CRUX_OK;
    }
    if (key == CRUX_TUNE_SHORT_PLANS) {
        *value = ctx->short_plans ? 1 : 0;
        return CRUX_OK;
    }
    if (key == CRUX_TUNE_SPLIT) {
        *value = ctx->split20;
        return CRUX_OK;
    }
    if (key != CRUX_TUNE_STACK_DEPTH && key != CRUX_TUNE_CHUNK && key != CRUX_TUNE_WAVES)
        return set_error(CRUX_EINVAL, "unknown tuning key %d", key);
    *value = *ctx_shape_field(const_cast<crux_ctx *>(ctx), key);
    return CRUX_OK;
}

// Class of (K, C) and slot of a shape key in the class table.
static int class_slot(int K, int C, int key, int *ci, int *ki)
{
    const int cm = traverse_cmax(K, C);
    *ci = cm > 0 ? class_index(K, cm) : -1;
    if (*ci < 0)
        return set_error(CRUX_EINVAL, "no traversal class for K=%d C=%d", K, C);
    *ki = key == CRUX_TUNE_STACK_DEPTH ? 0
          : key == CRUX_TUNE_CHUNK     ? 1
          : key == CRUX_TUNE_WAVES     ? 2
          : key == CRUX_TUNE_SPLIT     ? 3
                                       : -1;
    if (*ki < 0)
        return set_error(CRUX_EINVAL, "tuning key %d is not a launch-shape key", key);
    return CRUX_OK;
}

int crux_ctx_set_class_tuning(crux_ctx *ctx, int K, int C, int key, int value)
{
    if (!ctx)
        return set_error(CRUX_EINVAL, "ctx is NULL");
    int ci, ki;
    if (int rc = class_slot(K, C, key, &ci, &ki))
        return rc;
    if (value != -1)
        if (int rc = check_shape_value(key, value))
            return rc;
    ctx->ctune[ci][ki] = value;
    return CRUX_OK;
}

int crux_ctx_get_class_tuning(const crux_ctx *ctx, int K, int C, int key, int *value)
{
    if (!ctx || !value)
        return set_error(CRUX_EINVAL, "NULL argument");
    int ci, ki;
    if (int rc = class_slot(K, C, key, &ci, &ki))
        return rc;
    *value = ctx->ctune[ci][ki];
    return CRUX_OK;
}

int crux_ctx_launch_shape(const crux_ctx *ctx, int K, int C, int S, int *shape4)
{
    if (!ctx || !shape4)
        return set_error(CRUX_EINVAL, "NULL argument");
    const int cm = traverse_cmax(K, C);
    if (cm < 0 || class_index(K, cm) < 0)
        return set_error(CRUX_EINVAL, "no traversal class for K=%d C=%d", K, C);
    if (S < 1)
        return set_error(CRUX_EINVAL, "pattern count %d", S);
    const TravShape sh = trav_shape(ctx, K, cm, (S + tile_sites(K, cm) - 1) / tile_sites(K, cm));
    shape4[0] = sh.depth;
    shape4[1] = K == 4 ? trav4_chunk(cm, sh.depth, sh.chunk) : 0;
    shape4[2] = K == 4 ? 0 : sh.waves;
    shape4[3] = cm;
    return CRUX_OK;
}

int crux_ctx_set_timing(crux_ctx *ctx, int on)
{
    if (!ctx)
        return set_error(CRUX_EINVAL, "ctx is NULL");
    ctx->timing = on != 0;
    return CRUX_OK;
}

int crux_ctx_last_timing(const crux_ctx *ctx, double *ms4)
{
    if (!ctx || !ms4)
        return set_error(CRUX_EINVAL, "NULL argument");
    for (int i = 0; i < 4; i++)
        ms4[i] = ctx->last_ms[i];
    return CRUX_OK;
}

int crux_ctx_last_host_timing(const crux_ctx *ctx, double *us4)
{
    if (!ctx || !us4)
        return set_error(CRUX_EINVAL, "NULL argument");
    for (int i = 0; i < 4; i++)
        us4[i] = ctx->host_us[i];
    return CRUX_OK;
}

int crux_ctx_last_traffic(const crux_ctx *ctx, uint64_t *bytes, uint64_t *formula_bytes,
                          int *nodes)
{
    if (!ctx)
        return set_error(CRUX_EINVAL, "ctx is NULL");
    if (bytes)
        *bytes = ctx->last_bytes;
    if (formula_bytes)
        *formula_bytes = ctx->last_formula;
    if (nodes)
        *nodes = ctx->last_nodes;
    return CRUX_OK;
}

int crux_ctx_last_launches(const crux_ctx *ctx, int *launches, int *short_launches)
{
    if (!ctx)
        return set_error(CRUX_EINVAL, "ctx is NULL");
    if (launches)
        *launches = ctx->last_launches;
    if (short_launches)
        *short_launches = ctx->last_short;
    return CRUX_OK;
}

int crux_tree_create(crux_ctx *ctx, int dim, int nsites, int ntaxa, const uint8_t *codes,
                     int ncodes, const double *code_vecs, const uint32_t *weights,
                     crux_tree **out)
{
    if (!ctx || !out || !codes || !code_vecs || !weights)
        return set_error(CRUX_EINVAL, "NULL argument");
    if (dim < 2 || dim > kMaxStates || nsites < 1 || ntaxa < 1 || ncodes < 1 || ncodes > 256)
        return set_error(CRUX_EINVAL, "bad tree shape dim=%d nsites=%d ntaxa=%d ncodes=%d", dim,
                         nsites, ntaxa, ncodes);
    for (size_t i = 0; i < (size_t)ntaxa * nsites; i++)
        if (codes[i] >= ncodes)
            return set_error(CRUX_EINVAL, "code %d out of range at %zu", codes[i], i);
    crux_tree *t = new crux_tree;
    t->ctx = ctx;
    t->K = dim;
    t->S = nsites;
    t->ntaxa = ntaxa;
    t->ncodes = ncodes;
    int rc = dev_alloc(ctx, (void **)&t->d_codevecs, sizeof(double) * ncodes * dim);
    if (!rc)
        rc = dev_alloc(ctx, (void **)&t->d_weights, sizeof(uint32_t) * nsites);
    if (rc) {
        crux_tree_destroy(t);
        return rc;
    }
    hipError_t e = hipMemcpy(t->d_codevecs, code_vecs, sizeof(double) * ncodes * dim,
                             hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(t->d_weights, weights, sizeof(uint32_t) * nsites, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        // Tip rows padded to whole 64-site tiles (the kernels read four sites
        // per lane).  The 4-state kernel reads the leaf vector itself as a
        // 4-bit mask; the 20/22-state kernel the code, for its tip table.
        std::vector<uint8_t> mask_of(ncodes), masks((size_t)ntaxa * t->mask_stride(), 0);
        for (int c = 0; c < ncodes; c++) {
            uint8_t m = 0;
            for (int j = 0; j < 4 && dim == 4; j++)
                if (code_vecs[(size_t)c * 4 + j] != 0.0)
                    m |= (uint8_t)(1u << j);
            mask_of[c] = dim == 4 ? m : (uint8_t)c;
        }
        for (int x = 0; x < ntaxa; x++)
            for (int s = 0; s < nsites; s++)
                masks[(size_t)x * t->mask_stride() + s] = mask_of[codes[(size_t)x * nsites + s]];
        rc = dev_alloc(ctx, (void **)&t->d_masks, masks.size());
        if (rc) {
            crux_tree_destroy(t);
            return rc;
        }
        e = hipMemcpy(t->d_masks, masks.data(), masks.size(), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        crux_tree_destroy(t);
        return set_error(CRUX_EDEVICE, "tree upload: %s", hipGetErrorString(e));
    }
    {
        std::lock_guard<std::mutex> g(ctx->amu);
        ctx->ntrees++;
    }
    *out = t;
    return CRUX_OK;
}

int crux_tree_destroy(crux_tree *t)
{
    if (!t)
        return CRUX_OK;
    crux_ctx *ctx = t->ctx;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        (void)drain(ctx);
    }
    (void)hipStreamSynchronize(ctx->stream);
    arena_reset(t, 0);
    arena_reset(t, 1);
    dev_free(ctx, t->d_codevecs, sizeof(double) * t->ncodes * t->K);
    dev_free(ctx, t->d_masks, (size_t)t->ntaxa * t->mask_stride());
    dev_free(ctx, t->d_weights, sizeof(uint32_t) * t->S);
    // Likelihoods still alive lose their device buffers and their tree
    // (crux_lik_destroy then only frees the handle).
    for (crux_lik *L : t->liks) {
        dev_free(ctx, L->d_model, L->model_bytes);
        dev_free(ctx, L->d_piw, sizeof(double) * L->C * t->K);
        dev_free(ctx, L->d_site, sizeof(double) * t->S);
        L->d_model = nullptr;
        L->d_piw = nullptr;
        L->d_site = nullptr;
        L->tree = nullptr;
    }
    delete t;
    bool last;
    {
        std::lock_guard<std::mutex> g(ctx->amu);
        last = --ctx->ntrees == 0 && ctx->dying;
    }
    if (last)
        crux_ctx_destroy(ctx);
    return CRUX_OK;
}

int crux_tree_set_topology(crux_tree *t, int nnodes, int nedges, const int32_t *ring_node,
                           const int32_t *ring_next, const int32_t *node_ring,
                           const int32_t *node_taxon, const double *edge_len)
{
    if (!t || nnodes < 1 || nedges < 0 || !ring_node || !ring_next || !node_ring ||
        !node_taxon || (nedges > 0 && !edge_len))
        return set_error(CRUX_EINVAL, "bad topology arguments");
    const int nr = 2 * nedges;
    std::vector<int32_t> degree(nnodes, 0);
    for (int r = 0; r < nr; r++) {
        if (ring_node[r] < -1 || ring_node[r] >= nnodes || ring_next[r] < -1 ||
            ring_next[r] >= nr || (ring_node[r] >= 0) != (ring_next[r] >= 0) ||
            (ring_node[r] >= 0) != (ring_node[r ^ 1] >= 0))
            return set_error(CRUX_EINVAL, "bad ring %d", r);
        if (ring_node[r] >= 0) {
            if (ring_node[ring_next[r]] != ring_node[r])
                return set_error(CRUX_EINVAL, "ring %d: next ring on another node", r);
            degree[ring_node[r]]++;
        }
    }
    for (int n = 0; n < nnodes; n++) {
        if (node_ring[n] < -1 || node_ring[n] >= nr ||
            (node_ring[n] >= 0 && ring_node[node_ring[n]] != n))
            return set_error(CRUX_EINVAL, "bad node ring %d", n);
        // Every ring of the node must be reachable around its cycle.
        if (node_ring[n] >= 0) {
            int cnt = 0;
            int32_t r = node_ring[n];
            do {
                r = ring_next[r];
                if (++cnt > degree[n])
                    return set_error(CRUX_EINVAL, "ring cycle of node %d is broken", n);
            } while (r != node_ring[n]);
            if (cnt != degree[n])
                return set_error(CRUX_EINVAL, "ring cycle of node %d is broken", n);
        }
    }
    t->nnodes = nnodes;
    t->nedges = nedges;
    t->ring_node.assign(ring_node, ring_node + nr);
    t->ring_next.assign(ring_next, ring_next + nr);
    t->node_ring.assign(node_ring, node_ring + nnodes);
    t->node_taxon.assign(node_taxon, node_taxon + nnodes);
    t->edge_len.assign(edge_len, edge_len + nedges);
    t->degree = degree;
    for (int p = 0; p < 2; p++)
        t->clean[p].assign(nr, 0);
    for (int p = 0; p < 2; p++) {
        if ((int)t->cache[p].size() < nr)
            t->cache[p].resize(nr);
        else if ((int)t->cache[p].size() > nr) {
            for (int r = nr; r < (int)t->cache[p].size(); r++)
                arena_put(t, p, t->cache[p][r].slot);
            t->cache[p].resize(nr);
        }
    }
    for (int r = 0; r < nr; r++)
        if (ring_node[r] < 0)
            release_ring(t, r);
    t->has_topology = true;
    return CRUX_OK;
}

int crux_tree_update_rings(crux_tree *t, int nrings, const int32_t *rings,
                           const int32_t *ring_node, const int32_t *ring_next, int nnodes,
                           const int32_t *nodes, const int32_t *node_ring)
{
    if (!t || !t->has_topology || nrings < 0 || nnodes < 0 ||
        (nrings > 0 && (!rings || !ring_node || !ring_next)) ||
        (nnodes > 0 && (!nodes || !node_ring)))
        return set_error(CRUX_EINVAL, "bad ring update arguments");
    const int nr = 2 * t->nedges;
    for (int i = 0; i < nrings; i++)
        if (rings[i] < 0 || rings[i] >= nr || ring_node[i] < -1 || ring_node[i] >= t->nnodes ||
            ring_next[i] < -1 || ring_next[i] >= nr || (ring_node[i] >= 0) != (ring_next[i] >= 0))
            return set_error(CRUX_EINVAL, "bad ring update %d", i);
    for (int i = 0; i < nnodes; i++)
        if (nodes[i] < 0 || nodes[i] >= t->nnodes || node_ring[i] < -1 || node_ring[i] >= nr)
            return set_error(CRUX_EINVAL, "bad node update %d", i);
    // Apply to copies, check every touched node, then commit.
    std::vector<int32_t> rn = t->ring_node, rx = t->ring_next, nring = t->node_ring;
    std::vector<int32_t> touched;
    for (int i = 0; i < nrings; i++) {
        if (rn[rings[i]] >= 0)
            touched.push_back(rn[rings[i]]);
        rn[rings[i]] = ring_node[i];
        rx[rings[i]] = ring_next[i];
        if (ring_node[i] >= 0)
            touched.push_back(ring_node[i]);
    }
    for (int i = 0; i < nnodes; i++) {
        nring[nodes[i]] = node_ring[i];
        touched.push_back(nodes[i]);
    }
    for (int i = 0; i < nrings; i++) {
        const int32_t r = rings[i];
        if ((rn[r] >= 0) != (rn[r ^ 1] >= 0))
            return set_error(CRUX_EINVAL, "ring %d attached at one end only", r);
        if (rn[r] >= 0 && rn[rx[r]] != rn[r])
            return set_error(CRUX_EINVAL, "ring %d: next ring on another node", r);
    }
    std::sort(touched.begin(), touched.end());
    touched.erase(std::unique(touched.begin(), touched.end()), touched.end());
    std::vector<std::pair<int32_t, int>> deg;
    for (int32_t n : touched) {
        const int32_t h = nring[n];
        int cnt = 0;
        if (h >= 0) {
            if (rn[h] != n)
                return set_error(CRUX_EINVAL, "bad node ring %d", n);
            int32_t r = h;
            do {
                if (rn[r] != n || ++cnt > nr)
                    return set_error(CRUX_EINVAL, "ring cycle of node %d is broken", n);
                r = rx[r];
            } while (r != h);
        }
        deg.push_back({n, cnt});
    }
    // Every ring that claims a touched node must lie on its cycle.
    for (int i = 0; i < nrings; i++) {
        const int32_t n = rn[rings[i]];
        if (n < 0)
            continue;
        int32_t r = nring[n];
        bool found = false;
        if (r >= 0)
            do {
                found |= r == rings[i];
                r = rx[r];
            } while (r != nring[n] && !found);
        if (!found)
            return set_error(CRUX_EINVAL, "ring %d is not on node %d's cycle", rings[i], n);
    }
    t->ring_node.swap(rn);
    t->ring_next.swap(rx);
    t->node_ring.swap(nring);
    for (int i = 0; i < nrings; i++)
        for (int p = 0; p < 2; p++)
            t->clean[p][rings[i]] = 0;
    mark_nodes(t, touched);
    for (auto &d : deg)
        t->degree[d.first] = d.second;
    for (int i = 0; i < nrings; i++)
        if (t->ring_node[rings[i]] < 0)
            release_ring(t, rings[i]);
    return CRUX_OK;
}

int crux_tree_set_lengths(crux_tree *t, int n, const int32_t *edges, const double *lens)
{
    if (!t || n < 0 || (n > 0 && (!edges || !lens)))
        return set_error(CRUX_EINVAL, "bad arguments");
    for (int i = 0; i < n; i++) {
        if (edges[i] < 0 || edges[i] >= t->nedges)
            return set_error(CRUX_EINVAL, "edge %d out of range", edges[i]);
        // An equal length leaves every key and CL as it was.
        if (!(t->edge_len[edges[i]] == lens[i]) && t->has_topology)
            mark_edit(t, edges[i]);
        t->edge_len[edges[i]] = lens[i];
    }
    return CRUX_OK;
}

int crux_lik_create(crux_tree *t, int polarity, crux_lik **out)
{
    if (!t || !out || polarity < 0 || polarity > 1)
        return set_error(CRUX_EINVAL, "bad arguments");
    crux_lik *L = new crux_lik;
    L->tree = t;
    L->pol = polarity;
    int rc = dev_alloc(t->ctx, (void **)&L->d_site, sizeof(double) * t->S);
    if (rc) {
        delete L;
        return rc;
    }
    {
        std::lock_guard<std::mutex> g(t->ctx->mu);
        t->liks.push_back(L);
    }
    *out = L;
    return CRUX_OK;
}

int crux_lik_destroy(crux_lik *L)
{
    if (!L)
        return CRUX_OK;
    crux_tree *t = L->tree;
    if (t) {
        crux_ctx *ctx = t->ctx;
        {
            // Submitted batches may still read this lik's buffers.
            std::lock_guard<std::mutex> g(ctx->mu);
            (void)drain(ctx);
        }
        (void)hipStreamSynchronize(ctx->stream);
        if (t->owner[L->pol] == L)
            t->owner[L->pol] = nullptr;
        if (L->root_tmp_slot >= 0 && L->root_tmp_gen == t->arena[L->pol].gen)
            arena_put(t, L->pol, L->root_tmp_slot);
        {
            std::lock_guard<std::mutex> g(ctx->mu);
            t->liks.erase(std::remove(t->liks.begin(), t->liks.end(), L), t->liks.end());
        }
        dev_free(ctx, L->d_model, L->model_bytes);
        dev_free(ctx, L->d_piw, sizeof(double) * L->C * t->K);
        dev_free(ctx, L->d_site, sizeof(double) * t->S);
    }
    delete L;
    return CRUX_OK;
}

int crux_lik_set_model(crux_lik *L, int nmodels, const double *evec, const double *evec_inv,
                       const double *evals, const double *pi_norm, int ncomp,
                       const int32_t *comp_model, const double *comp_weight,
                       const double *comp_rate, int invalidate)
{
    if (!L || !L->tree || nmodels < 1 || ncomp < 1 || !evec || !evec_inv || !evals ||
        !pi_norm || !comp_model || !comp_weight || !comp_rate)
        return set_error(CRUX_EINVAL, "bad model arguments");
    crux_tree *t = L->tree;
    crux_ctx *ctx = t->ctx;
    const int K = t->K, st = model_stride(K);
    if (traverse_cmax(K, ncomp) < 0)
        return set_error(CRUX_EINVAL, "unsupported component count %d for %d states", ncomp, K);
    std::vector<double> blk((size_t)ncomp * st), piw((size_t)ncomp * K);
    for (int c = 0; c < ncomp; c++) {
        int m = comp_model[c];
        if (m < 0 || m >= nmodels)
            return set_error(CRUX_EINVAL, "component %d: bad model %d", c, m);
        if (!(comp_rate[c] >= 0.0) || !std::isfinite(comp_rate[c]))
            return set_error(CRUX_EINVAL, "component %d: bad rate", c);
        double *b = &blk[(size_t)c * st];
        memcpy(b, evec + (size_t)m * K * K, sizeof(double) * K * K);
        memcpy(b + K * K, evec_inv + (size_t)m * K * K, sizeof(double) * K * K);
        memcpy(b + 2 * K * K, evals + (size_t)m * K, sizeof(double) * K);
        b[2 * K * K + K] = comp_rate[c];
        for (int i = 0; i < K; i++)
            piw[(size_t)c * K + i] = pi_norm[(size_t)m * K + i] * comp_weight[c];
    }
    (void)hipStreamSynchronize(ctx->stream);
    size_t mb = blk.size() * sizeof(double);
    if (mb != L->model_bytes || ncomp != L->C) {
        dev_free(ctx, L->d_model, L->model_bytes);
        dev_free(ctx, L->d_piw, sizeof(double) * L->C * K);
        L->d_model = nullptr;
        L->d_piw = nullptr;
        L->model_bytes = 0;
        int rc = dev_alloc(ctx, (void **)&L->d_model, mb);
        if (!rc)
            rc = dev_alloc(ctx, (void **)&L->d_piw, sizeof(double) * ncomp * K);
        if (rc) {
            L->have_model = false;
            return rc;
        }
        L->model_bytes = mb;
    }
    HIP_TRY(hipMemcpy(L->d_model, blk.data(), mb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(L->d_piw, piw.data(), sizeof(double) * ncomp * K, hipMemcpyHostToDevice));
    if (ncomp != L->C)
        invalidate = 1;
    L->C = ncomp;
    L->have_model = true;
    if (invalidate)
        L->invalidate = true;
    return CRUX_OK;
}

static int run_one(crux_lik *L, int root)
{
    crux_ctx *ctx = L->tree->ctx;
    std::lock_guard<std::mutex> g(ctx->mu);
    int rc = plan_lik(L, root);
    if (rc)
        return rc;
    std::vector<crux_lik *> v{L};
    double out;
    rc = execute(ctx, v, &out);
    if (rc) {
        L->invalidate = true;
        return rc;
    }
    return CRUX_OK;
}

int crux_lik_lnl(crux_lik *L, int root, double *lnl)
{
    if (!L || !L->tree || !lnl)
        return set_error(CRUX_EINVAL, "bad arguments");
    int rc = run_one(L, root);
    if (rc)
        return rc;
    *lnl = L->last_lnl;
    return CRUX_OK;
}

int crux_lik_site_lnl(crux_lik *L, int root, double *out, int n)
{
    if (!L || !L->tree || !out || n < 0 || n > L->tree->S)
        return set_error(CRUX_EINVAL, "bad arguments");
    int rc = run_one(L, root);
    if (rc)
        return rc;
    HIP_TRY(hipMemcpy(out, L->d_site, sizeof(double) * n, hipMemcpyDeviceToHost));
    return CRUX_OK;
}

int crux_lik_flush(crux_lik *L)
{
    if (!L || !L->tree)
        return set_error(CRUX_EINVAL, "bad arguments");
    crux_tree *t = L->tree;
    std::lock_guard<std::mutex> g(t->ctx->mu);
    if (t->owner[L->pol] == L) {
        for (size_t r = 0; r < t->cache[L->pol].size(); r++) {
            RingCache &c = t->cache[L->pol][r];
            if (c.slot >= 0) {
                arena_put(t, L->pol, c.slot);
                c.slot = -1;
            }
            c.valid = false;
            c.parent = INT32_MIN;
        }
    }
    L->root_valid = false;
    return CRUX_OK;
}

}  // extern "C"

// Validate and plan one batch (caller holds ctx->mu).
static int batch_plan(crux_ctx *ctx, int n, crux_lik *const *liks, const int32_t *roots,
                      std::vector<crux_lik *> &v)
{
    v.assign(liks, liks + n);
    // Each tree-polarity may appear once: the plan mutates its cache state.
    for (int i = 0; i < n; i++) {
        if (!v[i] || !v[i]->tree || v[i]->tree->ctx != ctx)
            return set_error(CRUX_EINVAL, "lik %d: bad handle", i);
        for (int j = 0; j < i; j++)
            if (v[j]->tree == v[i]->tree && v[j]->pol == v[i]->pol)
                return set_error(CRUX_EINVAL, "liks %d and %d share a tree polarity", j, i);
    }
    for (int i = 0; i < n; i++) {
        int rc = plan_lik(v[i], roots[i]);
        if (rc) {
            for (int j = 0; j < i; j++)
                v[j]->invalidate = true;
            return set_error(rc, "lik %d: %s", i, crux_last_error());
        }
    }
    return CRUX_OK;
}

extern "C" {

int crux_batch_submit(crux_ctx *ctx, int n, crux_lik *const *liks, const int32_t *roots,
                      uint64_t *ticket)
{
    if (!ctx || n < 0 || !ticket || (n > 0 && (!liks || !roots)))
        return set_error(CRUX_EINVAL, "bad arguments");
    std::lock_guard<std::mutex> g(ctx->mu);
    std::vector<crux_lik *> v;
    int rc = batch_plan(ctx, n, liks, roots, v);
    if (rc)
        return rc;
    rc = submit(ctx, v, ticket);
    if (rc)
        for (crux_lik *L : v)
            L->invalidate = true;
    return rc;
}

int crux_batch_poll(crux_ctx *ctx, uint64_t ticket, int *done)
{
    if (!ctx || !done)
        return set_error(CRUX_EINVAL, "bad arguments");
    std::lock_guard<std::mutex> g(ctx->mu);
    return crux::poll_locked(ctx, ticket, done);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batches assembled one likelihood at a time (used by the serving loop).

namespace crux {

void ctx_lock(crux_ctx *ctx) { ctx->mu.lock(); }
void ctx_unlock(crux_ctx *ctx) { ctx->mu.unlock(); }

int plan_locked(crux_lik *L, int root)
{
    if (!L || !L->tree)
        return set_error(CRUX_EINVAL, "bad lik");
    return plan_lik(L, root);
}

void flush_locked(crux_lik *L)
{
    crux_tree *t = L->tree;
    if (t->owner[L->pol] == L) {
        for (RingCache &c : t->cache[L->pol]) {
            if (c.slot >= 0) {
                arena_put(t, L->pol, c.slot);
                c.slot = -1;
            }
            c.valid = false;
            c.parent = INT32_MIN;
        }
    }
    L->root_valid = false;
}

int submit_planned(crux_ctx *ctx, crux_lik *const *liks, int n, uint64_t *ticket)
{
    std::vector<crux_lik *> v(liks, liks + n);
    int rc = submit(ctx, v, ticket);
    if (rc)
        for (crux_lik *L : v)
            L->invalidate = true;
    return rc;
}

int collect_locked(crux_ctx *ctx, uint64_t ticket, double *out, size_t n)
{
    return collect(ctx, ticket, out, n);
}

void tree_release_clvs(crux_tree *t)
{
    crux_ctx *ctx = t->ctx;
    // Batches in flight may still write this tree's slots.
    (void)drain(ctx);
    (void)hipStreamSynchronize(ctx->stream);
    for (int p = 0; p < 2; p++) {
        arena_reset(t, p);
        if (t->owner[p])
            t->owner[p]->root_valid = false;
    }
}

uint64_t tree_clv_bytes(const crux_tree *t)
{
    uint64_t b = 0;
    for (int p = 0; p < 2; p++)
        b += (uint64_t)t->arena[p].chunks.size() * t->arena[p].chunk_slots * t->arena[p].slot_bytes;
    return b;
}

crux_tree *lik_tree(const crux_lik *L) { return L->tree; }
crux_ctx *tree_ctx(const crux_tree *t) { return t->ctx; }
int lik_polarity(const crux_lik *L) { return L->pol; }

int poll_locked(crux_ctx *ctx, uint64_t ticket, int *done)
{
    *done = 0;
    if (ctx->results.count(ticket)) {
        *done = 1;
        return CRUX_OK;
    }
    for (int k = 0; k < 2; k++) {
        Staging &S = ctx->stg[k];
        if (!S.pending || S.ticket != ticket)
            continue;
        const hipError_t e = hipEventQuery(S.done);
        if (e == hipErrorNotReady)
            return CRUX_OK;
        if (e != hipSuccess)
            return set_error(CRUX_EDEVICE, "batch %llu: %s", (unsigned long long)ticket,
                             hipGetErrorString(e));
        // Finished (and so is any older batch): hand the results over now.
        int rc = complete_slot(ctx, k);
        if (rc)
            return rc;
        *done = 1;
        return CRUX_OK;
    }
    return set_error(CRUX_EINVAL, "unknown or collected batch ticket %llu",
                     (unsigned long long)ticket);
}

}  // namespace crux

extern "C" {

int crux_lik_lnl_noalloc(crux_lik *L, int root, double *lnl)
{
    if (!L || !L->tree || !lnl)
        return set_error(CRUX_EINVAL, "bad arguments");
    crux_tree *t = L->tree;
    crux_ctx *ctx = t->ctx;
    std::lock_guard<std::mutex> g(ctx->mu);
    if (!ctx->scratch_ptr)
        return set_error(CRUX_ESTATE, "no scratch on device %d", ctx->device);
    if (t->arena[L->pol].C != L->C)
        return set_error(CRUX_ENOMEM,
                         "the tree's CLV pool is laid out for %d components, the model has %d; "
                         "re-laying it out frees and allocates (evaluate with crux_lik_lnl)",
                         t->arena[L->pol].C, L->C);
    std::lock_guard<std::mutex> gs(g_scratch_use[ctx->device]);
    // Batches in flight hand their results over first (they may hold this
    // likelihood, whose value must come out in call order).
    int rc = drain(ctx);
    if (rc)
        return rc;
    ctx->no_alloc = true;
    rc = plan_lik(L, root);
    ctx->no_alloc = false;
    if (rc)
        return rc;
    std::vector<crux_lik *> v{L};
    uint64_t ticket = 0;
    rc = submit_to(ctx, ctx->fixed_stg, true, v, &ticket);
    if (rc == CRUX_OK)
        rc = finish(ctx, ctx->fixed_stg);
    if (rc == CRUX_OK)
        rc = collect(ctx, ticket, lnl, 1);
    if (rc) {
        L->invalidate = true;
        return rc;
    }
    return CRUX_OK;
}

int crux_batch_wait(crux_ctx *ctx, uint64_t ticket, int n, double *out)
{
    if (!ctx || n < 0 || (n > 0 && !out))
        return set_error(CRUX_EINVAL, "bad arguments");
    std::lock_guard<std::mutex> g(ctx->mu);
    return collect(ctx, ticket, out, (size_t)n);
}

int crux_batch_lnl(crux_ctx *ctx, int n, crux_lik *const *liks, const int32_t *roots,
                   double *out)
{
    if (!ctx || n < 0 || (n > 0 && (!liks || !roots || !out)))
        return set_error(CRUX_EINVAL, "bad arguments");
    std::lock_guard<std::mutex> g(ctx->mu);
    std::vector<crux_lik *> v;
    int rc = batch_plan(ctx, n, liks, roots, v);
    if (rc)
        return rc;
    rc = execute(ctx, v, out);
    if (rc)
        for (crux_lik *L : v)
            L->invalidate = true;
    return rc;
}

}  // extern "C"
