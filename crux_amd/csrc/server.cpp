// Serving loop of libcrux_amd: continuous batching of likelihood requests on
// one device context.
//
// Callers (many MC3 chains, many clients) submit requests -- "apply these
// branch lengths to the tree, optionally drop its cached CLVs, and return
// lnL rooted at this node" -- from any thread.  A server thread forms a
// batch from every request that has arrived while the device runs the
// previous batch (up to two batches in flight), submits it as one fused
// device pass and hands each request its result and completion time.  The
// reference has no such layer: its callers evaluate one proposal per
// synchronous Lik.lnL call (pkg/Crux/Mc3/Chain.pyx:455, CxLik.c:914-968).
//
// Requests of one tree run in submission order.  A batch takes up to two
// requests of a tree, one per CLV polarity (a likelihood and its clone,
// Lik.pyx:835-883): edits, flushes and plans interleave, so each request is
// planned against the tree exactly as the requests before it leave it, and
// the two write different CLV slots.  Admission follows the context's HBM
// budget: when planning a request needs memory the budget does not allow,
// the CLV pools of the least recently used tree that is neither in flight
// nor in the batch being formed are released (that tree recomputes them
// when it is next used) and planning is retried; a request that still does
// not fit waits for the batches in flight, and fails with CRUX_ENOMEM only
// when nothing else holds memory.

#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "crux_lik.h"
#include "internal.h"

using namespace crux;

namespace {

struct Req {
    uint64_t id = 0;
    crux_lik *lik = nullptr;
    int root = 0;
    std::vector<int32_t> edges;
    std::vector<double> lens;
    bool flush = false;
    bool done = false;
    int status = CRUX_OK;
    std::string error;
    double lnl = 0.0;
    int64_t t_done = 0;        // steady clock, ns
};

struct Sess {
    crux_tree *tree = nullptr;
    crux_lik *lik[2] = {nullptr, nullptr};   // by polarity
    int queued = 0;            // requests not yet in a batch
    int in_flight = 0;         // requests in submitted batches
    uint64_t last_used = 0;
};

struct Batch {
    uint64_t ticket = 0;
    std::vector<Req *> reqs;
    std::vector<Sess *> sess;
};

int64_t now_ns()
{
    return std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

}  // namespace

struct crux_server {
    crux_ctx *ctx = nullptr;
    int max_batch = 256;
    std::mutex mu;
    std::condition_variable cv;        // work for the server thread
    std::condition_variable cv_done;   // a request finished
    std::unordered_map<const crux_tree *, Sess *> sess_of;
    std::deque<Req *> order;           // queued requests, arrival order
    std::unordered_map<uint64_t, Req *> reqs;
    std::deque<Batch> inflight;
    uint64_t next_id = 1;
    uint64_t tick = 0;
    bool stop = false;
    uint64_t stats[4] = {0, 0, 0, 0};  // batches, batched requests, evictions, failed
    std::thread th;

    void loop();
    void form_batch();
    void finish_oldest(bool block);
    void fail(Req *r, int status);
};

void crux_server::fail(Req *r, int status)
{
    r->status = status;
    r->error = crux_last_error();
    r->done = true;
    r->t_done = now_ns();
    stats[3]++;
}

// Called with mu held.  Takes queued requests in arrival order, at most one
// per (tree, polarity), none ahead of an earlier request of its tree.
void crux_server::form_batch()
{
    std::vector<crux_lik *> liks;
    Batch b;
    std::unordered_map<Sess *, int> used;
    std::unordered_set<Sess *> blocked;
    ctx_lock(ctx);
    for (auto it = order.begin(); it != order.end() && (int)liks.size() < max_batch;) {
        Req *r = *it;
        Sess *s = sess_of[lik_tree(r->lik)];
        if (blocked.count(s)) {
            ++it;
            continue;
        }
        int &mask = used[s];
        int p = lik_polarity(r->lik);
        crux_lik *L = nullptr;
        if (!(mask & (1 << p)))
            L = r->lik;
        else if (s->lik[p ^ 1] && !(mask & (1 << (p ^ 1)))) {
            p ^= 1;
            L = s->lik[p];
        }
        if (!L) {
            blocked.insert(s);
            ++it;
            continue;
        }
        // The request's edits, then its plan (absolute lengths and the
        // flush are idempotent, so a request deferred below is redone as is).
        int rc = CRUX_OK;
        if (!r->edges.empty())
            rc = crux_tree_set_lengths(s->tree, (int)r->edges.size(), r->edges.data(), r->lens.data());
        if (rc == CRUX_OK) {
            if (r->flush)
                flush_locked(L);
            rc = plan_locked(L, r->root);
        }
        while (rc == CRUX_ENOMEM) {
            // Release the least recently used idle tree that holds CLVs.
            Sess *victim = nullptr;
            for (auto &kv : sess_of) {
                Sess *v = kv.second;
                if (v == s || v->in_flight || (used.count(v) && used[v]) ||
                    tree_clv_bytes(v->tree) == 0)
                    continue;
                if (!victim || v->last_used < victim->last_used)
                    victim = v;
            }
            if (!victim)
                break;
            tree_release_clvs(victim->tree);
            stats[2]++;
            rc = plan_locked(L, r->root);
        }
        if (rc == CRUX_ENOMEM && (!liks.empty() || !inflight.empty())) {
            // Wait for memory held by this batch or the ones in flight.
            blocked.insert(s);
            ++it;
            continue;
        }
        if (rc != CRUX_OK) {
            fail(r, rc);
            s->queued--;
            it = order.erase(it);
            cv_done.notify_all();
            continue;
        }
        mask |= 1 << p;
        s->last_used = ++tick;
        liks.push_back(L);
        b.reqs.push_back(r);
        b.sess.push_back(s);
        it = order.erase(it);
    }
    if (!liks.empty()) {
        int rc = submit_planned(ctx, liks.data(), (int)liks.size(), &b.ticket);
        if (rc == CRUX_ENOMEM) {
            // The batch's staging buffers did not fit: release an idle tree
            // and queue the batch again, in order (its plans are redone).
            Sess *victim = nullptr;
            for (auto &kv : sess_of) {
                Sess *v = kv.second;
                if (v->in_flight || (used.count(v) && used[v]) || tree_clv_bytes(v->tree) == 0)
                    continue;
                if (!victim || v->last_used < victim->last_used)
                    victim = v;
            }
            if (victim) {
                tree_release_clvs(victim->tree);
                stats[2]++;
                for (auto r = b.reqs.rbegin(); r != b.reqs.rend(); ++r)
                    order.push_front(*r);
                ctx_unlock(ctx);
                return;
            }
        }
        if (rc != CRUX_OK) {
            for (size_t i = 0; i < b.reqs.size(); i++) {
                fail(b.reqs[i], rc);
                b.sess[i]->queued--;
            }
            cv_done.notify_all();
        } else {
            for (Sess *s : b.sess) {
                s->queued--;
                s->in_flight++;
            }
            stats[0]++;
            stats[1] += b.reqs.size();
            inflight.push_back(std::move(b));
        }
    }
    ctx_unlock(ctx);
}

// Called with mu held: collect the oldest batch if it is done (or wait for
// it when block).  The context lock is taken without mu, so that callers can
// submit while the device finishes.
void crux_server::finish_oldest(bool block)
{
    Batch &b = inflight.front();
    const uint64_t ticket = b.ticket;
    const size_t n = b.reqs.size();
    std::vector<double> out(n);
    int rc = CRUX_OK, done = 0;
    std::unique_lock<std::mutex> lk(mu, std::adopt_lock);
    lk.unlock();
    ctx_lock(ctx);
    if (!block)
        rc = poll_locked(ctx, ticket, &done);
    if (rc == CRUX_OK && (block || done))
        rc = collect_locked(ctx, ticket, out.data(), n);
    ctx_unlock(ctx);
    lk.lock();
    lk.release();
    if (rc == CRUX_OK && !block && !done)
        return;
    const int64_t t = now_ns();
    for (size_t i = 0; i < n; i++) {
        Req *r = b.reqs[i];
        if (rc == CRUX_OK) {
            r->lnl = out[i];
            r->done = true;
            r->t_done = t;
        } else {
            fail(r, rc);
        }
        b.sess[i]->in_flight--;
    }
    inflight.pop_front();
    cv_done.notify_all();
}

void crux_server::loop()
{
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        cv.wait(lk, [&] { return stop || !order.empty() || !inflight.empty(); });
        if (stop && order.empty() && inflight.empty())
            return;
        if (!order.empty() && inflight.size() < 2) {
            const size_t before = order.size();
            form_batch();
            if (order.size() < before || order.empty())
                continue;
            // Nothing could be taken (every queued request waits for memory
            // or for its tree): finish a batch first.  With none in flight
            // that cannot happen (a request that finds no memory then fails),
            // but the loop must not spin on it.
            if (inflight.empty()) {
                Req *r = order.front();
                set_error(CRUX_ESTATE, "request %llu cannot be scheduled", (unsigned long long)r->id);
                fail(r, CRUX_ESTATE);
                sess_of[lik_tree(r->lik)]->queued--;
                order.pop_front();
                cv_done.notify_all();
                continue;
            }
        }
        if (inflight.empty())
            continue;
        if (inflight.size() >= 2 || !order.empty()) {
            lk.release();
            finish_oldest(true);
            lk = std::unique_lock<std::mutex>(mu, std::adopt_lock);
            continue;
        }
        // One batch in flight and nothing queued: poll it, waking early
        // when a request arrives.
        lk.release();
        finish_oldest(false);
        lk = std::unique_lock<std::mutex>(mu, std::adopt_lock);
        if (!inflight.empty() && order.empty() && !stop)
            cv.wait_for(lk, std::chrono::microseconds(20),
                        [&] { return stop || !order.empty(); });
    }
}

extern "C" {

int crux_server_create(crux_ctx *ctx, int max_batch, crux_server **out)
{
    if (!ctx || !out || max_batch < 1)
        return set_error(CRUX_EINVAL, "bad server arguments");
    crux_server *srv = new crux_server;
    srv->ctx = ctx;
    srv->max_batch = max_batch;
    srv->th = std::thread([srv] { srv->loop(); });
    *out = srv;
    return CRUX_OK;
}

int crux_server_destroy(crux_server *srv)
{
    if (!srv)
        return CRUX_OK;
    {
        std::lock_guard<std::mutex> g(srv->mu);
        srv->stop = true;
    }
    srv->cv.notify_all();
    srv->th.join();
    for (auto &kv : srv->reqs)
        delete kv.second;
    for (auto &kv : srv->sess_of)
        delete kv.second;
    delete srv;
    return CRUX_OK;
}

int crux_server_add_lik(crux_server *srv, crux_lik *lik)
{
    if (!srv || !lik || !lik_tree(lik))
        return set_error(CRUX_EINVAL, "bad arguments");
    std::lock_guard<std::mutex> g(srv->mu);
    crux_tree *t = lik_tree(lik);
    if (tree_ctx(t) != srv->ctx)
        return set_error(CRUX_EINVAL, "likelihood of another context");
    Sess *&s = srv->sess_of[t];
    if (!s) {
        s = new Sess;
        s->tree = t;
    }
    const int p = lik_polarity(lik);
    if (s->lik[p] && s->lik[p] != lik)
        return set_error(CRUX_EINVAL, "polarity %d of this tree is served by another likelihood", p);
    s->lik[p] = lik;
    return CRUX_OK;
}

int crux_server_remove_lik(crux_server *srv, crux_lik *lik)
{
    if (!srv || !lik)
        return set_error(CRUX_EINVAL, "bad arguments");
    std::unique_lock<std::mutex> lk(srv->mu);
    auto f = srv->sess_of.find(lik_tree(lik));
    if (f == srv->sess_of.end())
        return CRUX_OK;
    Sess *s = f->second;
    srv->cv_done.wait(lk, [&] { return s->queued == 0 && s->in_flight == 0; });
    const int p = lik_polarity(lik);
    if (s->lik[p] == lik)
        s->lik[p] = nullptr;
    if (!s->lik[0] && !s->lik[1]) {
        srv->sess_of.erase(f);
        delete s;
    }
    return CRUX_OK;
}

int crux_server_submit(crux_server *srv, crux_lik *lik, int root_node, int nedges,
                       const int32_t *edges, const double *lens, int flush, uint64_t *req)
{
    if (!srv || !lik || !req || nedges < 0 || (nedges > 0 && (!edges || !lens)))
        return set_error(CRUX_EINVAL, "bad arguments");
    Req *r = new Req;
    r->lik = lik;
    r->root = root_node;
    r->edges.assign(edges, edges + nedges);
    r->lens.assign(lens, lens + nedges);
    r->flush = flush != 0;
    {
        std::lock_guard<std::mutex> g(srv->mu);
        auto f = srv->sess_of.find(lik_tree(lik));
        if (f == srv->sess_of.end() || f->second->lik[lik_polarity(lik)] != lik) {
            delete r;
            return set_error(CRUX_EINVAL, "likelihood not added to the server");
        }
        if (srv->stop) {
            delete r;
            return set_error(CRUX_ESTATE, "server is stopping");
        }
        r->id = srv->next_id++;
        f->second->queued++;
        srv->order.push_back(r);
        srv->reqs[r->id] = r;
        *req = r->id;
    }
    srv->cv.notify_one();
    return CRUX_OK;
}

int crux_server_poll(crux_server *srv, uint64_t req, int *done)
{
    if (!srv || !done)
        return set_error(CRUX_EINVAL, "bad arguments");
    std::lock_guard<std::mutex> g(srv->mu);
    auto f = srv->reqs.find(req);
    if (f == srv->reqs.end())
        return set_error(CRUX_EINVAL, "unknown or collected request %llu", (unsigned long long)req);
    *done = f->second->done ? 1 : 0;
    return CRUX_OK;
}

int crux_server_wait(crux_server *srv, uint64_t req, double *lnl, int64_t *t_done_ns)
{
    if (!srv || !lnl)
        return set_error(CRUX_EINVAL, "bad arguments");
    std::unique_lock<std::mutex> lk(srv->mu);
    auto f = srv->reqs.find(req);
    if (f == srv->reqs.end())
        return set_error(CRUX_EINVAL, "unknown or collected request %llu", (unsigned long long)req);
    Req *r = f->second;
    srv->cv_done.wait(lk, [&] { return r->done; });
    srv->reqs.erase(f);
    const int status = r->status;
    *lnl = r->lnl;
    if (t_done_ns)
        *t_done_ns = r->t_done;
    std::string err = r->error;
    delete r;
    if (status != CRUX_OK)
        return set_error(status, "%s", err.c_str());
    return CRUX_OK;
}

int crux_server_stats(crux_server *srv, uint64_t *stats4)
{
    if (!srv || !stats4)
        return set_error(CRUX_EINVAL, "bad arguments");
    std::lock_guard<std::mutex> g(srv->mu);
    for (int i = 0; i < 4; i++)
        stats4[i] = srv->stats[i];
    return CRUX_OK;
}

}  // extern "C"
