// Internal declarations shared by the host sources of libcrux_amd.
#pragma once

#include <cstdarg>
#include <cstddef>
#include <cstdint>

namespace crux {

constexpr int kMaxStates = 64;   // largest state count the host math accepts
constexpr int kMaxDevices = 64;  // devices the static scratch is set up for
// Static scratch of crux_lik_lnl_noalloc per device: plan, P matrices and
// per-wave partial sums of one evaluation.
constexpr size_t kScratchBytes = (size_t)16 << 20;
constexpr int kMaxChild = 3;     // children fused into one node update

// Record a formatted detail message for crux_last_error() and return status.
int set_error(int status, const char *fmt, ...);

int q_decomp(int n, const double *rtri, const double *pi, double *pi_norm,
             double *q_norm, double *evec, double *evec_inv, double *evals);
int pt(int n, const double *evec, const double *evec_inv, const double *evals,
       double t, double *P);
int gamma_rates(double alpha, int ncat, int median, double *cmult);

}  // namespace crux

struct crux_ctx;
struct crux_tree;
struct crux_lik;

namespace crux {

// Batches assembled one likelihood at a time (the serving loop).  With the
// context locked (ctx_lock .. ctx_unlock), edits of a tree, flushes and
// plans interleave, so each likelihood is planned against its tree as its
// own request leaves it; submit_planned then enqueues the planned batch.
void ctx_lock(crux_ctx *ctx);
void ctx_unlock(crux_ctx *ctx);
int plan_locked(crux_lik *L, int root);
void flush_locked(crux_lik *L);
int submit_planned(crux_ctx *ctx, crux_lik *const *liks, int n, uint64_t *ticket);
int poll_locked(crux_ctx *ctx, uint64_t ticket, int *done);
int collect_locked(crux_ctx *ctx, uint64_t ticket, double *out, size_t n);
// Release every CLV slot of both polarities of a tree and free their
// memory (the next evaluation recomputes everything); the tree's bytes of
// CLV pools.
void tree_release_clvs(crux_tree *t);
uint64_t tree_clv_bytes(const crux_tree *t);
crux_tree *lik_tree(const crux_lik *L);
crux_ctx *tree_ctx(const crux_tree *t);
int lik_polarity(const crux_lik *L);

}  // namespace crux
