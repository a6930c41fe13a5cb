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
