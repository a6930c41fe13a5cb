// Descriptors shared by the host planner and the gfx950 kernels.
//
// HBM layout (all FP32 CLVs are site-major like the reference's cLMat,
// lib/Crux/Tree/CxLik.h:10-36, but FP32 with an int32 power-of-two exponent
// per site instead of an FP64 lnScale):
//   CLV slot   : float  clv[S][C][K]   then   int32 exp[S]
//   tip codes  : uint8  codes[ntaxa][S]      index into the tree's code table
//   P block    : float  P[C][K][K]           row = parent state
//   tip table  : float  T[ncodes][C][K]      T = P * leaf vector of each code
//   root weight: double piw[C][K]            pi_norm * weightScaled
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace crux {

constexpr int kTile = 256;            // sites per block in the traversal
constexpr int kOpMaxChild = 3;        // children per fused node update

enum ChildKind : int32_t {
    kChildInternal = 0,   // P * child CLV
    kChildTip = 1,        // tip table lookup by code
    kChildIdent = 2,      // child CLV used as is (chained multifurcations)
};

struct OpChild {
    const float *clv;      // internal/ident: child CLV slot
    const int32_t *exps;   // internal/ident: child exponents
    const uint8_t *codes;  // tip: code row of the taxon
    const float *mat;      // internal: P block; tip: tip table
};

struct Op {
    float *parent;         // nullptr: root op, reduce to site lnLs
    int32_t *parent_exp;
    int32_t nchild;
    int32_t kinds;         // 2 bits per child (ChildKind)
    OpChild ch[kOpMaxChild];
};

struct Entry {
    const Op *ops;
    int32_t nops;
    int32_t S;
    int32_t C;
    int32_t K;
    int32_t tile0;         // first block of this entry in the launch
    int32_t ntiles;
    const double *piw;     // [C][K]
    const uint32_t *weights;
    double *site_out;      // [S] pattern-weighted site lnLs
    double *partials;      // [ntiles][2] compensated block sums
    double *lnl_out;       // final lnL
};

struct PJob {
    double len;            // branch length (times comp_rate[c] gives t)
    const double *model;   // per comp: V[K][K], Vinv[K][K], evals[K], rate
    const double *codevecs;// tip: [ncodes][K]
    float *out;
    int32_t tip;
    int32_t C;
    int32_t K;
    int32_t ncodes;
};

__host__ __device__ inline int model_stride(int K) { return 2 * K * K + K + 1; }

}  // namespace crux
