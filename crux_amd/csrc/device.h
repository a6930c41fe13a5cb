// Descriptors shared by the host planner and the gfx950 kernels.
//
// HBM layout (all FP32 CLVs are site-major like the reference's cLMat,
// lib/Crux/Tree/CxLik.h:10-36, but FP32 with an int32 power-of-two exponent
// per site instead of an FP64 lnScale):
//   CLV slot   : float  clv[S][C][K]   then   int32 exp[S]   (K != 4)
//                float  clv[C][S][4]   then   int32 exp[S]   (K == 4: component
//                planes, so a wave's float4 access covers 1 KiB contiguously)
//   tip codes  : uint8  codes[ntaxa][S]      index into the tree's code table
//   tip masks  : uint8  masks[ntaxa][S64]    K == 4: 4-bit state set, rows
//                padded to S64 = S rounded up to 64
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
    int32_t from_stack;    // bit j: child j is read from the lane's LDS stack
    // Four signed bytes: stack slot of child j in byte j, and in byte 3 the
    // slot the result is pushed to (-1: not pushed).  Packed in one dword so
    // the kernel reads it with a scalar load (a byte field would need a vector
    // load, whose wait also waits for the previous update's stores).
    int32_t pos;
    // 4-state kernel: before this update, stage the state masks of tip
    // children [tip_load, tip_load + kTipChunk) of the entry's tip list into
    // LDS (-1: none; the chunk at 0 is staged when the launch starts); a tip
    // child's byte in `pos` is its row in the staged chunk.
    int32_t tip_load;
    int32_t pad_;
    OpChild ch[kOpMaxChild];
};

constexpr int kTipChunk = 64;   // tip mask rows staged in LDS at a time (power of 2)

__host__ __device__ inline int op_pos(int32_t packed, int j)
{
    return (int)(int8_t)((uint32_t)packed >> (8 * j));
}

// Sites per block of the 4-state traversal (one wave) and of the generic one.
constexpr int kTile4 = 64;
__host__ __device__ inline int tile_sites(int K) { return K == 4 ? kTile4 : kTile; }

// Largest per-lane LDS stack depth of the 4-state traversal for a component
// class CM (keeps the stack at or below 32 KiB per block).
__host__ __device__ constexpr int stack_depth4_max(int CM) { return CM <= 4 ? 8 : (CM == 8 ? 4 : 2); }

// Stack depth actually used: the tuned value (0 = largest), rounded down to
// an instantiated depth (2, 4 or 8).
__host__ __device__ inline int stack_depth4(int CM, int tuned)
{
    // Untuned: 8 slots while a block's LDS stays small enough for every block
    // of a 60k-site evaluation to be resident at once, else 4.
    int d = tuned > 0 ? tuned : (CM <= 2 ? 8 : 4);
    if (d > stack_depth4_max(CM))
        d = stack_depth4_max(CM);
    return d >= 8 ? 8 : (d >= 4 ? 4 : 2);
}

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
    const uint8_t *const *tip_rows;  // 4-state: mask row of each tip child, plan order
    int32_t ntips;
    int32_t pad_;
    // 4-state: P blocks of the plan, dense [op][child][CM][16] with CM the
    // component class (rows of components >= C zero; child slots without a P
    // block are left unused), staged into LDS a chunk of updates at a time.
    const float *pdense;
    // 4-state: control words of each update, [op][8]: nchild | kinds << 8 |
    // from_stack << 16 | forwarded << 20 (bit j: child j is the previous
    // update's result, kept in registers), pos, tip_load, has_parent, parent
    // (2 words), parent_exp (2 words); staged with the P blocks.  The tip
    // list (tip_rows, ntips) is per entry, so a one-update entry starts its
    // own.
    const int32_t *ctl;
};

constexpr int kCtlWords = 8;

struct PJob {
    double len;            // branch length (times comp_rate[c] gives t)
    const double *model;   // per comp: V[K][K], Vinv[K][K], evals[K], rate
    const double *codevecs;// tip: [ncodes][K]
    float *out;
    int32_t tip;
    int32_t C;
    int32_t K;
    int32_t ncodes;
    int32_t cm;            // P blocks: rows written up to cm * K * K (zero past C)
    int32_t pad_;
};

__host__ __device__ inline int model_stride(int K) { return 2 * K * K + K + 1; }

}  // namespace crux
