// gfx950 kernels of the tree-likelihood hot path.
//
//  k_pmat      P(t) = max(0, V exp(L t) V^-1) in FP64 for every (edge, comp) of
//              a plan, stored FP32; for a tip child the product of P with the
//              leaf vector of every distinct code instead (tip table).  This
//              replaces the per-step, per-stripe CxLikPt calls of the
//              reference (lib/Crux/Tree/CxLik.c:334-347).
//  k_trav4h    The CLV updates of a whole 4-state plan: two waves per 64
//  k_trav4     sites with the components split over lane halves (k_trav4h,
//  k_trav20    two or more components), one wave per 64 sites (k_trav4, one
//              component) or per 64 (site, component) lanes (k_trav20, 20
//              and 22 states) walk every node update of the plan in
//              post-order, so one launch covers the whole traversal and
//              every evaluation of a batch.  Node update =
//              product over children of (P * child) (or a tip lookup), then a
//              power-of-two rescale of the site so its largest entry is in
//              [0.5, 1), the exponent added to the children's.  This is the
//              arithmetic of CxLikExecuteStripeDna / CxLikExecuteStripe
//              (CxLik.c:294-846) with FP32 storage and exact power-of-two
//              scaling in place of FP64 division by the max.  The root update
//              is not stored: it is reduced to the pattern-weighted site lnL
//              in FP64, NaN -> -inf (CxLik.c:598-632), and to a compensated
//              per-wave sum.
//              Each wave writes its compensated sum to host memory; the
//              host finishes the evaluation's sum once the batch is done.

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <type_traits>

#include "device.h"

namespace crux {

// Uniform descriptor and P-matrix data is read through the constant address
// space so it lands in scalar registers via s_load (a generic pointer would be
// fetched per lane with flat loads, each costing a full memory round trip);
// CLV, mask and exponent traffic goes through the global address space.
#define CRUX_CAS __attribute__((address_space(4)))
#define CRUX_GAS __attribute__((address_space(1)))

template <typename T>
__device__ __forceinline__ const CRUX_CAS T *cas(const T *p)
{
    return (const CRUX_CAS T *)p;
}

template <typename T>
__device__ __forceinline__ CRUX_GAS T *gas(T *p)
{
    return (CRUX_GAS T *)p;
}

template <typename T>
__device__ __forceinline__ const CRUX_GAS T *gas(const T *p)
{
    return (const CRUX_GAS T *)p;
}

// HIP's float4 class cannot be copied through an address-space-qualified
// reference, so global vector accesses go through clang vector types.
typedef float crux_f32x4 __attribute__((ext_vector_type(4)));
typedef float crux_f32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float4 ld4(const CRUX_GAS void *p)
{
    const crux_f32x4 v = *(const CRUX_GAS crux_f32x4 *)p;
    return make_float4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ float2 ld2(const CRUX_GAS void *p)
{
    const crux_f32x2 v = *(const CRUX_GAS crux_f32x2 *)p;
    return make_float2(v.x, v.y);
}

__device__ __forceinline__ void st4(CRUX_GAS void *p, float4 v)
{
    *(CRUX_GAS crux_f32x4 *)p = crux_f32x4{v.x, v.y, v.z, v.w};
}

__device__ __forceinline__ void st2(CRUX_GAS void *p, float2 v)
{
    *(CRUX_GAS crux_f32x2 *)p = crux_f32x2{v.x, v.y};
}

// Neumaier addition of x into (s, c); an infinite running sum drops the
// compensation so that -inf stays -inf.
__device__ __forceinline__ void nsum_add(double &s, double &c, double x)
{
    const double t = s + x;
    if (__builtin_isfinite(t)) {
        if (fabs(s) >= fabs(x))
            c += (s - t) + x;
        else
            c += (x - t) + s;
    } else {
        c = 0.0;
    }
    s = t;
}

__device__ __forceinline__ void nsum_merge(double &s, double &c, double s2, double c2)
{
    nsum_add(s, c, s2);
    if (__builtin_isfinite(s))
        c += c2;
}

#if !defined(CRUX_PART_K) || CRUX_PART_K == 0
// ---------------------------------------------------------------------------
// P matrices and tip tables.

// Blocks [0, njobs): one P job each.  Blocks past them copy the plan (n16
// 16-byte words) from pinned host memory to device memory for the
// traversal launch that follows.
__global__ void __launch_bounds__(256) k_pmat(const PJob *__restrict__ jobs, int njobs,
                                              const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                              int n16)
{
    extern __shared__ double sm[];
    if ((int)blockIdx.x >= njobs) {
        const int nb = (int)gridDim.x - njobs;
        for (int i = ((int)blockIdx.x - njobs) * 256 + (int)threadIdx.x; i < n16; i += nb * 256)
            dst[i] = src[i];
        return;
    }
    const PJob J = jobs[blockIdx.x];
    const int C = J.C, K = J.K, KK = K * K, st = model_stride(K);
    double *ex = sm;            // [C][K] exp(lambda_k * t_c)
    double *P = sm + C * K;     // [C][K][K], images and tip tables
    for (int idx = threadIdx.x; idx < C * K; idx += blockDim.x) {
        const int c = idx / K, k = idx - c * K;
        const double *m = J.model + (size_t)c * st;
        ex[idx] = exp(m[2 * KK + k] * (J.len * m[2 * KK + K]));
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < C * KK; idx += blockDim.x) {
        const int c = idx / KK, r = idx - c * KK, i = r / K, j = r - i * K;
        const double *V = J.model + (size_t)c * st, *Vi = V + KK, *e = ex + c * K;
        double p = 0.0;
        for (int k = 0; k < K; ++k)
            p += V[i * K + k] * e[k] * Vi[k * K + j];
        p = p < 0.0 ? 0.0 : p;
        if (J.layout == kPImage4)
            J.out[p4_index(J.child, J.cm, c, i, j)] = (float)p;
        else
            P[idx] = p;
    }
    if (J.layout == kPImage4) {
        // The columns of the idle components c in [C, cm) are zero.
        for (int idx = C * KK + (int)threadIdx.x; idx < J.cm * KK; idx += blockDim.x) {
            const int c = idx >> 4, i = (idx >> 2) & 3, k = idx & 3;
            J.out[p4_index(J.child, J.cm, c, i, k)] = 0.0f;
        }
        return;
    }
    __syncthreads();
    const int cm = J.cm, kp = J.kp;
    if (J.layout == kPImage && J.prec != kPrecF32) {
        // 16-bit image: dword v of lane l (dword v % 4 of the lane's 16 bytes
        // v / 4) holds P_c[4 I + i][(g cols + a) 4 + 2 p + e] in half e, for
        // v = (I groups + g) 2 + p and l = c (64 / cm) + 4 a + i; the split
        // precision's low image follows its high image.  P is rounded to
        // FP32, then to 16 bits (round to nearest even).
        const int G = 64 / cm, W = 16 / cm, NG = (kp + 4 * W - 1) / (4 * W), NV1 = (kp / 4) * NG * 2;
        const int NV1P = (NV1 + 3) & ~3;
        const int n = (J.prec == kPrecF16x3 ? 2 : 1) * NV1P * 64;
        uint32_t *o = reinterpret_cast<uint32_t *>(J.out);
        for (int idx = threadIdx.x; idx < n; idx += blockDim.x) {
            const int v = (idx >> 8) * 4 + (idx & 3), l = (idx >> 2) & 63;
            uint32_t word = 0;
            const int part = v / NV1P, vv = v - part * NV1P;
            if (vv < NV1) {
                const int pq = vv >> 1, I = pq / NG, g = pq - I * NG;
                const int c = l / G, a = (l % G) >> 2, row = 4 * I + (l & 3);
                const int col = (g * W + a) * 4 + 2 * (vv & 1);
                for (int e = 0; e < 2; ++e) {
                    float f = 0.0f;
                    if (c < C && row < K && col + e < K)
                        f = (float)P[(c * K + row) * K + col + e];
                    uint32_t h = 0;
                    if (J.prec == kPrecBF16) {
                        const uint32_t u = __float_as_uint(f);
                        h = (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
                    } else {
                        const _Float16 hi = (_Float16)f;
                        const _Float16 x = part == 0 ? hi : (_Float16)(f - (float)hi);
                        h = __builtin_bit_cast(uint16_t, x);
                    }
                    word |= h << (16 * e);
                }
            }
            o[idx] = word;
        }
        return;
    }
    if (J.layout == kPImage) {
        // Register v of lane l (float v % 4 of float4 v / 4 of the lane):
        // P_c[4 I + i][g * cols + a] for v = I * groups + g and
        // l = c * (64 / cm) + 4 a + i.
        const int G = 64 / cm, W = 16 / cm, NG = (kp + W - 1) / W, NV = (kp / 4) * NG;
        const int n = img_floats(kp, cm), RT = img_tail(kp, cm), nb = (NV / 4) * 256;
        for (int idx = threadIdx.x; idx < n; idx += blockDim.x) {
            // A float4 block, or the tail's [lane][RT] dwords.
            const int v = RT && idx >= nb ? 4 * (NV / 4) + (idx - nb) % RT : (idx >> 8) * 4 + (idx & 3);
            const int l = RT && idx >= nb ? (idx - nb) / RT : (idx >> 2) & 63;
            float val = 0.0f;
            if (v < NV) {
                const int I = v / NG, g = v - I * NG, c = l / G, a = (l % G) >> 2;
                const int row = 4 * I + (l & 3), col = g * W + a;
                if (c < C && row < K && col < K)
                    val = (float)P[(c * K + row) * K + col];
            }
            J.out[idx] = val;
        }
        return;
    }
    // Tip table: T[code][c][i] = sum_j P_c[i][j] leaf[code][j].
    const int n = J.ncodes * cm * kp;
    for (int idx = threadIdx.x; idx < n; idx += blockDim.x) {
        const int code = idx / (cm * kp), r = idx - code * cm * kp, c = r / kp, i = r - c * kp;
        double t = 0.0;
        if (c < C && i < K) {
            const double *row = P + (size_t)(c * K + i) * K, *v = J.codevecs + (size_t)code * K;
            for (int j = 0; j < K; ++j)
                t += row[j] * v[j];
        }
        J.out[idx] = (float)t;
    }
}

#endif

// ---------------------------------------------------------------------------
// Traversal helpers.

template <int N, int I = 0, typename F>
__device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<N, I + 1>(f);
    }
}

// Raw buffer access (SGPR descriptor + per-lane offset + SGPR offset): the
// per-load offset arithmetic stays on the scalar unit and no 64-bit address
// is formed per lane.  Every access here is in bounds by construction.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc_of(const void *p)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, -1, 0x00020000);
}

// A descriptor whose range ends after nbytes: accesses at or past it read
// zero and write nothing (buffer range checking), so lanes past the end of
// an array need no branch.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc_of(const void *p, uint32_t nbytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)nbytes, 0x00020000);
}

__device__ __forceinline__ crux_f32x4 bload4(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    return __builtin_bit_cast(crux_f32x4, __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, 0));
}

__device__ __forceinline__ int32_t bload1(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    return (int32_t)__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0);
}

// s_waitcnt immediate that waits for LDS and scalar loads only (vmcnt and
// expcnt at their maxima, lgkmcnt 0).
constexpr int kWaitLgkm0 = 0xC07F;
// ... and one that waits for vector memory only (vmcnt 0, expcnt and lgkmcnt
// at their maxima).
constexpr int kWaitVm0 = 0x0F70;

// Cache policy of a store: 0, or kNT (the non-temporal bit of gfx950's
// buffer instructions) for CLV stores, which a later evaluation reads back,
// if at all, long after the caches have turned over.
constexpr int kNT = 2;

template <int POL = 0>
__device__ __forceinline__ void bstore4(crux_f32x4 v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), r, (int)voff, (int)soff, POL);
}

template <int POL = 0>
__device__ __forceinline__ void bstore1(int32_t v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    __builtin_amdgcn_raw_buffer_store_b32((unsigned)v, r, (int)voff, (int)soff, POL);
}

// The values of a lane and of its partner lane ^ OFF, as an unordered pair
// (the pair is the same, in some order, in both lanes).  Lanes 32 or 16
// apart swap halves of the wave on the vector ALU (v_permlane32_swap /
// v_permlane16_swap); closer lanes go through ds_bpermute.
template <int OFF>
__device__ __forceinline__ void lane_pair(uint32_t v, uint32_t &a, uint32_t &b)
{
    if constexpr (OFF == 32) {
        const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        a = r[0];
        b = r[1];
    } else if constexpr (OFF == 16) {
        const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        a = r[0];
        b = r[1];
    } else {
        a = v;
        b = (uint32_t)__shfl_xor((int)v, OFF, 64);
    }
}

// Reductions over the lanes G, 2G, ... apart (the components of a site).
template <int G>
__device__ __forceinline__ int32_t comp_max(int32_t v)
{
    static_for<6>([&](auto K) {
        constexpr int off = 1 << decltype(K)::value;
        if constexpr (off >= G) {
            uint32_t a, b;
            lane_pair<off>((uint32_t)v, a, b);
            v = max((int32_t)a, (int32_t)b);
        }
    });
    return v;
}

template <int G>
__device__ __forceinline__ double comp_sum(double v)
{
    static_for<6>([&](auto K) {
        constexpr int off = 1 << decltype(K)::value;
        if constexpr (off >= G) {
            const uint64_t u = (uint64_t)__double_as_longlong(v);
            uint32_t al, bl, ah, bh;
            lane_pair<off>((uint32_t)u, al, bl);
            lane_pair<off>((uint32_t)(u >> 32), ah, bh);
            // Commutative, so both lanes of a pair get the same sum.
            v = __longlong_as_double((long long)(((uint64_t)ah << 32) | al)) +
                __longlong_as_double((long long)(((uint64_t)bh << 32) | bl));
        }
    });
    return v;
}

// ---------------------------------------------------------------------------
// 4-state helpers.

__device__ __forceinline__ float4 mul4(float4 a, float4 b)
{
    return make_float4(a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w);
}

// y = P x for the wave's 64 sites on the matrix core, FP32 in and out:
// v_mfma_f32_4x4x1f32 computes 16 independent 4x4 outer products, block b
// taking the A value of lanes 4b..4b+3 as a column and the B value of the
// same lanes as a row; with CBSZ 4 every block takes its A column from block
// ABID instead.  With A = column k of P (one block of an image register,
// broadcast) and B = each lane's own x[k], block b accumulates y[i] of site
// 4b+j in result register i of lane 4b+j, so after k = 0..3 every lane holds
// P x of its own site.  The sum is the same FMA chain P[i][0] x0 + ... +
// P[i][3] x3 as on the vector ALU, at twice the FP32 rate and off the vector
// pipeline, and one image register carries 16 columns of P.

// ---------------------------------------------------------------------------
// 4-state traversal, one component (the lane-halves form below takes two or
// more).  A block is one wave: 64 sites, one per lane, with all C mixture
// components of a site in that lane's registers.  The wave walks
// the whole plan (post-order list of fused node updates) for its sites: one
// launch covers a full traversal and every evaluation of a batch, and no
// wave waits for any other -- there is no barrier in the kernel.  Per update
// and component the wave forms P_c * child_c for every child on the matrix
// core (a tip contributes P_c * its 4-bit state mask, one byte per site),
// multiplies them, takes the site maximum over all components and states in
// registers, and rescales the site exactly by the power of two that brings it
// into [0.5, 1), adding the exponent to the children's.  The scaled component
// planes are stored to the HBM cache slot (layout [C][S][4], C = 1 where
// this kernel runs: 1 KiB contiguous per store) and, when the parent is in
// the same plan, pushed onto
// the wave's LDS stack; parents pop children from there, so a full traversal
// reads no CLV back from HBM.  The root update reduces to the
// pattern-weighted FP64 site lnL (NaN -> -inf, CxLik.c:598-632) and a
// compensated sum per wave.  Arithmetic follows CxLikExecuteStripeDna
// (CxLik.c:294-633) with FP32 storage.
//
// Decoding an update, reading a tip mask and finding the site's exponent
// are done once for all components, so the per-update control cost is shared
// C ways.  Latency: every wave walks the plan in step with every other, so a
// load issued when an update needs it meets the full L2 latency with nothing
// to overlap it; and on this GPU one counter tracks loads and stores alike,
// so a wave that waits for a load also waits for its own CLV stores to
// drain.  Hence the P images and control words are prefetched into registers
// a chunk of updates ahead and written to the wave's LDS double-buffered
// (the chunk after next is fetched when a chunk starts and written just
// before it is needed), the next update's control words and P image are read
// from LDS while the current one finishes, and an update reads all its
// children's stack entries and tip masks at once before picking by kind.
// Tip masks of up to kTipChunk tips are staged when the wave starts (four
// sites of a row per lane and load); a plan with more tips stages the next
// chunk where it starts.  CLV stores and cached-child loads go through raw
// buffer descriptors (scalar base and range): the address arithmetic stays
// on the scalar unit, and lanes past the last site fall outside a store's
// range instead of branching around it.
//
// PC: updates per staged chunk (control words of a chunk are one dword per
// lane).
template <int CM, int D, int PC>
__global__ void __launch_bounds__(kTile4)
k_trav4(const Entry *__restrict__ entries, const int32_t *__restrict__ block_entry, int blk0)
{
    // Two or more components run k_trav4h (the loop below is written for any
    // CM, but only the one-component class is launched).
    static_assert(CM == 1, "k_trav4 serves the one-component class");
    static_assert(PC * kCtlWords <= kTile4, "a chunk's control words fit one load per lane");
    constexpr int NR = p4_regs(CM);                 // P image registers per update
    constexpr int kChunkRows = PC * NR * 16;        // float4s of P per chunk
    __shared__ float4 stk[D][CM][kTile4];
    __shared__ int32_t stke[D][kTile4];
    __shared__ uint32_t tipm[kTipChunk][kTile4 / 4];
    __shared__ float4 pch[2][kChunkRows];
    __shared__ int32_t cch[2][PC * kCtlWords];
    const int b = blk0 + (int)blockIdx.x;
    const CRUX_CAS Entry *E = cas(entries) + cas(block_entry)[b];
    const int S = E->S, C = E->C, nops = E->nops, ntips = E->ntips;
    const int tile = b - E->tile0;
    const int lane = (int)threadIdx.x;
    const int s = tile * kTile4 + lane;
    const bool act = s < S;
    const int sc = act ? s : S - 1;       // in-bounds site for loads of idle lanes
    const CRUX_CAS Op *ops = cas(E->ops);
    double sum = 0.0;

    // ---- P / control-word staging ---------------------------------------
    constexpr int kPer = (kChunkRows + kTile4 - 1) / kTile4;   // float4 per lane
    const int total_r = nops * NR * 16;
    const CRUX_GAS float4 *pd = gas(reinterpret_cast<const float4 *>(E->pdense));
    const CRUX_GAS int32_t *cg = gas(E->ctl);
    float4 pf[kPer];
    int32_t cf = 0;
    auto fetch = [&](int k) {
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            // Rows past the chunk or the plan load a real row (never read).
            const int i = min(q * kTile4 + lane, kChunkRows - 1);
            const int g = min(k * kChunkRows + i, total_r - 1);
            const crux_f32x4 v = *(const CRUX_GAS crux_f32x4 *)(pd + g);
            pf[q] = make_float4(v.x, v.y, v.z, v.w);
        }
        if (lane < PC * kCtlWords)
            cf = (k * PC + lane / kCtlWords < nops) ? cg[k * PC * kCtlWords + lane] : 0;
    };
    auto commit = [&](int buf) {
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int i = q * kTile4 + lane;
            if (i < kChunkRows)
                pch[buf][i] = pf[q];
        }
        if (lane < PC * kCtlWords)
            cch[buf][lane] = cf;
    };

    // ---- tip staging: rows [base, base + n) of the entry's tip list --------
    // One dword load covers four sites of a row, so an instruction covers
    // four rows.  Mask rows are padded to whole tiles.
    auto stage_tips = [&](int base) {
        constexpr int kInstr = kTipChunk / 4;
        const int n = min(kTipChunk, ntips - base);
        const CRUX_CAS uint8_t *const CRUX_CAS *rows =
            (const CRUX_CAS uint8_t *const CRUX_CAS *)cas(E->tip_rows) + base;
        const int sub = lane >> 4;                            // row within the instruction
        const int off = tile * kTile4 + (lane & 15) * 4;      // first site of the lane
        uint32_t v[kInstr];
#pragma unroll
        for (int i = 0; i < kInstr; ++i) {
            const int r = 4 * i + sub;
            v[i] = 0u;
            if (4 * i < n) {
                const CRUX_CAS uint8_t *p0 = rows[4 * i];
                const CRUX_CAS uint8_t *p1 = 4 * i + 1 < n ? rows[4 * i + 1] : p0;
                const CRUX_CAS uint8_t *p2 = 4 * i + 2 < n ? rows[4 * i + 2] : p0;
                const CRUX_CAS uint8_t *p3 = 4 * i + 3 < n ? rows[4 * i + 3] : p0;
                const uint8_t *p = (const uint8_t *)(sub == 0 ? p0 : sub == 1 ? p1 : sub == 2 ? p2 : p3);
                if (r < n)
                    v[i] = *(const CRUX_GAS uint32_t *)gas(p + off);
            }
        }
#pragma unroll
        for (int i = 0; i < kInstr; ++i) {
            const int r = 4 * i + sub;
            if (r < n)
                tipm[r][lane & 15] = v[i];
        }
        // No load may stay outstanding past this point: the compiler would
        // otherwise make later updates wait for their stores as well.
        __builtin_amdgcn_s_waitcnt(0);
    };

    fetch(0);
    commit(0);
    if (nops > PC)
        fetch(1);
    if (ntips > 0)
        stage_tips(E->tip0);

    // Control words and P image (one float per lane and register) of the
    // update about to run.
    int4 cwa, cwb;
    float pr[NR];
    auto read_op = [&](int o) {
        const int rb = (o / PC) & 1, rq = o % PC;
        const int4 *cw = reinterpret_cast<const int4 *>(&cch[rb][rq * kCtlWords]);
        cwa = cw[0];
        cwb = cw[1];
        // Components c >= C have all-zero P columns, so every product below
        // runs over all CM components without a branch and leaves them 0.
        const float *img = reinterpret_cast<const float *>(&pch[rb][rq * NR * 16]);
#pragma unroll
        for (int v = 0; v < NR; ++v)
            pr[v] = img[v * 64 + lane];
    };
    read_op(0);

    // The previous update's scaled result and exponent: a child that is the
    // result of the update just before is taken from here (it was never
    // pushed).
    float4 yl[CM];
    int32_t el = 0;
#pragma unroll
    for (int c = 0; c < CM; ++c)
        yl[c] = make_float4(0.f, 0.f, 0.f, 0.f);

    for (int o = 0; o < nops; ++o) {
        const CRUX_CAS Op *op = ops + o;
        const int buf = (o / PC) & 1, oq = o % PC;
        if (oq == 0 && o > 0 && o + PC < nops)
            fetch(o / PC + 1);
        const int w0 = __builtin_amdgcn_readfirstlane(cwa.x);
        const int32_t pk = __builtin_amdgcn_readfirstlane(cwa.y);
        const int tl = __builtin_amdgcn_readfirstlane(cwa.z);
        const bool has_parent = __builtin_amdgcn_readfirstlane(cwa.w) != 0;
        const uint64_t par = (uint32_t)__builtin_amdgcn_readfirstlane(cwb.x) |
                             ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane(cwb.y) << 32);
        const uint64_t pex = (uint32_t)__builtin_amdgcn_readfirstlane(cwb.z) |
                             ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane(cwb.w) << 32);
        const int nch = w0 & 0xff, kinds = (w0 >> 8) & 0xff, fst = (w0 >> 16) & 0xf,
                  fwd = (w0 >> 20) & 0xf;
        if (tl > 0)
            stage_tips(tl);
        // The chunk after this one goes into the other buffer as this one
        // ends.
        if (oq == PC - 1 && o + 1 < nops)
            commit(buf ^ 1);

        // LDS reads of every child slot's stack entries and tip mask first,
        // whatever the kinds (in bounds whatever the slot holds): they are in
        // flight while the forwarded child is multiplied, and values defined
        // on every path cost no register copies.
        float4 sx[kOpMaxChild][CM];
        int32_t se[kOpMaxChild];
        uint32_t tm[kOpMaxChild];
#pragma unroll
        for (int j = 0; j < kOpMaxChild; ++j) {
            const int p = op_pos(pk, j);
#pragma unroll
            for (int c = 0; c < CM; ++c)
                sx[j][c] = stk[p & (D - 1)][c][lane];
            se[j] = stke[p & (D - 1)][lane];
            tm[j] = tipm[p & (kTipChunk - 1)][lane >> 2];
        }
        auto stack_of = [&](int j, int c) -> float4 { return sx[j][c]; };
        float4 acc[CM];
        int32_t ex = 0;
        // acc = (first child ? v : acc * v), v = P_c x_c (matrix core, the
        // components' chains interleaved; control flow is uniform across the
        // wave here, as the matrix instructions need) or x_c (ident child).
        // Column k of P_c comes from block q % 16 of image register q / 16,
        // broadcast to all 16 blocks (CBSZ 4, ABID q % 16), so block b
        // forms P_c[:][k] x_k for its four sites.
        auto take = [&](auto Jc, auto &&xc, bool mat) {
            constexpr int j = decltype(Jc)::value;
            if (mat) {
                crux_f32x4 d[CM];
                static_for<4>([&](auto Kc) {
                    constexpr int k = decltype(Kc)::value;
                    static_for<CM>([&](auto Cc) {
                        constexpr int c = decltype(Cc)::value;
                        constexpr int q = (j * CM + c) * 4 + k;
                        const float4 xv = xc(c);
                        const float xk = k == 0 ? xv.x : k == 1 ? xv.y : k == 2 ? xv.z : xv.w;
                        d[c] = __builtin_amdgcn_mfma_f32_4x4x1f32(
                            pr[q >> 4], xk, k == 0 ? crux_f32x4{0.f, 0.f, 0.f, 0.f} : d[c], 4, q & 15, 0);
                    });
                });
#pragma unroll
                for (int c = 0; c < CM; ++c) {
                    const float4 v = make_float4(d[c].x, d[c].y, d[c].z, d[c].w);
                    acc[c] = (j == 0) ? v : mul4(acc[c], v);
                }
            } else {
#pragma unroll
                for (int c = 0; c < CM; ++c)
                    acc[c] = (j == 0) ? xc(c) : mul4(acc[c], xc(c));
            }
        };
        static_for<kOpMaxChild>([&](auto Jc) {
            constexpr int j = decltype(Jc)::value;
            if (j >= nch)
                return;
            const int kind = (kinds >> (2 * j)) & 3;
            const bool mat = kind != kChildIdent;
            if (kind == kChildTip) {
                const uint32_t m = tm[j] >> (8 * (lane & 3));
                const float4 t = make_float4((float)(m & 1u), (float)((m >> 1) & 1u),
                                             (float)((m >> 2) & 1u), (float)((m >> 3) & 1u));
                take(Jc, [&](int) { return t; }, true);
            } else if ((fwd >> j) & 1) {
                take(Jc, [&](int c) { return yl[c]; }, mat);
                ex += el;
            } else if ((fst >> j) & 1) {
                take(Jc, [&](int c) { return stack_of(j, c); }, mat);
                ex += se[j];
            } else {
                // Cached by an earlier plan (partial re-evaluation).  The wait
                // stays inside this branch (see stage_tips).
                const CRUX_CAS OpChild &ch = op->ch[j];
                float4 x[CM];
                const auto rc = rsrc_of(ch.clv);
#pragma unroll
                for (int c = 0; c < CM; ++c) {
                    // Components past C read component C - 1 (in the slot).
                    const crux_f32x4 v = bload4(rc, (uint32_t)sc * 16, (uint32_t)(min(c, C - 1) * S) * 16);
                    x[c] = c < C ? make_float4(v.x, v.y, v.z, v.w) : make_float4(0.f, 0.f, 0.f, 0.f);
                }
                ex += bload1(rsrc_of(ch.exps), (uint32_t)sc * 4, 0);
                __builtin_amdgcn_s_waitcnt(0);
                take(Jc, [&](int c) { return x[c]; }, mat);
            }
        });
        // The next update's control words and P image, while this one ends.
        if (o + 1 < nops)
            read_op(o + 1);
        if (has_parent) {
            // Site maximum over every component (idle ones are 0).  All
            // values are >= 0, so the largest as a signed integer is the
            // largest float (NaN above +inf, -0 below everything).
            int32_t mb = 0;
#pragma unroll
            for (int c = 0; c < CM; ++c)
                mb = max(max(mb, max(__float_as_int(acc[c].x), __float_as_int(acc[c].y))),
                         max(__float_as_int(acc[c].z), __float_as_int(acc[c].w)));
            const float m = __int_as_float(mb);
            int e2 = 0;
            if (m > 0.0f)
                (void)frexpf(m, &e2);
            el = ex + e2;
#pragma unroll
            for (int c = 0; c < CM; ++c)
                yl[c] = make_float4(ldexpf(acc[c].x, -e2), ldexpf(acc[c].y, -e2),
                                    ldexpf(acc[c].z, -e2), ldexpf(acc[c].w, -e2));
            const int pp = op_pos(pk, kOpMaxChild);
            if (pp >= 0) {
#pragma unroll
                for (int c = 0; c < CM; ++c)
                    stk[pp][c][lane] = yl[c];
                stke[pp][lane] = el;
            }
            // Each component plane is its own range of S sites: lanes past the
            // last site and components past C (an empty range) store nothing.
            const char *pp0 = reinterpret_cast<const char *>(par);
#pragma unroll
            for (int c = 0; c < CM; ++c)
                bstore4<kNT>(crux_f32x4{yl[c].x, yl[c].y, yl[c].z, yl[c].w},
                        rsrc_of(pp0 + (size_t)c * S * 16, c < C ? (uint32_t)S * 16 : 0u),
                        (uint32_t)s * 16, 0);
            bstore1<kNT>(el, rsrc_of(reinterpret_cast<const void *>(pex), (uint32_t)S * 4), (uint32_t)s * 4, 0);
        } else {
            double L = 0.0;
#pragma unroll
            for (int c = 0; c < CM; ++c) {
                if (c < C) {
                    const CRUX_CAS double *piw = cas(E->piw) + 4 * c;
                    L += piw[0] * (double)acc[c].x + piw[1] * (double)acc[c].y +
                         piw[2] * (double)acc[c].z + piw[3] * (double)acc[c].w;
                }
            }
            double v = (log(L) + (double)ex * 0.69314718055994530942) *
                       (double)gas(E->weights)[sc];
            if (__builtin_isnan(v))
                v = -INFINITY;
            if (act) {
                gas(E->site_out)[s] = v;
                sum = v;
            }
        }
    }
    // Compensated wave reduction: both lanes of a pair merge in the same order.
    double cmp = 0.0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double s2 = __shfl_xor(sum, off, 64), c2 = __shfl_xor(cmp, off, 64);
        if ((lane & off) == 0) {
            nsum_merge(sum, cmp, s2, c2);
        } else {
            double a = s2, bb = c2;
            nsum_merge(a, bb, sum, cmp);
            sum = a;
            cmp = bb;
        }
    }
    if (lane == 0 && E->partials != nullptr) {
        gas(E->partials)[2 * tile] = sum;
        gas(E->partials)[2 * tile + 1] = cmp;
    }
}

// ---------------------------------------------------------------------------
// 4-state traversal with the components split over lane halves (CM >= 2).
// A block is two waves covering one 64-site tile, 32 sites each: lane l of
// wave w holds site 32 w + l % 32 of the tile and, in half h = l / 32,
// components h CM/2 .. h CM/2 + CM/2 - 1.  Each lane does half of k_trav4's
// per-site work and there are twice the waves, two per SIMD, so that one
// wave issues while the other waits.  The two waves share the staged P images
// and control words (one barrier per chunk of updates) and keep their own
// LDS stacks and tip masks.  P x runs on the matrix core with the A operand
// broadcast within each half (CBSZ 3: the blocks of lanes 0-31 take block
// ABID, those of lanes 32-63 block 8 + ABID), so an image register carries 8
// columns of each half's components; the site maximum pairs the halves with
// v_permlane32_swap.  Plan walk, forwarding, stack, rescaling and the root
// reduction follow k_trav4; at the root the halves' component sums are
// added, and the block's two wave sums are merged in site order.  CLV slots
// are tiles of 32 sites laid out as the lane halves hold them, so every
// store writes 1 KiB contiguously; stores are non-temporal.
//
// PC: updates per staged chunk.  Each chunk ends in a barrier at which both
// waves wait for their stores to drain, so the untuned length is as long as
// the LDS allows with four blocks per CU (at most 40 KiB per block: 11
// updates at four components with up to 4 stack slots, 14 at two).
template <int CM, int D, int PC>
__global__ void __launch_bounds__(2 * kTile4)
k_trav4h(const Entry *__restrict__ entries, const int32_t *__restrict__ block_entry, int blk0)
{
    static_assert(CM >= 2, "at least one component per lane half");
    static_assert(PC * kCtlWords <= 2 * kTile4, "a chunk's control words fit one load per thread");
    constexpr int CL = CM / 2;                    // components per lane
    constexpr int HS = kTile4 / 2;                // sites per wave
    constexpr int NR = p4_regs(CM);               // P image registers per update
    constexpr int kChunkRows = PC * NR * 16;      // float4s of P per chunk
    __shared__ float4 stk[2][D][CL][kTile4];
    __shared__ int32_t stke[2][D][kTile4];
    __shared__ uint32_t tipm[2][kTipChunk][HS / 4];
    __shared__ float4 pch[2][kChunkRows];
    __shared__ int32_t cch[2][PC * kCtlWords];
    __shared__ double red[2];
    const int b = blk0 + (int)blockIdx.x;
    const CRUX_CAS Entry *E = cas(entries) + cas(block_entry)[b];
    const int S = E->S, C = E->C, nops = E->nops, ntips = E->ntips;
    const int tile = b - E->tile0;
    const int tid = (int)threadIdx.x;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    const int h = lane >> 5;
    const int s = tile * kTile4 + wv * HS + (lane & 31);
    const bool act = s < S;
    const int sc = act ? s : S - 1;       // in-bounds site for loads of idle lanes
    const CRUX_CAS Op *ops = cas(E->ops);
    double sum = 0.0;
    // A slot holds tiles of 32 sites, each [CL][2][32][4] floats: component
    // h CL + m of site 32 t + i at ((t CL + m) 2 + h) 32 + i, so that one
    // store (fixed m) of a wave writes 1 KiB and an update 512 CM bytes
    // contiguously.  Lanes past the last site and components past C (lane
    // half 1 only) use an offset no buffer range covers: their stores write
    // nothing and their loads read zero.
    constexpr uint32_t kOut = 0xfffffff0u;
    uint32_t vo[CL], vol[CL];
#pragma unroll
    for (int m = 0; m < CL; ++m) {
        const bool comp = h == 0 || CL + m < C;
        const auto at = [&](int site) {
            return (uint32_t)((((site >> 5) * CL + m) * 2 + h) * 32 + (site & 31)) * 16;
        };
        vo[m] = act && comp ? at(s) : kOut;
        vol[m] = comp ? at(sc) : kOut;
    }
    const uint32_t voe = (act && h == 0) ? (uint32_t)s * 4 : kOut;

    // Both waves of the block order their LDS writes before a barrier; no
    // global memory is ordered by it.
    auto lds_barrier = [&]() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    };

    // ---- P / control-word staging, shared by the two waves ---------------
    constexpr int kPer = (kChunkRows + 2 * kTile4 - 1) / (2 * kTile4);
    const int total_r = nops * NR * 16;
    const CRUX_GAS float4 *pd = gas(reinterpret_cast<const float4 *>(E->pdense));
    const CRUX_GAS int32_t *cg = gas(E->ctl);
    float4 pf[kPer];
    int32_t cf = 0;
    auto fetch = [&](int k) {
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            // Rows past the chunk or the plan load a real row (never read).
            const int i = min(q * 2 * kTile4 + tid, kChunkRows - 1);
            const int g = min(k * kChunkRows + i, total_r - 1);
            const crux_f32x4 v = *(const CRUX_GAS crux_f32x4 *)(pd + g);
            pf[q] = make_float4(v.x, v.y, v.z, v.w);
        }
        if (tid < PC * kCtlWords)
            cf = (k * PC + tid / kCtlWords < nops) ? cg[k * PC * kCtlWords + tid] : 0;
    };
    auto commit = [&](int buf) {
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int i = q * 2 * kTile4 + tid;
            if (kChunkRows % (2 * kTile4) == 0 || i < kChunkRows)
                pch[buf][i] = pf[q];
        }
        if (tid < PC * kCtlWords)
            cch[buf][tid] = cf;
        lds_barrier();
    };

    // ---- tip staging: rows [base, base + n) of the entry's tip list --------
    // One dword covers four sites of a row: 8 rows of the wave's 32 sites
    // per load instruction.  Mask rows are padded to whole 64-site tiles.
    auto stage_tips = [&](int base) {
        constexpr int kDw = HS / 4;                  // dwords of a row per wave
        constexpr int kRows = 64 / kDw;              // rows per instruction
        constexpr int kInstr = kTipChunk / kRows;
        const int n = min(kTipChunk, ntips - base);
        // A chunk outside the list has no rows to read.
        if (n <= 0)
            return;
        const CRUX_GAS uint8_t *const CRUX_GAS *rows =
            (const CRUX_GAS uint8_t *const CRUX_GAS *)gas(E->tip_rows) + base;
        const int sub = lane / kDw, dw = lane % kDw;
        const int off = tile * kTile4 + wv * HS + dw * 4;
        // Row pointers first, then the rows (rows past the list read the
        // last row again and are not kept): two waits in all.
        const CRUX_GAS uint8_t *rp[kInstr];
#pragma unroll
        for (int i = 0; i < kInstr; ++i)
            rp[i] = rows[min(kRows * i + sub, n - 1)];
        uint32_t v[kInstr];
#pragma unroll
        for (int i = 0; i < kInstr; ++i)
            v[i] = *(const CRUX_GAS uint32_t *)(rp[i] + off);
#pragma unroll
        for (int i = 0; i < kInstr; ++i) {
            const int r = kRows * i + sub;
            if (r < n)
                tipm[wv][r][dw] = v[i];
        }
        // No load may stay outstanding past this point: the compiler would
        // otherwise make later updates wait for their stores as well.
        __builtin_amdgcn_s_waitcnt(0);
    };

    // Updates run in chunks of PC: chunk k's P images and control words are
    // fetched into registers while chunk k - 1 runs and written to LDS
    // buffer k % 2 when it ends.
    const int nchunks = (nops + PC - 1) / PC;
    fetch(0);
    commit(0);
    if (nchunks > 1)
        fetch(1);
    if (ntips > 0)
        stage_tips(E->tip0);

    // Control words and P image of the update about to run.
    int4 cwa, cwb;
    float pr[NR];
    auto read_op = [&](int rb, int rq) {
        const int4 *cw = reinterpret_cast<const int4 *>(&cch[rb][rq * kCtlWords]);
        cwa = cw[0];
        cwb = cw[1];
        const float *img = reinterpret_cast<const float *>(&pch[rb][rq * NR * 16]);
#pragma unroll
        for (int v = 0; v < NR; ++v)
            pr[v] = img[v * 64 + lane];
    };
    read_op(0, 0);

    float4 yl[CL];
    int32_t el = 0;
#pragma unroll
    for (int m = 0; m < CL; ++m)
        yl[m] = make_float4(0.f, 0.f, 0.f, 0.f);
    // Children of the next update cached in HBM by an earlier plan (partial
    // re-evaluations), loaded while the current update computes and before
    // it issues its stores: the next update then waits for these loads only,
    // not for the current update's stores as well (one counter tracks both,
    // in order).  npf: they were loaded (not for a chunk's first update).
    float4 xp[2][CL];
    int32_t epf[2] = {0, 0};
    bool npf = false;

    for (int k = 0; k < nchunks; ++k) {
    const int buf = k & 1, o0 = k * PC, n = min(PC, nops - o0);
    if (k > 0 && k + 1 < nchunks)
        fetch(k + 1);
    // At least one update per chunk, and every update issues its stores (the
    // root's to an empty range): the compiler's wait for the fetch at the
    // chunk's end then never counts the last update's stores.
    int oq = 0;
    do {
        const CRUX_CAS Op *op = ops + o0 + oq;
        const int w0 = __builtin_amdgcn_readfirstlane(cwa.x);
        const int32_t pk = __builtin_amdgcn_readfirstlane(cwa.y);
        const int tl = __builtin_amdgcn_readfirstlane(cwa.z);
        const uint64_t par = (uint32_t)__builtin_amdgcn_readfirstlane(cwb.x) |
                             ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane(cwb.y) << 32);
        // The exponents follow the CLV planes in the parent's slot.
        const uint64_t pex = par + (uint32_t)E->exp_off;
        const bool has_parent = (w0 >> 24) & 1;
        const int nch = w0 & 0xff, kinds = (w0 >> 8) & 0xff, fst = (w0 >> 16) & 0xf,
                  fwd = (w0 >> 20) & 0xf;
        if (tl > 0)
            stage_tips(tl);

        float4 sx[kOpMaxChild][CL];
        int32_t se[kOpMaxChild];
        uint32_t tm[kOpMaxChild];
        auto read_stack = [&](int j) {
            const int p = op_pos(pk, j);
#pragma unroll
            for (int m = 0; m < CL; ++m)
                sx[j][m] = stk[wv][p & (D - 1)][m][lane];
            se[j] = stke[wv][p & (D - 1)][lane];
        };
        auto read_tip = [&](int j) { tm[j] = tipm[wv][op_pos(pk, j) & (kTipChunk - 1)][(lane & 31) >> 2]; };
        float4 acc[CL];
        int32_t ex = 0;
        // d = P_c x_c on the matrix core for child j, the components' chains
        // interleaved (control flow is uniform across the wave here, as the
        // matrix instructions need).
        auto pmul = [&](auto Jc, auto &&xc, crux_f32x4 (&d)[CL]) {
            constexpr int j = decltype(Jc)::value;
            static_for<4>([&](auto Kc) {
                constexpr int k = decltype(Kc)::value;
                static_for<CL>([&](auto Mc) {
                    constexpr int m = decltype(Mc)::value;
                    constexpr int q = (j * CL + m) * 4 + k;
                    const float4 xv = xc(m);
                    const float xk = k == 0 ? xv.x : k == 1 ? xv.y : k == 2 ? xv.z : xv.w;
                    d[m] = __builtin_amdgcn_mfma_f32_4x4x1f32(
                        pr[q >> 3], xk, k == 0 ? crux_f32x4{0.f, 0.f, 0.f, 0.f} : d[m], 3, q & 7, 0);
                });
            });
        };
        auto tip_vec = [&](int j) {
            const uint32_t mk = tm[j] >> (8 * (lane & 3));
            return make_float4((float)(mk & 1u), (float)((mk >> 1) & 1u), (float)((mk >> 2) & 1u),
                               (float)((mk >> 3) & 1u));
        };
        // A child cached by an earlier plan (partial re-evaluation): planes
        // past the slot's C read as zero (buffer range).  The caller waits.
        auto load_child_of = [&](const CRUX_CAS Op *uop, int j, float4 (&x)[CL]) -> int32_t {
            const CRUX_CAS OpChild &ch = uop->ch[j];
            const char *cb = reinterpret_cast<const char *>(ch.clv);
            const auto rc = rsrc_of(cb, (uint32_t)E->exp_off);
#pragma unroll
            for (int m = 0; m < CL; ++m) {
                const crux_f32x4 v = bload4(rc, vol[m], 0);
                x[m] = make_float4(v.x, v.y, v.z, v.w);
            }
            return bload1(rsrc_of(ch.exps), (uint32_t)sc * 4, 0);
        };
        auto load_child = [&](int j, float4 (&x)[CL]) -> int32_t { return load_child_of(op, j, x); };
        // Two-child updates with a parent and no identity child (most
        // updates) run a straight-line body specialised for their children's
        // sources; the control word names the pair.  acc = d0 * d1 and the
        // exponent sum are formed as in the generic walk below.
        auto two = [&](auto S0c, auto S1c) {
            constexpr int s0 = decltype(S0c)::value, s1 = decltype(S1c)::value;
            if constexpr (s0 == kSrcStk)
                read_stack(0);
            if constexpr (s1 == kSrcStk)
                read_stack(1);
            if constexpr (s0 == kSrcTip)
                read_tip(0);
            if constexpr (s1 == kSrcTip)
                read_tip(1);
            float4 xh[2][CL];
            int32_t eh[2] = {0, 0};
            if constexpr (s0 == kSrcHbm || s1 == kSrcHbm) {
                if (npf) {
                    // Prefetched by the previous update.
#pragma unroll
                    for (int m = 0; m < CL; ++m) {
                        xh[0][m] = xp[0][m];
                        xh[1][m] = xp[1][m];
                    }
                    eh[0] = epf[0];
                    eh[1] = epf[1];
                } else {
                    if constexpr (s0 == kSrcHbm)
                        eh[0] = load_child(0, xh[0]);
                    if constexpr (s1 == kSrcHbm)
                        eh[1] = load_child(1, xh[1]);
                    __builtin_amdgcn_s_waitcnt(0);
                }
            }
            crux_f32x4 d[2][CL];
            static_for<2>([&](auto Jc) {
                constexpr int j = decltype(Jc)::value;
                constexpr int sj = j == 0 ? s0 : s1;
                if constexpr (sj == kSrcTip) {
                    const float4 t = tip_vec(j);
                    pmul(Jc, [&](int) { return t; }, d[j]);
                } else if constexpr (sj == kSrcFwd) {
                    pmul(Jc, [&](int m) { return yl[m]; }, d[j]);
                    ex += el;
                } else if constexpr (sj == kSrcStk) {
                    pmul(Jc, [&](int m) { return sx[j][m]; }, d[j]);
                    ex += se[j];
                } else {
                    pmul(Jc, [&](int m) { return xh[j][m]; }, d[j]);
                    ex += eh[j];
                }
            });
#pragma unroll
            for (int m = 0; m < CL; ++m)
                acc[m] = mul4(make_float4(d[0][m].x, d[0][m].y, d[0][m].z, d[0][m].w),
                              make_float4(d[1][m].x, d[1][m].y, d[1][m].z, d[1][m].w));
        };
        using Tip = std::integral_constant<int, kSrcTip>;
        using Fwd = std::integral_constant<int, kSrcFwd>;
        using Stk = std::integral_constant<int, kSrcStk>;
        using Hbm = std::integral_constant<int, kSrcHbm>;
        auto two_of = [&](auto S0c, int s1) {
            if (s1 == kSrcTip)
                two(S0c, Tip{});
            else if (s1 == kSrcStk)
                two(S0c, Stk{});
            else if (s1 == kSrcHbm)
                two(S0c, Hbm{});
            else if constexpr (decltype(S0c)::value != kSrcFwd)
                two(S0c, Fwd{});
        };
        const int shape = (w0 >> 25) & 31;
        if (shape != 0) {
            const int s0 = (shape - 1) >> 2, s1 = (shape - 1) & 3;
            if (s0 == kSrcFwd)
                two_of(Fwd{}, s1);
            else if (s0 == kSrcTip)
                two_of(Tip{}, s1);
            else if (s0 == kSrcStk)
                two_of(Stk{}, s1);
            else
                two_of(Hbm{}, s1);
        } else {
            // Generic walk: any child count, identity children, the root.
#pragma unroll
            for (int j = 0; j < kOpMaxChild; ++j) {
                read_stack(j);
                read_tip(j);
            }
            static_for<kOpMaxChild>([&](auto Jc) {
                constexpr int j = decltype(Jc)::value;
                if (j >= nch)
                    return;
                const int kind = (kinds >> (2 * j)) & 3;
                float4 x[CL];
                int32_t e = 0;
                if (kind == kChildTip) {
                    const float4 t = tip_vec(j);
#pragma unroll
                    for (int m = 0; m < CL; ++m)
                        x[m] = t;
                } else if ((fwd >> j) & 1) {
#pragma unroll
                    for (int m = 0; m < CL; ++m)
                        x[m] = yl[m];
                    e = el;
                } else if ((fst >> j) & 1) {
#pragma unroll
                    for (int m = 0; m < CL; ++m)
                        x[m] = sx[j][m];
                    e = se[j];
                } else {
                    e = load_child(j, x);
                    __builtin_amdgcn_s_waitcnt(0);
                }
                ex += e;
                float4 v[CL];
                if (kind != kChildIdent) {
                    crux_f32x4 d[CL];
                    pmul(Jc, [&](int m) { return x[m]; }, d);
#pragma unroll
                    for (int m = 0; m < CL; ++m)
                        v[m] = make_float4(d[m].x, d[m].y, d[m].z, d[m].w);
                } else {
#pragma unroll
                    for (int m = 0; m < CL; ++m)
                        v[m] = x[m];
                }
#pragma unroll
                for (int m = 0; m < CL; ++m)
                    acc[m] = (j == 0) ? v[m] : mul4(acc[m], v[m]);
            });
        }
        // The next update's control words and P image, while this one ends,
        // and the children it takes from HBM (two-child shapes).
        npf = false;
        if (oq + 1 < n) {
            read_op(buf, oq + 1);
            const int ns = (__builtin_amdgcn_readfirstlane(cwa.x) >> 25) & 31;
            if (ns != 0) {
                const CRUX_CAS Op *nop = op + 1;
                if (((ns - 1) >> 2) == kSrcHbm)
                    epf[0] = load_child_of(nop, 0, xp[0]);
                if (((ns - 1) & 3) == kSrcHbm)
                    epf[1] = load_child_of(nop, 1, xp[1]);
                npf = true;
            }
        }
        // Site maximum over the lane's components, then over the halves, and
        // the rescaled result (also formed at the root, where nothing uses
        // it: the loop-carried registers then have one definition).
        int32_t mb = 0;
#pragma unroll
        for (int m = 0; m < CL; ++m)
            mb = max(max(mb, max(__float_as_int(acc[m].x), __float_as_int(acc[m].y))),
                     max(__float_as_int(acc[m].z), __float_as_int(acc[m].w)));
        {
            uint32_t a0, a1;
            lane_pair<32>((uint32_t)mb, a0, a1);
            mb = max((int32_t)a0, (int32_t)a1);
        }
        const float mx = __int_as_float(mb);
        int e2 = 0;
        if (mx > 0.0f)
            (void)frexpf(mx, &e2);
        el = ex + e2;
#pragma unroll
        for (int m = 0; m < CL; ++m)
            yl[m] = make_float4(ldexpf(acc[m].x, -e2), ldexpf(acc[m].y, -e2),
                                ldexpf(acc[m].z, -e2), ldexpf(acc[m].w, -e2));
        {
            const char *pp0 = reinterpret_cast<const char *>(par);
            const auto rp = rsrc_of(pp0, has_parent ? (uint32_t)E->exp_off : 0u);
#pragma unroll
            for (int m = 0; m < CL; ++m)
                bstore4<kNT>(crux_f32x4{yl[m].x, yl[m].y, yl[m].z, yl[m].w}, rp, vo[m], 0);
            bstore1<kNT>(el, rsrc_of(reinterpret_cast<const void *>(pex), has_parent ? (uint32_t)S * 4 : 0u),
                         voe, 0);
        }
        if (has_parent) {
            const int pp = op_pos(pk, kOpMaxChild);
            if (pp >= 0) {
#pragma unroll
                for (int m = 0; m < CL; ++m)
                    stk[wv][pp][m][lane] = yl[m];
                stke[wv][pp][lane] = el;
            }
        } else {
            double L = 0.0;
#pragma unroll
            for (int m = 0; m < CL; ++m) {
                const CRUX_CAS double *w0p = cas(E->piw) + 4 * m;
                double wx = w0p[0], wy = w0p[1], wz = w0p[2], wq = w0p[3];
                if (CL + m < C) {
                    const CRUX_CAS double *w1p = cas(E->piw) + 4 * (CL + m);
                    if (h) {
                        wx = w1p[0];
                        wy = w1p[1];
                        wz = w1p[2];
                        wq = w1p[3];
                    }
                } else if (h) {
                    wx = wy = wz = wq = 0.0;
                }
                L += wx * (double)acc[m].x + wy * (double)acc[m].y +
                     wz * (double)acc[m].z + wq * (double)acc[m].w;
            }
            {
                const uint64_t u = (uint64_t)__double_as_longlong(L);
                uint32_t al, bl, ah, bh;
                lane_pair<32>((uint32_t)u, al, bl);
                lane_pair<32>((uint32_t)(u >> 32), ah, bh);
                // Commutative: both halves get the same sum.
                L = __longlong_as_double((long long)(((uint64_t)ah << 32) | al)) +
                    __longlong_as_double((long long)(((uint64_t)bh << 32) | bl));
            }
            double v = (log(L) + (double)ex * 0.69314718055994530942) *
                       (double)gas(E->weights)[sc];
            if (__builtin_isnan(v))
                v = -INFINITY;
            if (act && h == 0) {
                gas(E->site_out)[s] = v;
                sum = v;
            }
        }
    } while (++oq < n);
    // The next chunk goes into the other buffer; its first update is read
    // after the barrier.
    if (k + 1 < nchunks) {
        commit(buf ^ 1);
        read_op(buf ^ 1, 0);
    }
    }
    // Compensated wave reduction: both lanes of a pair merge in the same order.
    double cmp = 0.0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double s2 = __shfl_xor(sum, off, 64), c2 = __shfl_xor(cmp, off, 64);
        if ((lane & off) == 0) {
            nsum_merge(sum, cmp, s2, c2);
        } else {
            double a = s2, bb = c2;
            nsum_merge(a, bb, sum, cmp);
            sum = a;
            cmp = bb;
        }
    }
    // The tile's sum: wave 0's sites, then wave 1's.
    if (wv == 1 && lane == 0) {
        red[0] = sum;
        red[1] = cmp;
    }
    lds_barrier();
    if (wv == 0 && lane == 0 && E->partials != nullptr) {
        nsum_merge(sum, cmp, red[0], red[1]);
        gas(E->partials)[2 * tile] = sum;
        gas(E->partials)[2 * tile + 1] = cmp;
    }
}

// ---------------------------------------------------------------------------
// Short 4-state plans (partial re-evaluations, MC3 proposals: at most
// kPartMaxOps updates), two or four components.  The lanes, slots,
// arithmetic and results are k_trav4h's, bit for bit; what differs is where
// the inputs come from.  In a short plan almost every update takes a child
// cached in HBM by an earlier plan, and k_trav4h's one-update-ahead load of
// it leaves each update waiting for most of an HBM round trip.  Here:
// * the P images and control words of every update are staged into LDS
//   once, every load issued before the first wait, behind one barrier;
// * the children of update o + R that come from memory (tip rows, and CLVs
//   cached by an earlier plan: the host lists them, with their addresses, in
//   the control words) are loaded while update o runs, into a ring of R
//   register slots.  The update loop is unrolled R times so that each slot is
//   a fixed set of registers, and every update issues the same loads and
//   stores (a child with nothing to load reads an empty range), so that the
//   compiler's wait for a slot counts exactly the updates issued since.  The
//   first R updates' loads go out with the staging loads, from control words
//   read with scalar loads, and the plan's root takes its slot like any
//   other update;
// * the plan's own results re-read from HBM (after the stack overflowed) and
//   identity children are loaded where they are used, as in k_trav4h's
//   generic walk.
// The last update (the plan's root; the host checks that no other update
// lacks a parent) runs after the loop.
template <int CM, int D, int R>
__global__ void __launch_bounds__(2 * kTile4)
k_part4h(const Entry *__restrict__ entries, const int32_t *__restrict__ block_entry, int blk0)
{
    static_assert(CM >= 2, "at least one component per lane half");
    constexpr int CL = CM / 2;                    // components per lane
    constexpr int HS = kTile4 / 2;                // sites per wave
    constexpr int NT = 2 * kTile4;                // threads of a block
    constexpr int NC = kOpMaxChild;
    constexpr int NR = p4_regs(CM);               // P image registers per update
    constexpr int kRows = kPartMaxOps * NR * 16;  // float4s of P of a whole plan
    constexpr int kWords = kPartMaxOps * kCtlWordsPart;
    static_assert(kRows % NT == 0 && kWords % NT == 0, "staging: a whole number of loads per thread");
    __shared__ float4 stk[2][D][CL][kTile4];
    __shared__ int32_t stke[2][D][kTile4];
    __shared__ float4 pch[kRows];
    __shared__ int32_t cch[kWords];
    __shared__ double red[2];
    const int b = blk0 + (int)blockIdx.x;
    const CRUX_CAS Entry *E = cas(entries) + cas(block_entry)[b];
    const int S = E->S, C = E->C, nops = E->nops;
    const int tile = b - E->tile0;
    const int tid = (int)threadIdx.x;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    const int h = lane >> 5;
    const int s = tile * kTile4 + wv * HS + (lane & 31);
    const bool act = s < S;
    const int sc = act ? s : S - 1;       // in-bounds site for loads of idle lanes
    const uint32_t exp_off = (uint32_t)E->exp_off;
    // Bytes of a padded tip row (a multiple of the tile).
    const uint32_t row_bytes = (uint32_t)((S + kTile4 - 1) / kTile4 * kTile4);
    double sum = 0.0;
    // Slot offsets as in k_trav4h (kOut: outside every range).
    constexpr uint32_t kOut = 0xfffffff0u;
    uint32_t vo[CL], vol[CL];
#pragma unroll
    for (int m = 0; m < CL; ++m) {
        const bool comp = h == 0 || CL + m < C;
        const auto at = [&](int site) {
            return (uint32_t)((((site >> 5) * CL + m) * 2 + h) * 32 + (site & 31)) * 16;
        };
        vo[m] = act && comp ? at(s) : kOut;
        vol[m] = comp ? at(sc) : kOut;
    }
    const uint32_t voe = (act && h == 0) ? (uint32_t)s * 4 : kOut;
    // Offsets of this lane's exponent, and of the dword of its tip mask byte.
    const uint32_t voex = (uint32_t)sc * 4, votip = (uint32_t)sc & ~3u;

    auto lds_barrier = [&]() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    };

    // What update o's children load ahead, into a ring slot: CL float4 and a
    // dword each (a cached CLV and its exponent, or the dword of a tip row
    // holding this lane's mask byte), the same loads whatever the children
    // are (an empty range for a child with nothing to load, or o past the
    // plan's end).  cw0: the plan's control words, in LDS or in memory.
    auto load_ahead = [&](int o, const auto *cw0, float4 (&x)[NC][CL], int32_t (&e)[NC]) {
        const bool in = o < nops;
        const auto *cw = cw0 + (in ? o : 0) * kCtlWordsPart;
        const uint32_t w3 = in ? (uint32_t)__builtin_amdgcn_readfirstlane(cw[3]) : 0u;
        static_for<NC>([&](auto Jc) {
            constexpr int j = decltype(Jc)::value;
            const bool cl = (w3 >> (kCtlAheadShift + j)) & 1, tp = (w3 >> (kCtlTipShift + j)) & 1;
            const uint64_t p = (uint32_t)__builtin_amdgcn_readfirstlane(cw[kCtlAhead + 2 * j]) |
                               ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane(cw[kCtlAhead + 2 * j + 1]) << 32);
            const auto rc = rsrc_of(reinterpret_cast<const void *>(p), cl ? exp_off : 0u);
#pragma unroll
            for (int m = 0; m < CL; ++m) {
                const crux_f32x4 v = bload4(rc, vol[m], 0);
                x[j][m] = make_float4(v.x, v.y, v.z, v.w);
            }
            const auto re = rsrc_of(reinterpret_cast<const void *>(cl ? p + exp_off : p),
                                    cl ? (uint32_t)S * 4 : tp ? row_bytes : 0u);
            e[j] = bload1(re, tp ? votip : voex, 0);
        });
    };

    // Staging, with the first R updates' ring loads: every load issued, then
    // the staged data written to LDS behind one barrier (the ring loads stay
    // outstanding).
    float4 hx[R][NC][CL];
    int32_t hxe[R][NC];
    {
        const int nr = nops * NR * 16, nc = nops * kCtlWordsPart;
        const auto rp = rsrc_of(E->pdense, (uint32_t)nr * 16);
        const auto rw = rsrc_of(E->ctl, (uint32_t)nc * 4);
        crux_f32x4 pv[kRows / NT];
        int32_t cv[kWords / NT];
#pragma unroll
        for (int k = 0; k < kRows / NT; ++k)
            pv[k] = bload4(rp, (uint32_t)(tid + k * NT) * 16, 0);
#pragma unroll
        for (int k = 0; k < kWords / NT; ++k)
            cv[k] = bload1(rw, (uint32_t)(tid + k * NT) * 4, 0);
        static_for<R>([&](auto Kc) {
            constexpr int k = decltype(Kc)::value;
            load_ahead(k, cas(E->ctl), hx[k], hxe[k]);
        });
#pragma unroll
        for (int k = 0; k < kRows / NT; ++k)
            pch[tid + k * NT] = make_float4(pv[k].x, pv[k].y, pv[k].z, pv[k].w);
#pragma unroll
        for (int k = 0; k < kWords / NT; ++k)
            cch[tid + k * NT] = cv[k];
        lds_barrier();
    }

    float4 yl[CL];
    int32_t el = 0;
#pragma unroll
    for (int m = 0; m < CL; ++m)
        yl[m] = make_float4(0.f, 0.f, 0.f, 0.f);

    // Update o with what its children loaded ahead in hx/hxe (or loaded here:
    // own results, identity children).  Returns the product of the children
    // and the summed exponent.
    auto body = [&](int o, const float4 (&hx)[NC][CL], const int32_t (&hxe)[NC], float4 (&acc)[CL],
                    int32_t &ex, int &pk_out, int &w0_out) {
        const int32_t *cw = &cch[o * kCtlWordsPart];
        const int w0 = __builtin_amdgcn_readfirstlane(cw[0]);
        const int32_t pk = __builtin_amdgcn_readfirstlane(cw[1]);
        const int w3 = __builtin_amdgcn_readfirstlane(cw[3]);
        const int nch = w0 & 0xff, kinds = (w0 >> 8) & 0xff, fst = (w0 >> 16) & 0xf,
                  fwd = (w0 >> 20) & 0xf, ahead = w3 >> kCtlAheadShift;
        pk_out = pk;
        w0_out = w0;
        float pr[NR];
        {
            const float *img = reinterpret_cast<const float *>(&pch[o * NR * 16]);
#pragma unroll
            for (int v = 0; v < NR; ++v)
                pr[v] = img[v * 64 + lane];
        }
        ex = 0;
        static_for<NC>([&](auto Jc) {
            constexpr int j = decltype(Jc)::value;
            if (j >= nch)
                return;
            const int kind = (kinds >> (2 * j)) & 3;
            float4 x[CL];
            if (kind == kChildTip) {
                const uint32_t mk = (uint32_t)hxe[j] >> (8 * (lane & 3));
                const float4 t = make_float4((float)(mk & 1u), (float)((mk >> 1) & 1u),
                                             (float)((mk >> 2) & 1u), (float)((mk >> 3) & 1u));
#pragma unroll
                for (int m = 0; m < CL; ++m)
                    x[m] = t;
            } else if ((fwd >> j) & 1) {
#pragma unroll
                for (int m = 0; m < CL; ++m)
                    x[m] = yl[m];
                ex += el;
            } else if ((fst >> j) & 1) {
                const int p = op_pos(pk, j) & (D - 1);
#pragma unroll
                for (int m = 0; m < CL; ++m)
                    x[m] = stk[wv][p][m][lane];
                ex += stke[wv][p][lane];
            } else if ((ahead >> j) & 1) {
#pragma unroll
                for (int m = 0; m < CL; ++m)
                    x[m] = hx[j][m];
                ex += hxe[j];
            } else {
                // The child's slot, from the same words as loads ahead use.
                const uint64_t p = (uint32_t)__builtin_amdgcn_readfirstlane(cw[kCtlAhead + 2 * j]) |
                                   ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane(cw[kCtlAhead + 2 * j + 1])
                                    << 32);
                const auto rc = rsrc_of(reinterpret_cast<const void *>(p), exp_off);
#pragma unroll
                for (int m = 0; m < CL; ++m) {
                    const crux_f32x4 v = bload4(rc, vol[m], 0);
                    x[m] = make_float4(v.x, v.y, v.z, v.w);
                }
                ex += bload1(rsrc_of(reinterpret_cast<const void *>(p + exp_off), (uint32_t)S * 4), voex, 0);
                // Waited for here, so that where this branch joins the ring
                // slot's, only the slot's loads are outstanding and the wait
                // the compiler puts at the join counts them, not these.
                __builtin_amdgcn_s_waitcnt(kWaitVm0);
            }
            float4 v[CL];
            if (kind != kChildIdent) {
                crux_f32x4 d[CL];
                static_for<4>([&](auto Kc) {
                    constexpr int k = decltype(Kc)::value;
                    static_for<CL>([&](auto Mc) {
                        constexpr int m = decltype(Mc)::value;
                        constexpr int q = (j * CL + m) * 4 + k;
                        const float4 xv = x[m];
                        const float xk = k == 0 ? xv.x : k == 1 ? xv.y : k == 2 ? xv.z : xv.w;
                        d[m] = __builtin_amdgcn_mfma_f32_4x4x1f32(
                            pr[q >> 3], xk, k == 0 ? crux_f32x4{0.f, 0.f, 0.f, 0.f} : d[m], 3, q & 7, 0);
                    });
                });
#pragma unroll
                for (int m = 0; m < CL; ++m)
                    v[m] = make_float4(d[m].x, d[m].y, d[m].z, d[m].w);
            } else {
#pragma unroll
                for (int m = 0; m < CL; ++m)
                    v[m] = x[m];
            }
#pragma unroll
            for (int m = 0; m < CL; ++m)
                acc[m] = (j == 0) ? v[m] : mul4(acc[m], v[m]);
        });
    };
    // Rescale (as k_trav4h), store and push.
    auto finish = [&](int o, const float4 (&acc)[CL], int32_t ex, int32_t pk) {
        int32_t mb = 0;
#pragma unroll
        for (int m = 0; m < CL; ++m)
            mb = max(max(mb, max(__float_as_int(acc[m].x), __float_as_int(acc[m].y))),
                     max(__float_as_int(acc[m].z), __float_as_int(acc[m].w)));
        {
            uint32_t a0, a1;
            lane_pair<32>((uint32_t)mb, a0, a1);
            mb = max((int32_t)a0, (int32_t)a1);
        }
        const float mx = __int_as_float(mb);
        int e2 = 0;
        if (mx > 0.0f)
            (void)frexpf(mx, &e2);
        el = ex + e2;
#pragma unroll
        for (int m = 0; m < CL; ++m)
            yl[m] = make_float4(ldexpf(acc[m].x, -e2), ldexpf(acc[m].y, -e2), ldexpf(acc[m].z, -e2),
                                ldexpf(acc[m].w, -e2));
        const int32_t *cw = &cch[o * kCtlWordsPart];
        const uint64_t par = (uint32_t)__builtin_amdgcn_readfirstlane(cw[4]) |
                             ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane(cw[5]) << 32);
        const int pp = op_pos(pk, kOpMaxChild);
        if (pp >= 0) {
#pragma unroll
            for (int m = 0; m < CL; ++m)
                stk[wv][pp][m][lane] = yl[m];
            stke[wv][pp][lane] = el;
        }
        const auto rp = rsrc_of(reinterpret_cast<const char *>(par), exp_off);
#pragma unroll
        for (int m = 0; m < CL; ++m)
            bstore4<kNT>(crux_f32x4{yl[m].x, yl[m].y, yl[m].z, yl[m].w}, rp, vo[m], 0);
        bstore1<kNT>(el, rsrc_of(reinterpret_cast<const char *>(par) + exp_off, (uint32_t)S * 4), voe, 0);
    };

    // Update o with ring slot k = o % R: its children, then the loads of
    // update o + R into the slot it freed, then its result.
    auto step = [&](int o, auto Kc) {
        constexpr int k = decltype(Kc)::value;
        float4 acc[CL];
        int32_t ex = 0;
        int pk = 0, w0 = 0;
        body(o, hx[k], hxe[k], acc, ex, pk, w0);
        load_ahead(o + R, cch, hx[k], hxe[k]);
        finish(o, acc, ex, pk);
    };
    const int nin = nops - 1;
    int o = 0;
    for (; o + R <= nin; o += R)
        static_for<R>([&](auto Kc) { step(o + decltype(Kc)::value, Kc); });
    static_for<R>([&](auto Kc) {
        if (o + decltype(Kc)::value < nin)
            step(o + decltype(Kc)::value, Kc);
    });

    // The last update, from its ring slot nin - o: the root reduction.
    {
        const int ol = nin;
        float4 hz[NC][CL];
        int32_t hze[NC];
        static_for<R>([&](auto Kc) {
            constexpr int k = decltype(Kc)::value;
            if (k == ol - o) {
#pragma unroll
                for (int j = 0; j < NC; ++j) {
#pragma unroll
                    for (int m = 0; m < CL; ++m)
                        hz[j][m] = hx[k][j][m];
                    hze[j] = hxe[k][j];
                }
            }
        });
        float4 acc[CL];
        int32_t ex = 0;
        int pk = 0, w0 = 0;
        body(ol, hz, hze, acc, ex, pk, w0);
        if ((w0 >> 24) & 1) {
            finish(ol, acc, ex, pk);
        } else {
            double L = 0.0;
#pragma unroll
            for (int m = 0; m < CL; ++m) {
                const CRUX_CAS double *w0p = cas(E->piw) + 4 * m;
                double wx = w0p[0], wy = w0p[1], wz = w0p[2], wq = w0p[3];
                if (CL + m < C) {
                    const CRUX_CAS double *w1p = cas(E->piw) + 4 * (CL + m);
                    if (h) {
                        wx = w1p[0];
                        wy = w1p[1];
                        wz = w1p[2];
                        wq = w1p[3];
                    }
                } else if (h) {
                    wx = wy = wz = wq = 0.0;
                }
                L += wx * (double)acc[m].x + wy * (double)acc[m].y + wz * (double)acc[m].z +
                     wq * (double)acc[m].w;
            }
            {
                const uint64_t u = (uint64_t)__double_as_longlong(L);
                uint32_t al, bl, ah, bh;
                lane_pair<32>((uint32_t)u, al, bl);
                lane_pair<32>((uint32_t)(u >> 32), ah, bh);
                L = __longlong_as_double((long long)(((uint64_t)ah << 32) | al)) +
                    __longlong_as_double((long long)(((uint64_t)bh << 32) | bl));
            }
            double v = (log(L) + (double)ex * 0.69314718055994530942) * (double)gas(E->weights)[sc];
            if (__builtin_isnan(v))
                v = -INFINITY;
            if (act && h == 0) {
                gas(E->site_out)[s] = v;
                sum = v;
            }
        }
    }
    double cmp = 0.0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double s2 = __shfl_xor(sum, off, 64), c2 = __shfl_xor(cmp, off, 64);
        if ((lane & off) == 0) {
            nsum_merge(sum, cmp, s2, c2);
        } else {
            double a = s2, bb = c2;
            nsum_merge(a, bb, sum, cmp);
            sum = a;
            cmp = bb;
        }
    }
    if (wv == 1 && lane == 0) {
        red[0] = sum;
        red[1] = cmp;
    }
    lds_barrier();
    if (wv == 0 && lane == 0 && E->partials != nullptr) {
        nsum_merge(sum, cmp, red[0], red[1]);
        gas(E->partials)[2 * tile] = sum;
        gas(E->partials)[2 * tile + 1] = cmp;
    }
}

// ---------------------------------------------------------------------------
// 20/22-state traversal.  A block is one wave covering CM mixture components
// of G = 64 / CM consecutive sites: lane l holds component l / G of site
// l % G, all KP states of it in registers (states >= K stay 0).  Like the
// 4-state kernels, the wave walks the whole plan for its sites with no
// barrier; results of an update are stored to their HBM slot and handed to
// the consumer in registers (the update just before it) or through the
// wave's LDS stack, and the root update reduces to the pattern-weighted FP64
// site lnL (NaN -> -inf, CxLik.c:598-632) and a compensated sum per wave.  A
// slot holds the wave's registers as they are, [S / G][KP / 4][64 lanes][4]
// (lanes of components >= C unused), so every store and load of a CLV moves
// 1 KiB contiguously.  Arithmetic follows the generic stripe
// CxLikExecuteStripe (CxLik.c:635-846) with FP32 storage and power-of-two
// scaling.
//
// An internal child is y = P_c x on the matrix core: v_mfma_f32_4x4x1f32
// with lane l supplying B = its own x[col] and, broadcast from block ABID of
// its component's lane group, A = P_c[4 I + i][col] for the group's four
// rows i.  One MFMA per (row block I, column) gives each lane y[4 I .. 4 I +
// 3] of its own (site, component); the sum is the in-order FMA chain over
// the columns.  The A operands come from a P image laid out per lane by
// k_pmat, one float per register, so an update's P for all components is a
// few coalesced float4 loads per lane.  A tip child is a row of the tip
// table (P_c times the leaf vector of the site's code; the codes of up to
// kTipChunk tips are staged in LDS).
//
// A wave walks its plan alone on its SIMD more often than not (a launch has
// about one wave per SIMD at 20k patterns), so an update costs the wave's
// instruction count plus what it waits for; the structure keeps both small:
// * Control words are host-built records whose buffer descriptors load
//   straight into scalar registers (s_load; no per-lane decoding).  The
//   records of updates o + 1 and o + 2 are loaded when update o ends, once
//   every LDS read of update o has returned (scalar loads return out of
//   order and share a counter with LDS reads, so any wait for one of those
//   waits for them too), and waited for when update o + 1 starts.
// * Updates with two children and a parent (all but the root, chained
//   multifurcations, identity children and results re-read after the stack
//   overflowed) run a body specialised for their children's sources (tip,
//   forwarded, stack, cached in HBM by an earlier plan), named by the
//   update's shape code; other updates take a generic walk that loads what
//   it needs where it needs it.
// * What a child needs from memory is loaded one update ahead: a tip child's
//   tip row, a forwarded or stack child's P image, a cached child's CLV (its
//   P image is loaded in the update itself, from L2).  The loads are issued
//   at one place after the current update's products, the same loads for
//   every kind of child with only descriptor and offsets chosen by mask
//   arithmetic: loads issued on several paths, or inside the bodies, leave
//   the register allocator copying their targets where the paths join, and
//   a copy waits for its load.  The next update's tip codes are read from
//   LDS when the current update starts.
// * The plan's last update (its root) runs after the loop, and nothing
//   loaded before the loop stays outstanding into it, so that every way into
//   a wait in the loop has the same stores after the loads it waits for (a
//   wait count is the least over all ways in, and one counter tracks loads
//   and stores in order).
//
// Idle lanes (components >= C) hold zeros throughout: their P image rows and
// tip rows are zero, and their accesses to CLV slots use an offset outside
// every buffer range (loads read zero, stores write nothing).
//
// WV: waves per SIMD the register allocation is held to (1: up to 512
// registers per lane, 2: up to 256).
template <int KP, int CM, int D, int PREC, int WV>
__global__ void __launch_bounds__(64, WV)
k_trav20(const Entry *__restrict__ entries, const int32_t *__restrict__ block_entry, int blk0)
{
    using rsrc_t = __amdgpu_buffer_rsrc_t;
    constexpr int G = sites20(CM);                  // sites per wave
    constexpr int W = img_cols(CM);                 // P columns per image register
    constexpr int CBSZ = W == 16 ? 4 : W == 8 ? 3 : W == 4 ? 2 : W == 2 ? 1 : 0;
    constexpr int NI = KP / 4;                      // row blocks = float4 per CLV row
    constexpr int NG = img_groups(KP, CM);
    constexpr int NG16 = img16_groups(KP, CM);      // 16-bit images: column groups
    constexpr int NV16P = img16_nvp(KP, CM);        // dwords of one 16-bit image, padded
    // float4 per lane of a prefetched P image (the split precision's low
    // image is loaded where it is used, so that its registers stay within
    // two waves per SIMD).
    // Dwords per lane of an FP32 image's compact last block (0: the last
    // block is a padded float4 block like the others).
    constexpr int RT = PREC == kPrecF32 ? img_tail(KP, CM) : 0;
    constexpr int NQ = PREC == kPrecF16x3 ? NV16P / 4
                       : PREC == kPrecF32 ? (img_nv(KP, CM) + 3) / 4
                                          : img_words(KP, CM, PREC) / 256;
    constexpr int PQ = NQ > NI ? NQ : NI;           // float4 of a prefetched child
    constexpr int WPR = G / 4;                      // code dwords per staged tip row
    constexpr int RPI = 64 / WPR;                   // tip rows per staging load
    // Byte offset no buffer range covers (the largest offset added to it
    // stays below 2^32).
    constexpr uint32_t kOut = 0x80000000u;
    static_assert((1 << CBSZ) == W, "component lane groups are whole matrix blocks");
    __shared__ crux_f32x4 stk[D][NI][64];
    __shared__ int32_t stke[D][64];
    __shared__ uint32_t tipw[kTipChunk][WPR];

    const int b = blk0 + (int)blockIdx.x;
    const CRUX_CAS Entry *E = cas(entries) + cas(block_entry)[b];
    const int S = E->S, C = E->C, K = E->K, nops = E->nops, ntips = E->ntips;
    const uint32_t exp_off = (uint32_t)E->exp_off;
    const int tile = b - E->tile0;
    const int lane = (int)threadIdx.x;
    const int c = lane / G, t = lane % G;
    const int s = tile * G + t;
    const bool site_ok = s < S;
    const int sc = site_ok ? s : S - 1;     // in-bounds site for loads of idle lanes
    const uint32_t lane16 = (uint32_t)lane * 16;
    // Offsets within a CLV slot: the lane's first float4 (the others follow
    // 1 KiB apart; kOut for idle components); its site's exponent, stored
    // from component 0 of sites in range only.
    const uint32_t vcl = c < C ? (uint32_t)tile * NI * 1024 + lane16 : kOut;
    const uint32_t vexs = (c == 0 && site_ok) ? exp_off + (uint32_t)s * 4 : kOut;
    const uint32_t vexl = exp_off + (uint32_t)sc * 4;
    // A tip row of the lane's component: code * CM * KP floats further.
    const uint32_t vtip = (uint32_t)c * KP * 4;
    const CRUX_CAS uint32_t *ctlw = (const CRUX_CAS uint32_t *)cas(E->ctl);
    const uint8_t *tipb = reinterpret_cast<const uint8_t *>(&tipw[0][0]);
    double sum = 0.0;

    // One update's control record (scalar registers): flags, positions,
    // tip_load, has_parent, and of each buffer descriptor its base and range
    // (the descriptor is rebuilt from them where it is used, so that a
    // choice between two of them is two scalar selects).
    struct Rec {
        uint32_t w[kCtl20Slot + 8];
        __device__ uint32_t w0() const { return w[0]; }
        __device__ uint32_t pk() const { return w[1]; }
        __device__ uint32_t tl() const { return w[2]; }
        __device__ uint32_t hp() const { return w[3]; }
        // Base and range words of the parent's, child j's data and child j's
        // slot descriptors.
        __device__ const uint32_t (&par() const)[3] { return *reinterpret_cast<const uint32_t(*)[3]>(w + kCtl20Parent); }
        __device__ const uint32_t (&dat(int j) const)[3] { return *reinterpret_cast<const uint32_t(*)[3]>(w + kCtl20Data + 4 * j); }
        __device__ const uint32_t (&slot(int j) const)[3] { return *reinterpret_cast<const uint32_t(*)[3]>(w + kCtl20Slot + 4 * j); }
    };
    // Words 0 .. kCtl20Slot + 7 of update o's record, in two wide loads.
    auto rec = [&](int o) -> Rec {
        const CRUX_CAS uint32_t *w = ctlw + (size_t)o * kCtlWords20;
        Rec r;
#pragma unroll
        for (int k = 0; k < kCtl20Slot + 8; ++k)
            r.w[k] = w[k];
        return r;
    };
    auto mk = [](const uint32_t (&v)[3]) -> rsrc_t {
        const uint64_t base = (uint64_t)v[0] | ((uint64_t)v[1] << 32);
        return __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void *>(base), (short)0, (int)v[2],
                                                 (int)kVsharpFlags);
    };
    auto dat_of = [&](int o, int j) -> rsrc_t {
        return ((const CRUX_CAS rsrc_t *)(ctlw + (size_t)o * kCtlWords20))[kCtl20Data / 4 + j];
    };
    auto slot_of = [&](int o, int j) -> rsrc_t {
        return ((const CRUX_CAS rsrc_t *)(ctlw + (size_t)o * kCtlWords20))[kCtl20Slot / 4 + j];
    };
    auto shape_of = [](uint32_t w0) -> int { return (int)((w0 >> 25) & 31); };
    auto src_of = [](int sh, int j) -> int { return j == 0 ? (sh - 1) >> 2 : (sh - 1) & 3; };

    // Codes of tips [base, base + kTipChunk) of the entry's list, for this
    // wave's sites (rows are padded to whole 64-site tiles).
    auto stage_tips = [&](int base) {
        const int n = min(kTipChunk, ntips - base);
        const int sub = lane / WPR, w = lane % WPR;
        uint32_t v[kTipChunk / RPI];
#pragma unroll
        for (int i = 0; i < kTipChunk / RPI; ++i) {
            const int r = i * RPI + sub;
            v[i] = 0u;
            if (r < n) {
                const uint8_t *row = gas(E->tip_rows)[base + r];
                v[i] = *(const CRUX_GAS uint32_t *)gas(row + tile * G + 4 * w);
            }
        }
#pragma unroll
        for (int i = 0; i < kTipChunk / RPI; ++i) {
            const int r = i * RPI + sub;
            if (r < n)
                tipw[r][w] = v[i];
        }
        __builtin_amdgcn_s_waitcnt(0);
    };
    // Site code of the tip in row pos of the staged chunk (child j's byte of
    // the packed positions).
    auto code_of = [&](uint32_t pk, int j) -> uint32_t {
        return tipb[((pk >> (8 * j)) & (kTipChunk - 1)) * G + t];
    };

    // Loads: a P image (PQ float4, a wave apart), a tip row (NI consecutive
    // float4 of the lane's (code, component) row), a CLV (NI float4 1 KiB
    // apart) and its exponent.
    // The image's compact last block: RT dwords of the lane at voff (an
    // offset with kOut set reads nothing).
    auto load_tail = [&](rsrc_t r, uint32_t voff) -> crux_f32x4 {
        crux_f32x4 t = {0.f, 0.f, 0.f, 0.f};
        constexpr int soff = (NQ - 1) * 1024;
        if constexpr (RT == 1) {
            t[0] = __int_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, soff, 0));
        } else if constexpr (RT == 2) {
            const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, soff, 0);
            t[0] = __uint_as_float(v[0]);
            t[1] = __uint_as_float(v[1]);
        } else if constexpr (RT == 3) {
            const auto v = __builtin_amdgcn_raw_buffer_load_b96(r, (int)voff, soff, 0);
            t[0] = __uint_as_float(v[0]);
            t[1] = __uint_as_float(v[1]);
            t[2] = __uint_as_float(v[2]);
        }
        return t;
    };
    auto load_img = [&](rsrc_t r, crux_f32x4 (&pr)[PQ]) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            if (RT != 0 && q == NQ - 1)
                pr[q] = load_tail(r, (uint32_t)lane * RT * 4);
            else
                pr[q] = bload4(r, lane16, q * 1024);
        }
    };
    auto load_tip = [&](rsrc_t r, uint32_t code, crux_f32x4 (&pr)[PQ]) {
        const uint32_t voff = code * (CM * KP * 4) + vtip;
#pragma unroll
        for (int q = 0; q < NI; ++q)
            pr[q] = bload4(r, voff, q * 16);
    };
    auto load_clv = [&](rsrc_t r, crux_f32x4 (&x)[NI], int32_t &e) {
#pragma unroll
        for (int q = 0; q < NI; ++q)
            x[q] = bload4(r, vcl, q * 1024);
        e = bload1(r, vexl, 0);
    };

    // 16-bit products: four columns per instruction, the lane's own four
    // states as B and, broadcast from block ABID of its component group,
    // four columns of four P rows as A.
    typedef short crux_s16x4 __attribute__((ext_vector_type(4)));
    typedef _Float16 crux_f16x4 __attribute__((ext_vector_type(4)));
    typedef unsigned int crux_u32x2 __attribute__((ext_vector_type(2)));
    auto dword = [](const crux_f32x4 (&img)[PQ], int v) -> uint32_t {
        return __float_as_uint(img[v >> 2][v & 3]);
    };
    auto bf16x4 = [](crux_f32x4 v) -> crux_s16x4 {
        crux_s16x4 r;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const uint32_t u = __float_as_uint(v[e]);
            r[e] = (short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
        }
        return r;
    };
    auto pmul16 = [&](const crux_f32x4 (&img)[PQ], const crux_f32x4 (&x)[NI], crux_f32x4 (&y)[NI],
                      rsrc_t mat) {
        crux_f32x4 lo[PREC == kPrecF16x3 ? NV16P / 4 : 1];
        if constexpr (PREC == kPrecF16x3) {
#pragma unroll
            for (int q = 0; q < NV16P / 4; ++q)
                lo[q] = bload4(mat, lane16, (uint32_t)(NV16P / 4 + q) * 1024);
        }
        auto loword = [&](int v) -> uint32_t { return __float_as_uint(lo[v >> 2][v & 3]); };
#pragma unroll
        for (int I = 0; I < NI; ++I)
            y[I] = crux_f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int g = 0; g < NG16; ++g) {
            static_for<W>([&](auto Ac) {
                constexpr int a = decltype(Ac)::value;
                const int q = g * W + a;             // float4 of states (4 columns)
                if (q < NI) {
                    const crux_f32x4 xv = x[q];
                    if constexpr (PREC == kPrecBF16) {
                        const crux_s16x4 bx = bf16x4(xv);
#pragma unroll
                        for (int I = 0; I < NI; ++I) {
                            const int v = (I * NG16 + g) * 2;
                            const crux_s16x4 ap = __builtin_bit_cast(
                                crux_s16x4, crux_u32x2{dword(img, v), dword(img, v + 1)});
                            y[I] = __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(ap, bx, y[I], CBSZ, a, 0);
                        }
                    } else {
                        const crux_f16x4 hx = __builtin_convertvector(xv, crux_f16x4);
                        crux_f16x4 lx;
                        if constexpr (PREC == kPrecF16x3)
                            lx = __builtin_convertvector(xv - __builtin_convertvector(hx, crux_f32x4), crux_f16x4);
#pragma unroll
                        for (int I = 0; I < NI; ++I) {
                            const int v = (I * NG16 + g) * 2;
                            const crux_f16x4 ah = __builtin_bit_cast(
                                crux_f16x4, crux_u32x2{dword(img, v), dword(img, v + 1)});
                            if constexpr (PREC == kPrecF16x3) {
                                const crux_f16x4 al = __builtin_bit_cast(
                                    crux_f16x4, crux_u32x2{loword(v), loword(v + 1)});
                                // The small terms first, then the main one.
                                y[I] = __builtin_amdgcn_mfma_f32_4x4x4f16(al, hx, y[I], CBSZ, a, 0);
                                y[I] = __builtin_amdgcn_mfma_f32_4x4x4f16(ah, lx, y[I], CBSZ, a, 0);
                            }
                            y[I] = __builtin_amdgcn_mfma_f32_4x4x4f16(ah, hx, y[I], CBSZ, a, 0);
                        }
                    }
                }
            });
        }
    };

    // y = P_c x for every lane's own (site, component); mat: the P image's
    // descriptor (the split precision's low image is read from it).
    auto pmul = [&](const crux_f32x4 (&img)[PQ], const crux_f32x4 (&x)[NI], crux_f32x4 (&y)[NI],
                    rsrc_t mat) {
        if constexpr (PREC != kPrecF32) {
            pmul16(img, x, y, mat);
        } else {
            (void)mat;
#pragma unroll
            for (int I = 0; I < NI; ++I)
                y[I] = crux_f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                static_for<W>([&](auto A) {
                    constexpr int a = decltype(A)::value;
                    const int col = g * W + a;
                    if (col < KP) {
#pragma unroll
                        for (int I = 0; I < NI; ++I) {
                            const int v = I * NG + g;
                            y[I] = __builtin_amdgcn_mfma_f32_4x4x1f32(img[v >> 2][v & 3], x[col >> 2][col & 3],
                                                                      y[I], CBSZ, a, 0);
                        }
                    }
                });
            }
        }
    };

    // What the current update's children 0 and 1 took from memory ahead of
    // it (shape updates): P image, tip row, or CLV and exponent.
    crux_f32x4 pre[2][PQ];
    int32_t he[2] = {0, 0};
    // The previous update's scaled result and exponent (forwarded children).
    crux_f32x4 yl[NI];
    int32_t el = 0;
#pragma unroll
    for (int I = 0; I < NI; ++I)
        yl[I] = crux_f32x4{0.f, 0.f, 0.f, 0.f};

    // Child j of the next update (record rn, shape nsh != 0, tip code cd):
    // the same PQ float4 and exponent loads whatever the child is, only the
    // descriptor and offsets differ (so every way through writes the same
    // registers, and none of them is copied, which would wait for the
    // loads).
    auto fetch_next = [&](const Rec &rn, int nsh, uint32_t cd, int j) {
        const int src = src_of(nsh, j);
        const bool tip = src == kSrcTip, hbm = src == kSrcHbm;
        uint32_t v[3];
#pragma unroll
        for (int k = 0; k < 3; ++k)
            v[k] = hbm ? rn.slot(j)[k] : rn.dat(j)[k];
        const rsrc_t r = mk(v);
        // Selects by mask arithmetic, which the compiler keeps as such (a
        // select it turned into a branch would split the update's body).
        const uint32_t mtip = tip ? ~0u : 0u, mhbm = hbm ? ~0u : 0u;
        const uint32_t vimg = lane16 ^ ((lane16 ^ vcl) & mhbm);
        const uint32_t vrow = cd * (CM * KP * 4) + vtip;
        const uint32_t voff = vimg ^ ((vimg ^ vrow) & mtip);
        // Float4 past NI of a tip row or CLV: an offset outside the range
        // (no memory request).
        const uint32_t voff_hi = voff | ((mtip | mhbm) & kOut);
        const int sh = tip ? 4 : 10;
#pragma unroll
        for (int q = 0; q < PQ; ++q) {
            // An image's compact last block (past a tip row's or CLV's
            // float4s, where those read nothing).
            if (RT != 0 && q == NQ - 1)
                pre[j][q] = load_tail(r, ((uint32_t)lane * RT * 4) | ((mtip | mhbm) & kOut));
            else
                pre[j][q] = bload4(r, q < NI ? voff : voff_hi, (uint32_t)q << sh);
        }
        // The exponent of a cached child (an empty range otherwise).
        uint32_t ve[3] = {rn.slot(j)[0], rn.slot(j)[1], hbm ? rn.slot(j)[2] : 0u};
        he[j] = bload1(mk(ve), vexl, 0);
    };

    // Generic walk of update o (any child count and kind): every input is
    // loaded where it is used; cdc: the site codes of its tip children.
    auto generic = [&](int o, const Rec &r, const uint32_t (&cdc)[kOpMaxChild], crux_f32x4 (&acc)[NI],
                       int32_t &ex) {
        const int w0 = (int)r.w0();
        const int nch = w0 & 0xff, kinds = (w0 >> 8) & 0xff, fst = (w0 >> 16) & 0xf,
                  fwd = (w0 >> 20) & 0xf;
        static_for<kOpMaxChild>([&](auto J) {
            constexpr int j = decltype(J)::value;
            if (j >= nch)
                return;
            const int kind = (kinds >> (2 * j)) & 3;
            const rsrc_t mat = dat_of(o, j);
            crux_f32x4 y[NI];
            crux_f32x4 pr[PQ];
            if (kind == kChildTip) {
                load_tip(mat, cdc[j], pr);
#pragma unroll
                for (int I = 0; I < NI; ++I)
                    y[I] = pr[I];
            } else {
                crux_f32x4 x[NI];
                if ((fwd >> j) & 1) {
#pragma unroll
                    for (int I = 0; I < NI; ++I)
                        x[I] = yl[I];
                    ex += el;
                } else if ((fst >> j) & 1) {
                    const int p = op_pos((int32_t)r.pk(), j) & (D - 1);
#pragma unroll
                    for (int I = 0; I < NI; ++I)
                        x[I] = stk[p][I][lane];
                    ex += stke[p][lane];
                } else {
                    int32_t e = 0;
                    load_clv(slot_of(o, j), x, e);
                    ex += e;
                }
                if (kind == kChildIdent) {
                    // Used as is (idle components hold zeros).
#pragma unroll
                    for (int I = 0; I < NI; ++I)
                        y[I] = x[I];
                } else {
                    load_img(mat, pr);
                    pmul(pr, x, y, mat);
                }
            }
#pragma unroll
            for (int I = 0; I < NI; ++I)
                acc[I] = j == 0 ? y[I] : acc[I] * y[I];
        });
    };

    // Site maximum over the lane's states, then over the site's components
    // (lanes G apart); all values are >= 0, so the largest signed integer is
    // the largest float.  The result is scaled into [0.5, 1) by an exact
    // power of two, its exponent added to the children's.
    auto rescale = [&](const crux_f32x4 (&acc)[NI], int32_t ex) {
        int32_t mb = 0;
#pragma unroll
        for (int I = 0; I < NI; ++I)
            mb = max(max(mb, max(__float_as_int(acc[I].x), __float_as_int(acc[I].y))),
                     max(__float_as_int(acc[I].z), __float_as_int(acc[I].w)));
        mb = comp_max<G>(mb);
        const float m = __int_as_float(mb);
        int e2 = 0;
        if (m > 0.0f)
            (void)frexpf(m, &e2);
        el = ex + e2;
#pragma unroll
        for (int I = 0; I < NI; ++I)
            yl[I] = crux_f32x4{ldexpf(acc[I].x, -e2), ldexpf(acc[I].y, -e2), ldexpf(acc[I].z, -e2),
                               ldexpf(acc[I].w, -e2)};
    };
    // The result to the parent's slot (one float4 of every lane per store:
    // 1 KiB contiguous) and, if the plan says so, onto the stack.
    auto store = [&](rsrc_t par, uint32_t pk) {
#pragma unroll
        for (int I = 0; I < NI; ++I)
            bstore4<kNT>(yl[I], par, vcl, I * 1024);
        bstore1<kNT>(el, par, vexs, 0);
        const int pp = op_pos((int32_t)pk, kOpMaxChild);
        if (pp >= 0) {
#pragma unroll
            for (int I = 0; I < NI; ++I)
                stk[pp][I][lane] = yl[I];
            stke[pp][lane] = el;
        }
    };

    // Two-child update with a parent, children from sources s0 and s1: the
    // forwarded child's product first (its input is ready), then the stack's
    // (its LDS reads were issued at the start), then a cached one (its P
    // image is loaded at the start); a tip child is its row.  acc = y0 y1 as
    // in the generic walk.
    auto two = [&](auto S0c, auto S1c, const Rec &r, crux_f32x4 (&acc)[NI], int32_t &ex) {
        constexpr int s0 = decltype(S0c)::value, s1 = decltype(S1c)::value;
        crux_f32x4 xs[2][NI];
        int32_t es[2] = {0, 0};
        crux_f32x4 hi[2][PQ];
        static_for<2>([&](auto Jc) {
            constexpr int j = decltype(Jc)::value;
            constexpr int sj = j == 0 ? s0 : s1;
            if constexpr (sj == kSrcStk) {
                const int p = op_pos((int32_t)r.pk(), j) & (D - 1);
#pragma unroll
                for (int I = 0; I < NI; ++I)
                    xs[j][I] = stk[p][I][lane];
                es[j] = stke[p][lane];
            } else if constexpr (sj == kSrcHbm) {
                // The cached child's CLV came ahead; its P image now.
                load_img(mk(r.dat(j)), hi[j]);
#pragma unroll
                for (int I = 0; I < NI; ++I)
                    xs[j][I] = pre[j][I];
                es[j] = he[j];
            }
        });
        crux_f32x4 y[2][NI];
        auto child = [&](auto Jc) {
            constexpr int j = decltype(Jc)::value;
            constexpr int sj = j == 0 ? s0 : s1;
            if constexpr (sj == kSrcTip) {
#pragma unroll
                for (int I = 0; I < NI; ++I)
                    y[j][I] = pre[j][I];
            } else if constexpr (sj == kSrcFwd) {
                pmul(pre[j], yl, y[j], mk(r.dat(j)));
                ex += el;
            } else if constexpr (sj == kSrcStk) {
                pmul(pre[j], xs[j], y[j], mk(r.dat(j)));
                ex += es[j];
            } else {
                pmul(hi[j], xs[j], y[j], mk(r.dat(j)));
                ex += es[j];
            }
        };
        static_for<4>([&](auto Rc) {
            // Source order: forwarded, tip, stack, cached.
            constexpr int q = decltype(Rc)::value;
            constexpr int want = q == 0 ? kSrcFwd : q == 1 ? kSrcTip : q == 2 ? kSrcStk : kSrcHbm;
            if constexpr (s0 == want)
                child(std::integral_constant<int, 0>{});
            if constexpr (s1 == want)
                child(std::integral_constant<int, 1>{});
        });
#pragma unroll
        for (int I = 0; I < NI; ++I)
            acc[I] = y[0][I] * y[1][I];
    };
    using Tip = std::integral_constant<int, kSrcTip>;
    using Fwd = std::integral_constant<int, kSrcFwd>;
    using Stk = std::integral_constant<int, kSrcStk>;
    using Hbm = std::integral_constant<int, kSrcHbm>;
    auto two_of = [&](auto S0c, int s1, const Rec &r, crux_f32x4 (&acc)[NI], int32_t &ex) {
        if (s1 == kSrcTip)
            two(S0c, Tip{}, r, acc, ex);
        else if (s1 == kSrcStk)
            two(S0c, Stk{}, r, acc, ex);
        else if (s1 == kSrcHbm)
            two(S0c, Hbm{}, r, acc, ex);
        else if constexpr (decltype(S0c)::value != kSrcFwd)
            two(S0c, Fwd{}, r, acc, ex);
    };

    Rec r0 = rec(0);
    Rec r1 = rec(min(1, nops - 1));
    if (ntips > 0)
        stage_tips(E->tip0);
    if (nops > 1) {
        const int sh = shape_of(r0.w0());
        if (sh != 0) {
            const uint32_t cd[2] = {code_of(r0.pk(), 0), code_of(r0.pk(), 1)};
            fetch_next(r0, sh, cd[0], 0);
            fetch_next(r0, sh, cd[1], 1);
        }
    }
    // Nothing loaded before the loop stays outstanding into it (see above).
    __builtin_amdgcn_s_waitcnt(0);

    // Every update but the last has a parent (the last is the plan's root
    // update, or the one update of a one-update entry).
    for (int o = 0; o + 1 < nops; ++o) {
        const int shape = shape_of(r0.w0());
        // A generic update reads its tip codes before the next update's tips
        // may replace the staged chunk (a shape update's tip rows are in
        // registers already).
        uint32_t cdc[kOpMaxChild] = {0u, 0u, 0u};
        if (shape == 0) {
#pragma unroll
            for (int j = 0; j < kOpMaxChild; ++j)
                cdc[j] = code_of(r0.pk(), j);
        }
        // The next update: staged tips (its tip rows may start a new chunk),
        // shape (0 unless it is a shape update ahead of the last) and tip
        // codes.
        if ((int)r1.tl() > 0)
            stage_tips((int)r1.tl());
        const int nsh = o + 2 < nops ? shape_of(r1.w0()) : 0;
        const uint32_t cd[2] = {code_of(r1.pk(), 0), code_of(r1.pk(), 1)};
        crux_f32x4 acc[NI];
        int32_t ex = 0;
        if (shape != 0) {
            const int s0 = src_of(shape, 0), s1 = src_of(shape, 1);
            if (s0 == kSrcTip)
                two_of(Tip{}, s1, r0, acc, ex);
            else if (s0 == kSrcFwd)
                two_of(Fwd{}, s1, r0, acc, ex);
            else if (s0 == kSrcStk)
                two_of(Stk{}, s1, r0, acc, ex);
            else
                two_of(Hbm{}, s1, r0, acc, ex);
        } else {
            generic(o, r0, cdc, acc, ex);
        }
        // The next update's children, from one place for every kind of
        // update (the registers then need no copying where the paths join).
        if (nsh != 0) {
            fetch_next(r1, nsh, cd[0], 0);
            fetch_next(r1, nsh, cd[1], 1);
        }
        // Every LDS read of this update has returned by now; then the record
        // after next, waited for when the next update starts (nothing from
        // here to there waits for a scalar or LDS load, and scalar loads
        // return out of order, so any such wait would wait for it too).
        // Both records are loaded again (the next one from the scalar cache)
        // rather than moved: a register copy would wait for the load.
        const rsrc_t par = mk(r0.par());
        const uint32_t pk = r0.pk();
        __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
        r0 = rec(o + 1);
        r1 = rec(min(o + 2, nops - 1));
        rescale(acc, ex);
        store(par, pk);
    }

    // The last update: the root reduction (or one update with a parent).
    {
        crux_f32x4 acc[NI];
        int32_t ex = 0;
        uint32_t cdc[kOpMaxChild];
#pragma unroll
        for (int j = 0; j < kOpMaxChild; ++j)
            cdc[j] = code_of(r0.pk(), j);
        generic(nops - 1, r0, cdc, acc, ex);
        if ((r0.hp() & 1) != 0) {
            rescale(acc, ex);
            store(mk(r0.par()), r0.pk());
        } else {
            // Weights of the lane's component (all loads issued at once;
            // idle components read component 0's and add nothing).
            const CRUX_GAS double *pw = gas(E->piw) + (size_t)(c < C ? c : 0) * K;
            double L = 0.0;
#pragma unroll
            for (int i = 0; i < KP; ++i) {
                const double w = pw[i < K ? i : K - 1];
                L += (i < K && c < C ? w : 0.0) * (double)acc[i >> 2][i & 3];
            }
            // Sum over the site's components; every lane ends with the same
            // value (each pairwise sum is commutative).
            L = comp_sum<G>(L);
            double v = (log(L) + (double)ex * 0.69314718055994530942) *
                       (double)gas(E->weights)[sc];
            if (__builtin_isnan(v))
                v = -INFINITY;
            if (site_ok && c == 0) {
                gas(E->site_out)[s] = v;
                sum = v;
            }
        }
    }
    double cmp = 0.0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double s2 = __shfl_xor(sum, off, 64), c2 = __shfl_xor(cmp, off, 64);
        if ((lane & off) == 0) {
            nsum_merge(sum, cmp, s2, c2);
        } else {
            double a = s2, bb = c2;
            nsum_merge(a, bb, sum, cmp);
            sum = a;
            cmp = bb;
        }
    }
    // An entry that computes part of a plan for a later launch has no
    // partial sums.
    if (lane == 0 && E->partials != nullptr) {
        gas(E->partials)[2 * tile] = sum;
        gas(E->partials)[2 * tile + 1] = cmp;
    }
}

// ---------------------------------------------------------------------------
// Launchers.

// Launches take optional start and stop events (null: none), which the
// runtime stamps when the kernel itself starts and ends.

template <int CM, int D, int PC>
static void launch_trav4(const Entry *entries, const int32_t *block_entry, int blk0,
                         int nblocks, hipStream_t stream, hipEvent_t t0, hipEvent_t t1)
{
    if constexpr (CM == 1)
        hipExtLaunchKernelGGL((k_trav4<CM, D, PC>), dim3(nblocks), dim3(kTile4), 0, stream, t0, t1,
                              0, entries, block_entry, blk0);
    else
        hipExtLaunchKernelGGL((k_trav4h<CM, D, PC>), dim3(nblocks), dim3(2 * kTile4), 0, stream,
                              t0, t1, 0, entries, block_entry, blk0);
}

template <int CM, int D>
static void launch_trav4_chunk(const Entry *entries, const int32_t *block_entry, int blk0,
                               int nblocks, int chunk, hipStream_t stream, hipEvent_t t0,
                               hipEvent_t t1)
{
    if constexpr (D > stack_depth4_max(CM)) {
        return;
    } else {
        if (chunk == 0)
            launch_trav4<CM, D, trav4_chunk(CM, D, 0)>(entries, block_entry, blk0, nblocks, stream, t0, t1);
        else if (chunk == 2)
            launch_trav4<CM, D, trav4_chunk(CM, D, 2)>(entries, block_entry, blk0, nblocks, stream, t0, t1);
        else
            launch_trav4<CM, D, trav4_chunk(CM, D, 1)>(entries, block_entry, blk0, nblocks, stream, t0, t1);
    }
}

template <int KP, int CM, int D, int PREC, int WV>
static void launch_trav20(const Entry *entries, const int32_t *block_entry, int blk0,
                          int nblocks, hipStream_t stream, hipEvent_t t0, hipEvent_t t1)
{
    hipExtLaunchKernelGGL((k_trav20<KP, CM, D, PREC, WV>), dim3(nblocks), dim3(64), 0, stream, t0,
                          t1, 0, entries, block_entry, blk0);
}

template <int KP, int CM, int D>
static void launch_trav20_waves(const Entry *entries, const int32_t *block_entry, int blk0,
                                int nblocks, int waves, hipStream_t stream, hipEvent_t t0,
                                hipEvent_t t1)
{
    if constexpr (D < 8) {
        if (waves == 2) {
            launch_trav20<KP, CM, D, kPrecF32, 2>(entries, block_entry, blk0, nblocks, stream, t0, t1);
            return;
        }
    }
    launch_trav20<KP, CM, D, kPrecF32, 1>(entries, block_entry, blk0, nblocks, stream, t0, t1);
}

// The traversal launch of one (K, CM) class at a launch shape (depth, chunk
// and waves already resolved to instantiated values).
template <int K, int CM>
void launch_class(const Entry *entries, const int32_t *block_entry, int blk0, int nblocks,
                  const TravShape &sh, hipStream_t stream, hipEvent_t t0, hipEvent_t t1)
{
    // Stack depth: more slots keep more results in LDS, fewer let more
    // blocks share a CU.
    if constexpr (K == 4) {
        if constexpr (CM == 2 || CM == 4) {
            if (sh.part) {
                hipExtLaunchKernelGGL((k_part4h<CM, kPartDepth, kPartAhead>), dim3(nblocks), dim3(2 * kTile4),
                                      0, stream, t0, t1, 0, entries, block_entry, blk0);
                return;
            }
        }
        if (sh.depth <= 2)
            launch_trav4_chunk<CM, 2>(entries, block_entry, blk0, nblocks, sh.chunk, stream, t0, t1);
        else if (sh.depth <= 4)
            launch_trav4_chunk<CM, 4>(entries, block_entry, blk0, nblocks, sh.chunk, stream, t0, t1);
        else
            launch_trav4_chunk<CM, 8>(entries, block_entry, blk0, nblocks, sh.chunk, stream, t0, t1);
    } else {
        constexpr int KP = kpad(K);
        constexpr int WD = trav20_waves(KP, CM, 4, 0);
        // 16-bit inputs are instantiated at the default stack depth and
        // waves only.
        if (sh.prec == kPrecBF16)
            launch_trav20<KP, CM, 4, kPrecBF16, WD>(entries, block_entry, blk0, nblocks, stream, t0, t1);
        else if (sh.prec == kPrecF16)
            launch_trav20<KP, CM, 4, kPrecF16, WD>(entries, block_entry, blk0, nblocks, stream, t0, t1);
        else if (sh.prec == kPrecF16x3)
            launch_trav20<KP, CM, 4, kPrecF16x3, WD>(entries, block_entry, blk0, nblocks, stream, t0, t1);
        else if (sh.depth <= 2)
            launch_trav20_waves<KP, CM, 2>(entries, block_entry, blk0, nblocks, sh.waves, stream, t0, t1);
        else if (sh.depth <= 4)
            launch_trav20_waves<KP, CM, 4>(entries, block_entry, blk0, nblocks, sh.waves, stream, t0, t1);
        else
            launch_trav20_waves<KP, CM, 8>(entries, block_entry, blk0, nblocks, sh.waves, stream, t0, t1);
    }
}

// The instantiated classes.  Built in parts (CRUX_PART_K / CRUX_PART_CM name
// the one class a compilation of this file instantiates, CRUX_PART_K 0 the
// P-matrix kernel and the dispatch), so that the classes compile in
// parallel; without CRUX_PART_K the file holds everything.
#define CRUX_CLASSES(X)                                               \
    X(4, 1) X(4, 2) X(4, 4) X(4, 8) X(4, 16)                          \
    X(20, 1) X(20, 2) X(20, 4) X(20, 8)                               \
    X(22, 1) X(22, 2) X(22, 4) X(22, 8)
#define CRUX_LAUNCH_ARGS                                                                         \
    const Entry *, const int32_t *, int, int, const TravShape &, hipStream_t, hipEvent_t,       \
        hipEvent_t

#if !defined(CRUX_PART_K)
#define CRUX_INSTANTIATE(KK, CC) template void launch_class<KK, CC>(CRUX_LAUNCH_ARGS);
CRUX_CLASSES(CRUX_INSTANTIATE)
#undef CRUX_INSTANTIATE
#elif CRUX_PART_K != 0
template void launch_class<CRUX_PART_K, CRUX_PART_CM>(CRUX_LAUNCH_ARGS);
#else
#define CRUX_EXTERN(KK, CC) extern template void launch_class<KK, CC>(CRUX_LAUNCH_ARGS);
CRUX_CLASSES(CRUX_EXTERN)
#undef CRUX_EXTERN
#endif

#if !defined(CRUX_PART_K) || CRUX_PART_K == 0

// Component-count class of an evaluation: the smallest instantiated CM >= C.
int traverse_cmax(int K, int C)
{
    if (K == 4) {
        for (int cm : {1, 2, 4, 8, 16})
            if (C <= cm)
                return cm;
        return -1;
    }
    if (K == 20 || K == 22) {
        for (int cm : {1, 2, 4, 8})
            if (C <= cm)
                return cm;
        return -1;
    }
    return -1;
}

int launch_traverse(int K, int CM, const Entry *entries, const int32_t *block_entry,
                    int blk0, int nblocks, const TravShape &sh, hipStream_t stream, hipEvent_t t0,
                    hipEvent_t t1)
{
#define CRUX_CASE(KK, CC)                                                                    \
    if (K == KK && CM == CC) {                                                               \
        launch_class<KK, CC>(entries, block_entry, blk0, nblocks, sh, stream, t0, t1);       \
        return 0;                                                                            \
    }
    CRUX_CLASSES(CRUX_CASE)
#undef CRUX_CASE
    return -1;
}

void launch_pmat(const PJob *jobs, int njobs, size_t shmem, const void *src, void *dst,
                 size_t nbytes, hipStream_t stream, hipEvent_t t0, hipEvent_t t1)
{
    // One copy block per 4 KiB of plan, at most 64.
    const int n16 = (int)(nbytes / 16);
    const int ncopy = std::max(1, std::min(64, (n16 + 255) / 256));
    hipExtLaunchKernelGGL(k_pmat, dim3(njobs + ncopy), dim3(256), (uint32_t)shmem, stream, t0, t1, 0,
                          jobs, njobs, static_cast<const uint4 *>(src), static_cast<uint4 *>(dst), n16);
}

#endif

}  // namespace crux
