// gfx950 kernels of the tree-likelihood hot path.
//
//  k_pmat      P(t) = max(0, V exp(L t) V^-1) in FP64 for every (edge, comp) of
//              a plan, stored FP32; for a tip child the product of P with the
//              leaf vector of every distinct code instead (tip table).  This
//              replaces the per-step, per-stripe CxLikPt calls of the
//              reference (lib/Crux/Tree/CxLik.c:334-347).
//  k_traverse  The CLV updates of a whole plan.  One lane owns one site
//              across all components; a block owns 256 consecutive sites of
//              one evaluation and walks every node update of the plan in
//              post-order, so a lane only ever reads CLV entries it wrote
//              itself in an earlier update: one launch covers the whole
//              traversal and every evaluation of a batch with no barrier.
//              Node update = product over children of (P * child) (or a tip
//              table lookup), then a power-of-two rescale of the site so its
//              largest entry is in [0.5, 1), the exponent added to the
//              children's.  This is the arithmetic of CxLikExecuteStripe
//              (CxLik.c:294-846) with FP32 storage and exact power-of-two
//              scaling in place of FP64 division by the max.  The root
//              update is not stored: it is reduced to the pattern-weighted
//              site lnL in FP64, NaN -> -inf (CxLik.c:598-632), and to a
//              compensated per-block sum.
//  k_finalize  Compensated FP64 sum of the per-block partials per evaluation.

#include <hip/hip_runtime.h>

#include <cmath>

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

// ---------------------------------------------------------------------------
// P matrices and tip tables.

__global__ void __launch_bounds__(256) k_pmat(const PJob *__restrict__ jobs)
{
    extern __shared__ double sm[];
    const PJob J = jobs[blockIdx.x];
    const int C = J.C, K = J.K, KK = K * K, st = model_stride(K);
    double *ex = sm;            // [C][K] exp(lambda_k * t_c)
    double *P = sm + C * K;     // [C][K][K], tip jobs only
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
        if (J.tip)
            P[idx] = p;
        else
            J.out[idx] = (float)p;
    }
    if (!J.tip)
        return;
    __syncthreads();
    const int n = J.ncodes * C * K;
    for (int idx = threadIdx.x; idx < n; idx += blockDim.x) {
        const int code = idx / (C * K), r = idx - code * C * K;   // r = c*K + i
        const double *row = P + (size_t)r * K, *v = J.codevecs + (size_t)code * K;
        double t = 0.0;
        for (int j = 0; j < K; ++j)
            t += row[j] * v[j];
        J.out[idx] = (float)t;
    }
}

// ---------------------------------------------------------------------------
// Traversal.

template <int K>
__device__ __forceinline__ void load_vec(const CRUX_GAS float *__restrict__ p, float (&x)[K])
{
    if constexpr (K % 4 == 0) {
#pragma unroll
        for (int q = 0; q < K / 4; ++q) {
            const float4 v = ld4(p + 4 * q);
            x[4 * q + 0] = v.x;
            x[4 * q + 1] = v.y;
            x[4 * q + 2] = v.z;
            x[4 * q + 3] = v.w;
        }
    } else if constexpr (K % 2 == 0) {
#pragma unroll
        for (int q = 0; q < K / 2; ++q) {
            const float2 v = ld2(p + 2 * q);
            x[2 * q + 0] = v.x;
            x[2 * q + 1] = v.y;
        }
    } else {
#pragma unroll
        for (int q = 0; q < K; ++q)
            x[q] = p[q];
    }
}

template <int K>
__device__ __forceinline__ void store_vec(CRUX_GAS float *__restrict__ p, const float (&x)[K])
{
    if constexpr (K % 4 == 0) {
#pragma unroll
        for (int q = 0; q < K / 4; ++q)
            st4(p + 4 * q, make_float4(x[4 * q], x[4 * q + 1], x[4 * q + 2], x[4 * q + 3]));
    } else if constexpr (K % 2 == 0) {
#pragma unroll
        for (int q = 0; q < K / 2; ++q)
            st2(p + 2 * q, make_float2(x[2 * q], x[2 * q + 1]));
    } else {
#pragma unroll
        for (int q = 0; q < K; ++q)
            p[q] = x[q];
    }
}

template <int K, int CM>
__global__ void __launch_bounds__(kTile)
k_traverse(const Entry *__restrict__ entries, const int32_t *__restrict__ block_entry, int blk0)
{
    __shared__ double red_s[kTile];
    __shared__ double red_c[kTile];
    // block_entry is indexed by the block's position among all blocks of the
    // batch; this launch covers positions [blk0, blk0 + gridDim.x).
    const int b = blk0 + (int)blockIdx.x;
    const int e = cas(block_entry)[b];
    const CRUX_CAS Entry *E = cas(entries) + e;
    const int S = E->S, C = E->C, nops = E->nops;
    const int tile = b - E->tile0;
    const int s = tile * kTile + (int)threadIdx.x;
    const int CK = C * K;
    double sum = 0.0, cmp = 0.0;

    if (s < S) {
        for (int o = 0; o < nops; ++o) {
            const CRUX_CAS Op *op = cas(E->ops) + o;
            const int nch = op->nchild, kinds = op->kinds;
            float acc[CM][K];
            int32_t ex = 0;
#pragma unroll
            for (int j = 0; j < kOpMaxChild; ++j) {
                if (j >= nch)
                    break;
                const CRUX_CAS OpChild &ch = op->ch[j];
                const int kind = (kinds >> (2 * j)) & 3;
                if (kind == kChildTip) {
                    const int code = gas(ch.codes)[s];
                    const CRUX_GAS float *tb = gas(ch.mat) + (size_t)code * CK;
#pragma unroll
                    for (int c = 0; c < CM; ++c) {
                        if (c < C) {
                            float x[K];
                            load_vec<K>(tb + c * K, x);
#pragma unroll
                            for (int i = 0; i < K; ++i)
                                acc[c][i] = (j == 0) ? x[i] : acc[c][i] * x[i];
                        }
                    }
                } else {
                    const CRUX_GAS float *src = gas(ch.clv) + (size_t)s * CK;
                    ex += gas(ch.exps)[s];
                    if (kind == kChildIdent) {
#pragma unroll
                        for (int c = 0; c < CM; ++c) {
                            if (c < C) {
                                float x[K];
                                load_vec<K>(src + c * K, x);
#pragma unroll
                                for (int i = 0; i < K; ++i)
                                    acc[c][i] = (j == 0) ? x[i] : acc[c][i] * x[i];
                            }
                        }
                    } else {
                        const CRUX_CAS float *__restrict__ P = cas(ch.mat);
#pragma unroll
                        for (int c = 0; c < CM; ++c) {
                            if (c < C) {
                                float x[K];
                                load_vec<K>(src + c * K, x);
#pragma unroll
                                for (int i = 0; i < K; ++i) {
                                    const CRUX_CAS float *Pr = P + (c * K + i) * K;
                                    float v = 0.0f;
#pragma unroll
                                    for (int q = 0; q < K; ++q)
                                        v = fmaf(Pr[q], x[q], v);
                                    acc[c][i] = (j == 0) ? v : acc[c][i] * v;
                                }
                            }
                        }
                    }
                }
            }
            if (op->parent != nullptr) {
                float m = 0.0f;
#pragma unroll
                for (int c = 0; c < CM; ++c)
                    if (c < C)
#pragma unroll
                        for (int i = 0; i < K; ++i)
                            m = fmaxf(m, acc[c][i]);
                int e2 = 0;
                if (m > 0.0f)
                    (void)frexpf(m, &e2);
                CRUX_GAS float *dst = gas(op->parent) + (size_t)s * CK;
#pragma unroll
                for (int c = 0; c < CM; ++c) {
                    if (c < C) {
                        float y[K];
#pragma unroll
                        for (int i = 0; i < K; ++i)
                            y[i] = ldexpf(acc[c][i], -e2);
                        store_vec<K>(dst + c * K, y);
                    }
                }
                gas(op->parent_exp)[s] = ex + e2;
            } else {
                const CRUX_CAS double *piw = cas(E->piw);
                double L = 0.0;
#pragma unroll
                for (int c = 0; c < CM; ++c)
                    if (c < C)
#pragma unroll
                        for (int i = 0; i < K; ++i)
                            L += piw[c * K + i] * (double)acc[c][i];
                double v = (log(L) + (double)ex * 0.69314718055994530942) * (double)gas(E->weights)[s];
                if (__builtin_isnan(v))
                    v = -INFINITY;
                gas(E->site_out)[s] = v;
                sum = v;
            }
        }
    }
    red_s[threadIdx.x] = sum;
    red_c[threadIdx.x] = cmp;
    __syncthreads();
#pragma unroll
    for (int w = kTile / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            double a = red_s[threadIdx.x], b = red_c[threadIdx.x];
            nsum_merge(a, b, red_s[threadIdx.x + w], red_c[threadIdx.x + w]);
            red_s[threadIdx.x] = a;
            red_c[threadIdx.x] = b;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        gas(E->partials)[2 * tile] = red_s[0];
        gas(E->partials)[2 * tile + 1] = red_c[0];
    }
}

// ---------------------------------------------------------------------------
// 4-state traversal.  One lane owns one site across all components; a block is
// one wave of 64 sites.  A node update computed earlier in the same plan is
// handed to its parent through a per-lane stack in LDS (post-order makes the
// pending results LIFO), so a full traversal reads no CLV back from HBM: it
// only streams the cache writes.  Children cached by an earlier plan come
// from HBM.  A tip contributes P * leaf vector, formed in registers from the
// 4-bit state mask, so tips cost one byte per site.  The 16 P entries of a
// component are uniform across the block and live in scalar registers.

__device__ __forceinline__ float4 pmul4(const CRUX_CAS float *__restrict__ P, float4 x)
{
    float4 v;
    v.x = fmaf(P[0], x.x, fmaf(P[1], x.y, fmaf(P[2], x.z, P[3] * x.w)));
    v.y = fmaf(P[4], x.x, fmaf(P[5], x.y, fmaf(P[6], x.z, P[7] * x.w)));
    v.z = fmaf(P[8], x.x, fmaf(P[9], x.y, fmaf(P[10], x.z, P[11] * x.w)));
    v.w = fmaf(P[12], x.x, fmaf(P[13], x.y, fmaf(P[14], x.z, P[15] * x.w)));
    return v;
}

__device__ __forceinline__ float4 mul4(float4 a, float4 b)
{
    return make_float4(a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w);
}

__device__ __forceinline__ float max4(float4 a)
{
    return fmaxf(fmaxf(a.x, a.y), fmaxf(a.z, a.w));
}

// P block of component c for every child of an update, in scalar registers.
__device__ __forceinline__ void load_p4(const CRUX_CAS Op *op, int c, float (&p)[kOpMaxChild][16])
{
    const int nch = op->nchild, kinds = op->kinds;
#pragma unroll
    for (int j = 0; j < kOpMaxChild; ++j) {
        if (j < nch && ((kinds >> (2 * j)) & 3) != kChildIdent) {
            const CRUX_CAS float *P = cas(op->ch[j].mat) + 16 * c;
#pragma unroll
            for (int k = 0; k < 16; ++k)
                p[j][k] = P[k];
        }
    }
}

__device__ __forceinline__ float4 pmul4r(const float (&P)[16], float4 x)
{
    float4 v;
    v.x = fmaf(P[0], x.x, fmaf(P[1], x.y, fmaf(P[2], x.z, P[3] * x.w)));
    v.y = fmaf(P[4], x.x, fmaf(P[5], x.y, fmaf(P[6], x.z, P[7] * x.w)));
    v.z = fmaf(P[8], x.x, fmaf(P[9], x.y, fmaf(P[10], x.z, P[11] * x.w)));
    v.w = fmaf(P[12], x.x, fmaf(P[13], x.y, fmaf(P[14], x.z, P[15] * x.w)));
    return v;
}

// ---------------------------------------------------------------------------
// 4-state traversal.  A block covers 64 sites with one wave per mixture
// component (CM waves; waves c >= C idle), so the 16 P entries a wave needs
// are uniform and sit in scalar registers, and each wave does one
// component's share of the arithmetic.  The block walks the whole plan
// (post-order list of fused node updates) for its sites: one launch covers a
// full traversal and every evaluation of a batch, with no dependency between
// blocks.  Per update: each wave forms P_c * child_c for every child (a tip
// contributes P_c * its 4-bit state mask, one byte per site) and multiplies
// them; the per-site maximum over components is exchanged through LDS (one
// barrier), giving the exponent e with max in [0.5, 1) after an exact
// scaling by 2^-e.  The scaled component plane is stored to the HBM cache
// slot (layout [C][S][4], 1 KiB contiguous per wave store) and, when the
// parent is in the same plan, pushed onto the wave's LDS stack; parents pop
// children from there, so a full traversal reads no CLV back from HBM.  The
// root update reduces to the pattern-weighted FP64 site lnL (NaN -> -inf,
// CxLik.c:598-632) and a compensated per-block sum.  Arithmetic follows
// CxLikExecuteStripeDna (CxLik.c:294-633) with FP32 storage.
template <int CM, int D>
__global__ void __launch_bounds__(kTile4 * CM)
k_trav4(const Entry *__restrict__ entries, const int32_t *__restrict__ block_entry, int blk0)
{
    __shared__ float4 stk[D][CM][kTile4];
    __shared__ int32_t stke[D][CM][kTile4];
    __shared__ float smax[2][CM][kTile4];
    __shared__ double sL[CM][kTile4];
    __shared__ uint8_t tipm[kTipChunk][kTile4];
    const int b = blk0 + (int)blockIdx.x;
    const CRUX_CAS Entry *E = cas(entries) + cas(block_entry)[b];
    const int S = E->S, C = E->C, nops = E->nops;
    const int tile = b - E->tile0;
    const int lane = (int)threadIdx.x & (kTile4 - 1);
    const int c = __builtin_amdgcn_readfirstlane((int)threadIdx.x / kTile4);
    const int s = tile * kTile4 + lane;
    const bool act = s < S;
    const int sc = act ? s : S - 1;       // in-bounds site for loads of idle lanes
    const bool wc = c < C;
    const CRUX_CAS Op *ops = cas(E->ops);
    double sum = 0.0;

    // P blocks are staged into LDS kPChunk updates at a time, the next chunk
    // being fetched while the current one runs: every block walks the plan
    // in step with every other, so a per-update load would meet the full L2
    // latency with nothing to overlap it.
    __shared__ float pch[2][kPChunk * kOpMaxChild * CM * 16];
    constexpr int kPer = kPChunk * kOpMaxChild * 16 / kTile4;   // floats per thread
    const int tid = (int)threadIdx.x;
    const int chunk_f = kPChunk * kOpMaxChild * C * 16;
    const int total_f = nops * kOpMaxChild * C * 16;
    const CRUX_GAS float *pd = gas(E->pdense);
    // The updates' control words travel with the P chunks for the same reason.
    __shared__ int32_t cch[2][kPChunk * kCtlWords];
    const CRUX_GAS int32_t *cg = gas(E->ctl);
    float pf[kPer];
    int32_t cf = 0;
    auto fetch = [&](int k) {
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int i = tid * kPer + q, g = k * chunk_f + i;
            pf[q] = (i < chunk_f && g < total_f) ? pd[g] : 0.0f;
        }
        if (tid < kPChunk * kCtlWords)
            cf = (k * kPChunk + tid / kCtlWords < nops) ? cg[k * kPChunk * kCtlWords + tid] : 0;
    };
    auto commit = [&](int buf) {
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int i = tid * kPer + q;
            if (i < chunk_f)
                pch[buf][i] = pf[q];
        }
        if (tid < kPChunk * kCtlWords)
            cch[buf][tid] = cf;
    };
    fetch(0);
    commit(0);
    fetch(1);
    __syncthreads();

    for (int o = 0; o < nops; ++o) {
        const CRUX_CAS Op *op = ops + o;
        if (o > 0 && o % kPChunk == 0) {
            const int k = o / kPChunk;
            commit(k & 1);
            __syncthreads();
            fetch(k + 1);
        }
        const int buf = (o / kPChunk) & 1;
        const float *pchunk = &pch[buf][(o % kPChunk) * kOpMaxChild * C * 16];
        const int32_t *cw = &cch[buf][(o % kPChunk) * kCtlWords];
        const int w0 = __builtin_amdgcn_readfirstlane(cw[0]);
        const int32_t pk = __builtin_amdgcn_readfirstlane(cw[1]);
        const int tl = __builtin_amdgcn_readfirstlane(cw[2]);
        const bool has_parent = __builtin_amdgcn_readfirstlane(cw[3]) != 0;
        const int nch = w0 & 0xff, kinds = (w0 >> 8) & 0xff, fst = (w0 >> 16) & 0xff;
        if (tl >= 0) {
            // Stage the next chunk of tip state masks; waves split the rows.
            __syncthreads();
            const int n = min(kTipChunk, E->ntips - tl);
            const CRUX_CAS uint8_t *const CRUX_CAS *rows =
                (const CRUX_CAS uint8_t *const CRUX_CAS *)cas(E->tip_rows) + tl;
            for (int k = c; k < n; k += CM)
                tipm[k][lane] = gas((const uint8_t *)rows[k])[sc];
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
        }

        float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
        int32_t ex = 0;
        if (wc) {
#pragma unroll
            for (int j = 0; j < kOpMaxChild; ++j) {
                if (j >= nch)
                    break;
                const CRUX_CAS OpChild &ch = op->ch[j];
                const int kind = (kinds >> (2 * j)) & 3;
                const int p = op_pos(pk, j);
                float4 x;
                if (kind == kChildTip) {
                    const uint32_t m = tipm[p][lane];
                    x = make_float4((float)(m & 1u), (float)((m >> 1) & 1u),
                                    (float)((m >> 2) & 1u), (float)((m >> 3) & 1u));
                } else if ((fst >> j) & 1) {
                    x = stk[p][c][lane];
                    ex += stke[p][c][lane];
                } else {
                    // Cached by an earlier plan (partial re-evaluation).
                    x = ld4(gas(reinterpret_cast<const float4 *>(ch.clv)) + (size_t)c * S + sc);
                    ex += gas(ch.exps)[sc];
                    __builtin_amdgcn_s_waitcnt(0);
                }
                float4 v = x;
                if (kind != kChildIdent) {
                    const float4 *Pq = reinterpret_cast<const float4 *>(pchunk + (j * C + c) * 16);
                    const float4 r0 = Pq[0], r1 = Pq[1], r2 = Pq[2], r3 = Pq[3];
                    v.x = fmaf(r0.x, x.x, fmaf(r0.y, x.y, fmaf(r0.z, x.z, r0.w * x.w)));
                    v.y = fmaf(r1.x, x.x, fmaf(r1.y, x.y, fmaf(r1.z, x.z, r1.w * x.w)));
                    v.z = fmaf(r2.x, x.x, fmaf(r2.y, x.y, fmaf(r2.z, x.z, r2.w * x.w)));
                    v.w = fmaf(r3.x, x.x, fmaf(r3.y, x.y, fmaf(r3.z, x.z, r3.w * x.w)));
                }
                acc = (j == 0) ? v : mul4(acc, v);
            }
        }
        if (has_parent) {
            smax[o & 1][c][lane] = wc ? max4(acc) : 0.0f;
            __syncthreads();
            float m = smax[o & 1][0][lane];
#pragma unroll
            for (int q = 1; q < CM; ++q)
                if (q < C)
                    m = fmaxf(m, smax[o & 1][q][lane]);
            int e2 = 0;
            if (m > 0.0f)
                (void)frexpf(m, &e2);
            if (wc) {
                const float4 y = make_float4(ldexpf(acc.x, -e2), ldexpf(acc.y, -e2),
                                             ldexpf(acc.z, -e2), ldexpf(acc.w, -e2));
                const int pp = op_pos(pk, kOpMaxChild);
                const uint64_t par = (uint32_t)__builtin_amdgcn_readfirstlane(cw[4]) |
                                     ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane(cw[5]) << 32);
                if (act)
                    st4(gas(reinterpret_cast<float4 *>(par)) + (size_t)c * S + s, y);
                if (pp >= 0) {
                    stk[pp][c][lane] = y;
                    stke[pp][c][lane] = ex + e2;
                }
                if (c == 0 && act) {
                    const uint64_t pe = (uint32_t)__builtin_amdgcn_readfirstlane(cw[6]) |
                                        ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane(cw[7]) << 32);
                    gas(reinterpret_cast<int32_t *>(pe))[s] = ex + e2;
                }
            }
        } else {
            double Lc = 0.0;
            if (wc) {
                const CRUX_CAS double *piw = cas(E->piw) + 4 * c;
                Lc = piw[0] * (double)acc.x + piw[1] * (double)acc.y +
                     piw[2] * (double)acc.z + piw[3] * (double)acc.w;
            }
            sL[c][lane] = Lc;
            __syncthreads();
            if (c == 0) {
                double L = 0.0;
#pragma unroll
                for (int q = 0; q < CM; ++q)
                    if (q < C)
                        L += sL[q][lane];
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
    }
    if (c != 0)
        return;
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
    if (lane == 0) {
        gas(E->partials)[2 * tile] = sum;
        gas(E->partials)[2 * tile + 1] = cmp;
    }
}

__global__ void __launch_bounds__(64) k_finalize(const Entry *__restrict__ entries)
{
    const Entry *E = &entries[blockIdx.x];
    double s = 0.0, c = 0.0;
    for (int t = (int)threadIdx.x; t < E->ntiles; t += 64)
        nsum_merge(s, c, E->partials[2 * t], E->partials[2 * t + 1]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double s2 = __shfl_xor(s, off, 64), c2 = __shfl_xor(c, off, 64);
        // Both lanes of a pair combine in the same order, so every lane ends
        // with the same value.
        if (((int)threadIdx.x & off) == 0)
            nsum_merge(s, c, s2, c2);
        else {
            double a = s2, b = c2;
            nsum_merge(a, b, s, c);
            s = a;
            c = b;
        }
    }
    if (threadIdx.x == 0)
        *E->lnl_out = s + c;
}

// ---------------------------------------------------------------------------
// Launchers.

template <int CM, int D>
static void launch_trav4(const Entry *entries, const int32_t *block_entry, int blk0,
                         int nblocks, hipStream_t stream)
{
    if constexpr (D <= stack_depth4_max(CM))
        hipLaunchKernelGGL((k_trav4<CM, D>), dim3(nblocks), dim3(kTile4 * CM), 0, stream,
                           entries, block_entry, blk0);
}

template <int K, int CM>
static void launch_traverse(const Entry *entries, const int32_t *block_entry, int blk0,
                            int nblocks, int depth, hipStream_t stream)
{
    if constexpr (K == 4) {
        // Stack depth: more slots keep more results in LDS, fewer let more
        // blocks share a CU.
        if (depth <= 2)
            launch_trav4<CM, 2>(entries, block_entry, blk0, nblocks, stream);
        else if (depth <= 4)
            launch_trav4<CM, 4>(entries, block_entry, blk0, nblocks, stream);
        else
            launch_trav4<CM, 8>(entries, block_entry, blk0, nblocks, stream);
    } else {
        hipLaunchKernelGGL((k_traverse<K, CM>), dim3(nblocks), dim3(kTile), 0, stream, entries,
                           block_entry, blk0);
    }
}

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
                    int blk0, int nblocks, int depth, hipStream_t stream)
{
#define CRUX_CASE(KK, CC)                                                       \
    if (K == KK && CM == CC) {                                                  \
        launch_traverse<KK, CC>(entries, block_entry, blk0, nblocks, depth, stream); \
        return 0;                                                               \
    }
    CRUX_CASE(4, 1) CRUX_CASE(4, 2) CRUX_CASE(4, 4) CRUX_CASE(4, 8) CRUX_CASE(4, 16)
    CRUX_CASE(20, 1) CRUX_CASE(20, 2) CRUX_CASE(20, 4) CRUX_CASE(20, 8)
    CRUX_CASE(22, 1) CRUX_CASE(22, 2) CRUX_CASE(22, 4) CRUX_CASE(22, 8)
#undef CRUX_CASE
    return -1;
}

void launch_pmat(const PJob *jobs, int njobs, size_t shmem, hipStream_t stream)
{
    hipLaunchKernelGGL(k_pmat, dim3(njobs), dim3(256), shmem, stream, jobs);
}

void launch_finalize(const Entry *entries, int nentries, hipStream_t stream)
{
    hipLaunchKernelGGL(k_finalize, dim3(nentries), dim3(64), 0, stream, entries);
}

}  // namespace crux
