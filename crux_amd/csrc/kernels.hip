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
__device__ __forceinline__ void load_vec(const float *__restrict__ p, float (&x)[K])
{
    if constexpr (K % 4 == 0) {
#pragma unroll
        for (int q = 0; q < K / 4; ++q) {
            const float4 v = reinterpret_cast<const float4 *>(p)[q];
            x[4 * q + 0] = v.x;
            x[4 * q + 1] = v.y;
            x[4 * q + 2] = v.z;
            x[4 * q + 3] = v.w;
        }
    } else if constexpr (K % 2 == 0) {
#pragma unroll
        for (int q = 0; q < K / 2; ++q) {
            const float2 v = reinterpret_cast<const float2 *>(p)[q];
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
__device__ __forceinline__ void store_vec(float *__restrict__ p, const float (&x)[K])
{
    if constexpr (K % 4 == 0) {
#pragma unroll
        for (int q = 0; q < K / 4; ++q)
            reinterpret_cast<float4 *>(p)[q] =
                make_float4(x[4 * q], x[4 * q + 1], x[4 * q + 2], x[4 * q + 3]);
    } else if constexpr (K % 2 == 0) {
#pragma unroll
        for (int q = 0; q < K / 2; ++q)
            reinterpret_cast<float2 *>(p)[q] = make_float2(x[2 * q], x[2 * q + 1]);
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
    const int e = block_entry[b];
    const Entry *E = &entries[e];
    const int S = E->S, C = E->C, nops = E->nops;
    const int tile = b - E->tile0;
    const int s = tile * kTile + (int)threadIdx.x;
    const int CK = C * K;
    double sum = 0.0, cmp = 0.0;

    if (s < S) {
        for (int o = 0; o < nops; ++o) {
            const Op *op = &E->ops[o];
            const int nch = op->nchild, kinds = op->kinds;
            float acc[CM][K];
            int32_t ex = 0;
#pragma unroll
            for (int j = 0; j < kOpMaxChild; ++j) {
                if (j >= nch)
                    break;
                const OpChild &ch = op->ch[j];
                const int kind = (kinds >> (2 * j)) & 3;
                if (kind == kChildTip) {
                    const int code = ch.codes[s];
                    const float *tb = ch.mat + (size_t)code * CK;
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
                    const float *src = ch.clv + (size_t)s * CK;
                    ex += ch.exps[s];
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
                        const float *__restrict__ P = ch.mat;
#pragma unroll
                        for (int c = 0; c < CM; ++c) {
                            if (c < C) {
                                float x[K];
                                load_vec<K>(src + c * K, x);
#pragma unroll
                                for (int i = 0; i < K; ++i) {
                                    const float *Pr = P + (c * K + i) * K;
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
                float *dst = op->parent + (size_t)s * CK;
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
                op->parent_exp[s] = ex + e2;
            } else {
                const double *piw = E->piw;
                double L = 0.0;
#pragma unroll
                for (int c = 0; c < CM; ++c)
                    if (c < C)
#pragma unroll
                        for (int i = 0; i < K; ++i)
                            L += piw[c * K + i] * (double)acc[c][i];
                double v = (log(L) + (double)ex * 0.69314718055994530942) * (double)E->weights[s];
                if (__builtin_isnan(v))
                    v = -INFINITY;
                E->site_out[s] = v;
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
        E->partials[2 * tile] = red_s[0];
        E->partials[2 * tile + 1] = red_c[0];
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

template <int K, int CM>
static void launch_traverse(const Entry *entries, const int32_t *block_entry, int blk0,
                            int nblocks, hipStream_t stream)
{
    hipLaunchKernelGGL((k_traverse<K, CM>), dim3(nblocks), dim3(kTile), 0, stream, entries,
                       block_entry, blk0);
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
                    int blk0, int nblocks, hipStream_t stream)
{
#define CRUX_CASE(KK, CC)                                                       \
    if (K == KK && CM == CC) {                                                  \
        launch_traverse<KK, CC>(entries, block_entry, blk0, nblocks, stream);   \
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
