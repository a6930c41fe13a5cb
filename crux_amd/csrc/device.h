// Descriptors shared by the host planner and the gfx950 kernels.
//
// HBM layout (all FP32 CLVs are site-major like the reference's cLMat,
// lib/Crux/Tree/CxLik.h:10-36, but FP32 with an int32 power-of-two exponent
// per site instead of an FP64 lnScale):
//   CLV slot   : float  clv[S/G][KP/4][64][4]  then  int32 exp[S]   (K != 4;
//                KP = K rounded up to a multiple of 4, states >= K held at 0;
//                G = 64/CM sites per tile, lane c*G + s%G as in the kernel)
//                float  clv[S32/32][CM/2][2][32][4]  then  int32 exp[S]  (K == 4,
//                CM >= 2: tiles of 32 sites, component h*CM/2 + m of a site in
//                [m][h]; a wave's float4 store covers 1 KiB contiguously)
//                float  clv[S][4]  then  int32 exp[S]   (K == 4, one component)
//   tip rows   : uint8  rows[ntaxa][S64]     K == 4: 4-bit state set; K != 4:
//                the code index; rows padded to S64 = S rounded up to 64
//   P image 4  : float  img[p4_regs][64]     K == 4: an update's P matrices as
//                the matrix core's broadcast A operand (see p4_regs)
//   P image    : float  img[NVP/4][64][4]    K != 4: P laid out as the matrix
//                core's A operand (see img_nv)
//   tip table  : float  T[ncodes][CM][KP]    K != 4: T = P * leaf vector of each
//                code
//   root weight: double piw[C][K]            pi_norm * weightScaled
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace crux {

constexpr int kOpMaxChild = 3;        // children per fused node update

enum ChildKind : int32_t {
    kChildInternal = 0,   // P * child CLV
    kChildTip = 1,        // tip table lookup by code
    kChildIdent = 2,      // child CLV used as is (chained multifurcations)
};

struct OpChild {
    const float *clv;      // internal/ident: child CLV slot
    const int32_t *exps;   // internal/ident: child exponents
    const uint8_t *codes;  // tip: padded tip row of the taxon
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

// Short 4-state plans (partial re-evaluations, MC3 proposals) of at most
// kPartMaxOps updates, two or four components, run on the short-plan
// kernel: the whole plan's P images and control words staged in LDS at
// once, stack depth kPartDepth, tip rows and cached children loaded
// kPartAhead updates ahead (two: as fast as four or faster at 5k-100k
// patterns, with fewer registers and loads in flight).
constexpr int kPartMaxOps = 24;
constexpr int kPartDepth = 2;
constexpr int kPartAhead = 2;

__host__ __device__ inline int op_pos(int32_t packed, int j)
{
    return (int)(int8_t)((uint32_t)packed >> (8 * j));
}

// Sites per block (one wave) of the 4-state traversal.
constexpr int kTile4 = 64;

// ---- 20/22-state traversal geometry -------------------------------------
// A wave covers CM components of 64 / CM sites: lane l holds component
// l / (64/CM) of site l % (64/CM), all KP states in its registers.  P * x
// runs on v_mfma_f32_4x4x1f32 (16 blocks of 4 lanes, block b's output lands
// in lanes 4b..4b+3): a lane group of one component is 16/CM blocks, and the
// A operand of one block is broadcast to its whole group (CBSZ = log2(16/CM),
// ABID = block), so one register carries 16/CM columns of four P rows.
__host__ __device__ constexpr int kpad(int K) { return K == 4 ? 4 : (K + 3) & ~3; }
__host__ __device__ constexpr int sites20(int CM) { return 64 / CM; }
__host__ __device__ constexpr int img_cols(int CM) { return 16 / CM; }   // columns per register
__host__ __device__ constexpr int img_groups(int KP, int CM) { return (KP + img_cols(CM) - 1) / img_cols(CM); }
// Registers of one P image: register v = I * groups + g holds, in lane
// c * (64/CM) + 4 a + i, P_c[4 I + i][g * cols + a] (0 outside the matrix or
// for c >= C).
__host__ __device__ constexpr int img_nv(int KP, int CM) { return (KP / 4) * img_groups(KP, CM); }
// In memory, registers 4q .. 4q + 3 of every lane form 1 KiB blocks of
// float4 ([q][64 lanes][4]).  Where the traversal loads the last, partly
// used block past a CLV's float4s (img_tail > 0: the image's last
// img_nv % 4 registers), those registers follow as [64 lanes][img_tail]
// dwords instead, so that the load moves what it uses; otherwise the last
// block is padded.
__host__ __device__ constexpr int img_tail(int KP, int CM)
{
    return (img_nv(KP, CM) % 4 != 0 && (img_nv(KP, CM) + 3) / 4 - 1 >= KP / 4) ? img_nv(KP, CM) % 4 : 0;
}
__host__ __device__ constexpr int img_floats(int KP, int CM)
{
    return img_tail(KP, CM) ? (img_nv(KP, CM) / 4) * 256 + img_tail(KP, CM) * 64
                            : ((img_nv(KP, CM) + 3) & ~3) * 64;
}

__host__ __device__ inline int tile_sites(int K, int CM) { return K == 4 ? kTile4 : sites20(CM); }

// Input precision of the 20/22-state P * x products (the context's
// CRUX_TUNE_PRECISION).  FP32 inputs run on v_mfma_f32_4x4x1f32, one column
// per instruction; 16-bit inputs on v_mfma_f32_4x4x4 (bf16 or f16), four
// columns per instruction, with FP32 accumulation: bfloat16, float16, or
// float16 split into high and low parts (x = xh + xl and P = Ph + Pl with
// xl = f16(x - xh), P x ~ Ph xh + Ph xl + Pl xh).  CLVs are stored in FP32
// whatever the precision.
enum Prec20 : int32_t { kPrecF32 = 0, kPrecBF16 = 1, kPrecF16 = 2, kPrecF16x3 = 3 };

// 16-bit P images: the A operand of v_mfma_f32_4x4x4 is four values of a row
// (four consecutive columns) per lane, two per dword.  Dword v = (I * groups
// + g) * 2 + p of lane c * (64/CM) + 4 a + i holds P_c[4 I + i][(g * cols +
// a) * 4 + 2 p + e] in its half e (low half first), cols = 16/CM blocks per
// component group; the split image is the high image, padded to whole
// 16-byte lane quads, followed by the low.
__host__ __device__ constexpr int img16_groups(int KP, int CM) { return (KP + 4 * img_cols(CM) - 1) / (4 * img_cols(CM)); }
__host__ __device__ constexpr int img16_nv(int KP, int CM) { return (KP / 4) * img16_groups(KP, CM) * 2; }
// Dwords of one 16-bit image padded to whole 16-byte lane quads (the low
// image of the split precision starts there).
__host__ __device__ constexpr int img16_nvp(int KP, int CM) { return (img16_nv(KP, CM) + 3) & ~3; }
// 32-bit words of one P image at a precision.
__host__ __device__ constexpr int img_words(int KP, int CM, int prec)
{
    return prec == kPrecF32 ? img_floats(KP, CM)
                            : (prec == kPrecF16x3 ? 2 : 1) * img16_nvp(KP, CM) * 64;
}

// Largest per-lane LDS stack depth of the 4-state traversal for a component
// class CM (keeps the stack at or below 32 KiB per block).
__host__ __device__ constexpr int stack_depth4_max(int CM) { return CM <= 4 ? 8 : (CM == 8 ? 4 : 2); }

// Stack depth actually used: the tuned value, capped by what fits and
// rounded down to an instantiated depth (2, 4 or 8); untuned: the default.
__host__ __device__ inline int stack_depth4(int CM, int tuned, int dflt)
{
    int d = tuned > 0 ? tuned : dflt;
    if (d > stack_depth4_max(CM))
        d = stack_depth4_max(CM);
    return d >= 8 ? 8 : (d >= 4 ? 4 : 2);
}

// Stack depth of the 20/22-state traversal: a slot holds KP floats per lane
// (5 or 6 KiB per wave); 4 slots unless tuned, so that about seven waves
// share a CU's LDS.
__host__ __device__ inline int stack_depth20(int tuned)
{
    const int d = tuned > 0 ? tuned : 4;
    return d >= 8 ? 8 : (d >= 4 ? 4 : 2);
}


// Updates per staged chunk of the 4-state traversal, three instantiated
// lengths per (CM, D) (choice 0 short, 1 middle, 2 long).  A chunk ends in a
// barrier at which the block's waves wait for their stores to drain; a
// longer chunk has fewer such waits and more LDS per block.  Every length
// keeps a block within 64 KiB of LDS and its control words within one load
// per thread (k_trav4: 8 updates, k_trav4h: 16).
__host__ __device__ constexpr int trav4_chunk(int CM, int D, int choice)
{
    return CM == 1    ? (choice == 0 ? 4 : choice == 1 ? 6 : 8)
           : CM == 2  ? (choice == 0 ? 7 : choice == 1 ? 14 : 16)
           : CM == 4  ? (D <= 4 ? (choice == 0 ? 6 : choice == 1 ? 11 : 16)
                                : (choice == 0 ? 4 : choice == 1 ? 8 : 12))
           : CM == 8  ? (choice == 0 ? 2 : choice == 1 ? 4 : 6)
                      : (choice == 0 ? 1 : choice == 1 ? 2 : 3);
}
// Untuned 4-state launch shape of a class launch of `tiles` blocks on `ncu`
// CUs.  A block walks the whole plan for its sites, so a launch with more
// blocks than fit on the chip at once runs a second round of whole block
// lifetimes; the shapes below were the fastest (or within a few percent of
// it) on MI355X (256 CUs) for 128 taxa and 256-4096 tiles per launch:
//   CM 1:  depth 8, 8 updates per chunk; above 8 tiles per CU depth 4, 6
//   CM 2:  depth 4, 16; above 5 tiles per CU 7
//   CM 4:  depth 4, 11; up to 2 tiles per CU depth 8, 12 (two such blocks
//          fit on a CU); at 4-5 tiles per CU depth 2, 6 (six blocks per CU
//          fit where four do at depth 4)
//   CM 8:  depth 4, 4
//   CM 16: depth 2, 3
// (chunk lengths as choices of trav4_chunk).
__host__ __device__ inline void default_shape4(int CM, long tiles, int ncu, int *depth, int *chunk)
{
    const double t = (double)tiles / (ncu > 0 ? ncu : 256);
    if (CM == 1) {
        *depth = t <= 8 ? 8 : 4;
        *chunk = t <= 8 ? 2 : 1;
    } else if (CM == 2) {
        *depth = 4;
        *chunk = t <= 5 ? 2 : 0;
    } else if (CM == 4) {
        const bool few = t <= 2, mid = t > 4 && t <= 5;
        *depth = few ? 8 : mid ? 2 : 4;
        *chunk = few ? 2 : mid ? 0 : 1;
    } else if (CM == 8) {
        *depth = 4;
        *chunk = 1;
    } else {
        *depth = 2;
        *chunk = 2;
    }
}

// Launch shape of one traversal class: LDS stack depth, 4-state chunk
// choice (0-2), waves per SIMD the 20/22-state kernel's registers are held
// to (1 or 2), input precision (20/22 states).
struct TravShape {
    int depth;
    int chunk;
    int waves;
    int prec;
    bool part = false;   // 4 states: the short-plan kernel (depth kPartDepth)
};

// Traversal classes (K, CM) with a kernel: 4 states at 1-16 components, 20
// and 22 at 1-8; -1 for none.
constexpr int kNumClasses = 13;
__host__ __device__ constexpr int class_index(int K, int CM)
{
    const int c = CM == 1 ? 0 : CM == 2 ? 1 : CM == 4 ? 2 : CM == 8 ? 3 : CM == 16 ? 4 : -1;
    return c < 0 ? -1 : K == 4 ? c : c > 3 ? -1 : K == 20 ? 5 + c : K == 22 ? 9 + c : -1;
}

// Waves per SIMD the 20/22-state kernel's registers are held to: the tuned
// value (1 or 2), else two where the registers fit without spilling (up to
// 256 per lane: 20 states with up to 4 components), so that a chip's worth of
// waves is resident at once.  At stack depth 8 a block's LDS (over 40 KiB)
// admits less than one wave per SIMD, so 1.
__host__ __device__ constexpr int trav20_waves(int KP, int CM, int D, int tuned)
{
    return D >= 8 ? 1 : tuned > 0 ? tuned : (KP == 20 && CM <= 4 ? 2 : 1);
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
    double *partials;      // [ntiles][2] compensated per-wave sums (host memory)
    const uint8_t *const *tip_rows;  // padded tip row of each tip child, plan order
    int32_t ntips;
    int32_t exp_off;       // bytes from a CLV slot to its exponents
    // 20/22 states: the tip chunk staged when the entry starts (an entry
    // that starts inside a plan starts in that plan's chunk).
    int32_t tip0;
    // 4-state: P images of the plan, p4_floats(CM) per update with CM the
    // component class (columns of child slots without P are left unused),
    // staged into LDS a chunk of updates at a time.
    const float *pdense;
    // Control words of each update, kCtlWords (4-state) or kCtlWords20 per
    // update: nchild | kinds << 8 | from_stack << 16 | forwarded << 20 (bit
    // j: child j is the previous update's result, kept in registers) |
    // has_parent << 24 | shape << 25 (see ChildSrc), pos,
    // tip_load, has_parent; 4 states: then parent (2 words), parent_exp (2
    // words); K != 4: then buffer descriptors (kCtl20Parent, kCtl20Data,
    // kCtl20Slot).  The tip list (tip_rows, ntips) is per entry, so a
    // one-update entry starts its own.
    const int32_t *ctl;
};

// Child sources named by the control word's update shape: shape
// 1 + 4 s0 + s1 is a two-child update with a parent whose children come from
// sources s0 and s1 (no identity child); shape 0 is any other update.
enum ChildSrc : int {
    kSrcTip = 0,   // tip mask row
    kSrcFwd = 1,   // the previous update's result, in registers
    kSrcStk = 2,   // the wave's LDS stack
    kSrcHbm = 3,   // a CLV slot cached by an earlier plan
};

constexpr int kCtlWords = 8;
// Word 3 of the control words: bit 0 has_parent, bits 8-10 the children that
// are the plan's own results read back from HBM (not loadable ahead), bits
// 12-14 (short plans) the children loaded ahead that are CLVs cached by an
// earlier plan, bits 16-18 (short plans) those that are tips.
constexpr int kCtlOwnShift = 8;
constexpr int kCtlAheadShift = 12;
constexpr int kCtlTipShift = 16;
// Short 4-state plans: kCtlWordsPart words per update, the kCtlWords above
// and then, in words kCtlAhead + 2 j, child j's CLV slot or padded tip row
// (bits 12-18 of word 3 say which of them are loaded ahead).
constexpr int kCtlWordsPart = 16;
constexpr int kCtlAhead = 8;
constexpr int kCtlWords20 = 32;
// 20/22-state control words: words 0-3 as for 4 states (flags and shape,
// pos, tip_load, has_parent), then gfx950 buffer descriptors (4 words each)
// that the kernel loads into scalar registers as they are: the parent's CLV
// slot (CLV, then exponents at exp_off), the P image or tip table of each
// child, and the CLV slot of each child (internal and identity children).
constexpr int kCtl20Parent = 4;
constexpr int kCtl20Data = 8;
constexpr int kCtl20Slot = 20;
// Dword 3 of every buffer descriptor (the value __builtin_amdgcn_make_buffer_rsrc
// is given in the kernels).
constexpr uint32_t kVsharpFlags = 0x00020000u;
constexpr size_t kZeroBytes = 64 * 1024;

enum PLayout : int32_t {
    kPImage4 = 0,  // 4-state matrix-core image of one child (see p4_regs)
    kPImage = 1,   // matrix-core image, img_floats(kp, cm) floats
    kPTips = 2,    // tip table T[ncodes][cm][kp]
};

// 4 states: the P matrices of one update as broadcast A operands of
// v_mfma_f32_4x4x1f32, p4_regs(cm) registers of 64 lanes ([reg][lane]
// floats); rows of components c >= C are zero.  cm == 1 (a lane per site):
// column q = j * 4 + k (child j, column k) lives in register q / 16, lanes
// 4 (q % 16) + i holding P[i][k].  cm >= 2 (lane half h holds components
// h * cm/2 .. h * cm/2 + cm/2 - 1): column q = (j * cm/2 + m) * 4 + k of
// component c = h * cm/2 + m lives in register q / 8, lanes
// 32 h + 4 (q % 8) + i.
__host__ __device__ constexpr int p4_regs(int cm)
{
    return cm == 1 ? (kOpMaxChild * 4 + 15) / 16 : (kOpMaxChild * (cm / 2) * 4 + 7) / 8;
}
__host__ __device__ constexpr int p4_index(int j, int cm, int c, int i, int k)
{
    if (cm == 1) {
        const int q = j * 4 + k;
        return (q >> 4) * 64 + (q & 15) * 4 + i;
    }
    const int cl = cm / 2, h = c / cl, q = (j * cl + c % cl) * 4 + k;
    return (q >> 3) * 64 + h * 32 + (q & 7) * 4 + i;
}
__host__ __device__ constexpr int p4_floats(int cm) { return p4_regs(cm) * 64; }

struct PJob {
    double len;            // branch length (times comp_rate[c] gives t)
    const double *model;   // per comp: V[K][K], Vinv[K][K], evals[K], rate
    const double *codevecs;// tip: [ncodes][K]
    float *out;
    int32_t layout;        // PLayout
    int32_t C;
    int32_t K;
    int32_t ncodes;
    int32_t cm;            // component class
    int32_t kp;            // padded state count (K != 4)
    int32_t child;         // the child's index in its update (4-state image)
    int32_t prec;          // K != 4: Prec20 of the image
};

__host__ __device__ inline int model_stride(int K) { return 2 * K * K + K + 1; }

}  // namespace crux
