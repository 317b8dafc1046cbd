// Device-side layout and helpers of the batched heads-up sweeps, shared by
// the kernels compiled ahead of time (hu_batched.hip) and the per-template
// sweeps generated at run time (bt_jit.hip compiles this header's text
// together with the generated walks through hipRTC).  Everything here must
// compile both ways: with MTT_JIT defined (hipRTC, no system headers) and
// without.  The static_asserts pin the structs' layouts so that both
// compilations agree on the kernel arguments.
#pragma once
#ifdef MTT_JIT
typedef __hip_internal::uint8_t uint8_t;
typedef __hip_internal::int16_t int16_t;
typedef __hip_internal::uint16_t uint16_t;
typedef __hip_internal::uint32_t uint32_t;
#else
#include <cstddef>
#include <cstdint>

#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>
#endif

namespace mtt {
namespace bt {

constexpr int NPR = 1088;   // river combos padded (1081 live)
constexpr int NPT = 1152;   // turn combos padded (1128 live)
constexpr int NPF = 1216;   // flop combos padded (1176 live), the reference's padding
constexpr double FX = 281474976710656.0;  // 2^48: fixed-point scale of card sums

#ifdef MTT_JIT
struct bf16_t {
    unsigned short bits;
};
#else
typedef __hip_bfloat16 bf16_t;
#endif

enum : uint8_t { DN_DEC = 0, DN_FOLD = 1, DN_SD = 2, DN_LEAF = 3 };
enum : int { EV_ENTER = 0, EV_TERM = 1, EV_LEAVE = 2 };

struct DNode {
    uint8_t type, player, nact, depth;
    int16_t act, slot;     // action index from parent; first slot (decision)
    int16_t col, sidx;     // showdown column base; card-sum terminal index
    int16_t leaf, pad;     // leaf ordinal (turn level); parent's actor
    float pot, c0, c1;
};
static_assert(sizeof(DNode) == 28, "DNode layout");

// One DFS event of a template with its node record, 32 bytes.
struct EvRec {
    DNode nd;
    int kind;              // EV_*
};
static_assert(sizeof(EvRec) == 32, "EvRec layout");

struct LevelDev {
    int np, nslot, nterm, ncol, ncolp, ntmpl, nboards;
    int maxterm;               // largest terminal count of one template
    int maxcolt;               // largest showdown column count of one template
    int maxslot;               // largest action-slot count of one template
    const EvRec* ev;           // DFS events of every template, concatenated
    const int* tev0;           // [ntmpl+1]
    const EvRec* fev;          // the same without LEAVE events (preorder nodes)
    const int* tfev0;          // [ntmpl+1]
    const int* tslot0;         // [ntmpl+1] first slot of each template
    const int* tterm0;         // [ntmpl+1] first terminal of each template
    const int* tcol0;          // [ntmpl+1] first showdown column of each template
    float* R;                  // [board][nslot][np]
    float* S;
    unsigned long long* part;  // [board][nterm][half][2][52] card sums of reach, 2^48 fixed point
    const uint8_t* cc;         // [board][np][2] cards of combo
    const int* nlive;          // [board]
    const int* btmpl;          // flop instances: board -> its private template (else null)
    const float* bpot0;        // flop instances: starting pot of each board
};
static_assert(sizeof(LevelDev) == 152, "LevelDev layout");

struct BatchedDev {
    LevelDev L[3];             // L[0] turn (two-level games), L[1] river, L[2] flop instances
    int two_level;
    int nleaf;                 // turn template leaves (= river templates)
    const int* leaf_tmpl;      // turn leaf -> river template
    const int* tmpl_leaf;      // river template -> turn leaf
    float* leafreach;          // [49][nleaf][2][NPT]
    float* leafval;            // [49][nleaf][2][NPT] turn-leaf values (mean over rivers)
    float* rootval;            // [boards1][ntmpl1][2][NPR]
    float* rootval0;           // [49][2][NPT]
    const uint16_t* rmap;      // [boards1][NPR] river combo -> turn combo (padding 0)
    const uint16_t* imap;      // [boards1][NPT] turn combo -> river combo (0xFFFF dead)
    bf16_t* Bh;                // [boards1][ncolp][NPR]
    bf16_t* Bl;
    float* C;                  // [boards1][ncolp][NPR]
    float* Bf;                 // [boards1][ncolp][NPR] showdown columns in FP32 (strength-order sweep), else null
    const uint16_t* grank;     // [boards1][NPR] strength group, ascending
    const uint32_t* cx;        // [boards1][NPR] c1|c2<<8|c1<<16|c2<<24
    const uint32_t* cy;        // [boards1][NPR] c1|c1<<8|c2<<16|c2<<24
    float* B0;                 // [49][ncol0][NPT] turn showdown columns
    float* C0;
    const float* Eturn;        // [49][NPT][NPT] villain-major
    float pot0;
};
static_assert(sizeof(BatchedDev) == 616, "BatchedDev layout");

typedef const __attribute__((address_space(4))) int* cint_p;

// Wave-uniform reads of small per-level tables through the scalar data cache.
__device__ __forceinline__ int ldu(const int* p) { return *(cint_p)p; }

// r * 2^48 truncated to an integer (0 <= r < 2^15), without FP64: the
// integer and fraction parts of r * 2^16 (both exact in FP32) give the high
// and low words.
__device__ __forceinline__ unsigned long long fx48_of(float r) {
    const float t = r * 65536.0f;
    const uint32_t hi = (uint32_t)t;
    const uint32_t lo = (uint32_t)((t - (float)hi) * 4294967296.0f);
    return ((unsigned long long)hi << 32) | lo;
}

// v * 2^-48 for a signed 64-bit v, from its two 32-bit halves: the high
// half (signed) and the low half (unsigned) are converted to FP32, the low
// one scaled by 2^-48 (exact), and one explicit FMA adds them.  Every step is
// an IEEE operation with a fixed rounding, so the library and the sweeps
// generated at run time, which may be built by different compiler releases,
// produce identical bits; the result is within 2^-40 + half an ulp of
// v * 2^-48 (four VALU operations instead of a 64-bit normalisation).
__device__ __forceinline__ float fx48_to_float(long long v) {
    const float h = (float)(int)(v >> 32);
    const float l = (float)(uint32_t)v * 3.5527136788005009e-15f;   // 2^-48
    return fmaf(h, 1.52587890625e-05f, l);                           // 2^-16
}

// 1/n for an action count n in 1..4 (wave-uniform; exact constants).
__device__ __forceinline__ float uniform_share(int n) {
    return n == 2 ? 0.5f : (n == 3 ? 1.0f / 3.0f : (n == 4 ? 0.25f : 1.0f));
}

// Non-blocked reach mass of the other player for a combo holding cards
// a, b: total - sum[a] - sum[b] + own (the pair sharing both cards is the
// combo itself and was subtracted twice), exact in 2^48 fixed point.  The
// first difference can be negative (by at most own), so it is taken as a
// two's-complement signed value.
__device__ __forceinline__ float mass_u(const unsigned long long* s, int a, int b, float self) {
    const long long m = (long long)(s[52] - s[a] - s[b]) + (long long)fx48_of(self);
    return fx48_to_float(m);
}

__device__ __forceinline__ float root_reach(const BatchedDev& D, const float* __restrict__ ranges, int level, int b,
                                            int tm, int np, int k, int p) {
    if (level == 1 && D.two_level) {
        int t = b / 48, leaf = D.tmpl_leaf[tm];
        int ti = D.rmap[(size_t)b * NPR + k];
        return D.leafreach[(((size_t)t * D.nleaf + leaf) * 2 + p) * NPT + ti];
    }
    return ranges[((size_t)b * 2 + p) * np + k];
}

// Branch-free prologue pieces of the generated sweeps.  A load guarded by a
// per-lane condition becomes a branch around it, and the compiler then waits
// for each such load before the next; these forms load from an in-bounds
// address on every lane (the combo index clamped to the padded count) and
// dead lanes select zero afterwards, so all of a thread's prologue loads are
// in flight together.
//
// Where combo kc's root reach lives: its turn combo (two-level game, from the
// river-to-turn map, whose padding entries are 0) or kc itself.
__device__ __forceinline__ int root_index(const BatchedDev& D, int level, int b, int kc) {
    if (level == 1 && D.two_level) return D.rmap[(size_t)b * NPR + kc];
    return kc;
}

// Root reach of player p at index ri (from root_index), as root_reach.
__device__ __forceinline__ float root_reach_at(const BatchedDev& D, const float* __restrict__ ranges, int level,
                                               int b, int tm, int np, int ri, int p) {
    if (level == 1 && D.two_level) {
        const int t = b / 48, leaf = ldu(D.tmpl_leaf + tm);  // scalar load: no wait on the vector loads
        return D.leafreach[(((size_t)t * D.nleaf + leaf) * 2 + p) * NPT + ri];
    }
    return ranges[((size_t)b * 2 + p) * np + ri];
}

// BF16 of a finite float, rounded to nearest even, as raw bits.
__device__ __forceinline__ unsigned short bf16_rne(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_value(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

// v = hi + lo with hi = bf16(v) and lo = bf16(v - hi): 16 significant bits.
__device__ __forceinline__ void store_split(bf16_t* hi, bf16_t* lo, size_t o, float v) {
    const unsigned short h = bf16_rne(v);
    reinterpret_cast<unsigned short*>(hi)[o] = h;
    reinterpret_cast<unsigned short*>(lo)[o] = bf16_rne(v - bf16_value(h));
}

// Forward sweep: clear the workgroup's card-sum bins ([nt][2][52]).
template <int NT>
__device__ __forceinline__ void fwd_bins_clear(unsigned long long* bins, int nt) {
    for (int e = threadIdx.x; e < nt * 104; e += NT) bins[e] = 0ull;
}

// Forward sweep: after the walk, write this half's bins to the partials.
template <int NT>
__device__ __forceinline__ void fwd_bins_flush(const LevelDev& L, const unsigned long long* bins, int b, int tb,
                                               int half, int nt) {
    __syncthreads();
    unsigned long long* out = L.part + ((size_t)b * L.nterm + tb) * 208 + half * 104;
    for (int e = threadIdx.x; e < nt * 104; e += NT) out[(size_t)(e / 104) * 208 + e % 104] = bins[e];
}

// The same with the bins kept in two copies (lanes of even and odd index
// add into copy 0 and 1, halving same-address LDS atomics within a wave);
// the copies are added here (exact 64-bit sums).
template <int NT>
__device__ __forceinline__ void fwd_bins_flush2(const LevelDev& L, const unsigned long long* bins, int b, int tb,
                                                int half, int nt) {
    __syncthreads();
    unsigned long long* out = L.part + ((size_t)b * L.nterm + tb) * 208 + half * 104;
    for (int e = threadIdx.x; e < nt * 104; e += NT)
        out[(size_t)(e / 104) * 208 + e % 104] = bins[e] + bins[nt * 104 + e];
}

// Backward sweep, first part of the prologue: the card sums of the
// template's nt terminals (both forward halves added) into ssum
// ([nt][2][53], slot 52 left for the total).  No barrier.
template <int NTB>
__device__ __forceinline__ void bwd_sums(const LevelDev& L, unsigned long long* ssum, int b, int nt, int tb) {
    const unsigned long long* P = L.part + ((size_t)b * L.nterm + tb) * 208;
    for (int e = threadIdx.x; e < nt * 104; e += NTB) {
        const int tl = e / 104, r = e % 104;
        ssum[tl * 106 + (r / 52) * 53 + r % 52] = P[(size_t)tl * 208 + r] + P[(size_t)tl * 208 + 104 + r];
    }
}

// This thread's showdown values of the template's ncolt columns into Cs
// ([ncolt][NTB], Cs already offset by the thread; each thread reads back only
// its own entries, so no barrier is needed).
template <int LEVEL, int NTB>
__device__ __forceinline__ void bwd_cs(const BatchedDev& D, const LevelDev& L, float* Cs, int b, int k, bool live,
                                       int cbase, int ncolt) {
    constexpr int cstride = LEVEL == 2 ? NPF : (LEVEL == 1 ? NPR : NPT);
    const float* Cb = (LEVEL >= 1 ? D.C : D.C0) + ((size_t)b * L.ncolp + cbase) * cstride + k;
    for (int x = 0; x < ncolt; ++x) Cs[x * NTB] = live ? Cb[(size_t)x * cstride] : 0.0f;
}

// bwd_sums for a template of NT terminals known at compile time: all of the
// thread's partial loads first (indices clamped), then the LDS writes.
// HALVES = 1 when the forward wrote its sums to the first half only.
template <int NTB, int NT, int HALVES>
__device__ __forceinline__ void bwd_sums_n(const LevelDev& L, unsigned long long* ssum, int b, int tb) {
    static_assert(NT > 0, "terminals");
    constexpr int N = NT * 104, IT = (N + NTB - 1) / NTB;
    const unsigned long long* P = L.part + ((size_t)b * L.nterm + tb) * 208;
    unsigned long long s[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int e = min((int)threadIdx.x + i * NTB, N - 1);
        const int tl = e / 104, r = e % 104;
        s[i] = HALVES == 2 ? P[(size_t)tl * 208 + r] + P[(size_t)tl * 208 + 104 + r] : P[(size_t)tl * 208 + r];
    }
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int e = (int)threadIdx.x + i * NTB;
        if (e < N) {
            const int tl = e / 104, r = e % 104;
            ssum[tl * 106 + (r / 52) * 53 + r % 52] = s[i];
        }
    }
}

// bwd_cs for NCOLT columns known at compile time, from the clamped combo
// index kc: all loads first, then the LDS writes (zero on dead lanes).
template <int LEVEL, int NTB, int NCOLT>
__device__ __forceinline__ void bwd_cs_n(const BatchedDev& D, const LevelDev& L, float* Cs, int b, int kc,
                                         bool live, int cbase) {
    static_assert(NCOLT > 0, "columns");
    constexpr int cstride = LEVEL == 2 ? NPF : (LEVEL == 1 ? NPR : NPT);
    const float* Cb = (LEVEL >= 1 ? D.C : D.C0) + ((size_t)b * L.ncolp + cbase) * cstride + kc;
    float v[NCOLT];
#pragma unroll
    for (int x = 0; x < NCOLT; ++x) v[x] = Cb[(size_t)x * cstride];
#pragma unroll
    for (int x = 0; x < NCOLT; ++x) Cs[x * NTB] = live ? v[x] : 0.0f;
}

// Both parts (the interpreted kernels' prologue).  No barrier: the caller
// issues its own loads before bwd_totals.
template <int LEVEL, int NTB>
__device__ __forceinline__ void bwd_stage(const BatchedDev& D, const LevelDev& L, unsigned long long* ssum, float* Cs,
                                          int b, int k, bool live, int nt, int tb, int cbase, int ncolt) {
    bwd_sums<NTB>(L, ssum, b, nt, tb);
    bwd_cs<LEVEL, NTB>(D, L, Cs, b, k, live, cbase, ncolt);
}

// Backward sweep, second part: barrier, then every (terminal, player) total
// of the 52 card sums (each combo counted on two cards, so halved), one
// wave per total with a butterfly of 64-bit adds (exact, order independent),
// then a barrier.
template <int NTB>
__device__ __forceinline__ void bwd_totals(unsigned long long* ssum, int nt) {
    static_assert(NTB % 64 == 0, "whole waves");
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int e = wave; e < nt * 2; e += NTB / 64) {
        unsigned long long* sp = ssum + (size_t)(e >> 1) * 106 + (e & 1) * 53;
        unsigned long long v = lane < 52 ? sp[lane] : 0ull;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0) sp[52] = v >> 1;
    }
    __syncthreads();
}

// Backward sweep epilogue: the template root's values for both players.
template <int LEVEL>
__device__ __forceinline__ void bwd_root_out(const BatchedDev& D, const LevelDev& L, int b, int tm, int k, int np,
                                             bool live, float out0, float out1) {
    if (k >= np) return;
    if (LEVEL == 2) {
        if (D.rootval0) {
            D.rootval0[((size_t)b * 2 + 0) * NPF + k] = live ? out0 : 0.0f;
            D.rootval0[((size_t)b * 2 + 1) * NPF + k] = live ? out1 : 0.0f;
        }
    } else if (LEVEL == 1) {
        if (D.two_level) {
            float* rv = D.rootval + (((size_t)b * L.ntmpl + tm) * 2) * NPR + k;
            rv[0] = live ? out0 : 0.0f;
            rv[NPR] = live ? out1 : 0.0f;
        } else {
            D.rootval0[k] = live ? out0 : 0.0f;
            D.rootval0[np + k] = live ? out1 : 0.0f;
        }
    } else {
        D.rootval0[((size_t)b * 2 + 0) * NPT + k] = live ? out0 : 0.0f;
        D.rootval0[((size_t)b * 2 + 1) * NPT + k] = live ? out1 : 0.0f;
    }
}

}  // namespace bt
}  // namespace mtt
