// Batched heads-up path: per-street fused sweeps and the BF16 hi+lo split
// showdown GEMM on CDNA4 matrix cores.
//
// The tree is handled as street templates instantiated over boards (the
// turn street over 49 turn cards, each river line over 49x48 river boards):
//   * state is per (board, template) in compact per-board combo order, SoA
//     [board][action slot][combo], so every sweep streams coalesced rows;
//   * one forward kernel per street walks a template per workgroup (regret
//     matching fused in), writes the showdown reach columns and the card
//     sums that fold / blocker terms need;
//   * the river showdown of every board is one GEMM
//       C[col][i] = sum_j S[i][j] * r_col[j],  S = sign(s_i - s_j) * nb(i, j),
//     with S generated in registers (exact in BF16) and r split into BF16
//     hi + lo, accumulated in FP32 on v_mfma_f32_32x32x16_bf16;
//   * one backward kernel per street recomputes the forward reaches, takes
//     terminal values from the GEMM and the card sums, and applies the
//     DCFR-generalised regret / strategy update in the same pass.
// River combos are ordered by hand strength on each board.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include <functional>

#include <hip/hip_bf16.h>

#include "solver.h"

namespace mtt {

namespace {

constexpr int NPR = 1088;   // river combos padded (1081 live)
constexpr int NPT = 1152;   // turn combos padded (1128 live)

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

constexpr int MAXN = 128;      // nodes per template
constexpr int MAXE = 256;      // DFS events per template
constexpr int MAXSLOT = 64;    // action slots per template
constexpr double FX = 281474976710656.0;  // 2^48: fixed-point scale of card sums

struct LevelDev {
    int np, nslot, nterm, ncol, ncolp, ntmpl, nboards;
    int maxterm;               // largest terminal count of one template
    const DNode* nodes;        // level nodes, templates concatenated (DFS preorder)
    const int* tnode0;         // [ntmpl+1]
    const int* tev0;           // [ntmpl+1]
    const int* tslot0;         // [ntmpl+1] first slot of each template
    const int* tterm0;         // [ntmpl+1] first terminal of each template
    const int* events;         // (local node << 2) | kind
    float* R;                  // [board][nslot][np]
    float* S;
    double* sums;              // [board][nterm][2][56] card sums, slot 52 = total
    const uint8_t* cc;         // [board][np][2] cards of combo
    const int* nlive;          // [board]
};

struct BatchedDev {
    LevelDev L[2];             // L[0] turn (two-level games), L[1] river
    int two_level;
    int nleaf;                 // turn template leaves (= river templates)
    const int* leaf_tmpl;      // turn leaf -> river template
    const int* tmpl_leaf;      // river template -> turn leaf
    float* leafreach;          // [49][nleaf][2][NPT]
    float* rootval;            // [boards1][ntmpl1][2][NPR]
    float* rootval0;           // [49][2][NPT]
    const uint16_t* rmap;      // [boards1][NPR] river combo -> turn combo
    const uint16_t* imap;      // [boards1][NPT] turn combo -> river combo (0xFFFF dead)
    __hip_bfloat16* Bh;        // [boards1][ncolp][NPR]
    __hip_bfloat16* Bl;
    float* C;                  // [boards1][ncolp][NPR]
    const uint16_t* grank;     // [boards1][NPR] strength group, ascending
    const uint32_t* cx;        // [boards1][NPR] c1|c2<<8|c1<<16|c2<<24
    const uint32_t* cy;        // [boards1][NPR] c1|c1<<8|c2<<16|c2<<24
    float* B0;                 // [49][ncol0][NPT] turn showdown columns
    float* C0;
    const float* Eturn;        // [49][NPT][NPT] villain-major
    float pot0;
};

// Template tables are identical for every lane: after staging them in LDS,
// read them into wave-uniform registers so depth / slot indices select
// registers directly.
__device__ __forceinline__ DNode uniform_node(const DNode* p) {
    const int* w = reinterpret_cast<const int*>(p);
    int v[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) v[k] = __builtin_amdgcn_readfirstlane(w[k]);
    DNode d;
    __builtin_memcpy(&d, v, sizeof(d));
    return d;
}

// Select element q (wave-uniform) of two 32-register arrays: a switch on a
// uniform value becomes scalar branches and every array access keeps a
// constant index, so the arrays stay in registers.
__device__ __forceinline__ float sel64(const float (&a)[32], const float (&b)[32], int q) {
    switch (q) {
#define MTT_C(i) case i: return a[i]; case i + 32: return b[i];
        MTT_C(0) MTT_C(1) MTT_C(2) MTT_C(3) MTT_C(4) MTT_C(5) MTT_C(6) MTT_C(7)
        MTT_C(8) MTT_C(9) MTT_C(10) MTT_C(11) MTT_C(12) MTT_C(13) MTT_C(14) MTT_C(15)
        MTT_C(16) MTT_C(17) MTT_C(18) MTT_C(19) MTT_C(20) MTT_C(21) MTT_C(22) MTT_C(23)
        MTT_C(24) MTT_C(25) MTT_C(26) MTT_C(27) MTT_C(28) MTT_C(29) MTT_C(30) MTT_C(31)
#undef MTT_C
        default: return 0.0f;
    }
}

__device__ __forceinline__ float mass_of(const double* s, int a, int b, float self) {
    return (float)(s[52] - s[a] - s[b] + (double)self);
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

// ------------------------------------------------------------ forward ----
// One workgroup per (board, template); each thread owns two combos (tid and
// tid + blockDim), so the whole board's range is in one workgroup and the
// card sums of every terminal reach column are reduced in LDS with 64-bit
// fixed-point atomics (exact and order independent).  Reaches and
// strategies live in register stacks indexed by the uniform template depth.
template <int LEVEL, int DM>
__global__ __launch_bounds__(576) void k_bt_forward(BatchedDev D, const float* __restrict__ ranges) {
    const LevelDev& L = D.L[LEVEL];
    extern __shared__ unsigned long long bins[];  // [2*maxterm][52]
    __shared__ DNode tn[MAXN];
    const int inst = blockIdx.x;
    const int b = inst / L.ntmpl, tm = inst % L.ntmpl;
    const int tid = threadIdx.x, NT = blockDim.x;
    const int np = L.np;
    const int nlive = L.nlive[b];
    const int n0 = L.tnode0[tm], nn = L.tnode0[tm + 1] - n0;
    const int t0 = L.tterm0[tm], ncolt = 2 * (L.tterm0[tm + 1] - t0);
    for (int e = tid; e < nn; e += NT) tn[e] = L.nodes[n0 + e];
    for (int e = tid; e < ncolt * 52; e += NT) bins[e] = 0ull;
    const int kk[2] = {tid, tid + NT};
    const bool live[2] = {kk[0] < nlive, kk[1] < nlive};
    int cA[2], cB[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        cA[q] = L.cc[((size_t)b * np + kk[q]) * 2];
        cB[q] = L.cc[((size_t)b * np + kk[q]) * 2 + 1];
    }
    float rA0[DM], rA1[DM], rB0[DM], rB1[DM];
    float sA[DM * 4], sB[DM * 4];
    float root[2][2];
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int p = 0; p < 2; ++p) root[q][p] = live[q] ? root_reach(D, ranges, LEVEL, b, tm, np, kk[q], p) : 0.0f;
    __syncthreads();
    const float* Rb = L.R + (size_t)b * L.nslot * np;
    for (int n = 0; n < nn; ++n) {
        const DNode nd = uniform_node(&tn[n]);
        const int d = nd.depth;
        float r[2][2];
        if (n == 0) {
            r[0][0] = root[0][0]; r[0][1] = root[0][1]; r[1][0] = root[1][0]; r[1][1] = root[1][1];
        } else {
            const float pa0 = rA0[d - 1], pa1 = rA1[d - 1], pb0 = rB0[d - 1], pb1 = rB1[d - 1];
            const float sa = sA[(d - 1) * 4 + nd.act], sb = sB[(d - 1) * 4 + nd.act];
            r[0][0] = nd.pad == 0 ? pa0 * sa : pa0;
            r[0][1] = nd.pad == 1 ? pa1 * sa : pa1;
            r[1][0] = nd.pad == 0 ? pb0 * sb : pb0;
            r[1][1] = nd.pad == 1 ? pb1 * sb : pb1;
        }
        if (nd.type == DN_DEC) {
            rA0[d] = r[0][0]; rA1[d] = r[0][1]; rB0[d] = r[1][0]; rB1[d] = r[1][1];
            float pa[4] = {0, 0, 0, 0}, pb[4] = {0, 0, 0, 0}, suma = 0.0f, sumb = 0.0f;
            const float* Rk = Rb + (size_t)nd.slot * np;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                if (a < nd.nact) {
                    if (live[0]) pa[a] = fmaxf(Rk[(size_t)a * np + kk[0]], 0.0f);
                    if (live[1]) pb[a] = fmaxf(Rk[(size_t)a * np + kk[1]], 0.0f);
                }
                suma += pa[a];
                sumb += pb[a];
            }
            const float ia = suma > 1e-7f ? 1.0f / suma : 0.0f, ib = sumb > 1e-7f ? 1.0f / sumb : 0.0f;
            const float un = 1.0f / (float)nd.nact;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                sA[d * 4 + a] = suma > 1e-7f ? pa[a] * ia : un;
                sB[d * 4 + a] = sumb > 1e-7f ? pb[a] * ib : un;
            }
            continue;
        }
        if (nd.type == DN_LEAF) {
            if (LEVEL == 0) {
#pragma unroll
                for (int q = 0; q < 2; ++q)
#pragma unroll
                    for (int p = 0; p < 2; ++p)
                        D.leafreach[(((size_t)b * D.nleaf + nd.leaf) * 2 + p) * NPT + kk[q]] = live[q] ? r[q][p] : 0.0f;
            }
            continue;
        }
        if (nd.type == DN_SD) {
#pragma unroll
            for (int q = 0; q < 2; ++q)
#pragma unroll
                for (int p = 0; p < 2; ++p) {
                    const float ro = live[q] ? r[q][1 - p] : 0.0f;  // player p's column holds the opponent's reach
                    const size_t col = (size_t)nd.col + p;
                    if (LEVEL == 1) {
                        const __hip_bfloat16 hi = __float2bfloat16(ro);
                        D.Bh[((size_t)b * L.ncolp + col) * NPR + kk[q]] = hi;
                        D.Bl[((size_t)b * L.ncolp + col) * NPR + kk[q]] = __float2bfloat16(ro - __bfloat162float(hi));
                    } else {
                        D.B0[((size_t)b * L.ncolp + col) * NPT + kk[q]] = ro;
                    }
                }
        }
        // fold or showdown: card sums of both players' reaches at this node
        const int tl = nd.sidx - t0;
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                if (!live[q]) continue;
                const unsigned long long v = (unsigned long long)__double2ull_rn((double)r[q][p] * FX);
                if (v) {
                    atomicAdd(&bins[(2 * tl + p) * 52 + cA[q]], v);
                    atomicAdd(&bins[(2 * tl + p) * 52 + cB[q]], v);
                }
            }
    }
    __syncthreads();
    double* out = L.sums + ((size_t)b * L.nterm + t0) * 2 * 56;
    for (int e = tid; e < ncolt * 53; e += NT) {
        const int col = e / 53, c = e % 53;
        double v;
        if (c < 52) v = (double)bins[col * 52 + c] / FX;
        else {
            unsigned long long t = 0;
            for (int x = 0; x < 52; ++x) t += bins[col * 52 + x];
            v = 0.5 * ((double)t / FX);
        }
        out[(size_t)col * 56 + c] = v;
    }
}

// ----------------------------------------------------------- backward ----
// Two workgroups per (board, template), one combo per thread.  Template
// nodes, DFS events and the template's card sums are staged in LDS; the
// regrets of all the template's slots are loaded up front.  Values flow up
// through register stacks; regret matching, the DCFR-generalised regret
// update and the strategy sum happen in the same pass.
template <int LEVEL, int DM>
__global__ __launch_bounds__(576) void k_bt_backward(BatchedDev D, const float* __restrict__ ranges, float wp,
                                                     float wn, float ws) {
    const LevelDev& L = D.L[LEVEL];
    extern __shared__ double ssum[];  // [maxterm][2][56]
    __shared__ DNode tn[MAXN];
    __shared__ int tev[MAXE];
    const int np = L.np;
    const int NT = blockDim.x;
    const int inst = blockIdx.x >> 1;
    const int b = inst / L.ntmpl, tm = inst % L.ntmpl;
    const int tid = threadIdx.x;
    const int k = (blockIdx.x & 1) * NT + tid;
    const bool live = k < L.nlive[b];
    const int n0 = L.tnode0[tm], nn = L.tnode0[tm + 1] - n0;
    const int e0 = L.tev0[tm], ne = L.tev0[tm + 1] - e0;
    const int s0 = L.tslot0[tm], nsl = L.tslot0[tm + 1] - s0;
    const int t0 = L.tterm0[tm], nt = L.tterm0[tm + 1] - t0;
    for (int e = tid; e < nn; e += NT) tn[e] = L.nodes[n0 + e];
    for (int e = tid; e < ne; e += NT) tev[e] = L.events[e0 + e];
    const double* gs = L.sums + ((size_t)b * L.nterm + t0) * 2 * 56;
    for (int e = tid; e < nt * 2 * 56; e += NT) ssum[e] = gs[e];
    const int c1 = L.cc[((size_t)b * np + k) * 2], c2 = L.cc[((size_t)b * np + k) * 2 + 1];
    // regrets of every slot of the template (old values, also used by the update)
    float Ra[32], Rb2[32];
    float* Rg = L.R + ((size_t)b * L.nslot + s0) * np + k;
    float* Sg = L.S + ((size_t)b * L.nslot + s0) * np + k;
#pragma unroll
    for (int x = 0; x < 32; ++x) {
        Ra[x] = (live && x < nsl) ? Rg[(size_t)x * np] : 0.0f;
        Rb2[x] = (live && x + 32 < nsl) ? Rg[(size_t)(x + 32) * np] : 0.0f;
    }
    float root0 = 0.0f, root1 = 0.0f;
    if (live) {
        root0 = root_reach(D, ranges, LEVEL, b, tm, np, k, 0);
        root1 = root_reach(D, ranges, LEVEL, b, tm, np, k, 1);
    }
    __syncthreads();
    float r0[DM], r1[DM], sg[DM * 4], cv[DM * 4], op[DM], oS[DM * 4];
    float out0 = 0.0f, out1 = 0.0f;
    for (int e = 0; e < ne; ++e) {
        const int ev = __builtin_amdgcn_readfirstlane(tev[e]);
        const int kind = ev & 3;
        const DNode nd = uniform_node(&tn[ev >> 2]);
        const int d = nd.depth;
        float x0, x1;  // reach of the node
        if (d == 0) { x0 = root0; x1 = root1; }
        else {
            const float sv = sg[(d - 1) * 4 + nd.act];
            x0 = nd.pad == 0 ? r0[d - 1] * sv : r0[d - 1];
            x1 = nd.pad == 1 ? r1[d - 1] * sv : r1[d - 1];
        }
        if (kind == EV_ENTER) {
            r0[d] = x0;
            r1[d] = x1;
            const int ls = nd.slot - s0;
            float pos[4], sum = 0.0f;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int q = ls + a;
                const float rv = a < nd.nact ? sel64(Ra, Rb2, q) : 0.0f;
                pos[a] = fmaxf(rv, 0.0f);
                sum += pos[a];
                oS[d * 4 + a] = (live && a < nd.nact) ? Sg[(size_t)q * np] : 0.0f;
            }
            const float inv = sum > 1e-7f ? 1.0f / sum : 0.0f, un = 1.0f / (float)nd.nact;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                sg[d * 4 + a] = a < nd.nact ? (sum > 1e-7f ? pos[a] * inv : un) : 0.0f;
                cv[d * 4 + a] = 0.0f;
            }
            op[d] = 0.0f;
            continue;
        }
        float v0 = 0.0f, v1 = 0.0f;
        if (kind == EV_TERM) {
            if (live) {
                if (nd.type == DN_FOLD) {
                    const int w = nd.player, l = 1 - w;
                    const double* sw = ssum + ((size_t)(nd.sidx - t0) * 2 + w) * 56;
                    const double* sl = ssum + ((size_t)(nd.sidx - t0) * 2 + l) * 56;
                    const float cw = w == 0 ? nd.c0 : nd.c1, cl = w == 0 ? nd.c1 : nd.c0;
                    const float xw = w == 0 ? x0 : x1, xl = w == 0 ? x1 : x0;
                    const float vw = (nd.pot - cw) * mass_of(sl, c1, c2, xl);
                    const float vl = -cl * mass_of(sw, c1, c2, xw);
                    v0 = w == 0 ? vw : vl;
                    v1 = w == 0 ? vl : vw;
                } else if (nd.type == DN_SD) {
                    const double* sa = ssum + ((size_t)(nd.sidx - t0) * 2 + 0) * 56;
                    const double* sb = ssum + ((size_t)(nd.sidx - t0) * 2 + 1) * 56;
                    const float m0 = mass_of(sb, c1, c2, x1), m1 = mass_of(sa, c1, c2, x0);
                    if (LEVEL == 1) {
                        const float* Cb = D.C + ((size_t)b * L.ncolp + nd.col) * NPR + k;
                        v0 = 0.5f * nd.pot * Cb[0] + 0.5f * D.pot0 * m0;
                        v1 = 0.5f * nd.pot * Cb[NPR] + 0.5f * D.pot0 * m1;
                    } else {
                        const float* Cb = D.C0 + ((size_t)b * L.ncolp + nd.col) * NPT + k;
                        v0 = nd.pot * Cb[0] - nd.c0 * m0;
                        v1 = nd.pot * Cb[NPT] - nd.c1 * m1;
                    }
                } else if (LEVEL == 0) {  // leaf: mean of the river-line roots over the 44 rivers
                    const int rt = D.leaf_tmpl[nd.leaf];
                    const int nt1 = D.L[1].ntmpl;
                    float a0 = 0.0f, a1 = 0.0f;
                    for (int rk = 0; rk < 48; ++rk) {
                        const int bb = b * 48 + rk;
                        const int ri = D.imap[(size_t)bb * NPT + k];
                        if (ri == 0xFFFF) continue;
                        const float* rv = D.rootval + (((size_t)bb * nt1 + rt) * 2) * NPR + ri;
                        a0 += rv[0];
                        a1 += rv[NPR];
                    }
                    v0 = a0 * (1.0f / 44.0f);
                    v1 = a1 * (1.0f / 44.0f);
                }
            }
        } else {  // EV_LEAVE of a decision node
            const int p = nd.player;
            float node_v = 0.0f;
#pragma unroll
            for (int a = 0; a < 4; ++a)
                if (a < nd.nact) node_v = fmaf(sg[d * 4 + a], cv[d * 4 + a], node_v);
            if (live) {
                const float rp = p == 0 ? x0 : x1;
                const int ls = nd.slot - s0;
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    if (a >= nd.nact) break;
                    const int q = ls + a;
                    const float rold = sel64(Ra, Rb2, q);
                    const float qv = rold + (cv[d * 4 + a] - node_v);
                    Rg[(size_t)q * np] = qv * (qv > 0.0f ? wp : wn);
                    Sg[(size_t)q * np] = fmaf(ws * rp, sg[d * 4 + a], oS[d * 4 + a]);
                }
            }
            v0 = p == 0 ? node_v : op[d];
            v1 = p == 1 ? node_v : op[d];
        }
        if (d == 0) {
            out0 = v0;
            out1 = v1;
        } else {
            const int pp = nd.pad;  // parent's actor
            cv[(d - 1) * 4 + nd.act] = pp == 0 ? v0 : v1;
            op[d - 1] += pp == 0 ? v1 : v0;
        }
    }
    if (k >= np) return;
    if (LEVEL == 1) {
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

// ------------------------------------------------------- river GEMM -------
// C[col][i] = sum_j S[i][j] r_col[j] for one board and a 128-hero block.
// MFMA operands: A = reach columns (hi or lo, from LDS), B = S^T generated
// per lane; D rows are columns, D cols are heroes, so stores coalesce.
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

constexpr int GK = 32;          // K chunk (two MFMA k-steps)
constexpr int GCOLS = 288;      // columns per board, padded (9 tiles of 32)
constexpr int GROW = 40;        // LDS row stride in bf16 (80 bytes)

__device__ __forceinline__ uint32_t bf16pair(int sa, int sb) {
    // sign in {-1,0,1} -> bf16 bits (1.0 = 0x3F80, -1.0 = 0xBF80)
    uint32_t a = sa > 0 ? 0x3F80u : (sa < 0 ? 0xBF80u : 0u);
    uint32_t b2 = sb > 0 ? 0x3F80u : (sb < 0 ? 0xBF80u : 0u);
    return a | (b2 << 16);
}

__device__ __forceinline__ bool conflict(uint32_t x, uint32_t y) {
    uint32_t v = x ^ y;  // a zero byte means a shared card
    return ((v - 0x01010101u) & ~v & 0x80808080u) != 0;
}

template <int NTILE>
__global__ __launch_bounds__(256) void k_bt_gemm(BatchedDev D) {
    constexpr int ncolp = NTILE * 32;
    __shared__ __attribute__((aligned(16))) __bf16 Bs[2][GCOLS][GROW];
    __shared__ uint16_t gj_s[GK];
    __shared__ uint32_t cy_s[GK];
    const int b = blockIdx.y;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = blockIdx.x * 128 + wave * 32 + (lane & 31);
    const bool row_ok = i < NPR;
    const int ic = row_ok ? i : 0;
    const uint16_t gi = D.grank[(size_t)b * NPR + ic];
    const uint32_t xi = D.cx[(size_t)b * NPR + ic];
    const int half = lane >> 5;
    f32x16_t acc[NTILE];
#pragma unroll
    for (int n = 0; n < NTILE; ++n)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[n][r] = 0.0f;
    const __hip_bfloat16* Bh = D.Bh + (size_t)b * ncolp * NPR;
    const __hip_bfloat16* Bl = D.Bl + (size_t)b * ncolp * NPR;
    for (int j0 = 0; j0 < NPR; j0 += GK) {
        // stage 32 K x ncolp columns (hi, lo): 4 x 16-byte pieces per column
        for (int e = threadIdx.x; e < ncolp * 8; e += 256) {
            int col = e >> 3, part = e & 7;  // part: hl*4 + quarter
            int hl = part >> 2, qt = part & 3;
            const __hip_bfloat16* src = (hl ? Bl : Bh) + (size_t)col * NPR + j0 + qt * 8;
            *reinterpret_cast<uint4*>(&Bs[hl][col][qt * 8]) = *reinterpret_cast<const uint4*>(src);
        }
        if (threadIdx.x < GK) {
            gj_s[threadIdx.x] = D.grank[(size_t)b * NPR + j0 + threadIdx.x];
            cy_s[threadIdx.x] = D.cy[(size_t)b * NPR + j0 + threadIdx.x];
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            // B operand: S^T[j][i] for j = j0 + 16s + 8*half + e
            uint32_t w[4];
#pragma unroll
            for (int e2 = 0; e2 < 4; ++e2) {
                int ja = 16 * s + 8 * half + 2 * e2;
                int sa = conflict(xi, cy_s[ja]) ? 0 : ((int)gi > (int)gj_s[ja]) - ((int)gi < (int)gj_s[ja]);
                int sb = conflict(xi, cy_s[ja + 1]) ? 0 : ((int)gi > (int)gj_s[ja + 1]) - ((int)gi < (int)gj_s[ja + 1]);
                w[e2] = bf16pair(sa, sb);
            }
            bf16x8_t bop;
            __builtin_memcpy(&bop, w, 16);
#pragma unroll
            for (int n = 0; n < NTILE; ++n) {
                const int col = n * 32 + (lane & 31);
                bf16x8_t ah, al;
                __builtin_memcpy(&ah, &Bs[0][col][16 * s + 8 * half], 16);
                __builtin_memcpy(&al, &Bs[1][col][16 * s + 8 * half], 16);
                acc[n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bop, acc[n], 0, 0, 0);
                acc[n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bop, acc[n], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    if (!row_ok) return;
    float* Cb = D.C + (size_t)b * ncolp * NPR;
#pragma unroll
    for (int n = 0; n < NTILE; ++n) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            int col = n * 32 + (r >> 2) * 8 + half * 4 + (r & 3);
            Cb[(size_t)col * NPR + i] = acc[n][r];
        }
    }
}

// Sign-blocker matrix of every river board as bytes, precomputed once.
// Byte = high byte of the bf16 of S/2: 0x3F (+0.5, i beats j), 0xBF (-0.5,
// j beats i), 0 (tie, blocked, padding); the low byte of +-0.5 is zero, so a
// byte widens to bf16 with one permute.  Layout [board][chunk c][i][32]: the
// 32 villains j = 32c..32c+31 of hero i are contiguous, and one chunk of 128
// heroes is one contiguous 4 KiB block.
__global__ void k_bt_s8(BatchedDev D, uint8_t* __restrict__ S8) {
    const int b = blockIdx.y, i = blockIdx.x;
    const int nl = D.L[1].nlive[b];
    const uint16_t gi = D.grank[(size_t)b * NPR + i];
    const uint32_t xi = D.cx[(size_t)b * NPR + i];
    uint8_t* base = S8 + (size_t)b * NPR * NPR + (size_t)i * 32;
    for (int j = threadIdx.x; j < NPR; j += blockDim.x) {
        uint8_t v = 0;
        if (i < nl && j < nl && !conflict(xi, D.cy[(size_t)b * NPR + j])) {
            const uint16_t gj = D.grank[(size_t)b * NPR + j];
            v = gi > gj ? 0x3F : (gi < gj ? 0xBF : 0);
        }
        base[(size_t)(j >> 5) * NPR * 32 + (j & 31)] = v;
    }
}

constexpr int GHB = (NPR + 127) / 128;   // hero blocks of 128 per board
constexpr int GNCH = NPR / 32;           // K chunks of 32 villains

// River showdown GEMM: C[b][col][i] = 2 * sum_j (S/2)[i][j] * (Bh + Bl)[b][col][j].
// One workgroup = 128 heroes x all columns of one board; wave w owns the
// 32-column tile w (NW = column tiles) and all four 32-hero tiles.  S chunks
// (128 heroes x 32 villains) are widened to bf16 once per workgroup into LDS
// (double-buffered); each wave streams its own reach columns (hi and lo) from
// HBM straight into MFMA A-fragments, one chunk ahead.  Within a chunk, lane
// half h of a fragment covers villains 16h..16h+15 (8 per MFMA k-step), a
// permutation of K shared by both operands.  Workgroups are numbered so that
// the hero blocks of one board run on one XCD (dispatch is round-robin over
// the 8 XCDs), where they share that board's reach columns in L2.
template <int NW>
__global__ __launch_bounds__(NW * 64) void k_bt_gemm3(BatchedDev D, const uint8_t* __restrict__ S8, int nlogical) {
    constexpr int NT = NW * 64;
    constexpr int SPT = (256 + NT - 1) / NT;   // 16-byte S pieces staged per thread
    __shared__ uint4 Ssh[2][4][2][2][32];      // [buf][hero tile][h][k-step][hero in tile]
    const int per = (nlogical + 7) >> 3;
    const int logical = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (logical >= nlogical) return;
    const int b = logical / GHB, i0 = (logical % GHB) * 128;
    const int nht = min(4, (NPR - i0) >> 5);
    const int npieces = nht * 64;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, half = lane >> 5, l32 = lane & 31;
    const int ncolp = NW * 32;
    const int col = wave * 32 + l32;
    const uint4* ah_p = reinterpret_cast<const uint4*>(D.Bh + ((size_t)b * ncolp + col) * NPR + 16 * half);
    const uint4* al_p = reinterpret_cast<const uint4*>(D.Bl + ((size_t)b * ncolp + col) * NPR + 16 * half);
    const uint4* s_p = reinterpret_cast<const uint4*>(S8 + (size_t)b * NPR * NPR + (size_t)i0 * 32);
    constexpr int CH4 = NPR * 2;               // uint4 pieces per S chunk (all heroes)
    f32x16_t acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[t][r] = 0.0f;
    uint4 sv[SPT];
    auto load_s = [&](int c) {
#pragma unroll
        for (int q = 0; q < SPT; ++q) {
            const int e = threadIdx.x + NT * q;
            if (e < npieces) sv[q] = s_p[(size_t)c * CH4 + e];
        }
    };
    auto store_s = [&](int buf) {
#pragma unroll
        for (int q = 0; q < SPT; ++q) {
            const int e = threadIdx.x + NT * q;
            if (e < npieces) {
                const int hl = e >> 1, h = e & 1;
                const uint4 w = sv[q];
                Ssh[buf][hl >> 5][h][0][hl & 31] =
                    make_uint4(__builtin_amdgcn_perm(0u, w.x, 0x010C000Cu), __builtin_amdgcn_perm(0u, w.x, 0x030C020Cu),
                               __builtin_amdgcn_perm(0u, w.y, 0x010C000Cu), __builtin_amdgcn_perm(0u, w.y, 0x030C020Cu));
                Ssh[buf][hl >> 5][h][1][hl & 31] =
                    make_uint4(__builtin_amdgcn_perm(0u, w.z, 0x010C000Cu), __builtin_amdgcn_perm(0u, w.z, 0x030C020Cu),
                               __builtin_amdgcn_perm(0u, w.w, 0x010C000Cu), __builtin_amdgcn_perm(0u, w.w, 0x030C020Cu));
            }
        }
    };
    uint4 a_cur[4], a_nxt[4];   // hi s0, hi s1, lo s0, lo s1
    load_s(0);
    a_cur[0] = ah_p[0]; a_cur[1] = ah_p[1]; a_cur[2] = al_p[0]; a_cur[3] = al_p[1];
    store_s(0);
    __syncthreads();
    int buf = 0;
    for (int c = 0; c < GNCH; ++c) {
        const bool more = c + 1 < GNCH;
        if (more) {
            load_s(c + 1);
            const int o = (c + 1) * 4;   // 32 villains = 4 uint4 of bf16 per row; this half's 16 at +0, +1
            a_nxt[0] = ah_p[o]; a_nxt[1] = ah_p[o + 1]; a_nxt[2] = al_p[o]; a_nxt[3] = al_p[o + 1];
        }
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            bf16x8_t ah, al;
            __builtin_memcpy(&ah, &a_cur[ks], 16);
            __builtin_memcpy(&al, &a_cur[2 + ks], 16);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                if (t < nht) {
                    bf16x8_t bop;
                    __builtin_memcpy(&bop, &Ssh[buf][t][half][ks][l32], 16);
                    acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bop, acc[t], 0, 0, 0);
                    acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bop, acc[t], 0, 0, 0);
                }
            }
        }
        if (more) store_s(buf ^ 1);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) a_cur[q] = a_nxt[q];
        buf ^= 1;
    }
    float* Cb = D.C + (size_t)b * ncolp * NPR;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        if (t < nht) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int cc = wave * 32 + (r >> 2) * 8 + half * 4 + (r & 3);
                Cb[(size_t)cc * NPR + i0 + 32 * t + l32] = 2.0f * acc[t][r];
            }
        }
    }
}

// Turn all-in showdowns: C0[t][col][i] = sum_j E[t][j][i] * B0[t][col][j].
template <int MAXC>
__global__ __launch_bounds__(256) void k_bt_turn_sd(BatchedDev D, int ncol) {
    const int t = blockIdx.y;
    const int i = blockIdx.x * 256 + threadIdx.x;
    __shared__ float bcol[MAXC][64];
    float acc[MAXC];
#pragma unroll
    for (int c = 0; c < MAXC; ++c) acc[c] = 0.0f;
    const float* E = D.Eturn + (size_t)t * NPT * NPT;
    const float* B = D.B0 + (size_t)t * D.L[0].ncolp * NPT;
    for (int j0 = 0; j0 < NPT; j0 += 64) {
        for (int e = threadIdx.x; e < MAXC * 64; e += 256) {
            int c = e / 64, jj = e % 64;
            bcol[c][jj] = c < ncol ? B[(size_t)c * NPT + j0 + jj] : 0.0f;
        }
        __syncthreads();
        if (i < NPT)
            for (int jj = 0; jj < 64; ++jj) {
                float ev = E[(size_t)(j0 + jj) * NPT + i];
#pragma unroll
                for (int c = 0; c < MAXC; ++c) acc[c] = fmaf(ev, bcol[c][jj], acc[c]);
            }
        __syncthreads();
    }
    if (i >= NPT) return;
    for (int c = 0; c < ncol; ++c) D.C0[((size_t)t * D.L[0].ncolp + c) * NPT + i] = acc[c];
}

// Turn equity over the 44 rivers in the turn-compact layout (villain-major).
__global__ void k_bt_eturn(const uint16_t* __restrict__ imap, const uint16_t* __restrict__ grank,
                           const uint8_t* __restrict__ tcc, const int* __restrict__ tlive, float* __restrict__ E) {
    const int t = blockIdx.z, j = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NPT) return;
    float v = 0.0f;
    const int nl = tlive[t];
    if (i < nl && j < nl) {
        const uint8_t* ci = tcc + ((size_t)t * NPT + i) * 2;
        const uint8_t* cj = tcc + ((size_t)t * NPT + j) * 2;
        bool blk = ci[0] == cj[0] || ci[0] == cj[1] || ci[1] == cj[0] || ci[1] == cj[1];
        if (!blk) {
            int w = 0, ti = 0, n = 0;
            for (int rk = 0; rk < 48; ++rk) {
                int bb = t * 48 + rk;
                int ri = imap[(size_t)bb * NPT + i], rj = imap[(size_t)bb * NPT + j];
                if (ri == 0xFFFF || rj == 0xFFFF) continue;
                int gi = grank[(size_t)bb * NPR + ri], gj = grank[(size_t)bb * NPR + rj];
                w += gi > gj;
                ti += gi == gj;
                n++;
            }
            v = n ? ((float)w + 0.5f * (float)ti) / (float)n : 0.0f;
        }
    }
    E[((size_t)t * NPT + j) * NPT + i] = v;
}

// Strategy sums into the reference layout [infoset][MTT_MAX_ACTIONS][1216].
struct ExportRec { int info, level, board, slot, nact; };
__global__ void k_bt_export(BatchedDev D, const ExportRec* __restrict__ recs, int n,
                            const uint16_t* __restrict__ inv0, const uint16_t* __restrict__ inv1,
                            float* __restrict__ out, int which) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= kPadFlop) return;
    for (int r = blockIdx.y; r < n; r += gridDim.y) {
        ExportRec e = recs[r];
        const LevelDev& L = D.L[e.level];
        const uint16_t* inv = (e.level == 0 ? inv0 : inv1) + (size_t)e.board * kPadFlop;
        int k = inv[tid];
        for (int a = 0; a < MTT_MAX_ACTIONS; ++a) {
            float v = 0.0f;
            if (a < e.nact && k != 0xFFFF) v = (which ? L.R : L.S)[((size_t)e.board * L.nslot + e.slot + a) * L.np + k];
            out[((size_t)e.info * MTT_MAX_ACTIONS + a) * kPadFlop + tid] = v;
        }
    }
}

}  // namespace

// ================================================================ host ===
struct BatchedState {
    BatchedDev D;
    DeviceArena arena;
    int two_level = 0;
    int nb1 = 0;          // river boards
    std::vector<ExportRec> export_recs;
    ExportRec* d_export = nullptr;
    uint16_t *d_inv0 = nullptr, *d_inv1 = nullptr;
    float* d_ranges = nullptr;  // root reach of the first street [boards][2][np]
    uint8_t* d_S8 = nullptr;    // [boards1][NPR][NPR] sign-blocker bytes
    int dm[2] = {8, 8};         // register-stack depth per level (4, 6 or 8)
};

void batched_destroy(BatchedState* b) { delete b; }

namespace {

struct LevelHost {
    int maxdepth = 0;  // deepest decision node
    std::vector<DNode> nodes;
    std::vector<int> tnode0{0}, tev0{0}, events;
    int nslot = 0, nterm = 0, ncol = 0;
    std::vector<int> tmpl_ids;          // HuTree template ids at this level
    std::vector<int> slot0, term0, col0;
};

int add_template(const Template& T, LevelHost& L, int leaf_base) {
    int base_slot = L.nslot, base_term = L.nterm, base_col = L.ncol;
    for (const TNode& t : T.nodes)
        if (t.type == T_DECISION) L.maxdepth = std::max(L.maxdepth, (int)t.depth);
    L.slot0.push_back(base_slot);
    L.term0.push_back(base_term);
    L.col0.push_back(base_col);
    int nterm = 0;
    for (size_t n = 0; n < T.nodes.size(); ++n) {
        const TNode& t = T.nodes[n];
        DNode d;
        memset(&d, 0, sizeof(d));
        d.type = t.type == T_DECISION ? DN_DEC : t.type == T_FOLD ? DN_FOLD : t.type == T_SHOWDOWN ? DN_SD : DN_LEAF;
        d.player = t.player;
        d.nact = t.nact;
        d.depth = t.depth;
        if (t.depth >= 8) { set_error("template depth %d exceeds 7", t.depth); return MTT_ERR_TREE; }
        d.act = t.act;
        d.slot = (int16_t)(t.type == T_DECISION ? base_slot + t.slot : -1);
        d.col = -1;
        d.sidx = -1;
        d.leaf = -1;
        d.pad = (int16_t)(t.parent >= 0 ? T.nodes[t.parent].player : -1);
        if (t.type == T_SHOWDOWN) d.col = (int16_t)(base_col + 2 * t.term);
        if (t.type == T_FOLD || t.type == T_SHOWDOWN) d.sidx = (int16_t)(base_term + nterm++);
        if (t.type == T_LEAF) d.leaf = (int16_t)(leaf_base + t.term);
        d.pot = t.pot;
        d.c0 = t.c[0];
        d.c1 = t.c[1];
        L.nodes.push_back(d);
    }
    // DFS events in preorder (children contiguous after their parent)
    std::vector<int> ev;
    std::function<void(int)> dfs = [&](int n) {
        const TNode& t = T.nodes[n];
        if (t.type != T_DECISION) { ev.push_back((n << 2) | EV_TERM); return; }
        ev.push_back((n << 2) | EV_ENTER);
        for (int a = 0; a < t.nact; ++a) dfs(t.first_child + a);
        ev.push_back((n << 2) | EV_LEAVE);
    };
    dfs(0);
    // forward order: nodes must be in preorder with parents before children;
    // template nodes are allocated parent-first, and the forward walk needs a
    // node's parent to be the latest node on the stack at depth-1, i.e. DFS
    // preorder.  Reorder nodes into DFS preorder.
    std::vector<int> pre;
    for (int x : ev)
        if ((x & 3) != EV_LEAVE) pre.push_back(x >> 2);
    std::vector<int> newidx(T.nodes.size());
    for (size_t q = 0; q < pre.size(); ++q) newidx[pre[q]] = (int)q;
    int nbase = L.tnode0.back();
    std::vector<DNode> reordered(T.nodes.size());
    for (size_t n = 0; n < T.nodes.size(); ++n) reordered[newidx[n]] = L.nodes[nbase + n];
    std::copy(reordered.begin(), reordered.end(), L.nodes.begin() + nbase);
    for (int x : ev) L.events.push_back((newidx[x >> 2] << 2) | (x & 3));
    L.nslot += T.n_slots;
    L.nterm += nterm;
    L.ncol += 2 * T.n_showdown;
    L.tnode0.push_back((int)L.nodes.size());
    L.tev0.push_back((int)L.events.size());
    return MTT_OK;
}

}  // namespace

int batched_create(SolverImpl& s) {
    const mtt_hu_game& g = s.game;
    const HuTree& tr = s.tree;
    bool two = g.n_board == 3 && g.root_chance == 1 && g.last_street == 2;
    bool one = g.n_board == 5;
    MTT_ARG(two || one, "batched mode supports river games (5-card board) and turn+river games from a flop");
    auto st = new (std::nothrow) BatchedState;
    if (!st) { set_error("out of host memory"); return MTT_ERR_OOM; }
    s.batched = st;
    BatchedState& B = *st;
    B.two_level = two;
    memset(&B.D, 0, sizeof(B.D));
    B.D.two_level = two;
    B.D.pot0 = g.pot;
    DeviceArena& A = B.arena;
    // ---------------------------------------------------------- levels --
    LevelHost H0, H1;
    int rc = MTT_OK;
    std::vector<int> leaf_tmpl, tmpl_leaf;
    if (two) {
        const Template& T0 = tr.templates[0];
        rc = add_template(T0, H0, 0);
        if (rc) return rc;
        for (int l = 0; l < T0.n_leaf; ++l) {
            rc = add_template(tr.templates[T0.leaf_next[l]], H1, 0);
            if (rc) return rc;
            leaf_tmpl.push_back(l);
            tmpl_leaf.push_back(l);
        }
    } else {
        rc = add_template(tr.templates[0], H1, 0);
        if (rc) return rc;
    }
    // ---------------------------------------------------------- boards --
    const int nb1 = two ? 49 * 48 : 1;
    B.nb1 = nb1;
    std::vector<uint16_t> rank1((size_t)nb1 * NPR, 0), rmap((size_t)nb1 * NPR, 0xFFFF);
    std::vector<uint32_t> cx((size_t)nb1 * NPR, 0x80808080u), cy((size_t)nb1 * NPR, 0x40404040u);
    std::vector<uint8_t> cc1((size_t)nb1 * NPR * 2, 0);
    std::vector<int> nlive1(nb1);
    std::vector<uint16_t> inv1((size_t)nb1 * kPadFlop, 0xFFFF);
    // turn level (two-level games)
    std::vector<uint16_t> tcomp_of_root(two ? 49 * kPadFlop : 0, 0xFFFF);  // root-alive -> turn compact
    std::vector<uint16_t> troot((size_t)(two ? 49 : 0) * NPT, 0xFFFF);     // turn compact -> root-alive
    std::vector<uint8_t> cc0((size_t)(two ? 49 : 0) * NPT * 2, 0);
    std::vector<int> nlive0(two ? 49 : 0);
    std::vector<uint16_t> imap((size_t)(two ? nb1 : 0) * NPT, 0xFFFF);
    const int na = s.n_alive;
    if (two) {
        for (int t = 0; t < 49; ++t) {
            int tc = s.deck[t];
            int n = 0;
            std::vector<int> fill(52, 0);
            for (int i = 0; i < na; ++i) {
                if (s.c1[i] == tc || s.c2[i] == tc) continue;
                troot[(size_t)t * NPT + n] = (uint16_t)i;
                tcomp_of_root[(size_t)t * kPadFlop + i] = (uint16_t)n;
                cc0[((size_t)t * NPT + n) * 2] = s.c1[i];
                cc0[((size_t)t * NPT + n) * 2 + 1] = s.c2[i];
                n++;
            }
            nlive0[t] = n;
        }
    }
    for (int bi = 0; bi < nb1; ++bi) {
        uint8_t bd[5];
        int t = -1, r = -1;
        for (int k = 0; k < g.n_board; ++k) bd[k] = g.board[k];
        if (two) {
            t = bi / 48;
            int rk = bi % 48;
            r = rk < t ? rk : rk + 1;
            bd[3] = s.deck[t];
            bd[4] = s.deck[r];
        }
        uint64_t bm = 0, m16 = 0;
        for (int k = 0; k < 5; ++k) { bm |= 1ull << bd[k]; m16 |= card_bits16(bd[k]); }
        std::vector<std::pair<uint32_t, int>> sc;
        for (int i = 0; i < na; ++i) {
            uint64_t cm = (1ull << s.c1[i]) | (1ull << s.c2[i]);
            if (cm & bm) continue;
            sc.push_back({score_packed(m16 | card_bits16(s.c1[i]) | card_bits16(s.c2[i])), i});
        }
        std::stable_sort(sc.begin(), sc.end());
        int n = (int)sc.size();
        nlive1[bi] = n;
        std::vector<int> fill(52, 0);
        int grp = 0;
        for (int k = 0; k < n; ++k) {
            if (k > 0 && sc[k].first != sc[k - 1].first) grp++;
            int i = sc[k].second;
            size_t o = (size_t)bi * NPR + k;
            rank1[o] = (uint16_t)(grp + 1);
            uint32_t a = s.c1[i], c = s.c2[i];
            cx[o] = a | (c << 8) | (a << 16) | (c << 24);
            cy[o] = a | (a << 8) | (c << 16) | (c << 24);
            cc1[o * 2] = (uint8_t)a;
            cc1[o * 2 + 1] = (uint8_t)c;
            inv1[(size_t)bi * kPadFlop + i] = (uint16_t)k;
            if (two) {
                uint16_t tcpt = tcomp_of_root[(size_t)t * kPadFlop + i];
                rmap[o] = tcpt;
                imap[(size_t)bi * NPT + tcpt] = (uint16_t)k;
            }
        }
    }
    // --------------------------------------------------- device arrays --
    auto up_level = [&](LevelHost& H, LevelDev& L, int np, int nboards, const std::vector<uint8_t>& cc,
                        const std::vector<int>& nlive) -> int {
        L.np = np;
        L.nboards = nboards;
        L.nslot = H.nslot;
        L.nterm = H.nterm;
        L.ncol = H.ncol;
        L.ncolp = ((H.ncol + 31) / 32) * 32;
        if (L.ncolp == 0) L.ncolp = 32;
        L.ntmpl = (int)H.tnode0.size() - 1;
        int r = MTT_OK;
        DNode* dn;
        int *t0, *e0, *ev, *nl, *ts, *tt;
        uint8_t* c;
        std::vector<int> slot0 = H.slot0, term0 = H.term0;
        slot0.push_back(H.nslot);
        term0.push_back(H.nterm);
        L.maxterm = 0;
        for (int t = 0; t < L.ntmpl; ++t) {
            L.maxterm = std::max(L.maxterm, term0[t + 1] - term0[t]);
            MTT_ARG(H.tnode0[t + 1] - H.tnode0[t] <= MAXN, "template has more than %d nodes", MAXN);
            MTT_ARG(H.tev0[t + 1] - H.tev0[t] <= MAXE, "template has more than %d events", MAXE);
            MTT_ARG(slot0[t + 1] - slot0[t] <= MAXSLOT, "template has more than %d action slots", MAXSLOT);
        }
        r = r ? r : A.upload(H.nodes, &dn);
        r = r ? r : A.upload(H.tnode0, &t0);
        r = r ? r : A.upload(H.tev0, &e0);
        r = r ? r : A.upload(H.events, &ev);
        r = r ? r : A.upload(slot0, &ts);
        r = r ? r : A.upload(term0, &tt);
        r = r ? r : A.upload(cc, &c);
        r = r ? r : A.upload(nlive, &nl);
        if (r) return r;
        L.nodes = dn; L.tnode0 = t0; L.tev0 = e0; L.events = ev; L.tslot0 = ts; L.tterm0 = tt; L.cc = c; L.nlive = nl;
        size_t ns = (size_t)nboards * L.nslot * np;
        r = r ? r : A.alloc(ns, &L.R);
        r = r ? r : A.alloc(ns, &L.S);
        r = r ? r : A.alloc((size_t)nboards * std::max(L.nterm, 1) * 2 * 56, &L.sums);
        if (r) return r;
        MTT_HIP(hipMemsetAsync(L.R, 0, ns * sizeof(float), s.stream));
        MTT_HIP(hipMemsetAsync(L.S, 0, ns * sizeof(float), s.stream));
        return MTT_OK;
    };
    auto pick = [](int d) { return d + 1 <= 4 ? 4 : (d + 1 <= 6 ? 6 : 8); };
    B.dm[0] = pick(H0.maxdepth);
    B.dm[1] = pick(H1.maxdepth);
    rc = up_level(H1, B.D.L[1], NPR, nb1, cc1, nlive1);
    if (rc) return rc;
    LevelDev& L1 = B.D.L[1];
    MTT_ARG(L1.ncolp <= GCOLS, "river showdown columns %d exceed %d", L1.ncol, GCOLS);
    {
        uint16_t *d_rank, *d_rmap, *d_imap;
        uint32_t *d_cx, *d_cy;
        rc = rc ? rc : A.upload(rank1, &d_rank);
        rc = rc ? rc : A.upload(cx, &d_cx);
        rc = rc ? rc : A.upload(cy, &d_cy);
        rc = rc ? rc : A.upload(rmap, &d_rmap);
        if (rc) return rc;
        B.D.grank = d_rank; B.D.cx = d_cx; B.D.cy = d_cy; B.D.rmap = d_rmap;
        size_t nbc = (size_t)nb1 * L1.ncolp * NPR;
        rc = rc ? rc : A.alloc(nbc, &B.D.Bh);
        rc = rc ? rc : A.alloc(nbc, &B.D.Bl);
        rc = rc ? rc : A.alloc(nbc, &B.D.C);
        if (rc) return rc;
        rc = A.alloc((size_t)nb1 * NPR * NPR, &B.d_S8);
        if (rc) return rc;
        hipLaunchKernelGGL(k_bt_s8, dim3(NPR, nb1), dim3(256), 0, s.stream, B.D, B.d_S8);
        MTT_HIP(hipGetLastError());
        MTT_HIP(hipMemsetAsync(B.D.Bh, 0, nbc * 2, s.stream));
        MTT_HIP(hipMemsetAsync(B.D.Bl, 0, nbc * 2, s.stream));
        MTT_HIP(hipMemsetAsync(B.D.C, 0, nbc * 4, s.stream));
        rc = A.upload(inv1, &B.d_inv1);
        if (rc) return rc;
        if (two) {
            rc = A.upload(imap, &d_imap);
            if (rc) return rc;
            B.D.imap = d_imap;
        }
    }
    if (two) {
        rc = up_level(H0, B.D.L[0], NPT, 49, cc0, nlive0);
        if (rc) return rc;
        LevelDev& L0 = B.D.L[0];
        MTT_ARG(L0.ncol <= 16, "turn showdown columns %d exceed 16", L0.ncol);
        B.D.nleaf = (int)leaf_tmpl.size();
        int *d_lt, *d_tl;
        rc = rc ? rc : A.upload(leaf_tmpl, &d_lt);
        rc = rc ? rc : A.upload(tmpl_leaf, &d_tl);
        rc = rc ? rc : A.alloc((size_t)49 * B.D.nleaf * 2 * NPT, &B.D.leafreach);
        rc = rc ? rc : A.alloc((size_t)nb1 * L1.ntmpl * 2 * NPR, &B.D.rootval);
        rc = rc ? rc : A.alloc((size_t)49 * 2 * NPT, &B.D.rootval0);
        rc = rc ? rc : A.alloc((size_t)49 * L0.ncolp * NPT, &B.D.B0);
        rc = rc ? rc : A.alloc((size_t)49 * L0.ncolp * NPT, &B.D.C0);
        float* d_E;
        rc = rc ? rc : A.alloc((size_t)49 * NPT * NPT, &d_E);
        if (rc) return rc;
        B.D.leaf_tmpl = d_lt; B.D.tmpl_leaf = d_tl; B.D.Eturn = d_E;
        MTT_HIP(hipMemsetAsync(B.D.leafreach, 0, (size_t)49 * B.D.nleaf * 2 * NPT * 4, s.stream));
        MTT_HIP(hipMemsetAsync(B.D.B0, 0, (size_t)49 * L0.ncolp * NPT * 4, s.stream));
        // turn-compact -> root inverse, and the turn equity matrices
        std::vector<uint16_t> inv0((size_t)49 * kPadFlop, 0xFFFF);
        for (int t = 0; t < 49; ++t)
            for (int i = 0; i < kPadFlop; ++i) inv0[(size_t)t * kPadFlop + i] = tcomp_of_root[(size_t)t * kPadFlop + i];
        rc = A.upload(inv0, &B.d_inv0);
        if (rc) return rc;
        // root reach of each turn board: the ranges in turn-compact order
        std::vector<float> rr((size_t)49 * 2 * NPT, 0.0f);
        for (int t = 0; t < 49; ++t)
            for (int k = 0; k < nlive0[t]; ++k)
                for (int p = 0; p < 2; ++p) rr[((size_t)t * 2 + p) * NPT + k] = s.range[p][troot[(size_t)t * NPT + k]];
        rc = A.upload(rr, &B.d_ranges);
        if (rc) return rc;
        int* d_nl0 = const_cast<int*>(L0.nlive);
        hipLaunchKernelGGL(k_bt_eturn, dim3(NPT / 256 + 1, NPT, 49), dim3(256), 0, s.stream, B.D.imap, B.D.grank,
                           L0.cc, d_nl0, d_E);
        MTT_HIP(hipGetLastError());
    } else {
        // single board: ranges in the strength-sorted compact order, root values
        std::vector<float> rr(2 * NPR, 0.0f);
        for (int p = 0; p < 2; ++p)
            for (int i = 0; i < na; ++i) {
                int k = inv1[i];
                if (k != 0xFFFF) rr[p * NPR + k] = s.range[p][i];
            }
        rc = A.upload(rr, &B.d_ranges);
        rc = rc ? rc : A.alloc((size_t)2 * NPR, &B.D.rootval0);
        if (rc) return rc;
    }
    // ------------------------------------------------- strategy export --
    // map every flat decision node to (level, board, slot)
    const auto& F = tr.flat;
    for (int id = 0; id < (int)F.size(); ++id) {
        if (F[id].type != MTT_NODE_DECISION) continue;
        int tmpl = tr.flat_tmpl[id], tn = tr.flat_tnode[id];
        const Template& T = tr.templates[tmpl];
        int code = F[id].river_idx;
        int d1 = (code & 63) - 1, d2 = (code >> 6) - 1;
        ExportRec e;
        e.info = F[id].infoset_offset;
        e.nact = F[id].num_actions;
        if (two && T.street == 1) {
            e.level = 0;
            e.board = d1;
            e.slot = H0.slot0[0] + T.nodes[tn].slot;
        } else {
            e.level = 1;
            int li = 0;
            if (two) {
                int rk = d2 < d1 ? d2 : d2 - 1;
                e.board = d1 * 48 + rk;
                // template index among river templates
                const Template& T0 = tr.templates[0];
                for (int l = 0; l < T0.n_leaf; ++l)
                    if (T0.leaf_next[l] == tmpl) li = l;
            } else {
                e.board = 0;
            }
            e.slot = H1.slot0[li] + T.nodes[tn].slot;
        }
        B.export_recs.push_back(e);
    }
    rc = A.upload(B.export_recs, &B.d_export);
    if (rc) return rc;
    MTT_HIP(hipStreamSynchronize(s.stream));
    return MTT_OK;
}

int batched_run(SolverImpl& s, int iters) {
    BatchedState& B = *s.batched;
    const LevelDev& L1 = B.D.L[1];
    for (int it = 0; it < iters; ++it) {
        int t = s.iters_done + 1;
        float wp, wn, ws;
        iteration_weights(s.cfg, t, wp, wn, ws);
        const float* ranges = B.d_ranges;
        if (B.two_level) {
            const LevelDev& L0 = B.D.L[0];
            const size_t sh = (size_t)2 * L0.maxterm * 52 * 8;
            MTT_TIMED(s, "turn_forward",
                      if (B.dm[0] == 4) hipLaunchKernelGGL((k_bt_forward<0, 4>), dim3(49 * L0.ntmpl), dim3(NPT / 2), sh, s.stream, B.D, ranges);
                      else if (B.dm[0] == 6) hipLaunchKernelGGL((k_bt_forward<0, 6>), dim3(49 * L0.ntmpl), dim3(NPT / 2), sh, s.stream, B.D,
                                                                 ranges);
                      else hipLaunchKernelGGL((k_bt_forward<0, 8>), dim3(49 * L0.ntmpl), dim3(NPT / 2), sh, s.stream, B.D, ranges));
        }
        {
            const size_t sh = (size_t)2 * L1.maxterm * 52 * 8;
            MTT_TIMED(s, "river_forward",
                      if (B.dm[1] == 4) hipLaunchKernelGGL((k_bt_forward<1, 4>), dim3(B.nb1 * L1.ntmpl), dim3(NPR / 2), sh, s.stream, B.D, ranges);
                      else if (B.dm[1] == 6) hipLaunchKernelGGL((k_bt_forward<1, 6>), dim3(B.nb1 * L1.ntmpl), dim3(NPR / 2), sh, s.stream, B.D,
                                                                 ranges);
                      else hipLaunchKernelGGL((k_bt_forward<1, 8>), dim3(B.nb1 * L1.ntmpl), dim3(NPR / 2), sh, s.stream, B.D, ranges));
        }
        ::mtt::KernelTimer::Mark gm;
        if (s.timer.enabled) s.timer.begin("river_showdown_gemm", s.stream, gm);
        {
            const int nlog = B.nb1 * GHB;
            const dim3 g(((nlog + 7) / 8) * 8);
            switch (L1.ncolp / 32) {
            case 1: hipLaunchKernelGGL(k_bt_gemm3<1>, g, dim3(64), 0, s.stream, B.D, B.d_S8, nlog); break;
            case 2: hipLaunchKernelGGL(k_bt_gemm3<2>, g, dim3(128), 0, s.stream, B.D, B.d_S8, nlog); break;
            case 3: hipLaunchKernelGGL(k_bt_gemm3<3>, g, dim3(192), 0, s.stream, B.D, B.d_S8, nlog); break;
            case 4: hipLaunchKernelGGL(k_bt_gemm3<4>, g, dim3(256), 0, s.stream, B.D, B.d_S8, nlog); break;
            case 5: hipLaunchKernelGGL(k_bt_gemm3<5>, g, dim3(320), 0, s.stream, B.D, B.d_S8, nlog); break;
            case 6: hipLaunchKernelGGL(k_bt_gemm3<6>, g, dim3(384), 0, s.stream, B.D, B.d_S8, nlog); break;
            case 7: hipLaunchKernelGGL(k_bt_gemm3<7>, g, dim3(448), 0, s.stream, B.D, B.d_S8, nlog); break;
            case 8: hipLaunchKernelGGL(k_bt_gemm3<8>, g, dim3(512), 0, s.stream, B.D, B.d_S8, nlog); break;
            default: hipLaunchKernelGGL(k_bt_gemm3<9>, g, dim3(576), 0, s.stream, B.D, B.d_S8, nlog); break;
            }
        }
        if (s.timer.enabled) s.timer.end(gm, s.stream);
        {
            const size_t sh = (size_t)L1.maxterm * 2 * 56 * 8;
            MTT_TIMED(s, "river_backward",
                      if (B.dm[1] == 4) hipLaunchKernelGGL((k_bt_backward<1, 4>), dim3(B.nb1 * L1.ntmpl * 2), dim3(NPR / 2), sh, s.stream, B.D, ranges, wp,
                                                           wn, ws);
                      else if (B.dm[1] == 6) hipLaunchKernelGGL((k_bt_backward<1, 6>), dim3(B.nb1 * L1.ntmpl * 2), dim3(NPR / 2), sh, s.stream, B.D,
                                                                 ranges, wp, wn, ws);
                      else hipLaunchKernelGGL((k_bt_backward<1, 8>), dim3(B.nb1 * L1.ntmpl * 2), dim3(NPR / 2), sh, s.stream, B.D, ranges, wp, wn,
                                              ws));
        }
        if (B.two_level) {
            const LevelDev& L0 = B.D.L[0];
            MTT_TIMED(s, "turn_showdown",
                      hipLaunchKernelGGL(k_bt_turn_sd<16>, dim3(NPT / 256 + 1, 49), dim3(256), 0, s.stream, B.D,
                                         L0.ncol));
            const size_t sh = (size_t)L0.maxterm * 2 * 56 * 8;
            MTT_TIMED(s, "turn_backward",
                      if (B.dm[0] == 4) hipLaunchKernelGGL((k_bt_backward<0, 4>), dim3(49 * L0.ntmpl * 2), dim3(NPT / 2), sh, s.stream, B.D, ranges, wp,
                                                           wn, ws);
                      else if (B.dm[0] == 6) hipLaunchKernelGGL((k_bt_backward<0, 6>), dim3(49 * L0.ntmpl * 2), dim3(NPT / 2), sh, s.stream, B.D,
                                                                 ranges, wp, wn, ws);
                      else hipLaunchKernelGGL((k_bt_backward<0, 8>), dim3(49 * L0.ntmpl * 2), dim3(NPT / 2), sh, s.stream, B.D, ranges, wp, wn,
                                              ws));
        }
        MTT_HIP(hipGetLastError());
        s.iters_done = t;
    }
    return MTT_OK;
}

int batched_export(SolverImpl& s, float* d_out, int which) {
    BatchedState& B = *s.batched;
    int n = (int)B.export_recs.size();
    if (!n) return MTT_OK;
    MTT_HIP(hipMemsetAsync(d_out, 0, (size_t)s.tree.n_infosets * MTT_MAX_ACTIONS * kPadFlop * sizeof(float), s.stream));
    hipLaunchKernelGGL(k_bt_export, dim3(kPadFlop / 256 + 1, std::min(n, 65535)), dim3(256), 0, s.stream, B.D,
                       B.d_export, n, B.d_inv0, B.d_inv1, d_out, which);
    MTT_HIP(hipGetLastError());
    MTT_HIP(hipStreamSynchronize(s.stream));
    return MTT_OK;
}

int batched_strategy_sums(SolverImpl& s, float* d_S_ref) { return batched_export(s, d_S_ref, 0); }

}  // namespace mtt
