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
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>

#include <functional>

#include <hip/hip_bf16.h>

#include "bt_dev.h"
#include "bt_jit.h"
#include "bt_showdown.h"
#include "solver.h"

namespace mtt {

using namespace bt;

namespace {

constexpr int MAXN = 128;      // nodes per template
constexpr int MAXE = 256;      // DFS events per template
constexpr int MAXSLOT = 96;    // action slots per template (three 32-register vectors)

typedef float f32x32_t __attribute__((ext_vector_type(32)));

// Event records are staged in LDS per workgroup; the walk keeps the next
// record in vector registers (issued one event ahead) and makes the current
// one wave-uniform with readfirstlane.
struct EvRaw {
    int4 a, b;
};
__device__ __forceinline__ EvRaw ev_fetch(const EvRec* lds, int e) {
    const int4* q = reinterpret_cast<const int4*>(lds + e);
    EvRaw r;
    r.a = q[0];
    r.b = q[1];
    return r;
}
__device__ __forceinline__ EvRec ev_uniform(const EvRaw& r) {
    int v[8] = {r.a.x, r.a.y, r.a.z, r.a.w, r.b.x, r.b.y, r.b.z, r.b.w};
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = __builtin_amdgcn_readfirstlane(v[k]);
    EvRec e;
    __builtin_memcpy(&e, v, sizeof(e));
    return e;
}
__device__ __forceinline__ void ev_stage(EvRec* lds, const EvRec* __restrict__ g, int n) {
    const int4* src = reinterpret_cast<const int4*>(g);
    int4* dst = reinterpret_cast<int4*>(lds);
    for (int e = threadIdx.x; e < 2 * n; e += blockDim.x) dst[e] = src[e];
}

// Element q (wave-uniform, < 32 * NRV) of a register file held as NRV
// 32-register vectors a, b (, c): the reads share one indexed-register
// window.
template <int NRV>
__device__ __forceinline__ float rselv(const f32x32_t& a, const f32x32_t& b, const f32x32_t& c, int q) {
    const float va = a[q & 31], vb = b[q & 31];
    if constexpr (NRV == 3) {
        const float vc = c[q & 31];
        return q < 32 ? va : (q < 64 ? vb : vc);
    } else {
        return q < 32 ? va : vb;
    }
}

constexpr int SWEEP_NT = 576;   // threads per sweep workgroup; two workgroups cover one board's combos

// ------------------------------------------------------------ forward ----
// Two workgroups per (board, template), one combo per thread.  The regrets
// of every slot of the template are loaded up front into an indexed
// register file; the walk over the template's preorder events reads node
// records through scalar loads (one event ahead).  At each decision node:
// regret matching, reach propagation, and the strategy-sum update
// S += w * reach * sigma as a no-return float atomic (one writer per entry).
// At each terminal the reach of both players is added into per-card bins in
// LDS with 64-bit fixed-point atomics (exact, order independent); each
// workgroup writes its half's bins, and the backward adds the two halves.
//
// The body takes the (board, template) instance and the half explicitly, and
// walks its HALF combos (NT, or 640 in the persistent flop kernel) in PASSES
// slices of NT into the same bins (the persistent flop kernel runs a 640-combo half as three
// 256-thread passes, the last one partly masked; the bins are exact integer
// sums, so the result does not depend on it).
template <int LEVEL, int DM, int NT, int NRV, int PASSES>
__device__ __forceinline__ void bt_forward_body(const BatchedDev& D, const float* __restrict__ ranges, float ws,
                                                const float* __restrict__ wdev, int inst, int half, EvRec* sev,
                                                unsigned long long* bins /* [nt][2][52] */) {
    typedef float vS __attribute__((ext_vector_type(DM * 4)));
    typedef float vD __attribute__((ext_vector_type(DM)));
    const LevelDev& L = D.L[LEVEL];
    const int b = LEVEL == 2 ? inst : inst / L.ntmpl;
    const int tm = LEVEL == 2 ? ldu(L.btmpl + b) : inst % L.ntmpl;
    if (wdev) ws = wdev[2];
    const int tid = threadIdx.x;
    const int np = L.np;
    const int e0 = ldu(L.tfev0 + tm), e1 = ldu(L.tfev0 + tm + 1);
    const int s0 = ldu(L.tslot0 + tm), nsl = ldu(L.tslot0 + tm + 1) - s0;
    const int t0 = ldu(L.tterm0 + tm), nt = ldu(L.tterm0 + tm + 1) - t0;
    // flop instances keep per-board state for their own template only
    const int sb = LEVEL == 2 ? 0 : s0, tb = LEVEL == 2 ? 0 : t0, cb0 = LEVEL == 2 ? ldu(L.tcol0 + tm) : 0;
    fwd_bins_clear<NT>(bins, nt);
    ev_stage(sev, L.fev + e0, e1 - e0);
    constexpr int HALF = PASSES == 1 ? NT : 640;   // combos per half
    static_assert(PASSES * NT >= HALF, "passes cover the half");
    for (int pass = 0; pass < PASSES; ++pass) {
    const bool mine = pass * NT + tid < HALF;
    const int k = half * HALF + (mine ? pass * NT + tid : 0);
    const bool inb = mine && k < np;
    const bool live = mine && k < ldu(L.nlive + b);
    int c1 = 0, c2 = 0;
    float root0 = 0.0f, root1 = 0.0f;
    f32x32_t Ra, Rb, Rc;
    const size_t rbase = ((size_t)b * L.nslot + sb) * np + k;
#pragma unroll
    for (int x = 0; x < 32; ++x) {
        Ra[x] = (live && x < nsl) ? L.R[rbase + (size_t)x * np] : 0.0f;
        Rb[x] = (live && x + 32 < nsl) ? L.R[rbase + (size_t)(x + 32) * np] : 0.0f;
        if constexpr (NRV == 3) Rc[x] = (live && x + 64 < nsl) ? L.R[rbase + (size_t)(x + 64) * np] : 0.0f;
    }
    if (live) {
        c1 = L.cc[((size_t)b * np + k) * 2];
        c2 = L.cc[((size_t)b * np + k) * 2 + 1];
        root0 = root_reach(D, ranges, LEVEL, b, tm, np, k, 0);
        root1 = root_reach(D, ranges, LEVEL, b, tm, np, k, 1);
    }
    if (pass == 0) __syncthreads();   // bins cleared, events staged
    vD rs0, rs1;
    vS sg;
    const int ne = e1 - e0;
    EvRaw nx = ev_fetch(sev, 0);
    for (int e = 0; e < ne; ++e) {
        const EvRec ec = ev_uniform(nx);
        if (e + 1 < ne) nx = ev_fetch(sev, e + 1);
        const DNode& nd = ec.nd;
        const int d = nd.depth;
        float x0, x1;
        if (d == 0) {
            x0 = root0;
            x1 = root1;
        } else {
            const float sv = sg[(d - 1) * 4 + nd.act];
            const float p0 = rs0[d - 1], p1 = rs1[d - 1];
            x0 = nd.pad == 0 ? p0 * sv : p0;
            x1 = nd.pad == 1 ? p1 * sv : p1;
        }
        if (nd.type == DN_DEC) {
            rs0[d] = x0;
            rs1[d] = x1;
            const int ls = nd.slot - s0;
            float pos[4], sum = 0.0f;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                pos[a] = a < nd.nact ? fmaxf(rselv<NRV>(Ra, Rb, Rc, ls + a), 0.0f) : 0.0f;
                sum += pos[a];
            }
            const float inv = sum > 1e-7f ? __builtin_amdgcn_rcpf(sum) : 0.0f, un = uniform_share(nd.nact);
            const float wr = ws * (nd.player == 0 ? x0 : x1);
            float* Sg = L.S + ((size_t)b * L.nslot + (nd.slot - s0 + sb)) * np + k;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const float sa = a < nd.nact ? (sum > 1e-7f ? pos[a] * inv : un) : 0.0f;
                sg[d * 4 + a] = sa;
                if (live && a < nd.nact) unsafeAtomicAdd(Sg + (size_t)a * np, wr * sa);
            }
            continue;
        }
        if (nd.type == DN_LEAF) {
            if (LEVEL == 0 && inb) {
                float* lr = D.leafreach + ((size_t)b * D.nleaf + nd.leaf) * 2 * NPT + k;
                lr[0] = live ? x0 : 0.0f;
                lr[NPT] = live ? x1 : 0.0f;
            }
            continue;
        }
        if (nd.type == DN_SD && inb) {
            // player p's column holds the opponent's reach
            const float o0 = live ? x1 : 0.0f, o1 = live ? x0 : 0.0f;
            if (LEVEL >= 1) {
                constexpr int NX = LEVEL == 2 ? NPF : NPR;
                const size_t cb = ((size_t)b * L.ncolp + (nd.col - cb0)) * NX + k;
                if (LEVEL == 1 && D.Bf) {
                    D.Bf[cb] = o0;
                    D.Bf[cb + NX] = o1;
                } else {
                    store_split(D.Bh, D.Bl, cb, o0);
                    store_split(D.Bh, D.Bl, cb + NX, o1);
                }
            } else {
                const size_t cb = ((size_t)b * L.ncolp + nd.col) * NPT + k;
                D.B0[cb] = o0;
                D.B0[cb + NPT] = o1;
            }
        }
        if (live) {  // fold or showdown: card sums of both players' reaches
            unsigned long long* bn = bins + (size_t)(nd.sidx - t0) * 104;
            const unsigned long long v0 = fx48_of(x0);
            const unsigned long long v1 = fx48_of(x1);
            if (v0) {
                atomicAdd(&bn[c1], v0);
                atomicAdd(&bn[c2], v0);
            }
            if (v1) {
                atomicAdd(&bn[52 + c1], v1);
                atomicAdd(&bn[52 + c2], v1);
            }
        }
    }
    }  // pass
    fwd_bins_flush<NT>(L, bins, b, tb, half, nt);
}

template <int LEVEL, int DM, int NT, int NRV = 2>
__global__ __launch_bounds__(NT) void k_bt_forward(BatchedDev D, const float* __restrict__ ranges, float ws,
                                                   const float* __restrict__ wdev) {
    extern __shared__ unsigned long long bins[];  // [nt][2][52]
    __shared__ EvRec sev[MAXE];
    bt_forward_body<LEVEL, DM, NT, NRV, 1>(D, ranges, ws, wdev, (int)(blockIdx.x >> 1), (int)(blockIdx.x & 1), sev, bins);
}

// ----------------------------------------------------------- backward ----
// NH workgroups of NTB threads per (board, template), one combo per
// thread.  The card sums of the template's terminals (both forward halves
// added) and the showdown values of the template's columns (this
// workgroup's combos) are
// staged in LDS; the regrets of every slot of the template are loaded up
// front into an indexed register file.  The walk over the template's DFS
// events recomputes reaches top-down, takes terminal values from the
// showdown values and the card sums, and applies the DCFR-generalised
// regret update R <- (R + r) * (R + r > 0 ? wp : wn) at each decision node
// on the way up.
template <int LEVEL, int DM, int NTB, int NRV>
__device__ __forceinline__ void bt_backward_body(const BatchedDev& D, const float* __restrict__ ranges, float wp,
                                                 float wn, const float* __restrict__ wdev, int inst, int half,
                                                 EvRec* sev,
                                                 unsigned long long* ssum /* [nt][2][53], then float [ncolt][NTB] */) {
    typedef float vS __attribute__((ext_vector_type(DM * 4)));
    typedef float vD __attribute__((ext_vector_type(DM)));
    const LevelDev& L = D.L[LEVEL];
    const int b = LEVEL == 2 ? inst : inst / L.ntmpl;
    const int tm = LEVEL == 2 ? ldu(L.btmpl + b) : inst % L.ntmpl;
    if (wdev) {
        wp = wdev[0];
        wn = wdev[1];
    }
    const int tid = threadIdx.x;
    const int np = L.np;
    const int k = half * NTB + tid;
    const bool live = k < ldu(L.nlive + b);
    const int e0 = ldu(L.tev0 + tm), e1 = ldu(L.tev0 + tm + 1);
    const int s0 = ldu(L.tslot0 + tm), nsl = ldu(L.tslot0 + tm + 1) - s0;
    const int t0 = ldu(L.tterm0 + tm), nt = ldu(L.tterm0 + tm + 1) - t0;
    const int col0 = ldu(L.tcol0 + tm), ncolt = ldu(L.tcol0 + tm + 1) - col0;
    // flop instances keep per-board state for their own template only
    const int sb = LEVEL == 2 ? 0 : s0, tb = LEVEL == 2 ? 0 : t0, cbase = LEVEL == 2 ? 0 : col0;
    float* Cs = reinterpret_cast<float*>(ssum + (size_t)L.maxterm * 106) + tid;
    bwd_stage<LEVEL, NTB>(D, L, ssum, Cs, b, k, live, nt, tb, cbase, ncolt);
    ev_stage(sev, L.ev + e0, e1 - e0);
    int c1 = 0, c2 = 0;
    float root0 = 0.0f, root1 = 0.0f;
    f32x32_t Ra, Rb, Rc;
    const size_t rbase = ((size_t)b * L.nslot + sb) * np + k;
#pragma unroll
    for (int x = 0; x < 32; ++x) {
        Ra[x] = (live && x < nsl) ? L.R[rbase + (size_t)x * np] : 0.0f;
        Rb[x] = (live && x + 32 < nsl) ? L.R[rbase + (size_t)(x + 32) * np] : 0.0f;
        if constexpr (NRV == 3) Rc[x] = (live && x + 64 < nsl) ? L.R[rbase + (size_t)(x + 64) * np] : 0.0f;
    }
    if (live) {
        c1 = L.cc[((size_t)b * np + k) * 2];
        c2 = L.cc[((size_t)b * np + k) * 2 + 1];
        root0 = root_reach(D, ranges, LEVEL, b, tm, np, k, 0);
        root1 = root_reach(D, ranges, LEVEL, b, tm, np, k, 1);
    }
    bwd_totals<NTB>(ssum, nt);
    vD rs0, rs1, op;
    vS sg, cv;
    float out0 = 0.0f, out1 = 0.0f;
    float* Rg = L.R + rbase;
    const int ne = e1 - e0;
    EvRaw nx = ev_fetch(sev, 0);
    for (int e = 0; e < ne; ++e) {
        const EvRec ec = ev_uniform(nx);
        if (e + 1 < ne) nx = ev_fetch(sev, e + 1);
        const int kind = ec.kind;
        const DNode& nd = ec.nd;
        const int d = nd.depth;
        float x0, x1;  // reach of the node
        if (d == 0) {
            x0 = root0;
            x1 = root1;
        } else {
            const float sv = sg[(d - 1) * 4 + nd.act];
            const float p0 = rs0[d - 1], p1 = rs1[d - 1];
            x0 = nd.pad == 0 ? p0 * sv : p0;
            x1 = nd.pad == 1 ? p1 * sv : p1;
        }
        if (kind == EV_ENTER) {
            rs0[d] = x0;
            rs1[d] = x1;
            const int ls = nd.slot - s0;
            float pos[4], sum = 0.0f;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                pos[a] = a < nd.nact ? fmaxf(rselv<NRV>(Ra, Rb, Rc, ls + a), 0.0f) : 0.0f;
                sum += pos[a];
            }
            const float inv = sum > 1e-7f ? __builtin_amdgcn_rcpf(sum) : 0.0f, un = uniform_share(nd.nact);
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
            if (nd.type == DN_FOLD) {
                const int w = nd.player;
                const unsigned long long* sw = ssum + (size_t)(nd.sidx - t0) * 106 + w * 53;
                const unsigned long long* sl = ssum + (size_t)(nd.sidx - t0) * 106 + (1 - w) * 53;
                const float cw = w == 0 ? nd.c0 : nd.c1, cl = w == 0 ? nd.c1 : nd.c0;
                const float xw = w == 0 ? x0 : x1, xl = w == 0 ? x1 : x0;
                const float vw = (nd.pot - cw) * mass_u(sl, c1, c2, xl);
                const float vl = -cl * mass_u(sw, c1, c2, xw);
                v0 = w == 0 ? vw : vl;
                v1 = w == 0 ? vl : vw;
            } else if (nd.type == DN_SD) {
                const unsigned long long* sa = ssum + (size_t)(nd.sidx - t0) * 106;
                const float m0 = mass_u(sa + 53, c1, c2, x1), m1 = mass_u(sa, c1, c2, x0);
                const int lc = nd.col - col0;
                const float q0 = Cs[lc * NTB], q1 = Cs[(lc + 1) * NTB];
                if (LEVEL == 1) {
                    v0 = 0.5f * nd.pot * q0 + 0.5f * D.pot0 * m0;
                    v1 = 0.5f * nd.pot * q1 + 0.5f * D.pot0 * m1;
                } else if (LEVEL == 2) {  // C = sum (W - L) r over the 990 runouts: E = 1/2 + (W - L)/1980
                    const float hp = 0.5f * L.bpot0[b];
                    v0 = nd.pot * (1.0f / 1980.0f) * q0 + hp * m0;
                    v1 = nd.pot * (1.0f / 1980.0f) * q1 + hp * m1;
                } else {
                    v0 = nd.pot * q0 - nd.c0 * m0;
                    v1 = nd.pot * q1 - nd.c1 * m1;
                }
            } else if (LEVEL == 0 && live) {  // leaf: mean of the river-line roots over the rivers
                const float* lv = D.leafval + ((size_t)b * D.nleaf + nd.leaf) * 2 * NPT + k;
                v0 = lv[0];
                v1 = lv[NPT];
            }
            if (!live) {
                v0 = 0.0f;
                v1 = 0.0f;
            }
        } else {  // EV_LEAVE of a decision node
            const int p = nd.player;
            float node_v = 0.0f;
#pragma unroll
            for (int a = 0; a < 4; ++a)
                if (a < nd.nact) node_v = fmaf(sg[d * 4 + a], cv[d * 4 + a], node_v);
            if (live) {
                const int ls = nd.slot - s0;
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    if (a < nd.nact) {
                        const float qv = rselv<NRV>(Ra, Rb, Rc, ls + a) + (cv[d * 4 + a] - node_v);
                        Rg[(size_t)(ls + a) * np] = qv * (qv > 0.0f ? wp : wn);
                    }
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
    bwd_root_out<LEVEL>(D, L, b, tm, k, np, live, out0, out1);
}

template <int LEVEL, int DM, int NTB, int NRV = 2>
__global__ __launch_bounds__(NTB) void k_bt_backward(BatchedDev D, const float* __restrict__ ranges, float wp, float wn,
                                                     const float* __restrict__ wdev) {
    constexpr int NH = ((LEVEL == 2 ? NPF : NPT) + NTB - 1) / NTB;  // workgroups per (board, template)
    extern __shared__ unsigned long long ssum[];  // [nt][2][53] (slot 52 = total), then float [ncolt][NTB]
    __shared__ EvRec sev[MAXE];
    bt_backward_body<LEVEL, DM, NTB, NRV>(D, ranges, wp, wn, wdev, (int)(blockIdx.x / NH), (int)(blockIdx.x % NH), sev,
                                          ssum);
}

// Turn-leaf values: mean over the 44 rivers that leave the combo alive of
// the river-line root values (rivers are summed in deck order).  Rivers go
// 16 at a time: their map entries, then (from an in-bounds index on every
// lane) their root values, then the sums, so no load waits alone.
__global__ void k_bt_leafval(BatchedDev D) {
    const int t = blockIdx.y / D.nleaf, leaf = blockIdx.y % D.nleaf;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= NPT) return;
    const int rt = D.leaf_tmpl[leaf];
    const int nt1 = D.L[1].ntmpl;
    float a0 = 0.0f, a1 = 0.0f;
    constexpr int RG = 16;
    for (int r0 = 0; r0 < 48; r0 += RG) {
        int ri[RG];
        float v0[RG], v1[RG];
#pragma unroll
        for (int q = 0; q < RG; ++q) ri[q] = D.imap[(size_t)(t * 48 + r0 + q) * NPT + k];
#pragma unroll
        for (int q = 0; q < RG; ++q) {
            const float* rv = D.rootval + (((size_t)(t * 48 + r0 + q) * nt1 + rt) * 2) * NPR + (ri[q] == 0xFFFF ? 0 : ri[q]);
            v0[q] = rv[0];
            v1[q] = rv[NPR];
        }
#pragma unroll
        for (int q = 0; q < RG; ++q) {
            const bool on = ri[q] != 0xFFFF;
            a0 = on ? a0 + v0[q] : a0;
            a1 = on ? a1 + v1[q] : a1;
        }
    }
    float* lv = D.leafval + ((size_t)t * D.nleaf + leaf) * 2 * NPT + k;
    lv[0] = a0 * (1.0f / 44.0f);
    lv[NPT] = a1 * (1.0f / 44.0f);
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

__device__ __forceinline__ bool conflict(uint32_t x, uint32_t y) {
    uint32_t v = x ^ y;  // a zero byte means a shared card
    return ((v - 0x01010101u) & ~v & 0x80808080u) != 0;
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

constexpr int GNCH = NPR / 32;           // K chunks of 32 villains

// The same signs at two bits per pair: word (b, chunk c, hero i, half h)
// holds the codes of villains 32c + 16h + 8k + e (K step k = 0, 1; e =
// 0..7) with code 1 = +1/2 (hero ranks higher), 2 = -1/2, 0 = tie, blocked
// or dead.  The bits are placed so that a fragment decodes with one shift
// and one masked OR per pair of elements: with m = e >> 1, step 0 keeps
// even e at bit 8 + 2m and odd e at bit 24 + 2m, step 1 even e at bit 2m and
// odd e at bit 16 + 2m.  One thread per word.
__device__ __forceinline__ int s2_pos(int e16) {
    const int k = e16 >> 3, e = e16 & 7;
    return (k ? 0 : 8) + ((e & 1) ? 16 : 0) + 2 * (e >> 1);
}

__global__ void k_bt_s2(BatchedDev D, uint32_t* __restrict__ S2) {
    const int b = blockIdx.y, i = blockIdx.x, q = threadIdx.x;  // q = c * 2 + h
    if (q >= GNCH * 2) return;
    const int nl = D.L[1].nlive[b];
    const uint16_t gi = D.grank[(size_t)b * NPR + i];
    const uint32_t xi = D.cx[(size_t)b * NPR + i];
    const int c = q >> 1, h = q & 1;
    uint32_t w = 0;
    for (int e = 0; e < 16; ++e) {
        const int j = c * 32 + 16 * h + e;
        if (i < nl && j < nl && !conflict(xi, D.cy[(size_t)b * NPR + j])) {
            const uint16_t gj = D.grank[(size_t)b * NPR + j];
            w |= (gi > gj ? 1u : (gi < gj ? 2u : 0u)) << s2_pos(e);
        }
    }
    S2[(((size_t)b * GNCH + c) * NPR + i) * 2 + h] = w;
}

// The eight codes of K step k of an S word to the bf16 B-fragment of S/2:
// per output dword (elements 2m, 2m+1) the two codes are moved to bits 8 and
// 24, which makes them the byte selectors of one v_perm over the table bytes
// 0x00, 0x3F, 0xBF (the high byte of bf16 0, +1/2, -1/2); the masked-off
// selector bytes 0 and 2 are 0, which selects the table's zero byte 0 for
// the low bytes.  One shift (none for m = 0 of step 0), one AND, one perm.
__device__ __forceinline__ bf16x8_t s2_frag(uint32_t w, int k) {
    constexpr uint32_t T = 0x00BF3F00u;
    uint32_t o[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const uint32_t x = k ? (w << (8 - 2 * m)) : (w >> (2 * m));
        o[m] = __builtin_amdgcn_perm(0u, T, x & 0x03000300u);
    }
    bf16x8_t r;
    __builtin_memcpy(&r, o, 16);
    return r;
}

// River showdown GEMM: C[b][col][i] = 2 * sum_j (S/2)[i][j] * (Bh + Bl)[b][col][j].
// One workgroup = 4 32-hero tiles x NW 32-column tiles of one board (the
// board's column tiles are split into ncg groups of NW); wave w owns column
// tile cg * NW + w and all four hero tiles (waves past the board's tiles only
// help stage S).  The K loop runs over chunks of 32 villains: S chunks
// (128 heroes x 32 villains, bytes of bf16(S/2)) are widened to bf16 once per
// workgroup into LDS (double-buffered, one barrier per chunk); each wave
// streams its own reach columns (hi and lo) straight into MFMA A-fragments.
// Both streams are fetched PD chunks ahead into register rings, so the HBM
// latency of a chunk is covered by PD chunks of MFMA work.  Within a chunk,
// lane half h of a fragment covers villains 16h..16h+15 (8 per MFMA k-step),
// a permutation of K shared by both operands.  Workgroups are numbered so
// that all workgroups of one board run on one XCD (dispatch is round-robin
// over the 8 XCDs), where they share that board's reach and S bytes in L2.
constexpr int GHT = 4;                                       // hero tiles per workgroup
constexpr int GHB = (NPR + 32 * GHT - 1) / (32 * GHT);       // hero blocks per board

template <int NW, int PD>
__global__ __launch_bounds__(NW * 64) void k_bt_gemm5(BatchedDev D, const uint8_t* __restrict__ S8, int nlogical,
                                                      int ncg) {
    constexpr int NT = NW * 64;
    constexpr int PIECES = GHT * 64;                 // 16-byte S pieces of one chunk
    constexpr int SPT = (PIECES + NT - 1) / NT;      // pieces staged per thread
    __shared__ uint4 Ssh[2][GHT][2][2][32];          // [buf][hero tile][h][k-step][hero in tile]
    const int per = (nlogical + 7) >> 3;
    const int logical = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (logical >= nlogical) return;
    const int b = logical / (GHB * ncg), rem = logical % (GHB * ncg);
    const int hb = rem / ncg, cg = rem % ncg;
    const int i0 = hb * 32 * GHT;
    const int nht = min(GHT, (NPR - i0) >> 5);
    const int npieces = nht * 64;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, half = lane >> 5, l32 = lane & 31;
    const int ncolp = D.L[1].ncolp;
    const int tile = cg * NW + wave;
    const bool on = tile * 32 < ncolp;
    const int col = (on ? tile : 0) * 32 + l32;
    const uint4* ah_p = reinterpret_cast<const uint4*>(D.Bh + ((size_t)b * ncolp + col) * NPR + 16 * half);
    const uint4* al_p = reinterpret_cast<const uint4*>(D.Bl + ((size_t)b * ncolp + col) * NPR + 16 * half);
    const uint4* s_p = reinterpret_cast<const uint4*>(S8 + (size_t)b * NPR * NPR + (size_t)i0 * 32);
    constexpr int CH4 = NPR * 2;               // uint4 pieces per S chunk (all heroes)
    f32x16_t acc[GHT];
#pragma unroll
    for (int t = 0; t < GHT; ++t)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[t][r] = 0.0f;
    uint4 sv[PD][SPT], av[PD][4];   // rings: S pieces; reach hi s0, hi s1, lo s0, lo s1
    auto load = [&](int c, uint4 (&sr)[SPT], uint4 (&ar)[4]) {
#pragma unroll
        for (int q = 0; q < SPT; ++q) {
            const int e = threadIdx.x + NT * q;
            if (e < npieces) sr[q] = s_p[(size_t)c * CH4 + e];
        }
        const int o = c * 4;        // 32 villains = 4 uint4 of bf16 per row; this half's 16 at +0, +1
        ar[0] = ah_p[o];
        ar[1] = ah_p[o + 1];
        ar[2] = al_p[o];
        ar[3] = al_p[o + 1];
    };
    auto store_s = [&](int buf, const uint4 (&sr)[SPT]) {
#pragma unroll
        for (int q = 0; q < SPT; ++q) {
            const int e = threadIdx.x + NT * q;
            if (e < npieces) {
                const int hl = e >> 1, h = e & 1;
                const uint4 w = sr[q];
                Ssh[buf][hl >> 5][h][0][hl & 31] =
                    make_uint4(__builtin_amdgcn_perm(0u, w.x, 0x010C000Cu), __builtin_amdgcn_perm(0u, w.x, 0x030C020Cu),
                               __builtin_amdgcn_perm(0u, w.y, 0x010C000Cu), __builtin_amdgcn_perm(0u, w.y, 0x030C020Cu));
                Ssh[buf][hl >> 5][h][1][hl & 31] =
                    make_uint4(__builtin_amdgcn_perm(0u, w.z, 0x010C000Cu), __builtin_amdgcn_perm(0u, w.z, 0x030C020Cu),
                               __builtin_amdgcn_perm(0u, w.w, 0x010C000Cu), __builtin_amdgcn_perm(0u, w.w, 0x030C020Cu));
            }
        }
    };
    // ring slot x % PD holds chunk x; at step c the slots hold chunks c .. c + PD - 1, so chunk c + 1 is
    // already there when step c stages it (PD >= 2)
    static_assert(PD >= 2 && GNCH >= PD, "prefetch depth");
#pragma unroll
    for (int p = 0; p < PD; ++p) load(p, sv[p], av[p]);
    store_s(0, sv[0]);
    __syncthreads();
    for (int c0 = 0; c0 < GNCH; c0 += PD) {
#pragma unroll
        for (int p = 0; p < PD; ++p) {
            const int c = c0 + p;
            if (c < GNCH) {
                // chunk c + 1 into the other LDS buffer (its readers passed the last barrier)
                if (c + 1 < GNCH) store_s((c + 1) & 1, sv[(p + 1) % PD]);
                if (on) {
#pragma unroll
                    for (int ks = 0; ks < 2; ++ks) {
                        bf16x8_t ah, al;
                        __builtin_memcpy(&ah, &av[p][ks], 16);
                        __builtin_memcpy(&al, &av[p][2 + ks], 16);
#pragma unroll
                        for (int t = 0; t < GHT; ++t) {
                            if (t < nht) {
                                bf16x8_t bop;
                                __builtin_memcpy(&bop, &Ssh[c & 1][t][half][ks][l32], 16);
                                acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bop, acc[t], 0, 0, 0);
                                acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bop, acc[t], 0, 0, 0);
                            }
                        }
                    }
                }
                // ring slot p is free again: fetch chunk c + PD into it
                if (c + PD < GNCH) load(c + PD, sv[p], av[p]);
                __syncthreads();
            }
        }
    }
    if (!on) return;
    float* Cb = D.C + (size_t)b * ncolp * NPR;
#pragma unroll
    for (int t = 0; t < GHT; ++t) {
        if (t < nht) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int cc = tile * 32 + (r >> 2) * 8 + half * 4 + (r & 3);
                Cb[(size_t)cc * NPR + i0 + 32 * t + l32] = 2.0f * acc[t][r];
            }
        }
    }
}

// River showdown GEMM, hero-major form: one workgroup of 8 waves = 256
// heroes (8 tiles) x all of the board's column tiles.  Wave w owns hero
// tiles 2(w%4), 2(w%4)+1 and the first (w < 4) or second half of the column
// tiles; waves w and w+4 share a SIMD, so every SIMD does the same work.
// Per K chunk of 32 villains, the reach hi/lo of all columns is staged once
// into LDS in MFMA A-fragment order ([tile][hi/lo][k-step][lane]; each wave
// writes whole 1 KB fragment blocks, free of bank conflicts), and every wave
// reads its column tiles' fragments from there; each wave loads its own S
// rows (bytes of bf16(S/2), one contiguous 1 KB per hero tile and chunk)
// straight into registers and widens them into B-fragments.  Each reach
// element is fetched by 5 workgroups per board (instead of one per 128
// heroes).  Chunk c + 1 is fetched while chunk c is multiplied.
// Bytes of bf16(S/2) (16 villains) widened into the two k-step B-fragments.
__device__ __forceinline__ void gemm_widen(const uint4& w, bf16x8_t& k0, bf16x8_t& k1) {
    const uint4 a = make_uint4(__builtin_amdgcn_perm(0u, w.x, 0x010C000Cu), __builtin_amdgcn_perm(0u, w.x, 0x030C020Cu),
                               __builtin_amdgcn_perm(0u, w.y, 0x010C000Cu), __builtin_amdgcn_perm(0u, w.y, 0x030C020Cu));
    const uint4 c = make_uint4(__builtin_amdgcn_perm(0u, w.z, 0x010C000Cu), __builtin_amdgcn_perm(0u, w.z, 0x030C020Cu),
                               __builtin_amdgcn_perm(0u, w.w, 0x010C000Cu), __builtin_amdgcn_perm(0u, w.w, 0x030C020Cu));
    __builtin_memcpy(&k0, &a, 16);
    __builtin_memcpy(&k1, &c, 16);
}

constexpr int G6_HT = 8;                                      // hero tiles per workgroup
constexpr int G6_HB = (NPR / 32 + G6_HT - 1) / G6_HT;         // hero blocks per board

template <int NTL>
__global__ __launch_bounds__(512) void k_bt_gemm6(BatchedDev D, const uint8_t* __restrict__ S8, int nlogical) {
    constexpr int NTILE_H = NPR / 32;                         // hero tiles of a board
    constexpr int APT = (9 * 256 + 511) / 512;                // staged reach pieces per thread (at most)
    __shared__ uint4 As[2][9][2][2][64];                      // [buf][column tile][hi/lo][k-step][lane]
    const int per = (nlogical + 7) >> 3;
    const int logical = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (logical >= nlogical) return;
    const int b = logical / G6_HB, hb = logical % G6_HB;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, half = lane >> 5, l32 = lane & 31;
    const int ncolp = D.L[1].ncolp, ntiles = ncolp >> 5;
    const int nfirst = (ntiles + 1) >> 1;
    const int ct0 = wave < 4 ? 0 : nfirst;
    const int nct = wave < 4 ? nfirst : ntiles - nfirst;      // <= NTL
    const int ht0 = hb * G6_HT + 2 * (wave & 3);
    const bool on0 = ht0 < NTILE_H, on1 = ht0 + 1 < NTILE_H;
    // S rows of this lane: hero (tile * 32 + l32), villains 16 * half .. + 15 of each chunk
    const uint4* sb = reinterpret_cast<const uint4*>(S8 + (size_t)b * NPR * NPR);
    const size_t so0 = (size_t)((on0 ? ht0 : 0) * 32 + l32) * 2 + half;
    const size_t so1 = (size_t)((on1 ? ht0 + 1 : 0) * 32 + l32) * 2 + half;
    constexpr int CH4 = NPR * 2;                              // uint4 per S chunk (all heroes)
    // reach rows of the board: NPR bf16 = NPR / 8 uint4 per column
    const uint4* bh = reinterpret_cast<const uint4*>(D.Bh + (size_t)b * ncolp * NPR);
    const uint4* bl = reinterpret_cast<const uint4*>(D.Bl + (size_t)b * ncolp * NPR);
    const int npieces = ntiles * 256;                         // 32 columns x 2 planes x 4 uint4 per tile
    f32x16_t acc[NTL][2];
#pragma unroll
    for (int i = 0; i < NTL; ++i)
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][t][r] = 0.0f;
    uint4 ap0, ap1, ap2, ap3, ap4;   // staged reach pieces (APT == 5 named registers)
    static_assert(APT == 5, "reach staging registers");
    uint4 s0, s1;
#define G6_LOAD_A(C_)                                                                            \
    do {                                                                                         \
        const int c_ = (C_);                                                                     \
        G6_PIECE_LOAD(0, ap0, c_) G6_PIECE_LOAD(1, ap1, c_) G6_PIECE_LOAD(2, ap2, c_)           \
        G6_PIECE_LOAD(3, ap3, c_) G6_PIECE_LOAD(4, ap4, c_)                                      \
    } while (0)
#define G6_PIECE_LOAD(Q_, R_, C_)                                                                \
    {                                                                                            \
        const int e = threadIdx.x + 512 * (Q_);                                                  \
        if (e < npieces) {                                                                       \
            const int blk = e >> 6, ln = e & 63;  /* fragment block (tile, hi/lo, k-step), lane */ \
            const int col = (blk >> 2) * 32 + (ln & 31), j = 2 * (ln >> 5) + (blk & 1);         \
            R_ = ((blk >> 1) & 1 ? bl : bh)[(size_t)col * (NPR / 8) + (C_) * 4 + j];             \
        }                                                                                        \
    }
#define G6_STORE_A(BUF_)                                                                         \
    do {                                                                                         \
        G6_PIECE_STORE(0, ap0, BUF_) G6_PIECE_STORE(1, ap1, BUF_) G6_PIECE_STORE(2, ap2, BUF_)  \
        G6_PIECE_STORE(3, ap3, BUF_) G6_PIECE_STORE(4, ap4, BUF_)                                \
    } while (0)
#define G6_PIECE_STORE(Q_, R_, BUF_)                                                             \
    {                                                                                            \
        const int e = threadIdx.x + 512 * (Q_);                                                  \
        if (e < npieces) (&As[BUF_][0][0][0][0])[e] = R_;  /* whole 1 KB blocks per wave */      \
    }
    auto widen = [](const uint4& w, bf16x8_t& k0, bf16x8_t& k1) {
        const uint4 a = make_uint4(__builtin_amdgcn_perm(0u, w.x, 0x010C000Cu), __builtin_amdgcn_perm(0u, w.x, 0x030C020Cu),
                                   __builtin_amdgcn_perm(0u, w.y, 0x010C000Cu), __builtin_amdgcn_perm(0u, w.y, 0x030C020Cu));
        const uint4 c = make_uint4(__builtin_amdgcn_perm(0u, w.z, 0x010C000Cu), __builtin_amdgcn_perm(0u, w.z, 0x030C020Cu),
                                   __builtin_amdgcn_perm(0u, w.w, 0x010C000Cu), __builtin_amdgcn_perm(0u, w.w, 0x030C020Cu));
        __builtin_memcpy(&k0, &a, 16);
        __builtin_memcpy(&k1, &c, 16);
    };
    ap0 = ap1 = ap2 = ap3 = ap4 = make_uint4(0u, 0u, 0u, 0u);
    G6_LOAD_A(0);
    s0 = sb[so0];
    s1 = sb[so1];
    G6_STORE_A(0);
    __syncthreads();
    for (int c = 0; c < GNCH; ++c) {
        const bool more = c + 1 < GNCH;
        uint4 n0 = s0, n1 = s1;
        if (more) {
            G6_LOAD_A(c + 1);
            n0 = sb[(size_t)(c + 1) * CH4 + so0];
            n1 = sb[(size_t)(c + 1) * CH4 + so1];
        }
        bf16x8_t b00, b01, b10, b11;   // [hero tile][k-step]
        widen(s0, b00, b01);
        widen(s1, b10, b11);
        const int buf = c & 1;
#pragma unroll
        for (int i = 0; i < NTL; ++i) {
            if (on0 && i < nct) {
                const int ct = ct0 + i;
#pragma unroll
                for (int ks = 0; ks < 2; ++ks) {
                    bf16x8_t ah, al;
                    __builtin_memcpy(&ah, &As[buf][ct][0][ks][lane], 16);
                    __builtin_memcpy(&al, &As[buf][ct][1][ks][lane], 16);
                    const bf16x8_t& bq0 = ks ? b01 : b00;
                    const bf16x8_t& bq1 = ks ? b11 : b10;
                    acc[i][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bq0, acc[i][0], 0, 0, 0);
                    acc[i][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bq1, acc[i][1], 0, 0, 0);
                    acc[i][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bq0, acc[i][0], 0, 0, 0);
                    acc[i][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bq1, acc[i][1], 0, 0, 0);
                }
            }
        }
        if (more) G6_STORE_A(buf ^ 1);
        __syncthreads();
        s0 = n0;
        s1 = n1;
    }
#undef G6_LOAD_A
#undef G6_PIECE_LOAD
#undef G6_STORE_A
#undef G6_PIECE_STORE
    float* Cb = D.C + (size_t)b * ncolp * NPR;
#pragma unroll
    for (int i = 0; i < NTL; ++i) {
        if (i < nct) {
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                if (t == 0 ? on0 : on1) {
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const int cc = (ct0 + i) * 32 + (r >> 2) * 8 + half * 4 + (r & 3);
                        Cb[(size_t)cc * NPR + (ht0 + t) * 32 + l32] = 2.0f * acc[i][t][r];
                    }
                }
            }
        }
    }
}

// The hero-major river GEMM with S read at two bits per pair (k_bt_s2): a
// lane's 16 villains of one hero row are one dword per chunk, so the S
// stream is a quarter of the byte form and is fetched three chunks ahead in
// a register shift ring; the reach stream and everything else are as in
// k_bt_gemm6.
template <int NTL>
__global__ __launch_bounds__(512) void k_bt_gemm6s(BatchedDev D, const uint32_t* __restrict__ S2, int nlogical) {
    constexpr int NTILE_H = NPR / 32;
    constexpr int APT = (9 * 256 + 511) / 512;
    __shared__ uint4 As[2][9][2][2][64];
    const int per = (nlogical + 7) >> 3;
    const int logical = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (logical >= nlogical) return;
    const int b = logical / G6_HB, hb = logical % G6_HB;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, half = lane >> 5, l32 = lane & 31;
    const int ncolp = D.L[1].ncolp, ntiles = ncolp >> 5;
    const int nfirst = (ntiles + 1) >> 1;
    const int ct0 = wave < 4 ? 0 : nfirst;
    const int nct = wave < 4 ? nfirst : ntiles - nfirst;
    const int ht0 = hb * G6_HT + 2 * (wave & 3);
    const bool on0 = ht0 < NTILE_H, on1 = ht0 + 1 < NTILE_H;
    // S words of this lane: [b][chunk][hero][half]
    const uint32_t* sw = S2 + (size_t)b * GNCH * NPR * 2;
    const size_t so0 = (size_t)((on0 ? ht0 : 0) * 32 + l32) * 2 + half;
    const size_t so1 = (size_t)((on1 ? ht0 + 1 : 0) * 32 + l32) * 2 + half;
    constexpr int CHW = NPR * 2;                               // S words per chunk (all heroes)
    const uint4* bh = reinterpret_cast<const uint4*>(D.Bh + (size_t)b * ncolp * NPR);
    const uint4* bl = reinterpret_cast<const uint4*>(D.Bl + (size_t)b * ncolp * NPR);
    const int npieces = ntiles * 256;
    f32x16_t acc[NTL][2];
#pragma unroll
    for (int i = 0; i < NTL; ++i)
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][t][r] = 0.0f;
    uint4 ap0, ap1, ap2, ap3, ap4;
    static_assert(APT == 5, "reach staging registers");
#define G6S_PIECE_LOAD(Q_, R_, C_)                                                                \
    {                                                                                             \
        const int e = threadIdx.x + 512 * (Q_);                                                   \
        if (e < npieces) {                                                                        \
            const int blk = e >> 6, ln = e & 63;                                                  \
            const int col = (blk >> 2) * 32 + (ln & 31), j = 2 * (ln >> 5) + (blk & 1);          \
            R_ = ((blk >> 1) & 1 ? bl : bh)[(size_t)col * (NPR / 8) + (C_) * 4 + j];              \
        }                                                                                         \
    }
#define G6S_PIECE_STORE(Q_, R_, BUF_)                                                             \
    {                                                                                             \
        const int e = threadIdx.x + 512 * (Q_);                                                   \
        if (e < npieces) (&As[BUF_][0][0][0][0])[e] = R_;                                         \
    }
#define G6S_LOAD_A(C_)                                                                            \
    do {                                                                                          \
        const int c_ = (C_);                                                                      \
        G6S_PIECE_LOAD(0, ap0, c_) G6S_PIECE_LOAD(1, ap1, c_) G6S_PIECE_LOAD(2, ap2, c_)         \
        G6S_PIECE_LOAD(3, ap3, c_) G6S_PIECE_LOAD(4, ap4, c_)                                     \
    } while (0)
#define G6S_STORE_A(BUF_)                                                                         \
    do {                                                                                          \
        G6S_PIECE_STORE(0, ap0, BUF_) G6S_PIECE_STORE(1, ap1, BUF_) G6S_PIECE_STORE(2, ap2, BUF_) \
        G6S_PIECE_STORE(3, ap3, BUF_) G6S_PIECE_STORE(4, ap4, BUF_)                               \
    } while (0)
    ap0 = ap1 = ap2 = ap3 = ap4 = make_uint4(0u, 0u, 0u, 0u);
    static_assert(GNCH >= 3, "S ring depth");
    // S shift ring: (a0, a1) chunk c, (b0, b1) chunk c + 1, (c0, c1) chunk c + 2
    uint32_t sa0 = sw[so0], sa1 = sw[so1];
    uint32_t sb0 = sw[CHW + so0], sb1 = sw[CHW + so1];
    uint32_t sc0 = sw[2 * CHW + so0], sc1 = sw[2 * CHW + so1];
    G6S_LOAD_A(0);
    G6S_STORE_A(0);
    __syncthreads();
    for (int c = 0; c < GNCH; ++c) {
        const bool more = c + 1 < GNCH;
        if (more) G6S_LOAD_A(c + 1);
        uint32_t sd0 = 0u, sd1 = 0u;   // chunk c + 3
        if (c + 3 < GNCH) {
            sd0 = sw[(size_t)(c + 3) * CHW + so0];
            sd1 = sw[(size_t)(c + 3) * CHW + so1];
        }
        const int buf = c & 1;
        if (on0) {
            const bf16x8_t b00 = s2_frag(sa0, 0), b01 = s2_frag(sa0, 1);
            const bf16x8_t b10 = s2_frag(sa1, 0), b11 = s2_frag(sa1, 1);
#pragma unroll
            for (int i = 0; i < NTL; ++i) {
                if (i < nct) {
                    const int ct = ct0 + i;
#pragma unroll
                    for (int ks = 0; ks < 2; ++ks) {
                        bf16x8_t ah, al;
                        __builtin_memcpy(&ah, &As[buf][ct][0][ks][lane], 16);
                        __builtin_memcpy(&al, &As[buf][ct][1][ks][lane], 16);
                        const bf16x8_t& bq0 = ks ? b01 : b00;
                        const bf16x8_t& bq1 = ks ? b11 : b10;
                        acc[i][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bq0, acc[i][0], 0, 0, 0);
                        acc[i][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bq1, acc[i][1], 0, 0, 0);
                        acc[i][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bq0, acc[i][0], 0, 0, 0);
                        acc[i][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bq1, acc[i][1], 0, 0, 0);
                    }
                }
            }
        }
        if (more) G6S_STORE_A(buf ^ 1);
        __syncthreads();
        sa0 = sb0; sa1 = sb1;
        sb0 = sc0; sb1 = sc1;
        sc0 = sd0; sc1 = sd1;
    }
#undef G6S_LOAD_A
#undef G6S_PIECE_LOAD
#undef G6S_STORE_A
#undef G6S_PIECE_STORE
    float* Cb = D.C + (size_t)b * ncolp * NPR;
#pragma unroll
    for (int i = 0; i < NTL; ++i) {
        if (i < nct) {
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                if (t == 0 ? on0 : on1) {
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const int cc = (ct0 + i) * 32 + (r >> 2) * 8 + half * 4 + (r & 3);
                        Cb[(size_t)cc * NPR + (ht0 + t) * 32 + l32] = 2.0f * acc[i][t][r];
                    }
                }
            }
        }
    }
}

// k_bt_gemm6s with the reach stream fetched two chunks ahead: its 236
// registers leave one workgroup per CU, so the two waves of a SIMD belong to
// the same workgroup and meet at every chunk's barrier; a chunk's L2 latency
// is then covered by two chunks of MFMA work instead of one.  The staged
// pieces alternate between two register sets (P: even chunks, Q: odd), as do
// the S words (E, O; each also two chunks ahead), and the loop is unrolled by
// two, so every set is named statically and no register is copied between
// iterations (a copy of a register still being loaded would wait for every
// load issued before it); all loads are issued unconditionally (clamped chunk
// and piece indices, the surplus never stored), so the wait before each LDS
// store counts exactly the loads issued after it (254 registers, no spills).
// Accumulation order is that of k_bt_gemm6s, so the results are identical.
template <int NTL>
__global__ __launch_bounds__(512) void k_bt_gemm6d(BatchedDev D, const uint32_t* __restrict__ S2, int nlogical) {
    constexpr int NTILE_H = NPR / 32;
    constexpr int NT = 512;
    constexpr int APT = (9 * 256 + NT - 1) / NT;
    static_assert(APT == 5, "reach staging registers");
    static_assert(GNCH >= 4 && GNCH % 2 == 0, "chunk pairs");
    __shared__ uint4 As[2][9][2][2][64];
    const int per = (nlogical + 7) >> 3;
    const int logical = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (logical >= nlogical) return;
    const int b = logical / G6_HB, hb = logical % G6_HB;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, half = lane >> 5, l32 = lane & 31;
    const int ncolp = D.L[1].ncolp, ntiles = ncolp >> 5;
    const int nfirst = (ntiles + 1) >> 1;
    const int ct0 = wave < 4 ? 0 : nfirst;
    const int nct = wave < 4 ? nfirst : ntiles - nfirst;       // NTL or NTL - 1
    const int ht0 = hb * G6_HT + 2 * (wave & 3);
    const bool on0 = __builtin_amdgcn_readfirstlane(ht0 < NTILE_H ? 1 : 0) != 0;   // same across the wave
    const bool on1 = __builtin_amdgcn_readfirstlane(ht0 + 1 < NTILE_H ? 1 : 0) != 0;
    const uint32_t* sw = S2 + (size_t)b * GNCH * NPR * 2;
    const size_t so0 = (size_t)((on0 ? ht0 : 0) * 32 + l32) * 2 + half;
    const size_t so1 = (size_t)((on1 ? ht0 + 1 : 0) * 32 + l32) * 2 + half;
    constexpr int CHW = NPR * 2;
    const uint4* bh = reinterpret_cast<const uint4*>(D.Bh + (size_t)b * ncolp * NPR);
    const uint4* bl = reinterpret_cast<const uint4*>(D.Bl + (size_t)b * ncolp * NPR);
    const int npieces = ntiles * 256;
    // per staged piece q: its source in chunk 0, as a plane (hi or lo) and a
    // uint4 offset, and whether it is stored (a piece past the board's tiles
    // reads piece 0's source).  Pieces come in whole 64-piece blocks, so the
    // plane and the flag are the same across a wave.
    const uint4* pbase[APT];
    uint32_t poff[APT];
    bool pon[APT];
#pragma unroll
    for (int q = 0; q < APT; ++q) {
        const int e = threadIdx.x + NT * q;
        pon[q] = __builtin_amdgcn_readfirstlane(e < npieces ? 1 : 0) != 0;
        const int ee = pon[q] ? e : (int)threadIdx.x;
        const int blk = ee >> 6, ln = ee & 63;
        const int col = (blk >> 2) * 32 + (ln & 31), j = 2 * (ln >> 5) + (blk & 1);
        pbase[q] = __builtin_amdgcn_readfirstlane((blk >> 1) & 1) ? bl : bh;
        poff[q] = (uint32_t)(col * (NPR / 8) + j);
    }
    f32x16_t acc[NTL][2];
#pragma unroll
    for (int i = 0; i < NTL; ++i)
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][t][r] = 0.0f;
    uint4 p0, p1, p2, p3, p4, q0, q1, q2, q3, q4;
#define G6D_LOAD(C_, R0, R1, R2, R3, R4)                                                          \
    do {                                                                                          \
        const uint32_t c4_ = (uint32_t)(C_) * 4u;                                                 \
        R0 = pbase[0][poff[0] + c4_]; R1 = pbase[1][poff[1] + c4_]; R2 = pbase[2][poff[2] + c4_]; \
        R3 = pbase[3][poff[3] + c4_]; R4 = pbase[4][poff[4] + c4_];                               \
    } while (0)
#define G6D_STORE(BUF_, R0, R1, R2, R3, R4)                                                       \
    do {                                                                                          \
        uint4* as_ = &As[BUF_][0][0][0][0] + threadIdx.x;                                         \
        if (pon[0]) as_[0] = R0;                                                                  \
        if (pon[1]) as_[NT] = R1;                                                                 \
        if (pon[2]) as_[2 * NT] = R2;                                                             \
        if (pon[3]) as_[3 * NT] = R3;                                                             \
        if (pon[4]) as_[4 * NT] = R4;                                                             \
    } while (0)
#define G6D_MMA(BUF_, W0_, W1_)                                                                   \
    do {                                                                                          \
        if (on0) {                                                                                \
            const bf16x8_t b00 = s2_frag(W0_, 0), b01 = s2_frag(W0_, 1);                          \
            const bf16x8_t b10 = s2_frag(W1_, 0), b11 = s2_frag(W1_, 1);                          \
            _Pragma("unroll") for (int i = 0; i < NTL; ++i) {                                     \
                if (i < NTL - 1 || i < nct) {  /* only the last tile can be absent */             \
                    const int ct = ct0 + i;                                                       \
                    _Pragma("unroll") for (int ks = 0; ks < 2; ++ks) {                            \
                        bf16x8_t ah, al;                                                          \
                        __builtin_memcpy(&ah, &As[BUF_][ct][0][ks][lane], 16);                    \
                        __builtin_memcpy(&al, &As[BUF_][ct][1][ks][lane], 16);                    \
                        const bf16x8_t& bq0 = ks ? b01 : b00;                                     \
                        const bf16x8_t& bq1 = ks ? b11 : b10;                                     \
                        acc[i][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bq0, acc[i][0], 0, 0, 0); \
                        acc[i][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bq1, acc[i][1], 0, 0, 0); \
                        acc[i][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bq0, acc[i][0], 0, 0, 0); \
                        acc[i][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bq1, acc[i][1], 0, 0, 0); \
                    }                                                                             \
                }                                                                                 \
            }                                                                                     \
        }                                                                                         \
    } while (0)
    // S words of even (E) and odd (O) chunks, each fetched two chunks ahead
    uint32_t wE0 = sw[so0], wE1 = sw[so1];
    uint32_t wO0 = sw[CHW + so0], wO1 = sw[CHW + so1];
    G6D_LOAD(0, p0, p1, p2, p3, p4);
    G6D_LOAD(1, q0, q1, q2, q3, q4);
    G6D_STORE(0, p0, p1, p2, p3, p4);
    __syncthreads();
    for (int c = 0; c < GNCH; c += 2) {
        // even step: multiply chunk c (buffer 0); P and E take chunk c + 2; Q (chunk c + 1) into buffer 1
        const uint32_t ce = (uint32_t)min(c + 2, GNCH - 1);
        G6D_LOAD(ce, p0, p1, p2, p3, p4);
        G6D_MMA(0, wE0, wE1);
        wE0 = sw[(size_t)ce * CHW + so0];
        wE1 = sw[(size_t)ce * CHW + so1];
        G6D_STORE(1, q0, q1, q2, q3, q4);
        __syncthreads();
        // odd step: multiply chunk c + 1 (buffer 1); Q and O take chunk c + 3; P (chunk c + 2) into buffer 0
        const uint32_t co = (uint32_t)min(c + 3, GNCH - 1);
        G6D_LOAD(co, q0, q1, q2, q3, q4);
        G6D_MMA(1, wO0, wO1);
        wO0 = sw[(size_t)co * CHW + so0];
        wO1 = sw[(size_t)co * CHW + so1];
        if (c + 2 < GNCH) G6D_STORE(0, p0, p1, p2, p3, p4);
        __syncthreads();
    }
#undef G6D_LOAD
#undef G6D_STORE
#undef G6D_MMA
    float* Cb = D.C + (size_t)b * ncolp * NPR;
#pragma unroll
    for (int i = 0; i < NTL; ++i) {
        if (i < nct) {
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                if (t == 0 ? on0 : on1) {
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const int cc = (ct0 + i) * 32 + (r >> 2) * 8 + half * 4 + (r & 3);
                        Cb[(size_t)cc * NPR + (ht0 + t) * 32 + l32] = 2.0f * acc[i][t][r];
                    }
                }
            }
        }
    }
}

// Register-streamed river GEMM: one single-wave workgroup per (board,
// 32-column tile, group of GH hero tiles), no LDS and no barrier.  The K loop
// runs over chunks of 64 villains; per chunk a lane loads, straight into
// registers, 64 contiguous bytes of its column's reach in each plane
// (villains 64c + 32h .. +31 for lane half h) and, per hero tile, the two S
// words of its hero row for the same villains (8 contiguous bytes), one
// chunk ahead in a second register set.  K step ks of a chunk pairs lane
// half h's elements 8ks .. 8ks+7 of those 32 villains in both operands (a
// permutation of K shared by A and B), so each reach fragment feeds GH hero
// tiles and each decoded S fragment both planes.  Per accumulator the order
// is chunk, K step, high plane, low plane.  Workgroups are numbered (board,
// column tile, hero group) with the hero group fastest and all of a board's
// workgroups on one XCD, so the waves that share a column's reach run
// together and find it in L2.
constexpr int G7_NCH = NPR / 64;                 // K chunks of 64 villains
static_assert(NPR % 64 == 0, "64-villain chunks");

template <int GH, int NT>
__device__ __forceinline__ void gemm7_tile(const BatchedDev& D, const uint32_t* __restrict__ S2, int b, int ct, int hg) {
    static_assert(NT >= 1 && NT <= GH, "tile count");
    const int lane = threadIdx.x, half = lane >> 5, l32 = lane & 31;
    const int ht0 = hg * GH;
    const int ncolp = D.L[1].ncolp;
    const size_t row = ((size_t)b * ncolp + ct * 32 + l32) * NPR;
    const uint4* rh = reinterpret_cast<const uint4*>(D.Bh + row) + 4 * half;
    const uint4* rl = reinterpret_cast<const uint4*>(D.Bl + row) + 4 * half;
    // S words of hero tile t: uint2 [b][32-villain chunk][hero], chunk 2c + half
    const uint2* sw = reinterpret_cast<const uint2*>(S2) + (size_t)b * GNCH * NPR + (size_t)half * NPR + ht0 * 32 + l32;
    f32x16_t acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[t][r] = 0.0f;
    uint4 ah[4], al[4], bh[4], bl[4];
    uint2 as[NT], bs[NT];
#define G7_LOAD(C_, H_, L_, S_)                                                         \
    do {                                                                                \
        const int c_ = (C_);                                                            \
        _Pragma("unroll") for (int q = 0; q < 4; ++q) {                                 \
            H_[q] = rh[8 * c_ + q];                                                     \
            L_[q] = rl[8 * c_ + q];                                                     \
        }                                                                               \
        _Pragma("unroll") for (int t = 0; t < NT; ++t) S_[t] = sw[(size_t)(2 * c_) * NPR + 32 * t]; \
    } while (0)
    // per K step: the high-plane products as each tile's fragment is decoded,
    // then the low-plane ones; the scheduler is asked to spread the decode
    // VALU between the MFMAs
#define G7_MMA(H_, L_, S_)                                                              \
    do {                                                                                \
        _Pragma("unroll") for (int ks = 0; ks < 4; ++ks) {                              \
            bf16x8_t fh, fl, sf[NT];                                                    \
            __builtin_memcpy(&fh, &H_[ks], 16);                                         \
            __builtin_memcpy(&fl, &L_[ks], 16);                                         \
            _Pragma("unroll") for (int t = 0; t < NT; ++t) {                            \
                const uint32_t w = ks < 2 ? S_[t].x : S_[t].y;                          \
                sf[t] = s2_frag(w, ks & 1);                                             \
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fh, sf[t], acc[t], 0, 0, 0); \
            }                                                                           \
            _Pragma("unroll") for (int t = 0; t < NT; ++t)                              \
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fl, sf[t], acc[t], 0, 0, 0); \
        }                                                                               \
        _Pragma("unroll") for (int i_ = 0; i_ < 8 * NT; ++i_) {                         \
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);                          \
            __builtin_amdgcn_sched_group_barrier(0x002, 8, 0);                          \
        }                                                                               \
    } while (0)
    // the next chunk's loads are issued before the current chunk's products
    // (scheduling barriers keep the compiler from sinking them to their use);
    // a wave raises its issue priority for its products, so of a SIMD's two
    // waves the one with MFMA work ready is served first
    G7_LOAD(0, ah, al, as);
    int c = 0;
    for (; c + 1 < G7_NCH; c += 2) {
        G7_LOAD(c + 1, bh, bl, bs);
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_setprio(1);
        G7_MMA(ah, al, as);
        __builtin_amdgcn_s_setprio(0);
        G7_LOAD(min(c + 2, G7_NCH - 1), ah, al, as);
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_setprio(1);
        G7_MMA(bh, bl, bs);
        __builtin_amdgcn_s_setprio(0);
    }
    if (c < G7_NCH) G7_MMA(ah, al, as);
#undef G7_LOAD
#undef G7_MMA
    float* Cb = D.C + ((size_t)b * ncolp + ct * 32) * NPR;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int cc = (r >> 2) * 8 + half * 4 + (r & 3);
            Cb[(size_t)cc * NPR + (ht0 + t) * 32 + l32] = 2.0f * acc[t][r];
        }
    }
}


// The last column tile of a board when it holds at most 16 live columns
// (262 = 8 x 32 + 6 on the benchmark game), as 16-column products
// v_mfma_f32_16x16x32_bf16 over 16-hero tiles: half the MFMA work of the
// 32-column tile, whose other 16 columns stay zero.  One single-wave
// workgroup per (board, the 2 GH 16-hero tiles of a hero group).  Lane l:
// row or column l & 15 of a fragment, K group g = l >> 4 (villains
// 8g .. 8g+7 of a 32-villain K step).  Per 64-villain chunk, K step ks
// (0, 1) covers 32-villain chunk 2c + ks: the reach of villains 8g .. +7
// (one uint4 per plane) and, per 16-hero tile, the S word of (chunk, hero,
// half g >> 1) decoded at its step g & 1.  Per accumulator the order is
// chunk, K step, high plane, low plane, whatever GH is.
typedef float f32x4_t __attribute__((ext_vector_type(4)));

template <int GH, int NT>
__device__ __forceinline__ void gemm7t_tile(const BatchedDev& D, const uint32_t* __restrict__ S2, int b, int ct, int hg) {
    static_assert(NT >= 1 && NT <= 2 * GH, "tile count");
    const int lane = threadIdx.x, g = lane >> 4, l16 = lane & 15;
    const int h0 = hg * 2 * GH;                   // first 16-hero tile of the group
    const int ncolp = D.L[1].ncolp;
    const size_t row = ((size_t)b * ncolp + ct * 32 + l16) * NPR;
    const uint4* rh = reinterpret_cast<const uint4*>(D.Bh + row) + g;   // + 4 per 32-villain chunk
    const uint4* rl = reinterpret_cast<const uint4*>(D.Bl + row) + g;
    const uint32_t* sw = S2 + (size_t)b * GNCH * NPR * 2 + (g >> 1) + (h0 * 16 + l16) * 2;
    f32x4_t acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[t][r] = 0.0f;
    uint4 ah[2], al[2], bh[2], bl[2];
    uint32_t as[2][NT], bs[2][NT];
#define G7T_LOAD(C_, H_, L_, S_)                                                        \
    do {                                                                                \
        const int c_ = (C_);                                                            \
        _Pragma("unroll") for (int ks = 0; ks < 2; ++ks) {                              \
            H_[ks] = rh[4 * (2 * c_ + ks)];                                             \
            L_[ks] = rl[4 * (2 * c_ + ks)];                                             \
            _Pragma("unroll") for (int t = 0; t < NT; ++t)                              \
                S_[ks][t] = sw[(size_t)(2 * c_ + ks) * NPR * 2 + 32 * t];               \
        }                                                                               \
    } while (0)
#define G7T_MMA(H_, L_, S_)                                                             \
    do {                                                                                \
        _Pragma("unroll") for (int ks = 0; ks < 2; ++ks) {                              \
            bf16x8_t fh, fl, sf[NT];                                                    \
            __builtin_memcpy(&fh, &H_[ks], 16);                                         \
            __builtin_memcpy(&fl, &L_[ks], 16);                                         \
            _Pragma("unroll") for (int t = 0; t < NT; ++t) {                            \
                sf[t] = s2_frag(S_[ks][t], g & 1);                                      \
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fh, sf[t], acc[t], 0, 0, 0); \
            }                                                                           \
            _Pragma("unroll") for (int t = 0; t < NT; ++t)                              \
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fl, sf[t], acc[t], 0, 0, 0); \
        }                                                                               \
        _Pragma("unroll") for (int i_ = 0; i_ < 4 * NT; ++i_) {                         \
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);                          \
            __builtin_amdgcn_sched_group_barrier(0x002, 4, 0);                          \
        }                                                                               \
    } while (0)
    G7T_LOAD(0, ah, al, as);
    int c = 0;
    for (; c + 1 < G7_NCH; c += 2) {
        G7T_LOAD(c + 1, bh, bl, bs);
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_setprio(1);
        G7T_MMA(ah, al, as);
        __builtin_amdgcn_s_setprio(0);
        G7T_LOAD(min(c + 2, G7_NCH - 1), ah, al, as);
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_setprio(1);
        G7T_MMA(bh, bl, bs);
        __builtin_amdgcn_s_setprio(0);
    }
    if (c < G7_NCH) G7T_MMA(ah, al, as);
#undef G7T_LOAD
#undef G7T_MMA
    float* Cb = D.C + ((size_t)b * ncolp + ct * 32) * NPR;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
#pragma unroll
        for (int r = 0; r < 4; ++r) Cb[(size_t)(4 * g + r) * NPR + (h0 + t) * 16 + l16] = 2.0f * acc[t][r];
    }
}


// The register-streamed river GEMM launch: per board, nmain 32-column tiles
// x NHG hero groups of GH 32-hero tiles, then, when the board's last column
// tile holds at most 16 live columns (ntail = 1), that tile as NHGT groups
// of 2 x G7T_GH 16-hero tiles.  Workgroups are numbered (board, item) with
// all of a board's items on one XCD, the tail items last.
constexpr int G7T_GH = 5;   // 10 16-hero tiles per tail wave: 206 VGPRs, no spill

// WPE: waves per SIMD the register budget is sized for (2: up to 256 registers;
// 1: up to 512, the accumulators of 12 hero tiles in AGPRs, 3 hero groups per
// column tile instead of 5, so each reach chunk feeds more products).
template <int GH, int WPE = 2>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE)))
void k_bt_gemm7(BatchedDev D, const uint32_t* __restrict__ S2, int nlogical, int nmain, int ntail) {
    constexpr int NTH = NPR / 32;
    constexpr int NHG = (NTH + GH - 1) / GH;
    constexpr int NHGT = (NTH + G7T_GH - 1) / G7T_GH;
    const int per = (nlogical + 7) >> 3;
    const int logical = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (logical >= nlogical) return;
    const int nm = nmain * NHG, pb = nm + ntail * NHGT;
    const int b = logical / pb, rem = logical % pb;
    constexpr int NTL = NTH - GH * (NHG - 1);    // tiles of the last hero group
    if (rem < nm) {
        const int hg = rem % NHG;
        if (NTL == GH || hg < NHG - 1) gemm7_tile<GH, GH>(D, S2, b, rem / NHG, hg);
        else gemm7_tile<GH, NTL>(D, S2, b, rem / NHG, hg);
    } else {
        constexpr int NT16 = NPR / 16, NTLT = NT16 - 2 * G7T_GH * (NHGT - 1);   // 16-hero tiles of the last group
        const int hg = rem - nm;
        if (NTLT == 2 * G7T_GH || hg < NHGT - 1) gemm7t_tile<G7T_GH, 2 * G7T_GH>(D, S2, b, nmain, hg);
        else gemm7t_tile<G7T_GH, NTLT>(D, S2, b, nmain, hg);
    }
}

// Turn all-in showdowns: C0[t][col][i] = sum_j E[t][j][i] * B0[t][col][j].
// One workgroup = 64 heroes of one turn; its four waves each take a quarter
// of the villains (882 workgroups for 49 turns, several waves per SIMD),
// stage their columns' reach 32 villains at a time in LDS (every load issued
// from an in-bounds column before the LDS writes), and the quarters are
// added in a fixed order at the end, so the result is deterministic.
template <int MAXC>
__global__ __launch_bounds__(256) void k_bt_turn_sd(BatchedDev D, int ncol) {
    constexpr int KQ = NPT / 4, CH = 32, NST = 4 * MAXC * CH / 256;
    static_assert(NPT % 64 == 0 && KQ % CH == 0 && (4 * MAXC * CH) % 256 == 0, "turn showdown tiling");
    if (ncol <= 0) return;
    const int t = blockIdx.y;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = blockIdx.x * 64 + lane;
    __shared__ float bcol[4][MAXC][CH];
    __shared__ float red[3][MAXC][64];
    float acc[MAXC];
#pragma unroll
    for (int c = 0; c < MAXC; ++c) acc[c] = 0.0f;
    const float* E = D.Eturn + (size_t)t * NPT * NPT;
    const float* B = D.B0 + (size_t)t * D.L[0].ncolp * NPT;
    for (int st = 0; st < KQ / CH; ++st) {
        float v[NST];
#pragma unroll
        for (int q = 0; q < NST; ++q) {
            const int e = threadIdx.x + 256 * q;
            const int w = e / (MAXC * CH), c = (e / CH) % MAXC, jj = e % CH;
            v[q] = B[(size_t)min(c, ncol - 1) * NPT + w * KQ + st * CH + jj];
        }
#pragma unroll
        for (int q = 0; q < NST; ++q) {
            const int e = threadIdx.x + 256 * q;
            const int c = (e / CH) % MAXC;
            (&bcol[0][0][0])[e] = c < ncol ? v[q] : 0.0f;
        }
        __syncthreads();
        const float* ew = E + (size_t)(wave * KQ + st * CH) * NPT + i;
#pragma unroll 8
        for (int jj = 0; jj < CH; ++jj) {
            const float ev = ew[(size_t)jj * NPT];
#pragma unroll
            for (int c = 0; c < MAXC; ++c) acc[c] = fmaf(ev, bcol[wave][c][jj], acc[c]);
        }
        __syncthreads();
    }
    if (wave) {
#pragma unroll
        for (int c = 0; c < MAXC; ++c) red[wave - 1][c][lane] = acc[c];
    }
    __syncthreads();
    if (wave) return;
    for (int c = 0; c < ncol; ++c)
        D.C0[((size_t)t * D.L[0].ncolp + c) * NPT + i] = ((acc[c] + red[0][c][lane]) + red[1][c][lane]) + red[2][c][lane];
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
// Backward sweep workgroup size: two workgroups of 576 cover the 1152
// padded turn combos (and the 1088 river combos).  (384, three per board:
// two fit a CU, but the extra staging of card sums costs more than the
// overlap gains.)
constexpr int BWD_NT = 576;
constexpr int BWD_NH = (NPT + BWD_NT - 1) / BWD_NT;
// LDS of the backward sweep: card sums, then the showdown values.
static size_t bwd_lds(const LevelDev& L, int nt = BWD_NT) {
    return (size_t)L.maxterm * 106 * 8 + (size_t)L.maxcolt * nt * 4;
}

struct BatchedState {
    BatchedDev D{};            // value-initialised: pointers a mode does not use stay null
    DeviceArena arena;
    int two_level = 0;
    int nb1 = 0;          // river boards
    std::vector<ExportRec> export_recs;
    ExportRec* d_export = nullptr;
    uint16_t *d_inv0 = nullptr, *d_inv1 = nullptr;
    float* d_ranges = nullptr;  // root reach of the first street [boards][2][np]
    uint8_t* d_S8 = nullptr;    // [boards1][NPR][NPR] sign-blocker bytes
    uint32_t* d_S2 = nullptr;   // the same signs at two bits per pair ([boards1][chunk][NPR][2] words)
    int dm[2] = {8, 8};         // register-stack depth per level (4, 6 or 8)
    JitSweeps jit;              // generated per-template sweeps (when enabled)
    int gemm_shape = 100;       // river GEMM form and shape (autotuned; see river_gemm)
    int sd_mode = 0;            // river showdown: 0 GEMM, 1 strength-order sweep, 2 prefix scan
    uint32_t* d_schtab = nullptr;   // prefix scan: [boards1][3][NPR] per-combo group and list indices
    uint32_t* d_scltab = nullptr;   // prefix scan: [boards1][2][12][64] card lists
    uint32_t* d_sdprog = nullptr;   // sweep programs of the river boards (SD_REC-word records)
    uint32_t* d_sdhdr = nullptr;    // [boards1][4] first record and first three blocks of each board's program
    int* d_colinfo = nullptr;       // [ncol] terminal | player << 31 of each river showdown column
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

// Uploads one level's templates, tables and state.  priv: every board runs
// its own private template (flop instances), so per-board state is sized by
// the largest template instead of the sum over templates.
int upload_level(DeviceArena& A, hipStream_t st, LevelHost& H, LevelDev& L, int np, int nboards,
                 const std::vector<uint8_t>& cc, const std::vector<int>& nlive, bool priv, int bwd_nt) {
    L.np = np;
    L.nboards = nboards;
    L.ntmpl = (int)H.tnode0.size() - 1;
    int r = MTT_OK;
    std::vector<int> slot0 = H.slot0, term0 = H.term0, col0 = H.col0;
    slot0.push_back(H.nslot);
    term0.push_back(H.nterm);
    col0.push_back(H.ncol);
    if (priv) {  // each board holds the state of its own template only
        L.nslot = L.nterm = L.ncol = 0;
        for (int t = 0; t < L.ntmpl; ++t) {
            L.nslot = std::max(L.nslot, slot0[t + 1] - slot0[t]);
            L.nterm = std::max(L.nterm, term0[t + 1] - term0[t]);
            L.ncol = std::max(L.ncol, col0[t + 1] - col0[t]);
        }
    } else {
        L.nslot = H.nslot;
        L.nterm = H.nterm;
        L.ncol = H.ncol;
    }
    L.ncolp = ((L.ncol + 31) / 32) * 32;
    if (L.ncolp == 0) L.ncolp = 32;
    L.maxterm = 0;
    L.maxcolt = 0;
    L.maxslot = 0;
    for (int t = 0; t < L.ntmpl; ++t) {
        L.maxterm = std::max(L.maxterm, term0[t + 1] - term0[t]);
        L.maxcolt = std::max(L.maxcolt, col0[t + 1] - col0[t]);
        L.maxslot = std::max(L.maxslot, slot0[t + 1] - slot0[t]);
        MTT_ARG(H.tnode0[t + 1] - H.tnode0[t] <= MAXN, "template has more than %d nodes", MAXN);
        MTT_ARG(H.tev0[t + 1] - H.tev0[t] <= MAXE, "template has more than %d events", MAXE);
        MTT_ARG(slot0[t + 1] - slot0[t] <= MAXSLOT, "template has more than %d action slots", MAXSLOT);
    }
    // event records: node + kind, in DFS order per template
    std::vector<EvRec> recs(H.events.size());
    for (int t = 0; t < L.ntmpl; ++t)
        for (int e = H.tev0[t]; e < H.tev0[t + 1]; ++e) {
            memset(&recs[e], 0, sizeof(EvRec));
            recs[e].nd = H.nodes[H.tnode0[t] + (H.events[e] >> 2)];
            recs[e].kind = H.events[e] & 3;
        }
    std::vector<EvRec> frecs;
    std::vector<int> tf0{0};
    for (int t = 0; t < L.ntmpl; ++t) {
        for (int e = H.tev0[t]; e < H.tev0[t + 1]; ++e)
            if (recs[e].kind != EV_LEAVE) frecs.push_back(recs[e]);
        tf0.push_back((int)frecs.size());
    }
    EvRec *dev, *dfev;
    int *e0, *f0, *nl, *ts, *tt, *tc;
    uint8_t* c;
    r = r ? r : A.upload(recs, &dev);
    r = r ? r : A.upload(frecs, &dfev);
    r = r ? r : A.upload(tf0, &f0);
    r = r ? r : A.upload(H.tev0, &e0);
    r = r ? r : A.upload(slot0, &ts);
    r = r ? r : A.upload(term0, &tt);
    r = r ? r : A.upload(col0, &tc);
    r = r ? r : A.upload(cc, &c);
    r = r ? r : A.upload(nlive, &nl);
    if (r) return r;
    MTT_ARG(bwd_lds(L, bwd_nt) <= 152 * 1024, "template needs %zu bytes of LDS in the backward sweep",
            bwd_lds(L, bwd_nt));
    L.ev = dev; L.tev0 = e0; L.fev = dfev; L.tfev0 = f0; L.tslot0 = ts; L.tterm0 = tt; L.tcol0 = tc; L.cc = c; L.nlive = nl;
    size_t ns = (size_t)nboards * L.nslot * np;
    r = r ? r : A.alloc(ns, &L.R);
    r = r ? r : A.alloc(ns, &L.S);
    r = r ? r : A.alloc((size_t)nboards * std::max(L.nterm, 1) * 208, &L.part);
    if (r) return r;
    MTT_HIP(hipMemsetAsync(L.R, 0, ns * sizeof(float), st));
    MTT_HIP(hipMemsetAsync(L.S, 0, ns * sizeof(float), st));
    // a forward walking the combos in slices writes the first half of the
    // partials only; the second stays zero for the readers that add both
    MTT_HIP(hipMemsetAsync(L.part, 0, (size_t)nboards * std::max(L.nterm, 1) * 208 * 8, st));
    return MTT_OK;
}

// Tuning file: small JSON of "key": integer entries, $MTT_AUTOTUNE_FILE,
// else $HOME/.cache/mtt/autotune.json.  A file that cannot be read or
// written only means tuning again.
std::string tune_path() {
    const char* e = getenv("MTT_AUTOTUNE_FILE");
    if (e && *e) return e;
    const char* h = getenv("HOME");
    return std::string(h && *h ? h : "/tmp") + "/.cache/mtt/autotune.json";
}

int tune_lookup(const std::string& key) {
    FILE* f = fopen(tune_path().c_str(), "r");
    if (!f) return 0;
    char buf[8192];
    size_t n = fread(buf, 1, sizeof(buf) - 1, f);
    fclose(f);
    buf[n] = 0;
    std::string s(buf);
    size_t p = s.find("\"" + key + "\"");
    if (p == std::string::npos) return 0;
    p = s.find(':', p);
    if (p == std::string::npos) return 0;
    return atoi(s.c_str() + p + 1);
}

void tune_store(const std::string& key, int v) {
    std::string path = tune_path();
    std::string dir = path.substr(0, path.rfind('/'));
    std::string mk = "mkdir -p '" + dir + "'";
    if (dir.find('\'') == std::string::npos) (void)system(mk.c_str());
    std::string s;
    if (FILE* f = fopen(path.c_str(), "r")) {
        char buf[8192];
        size_t n = fread(buf, 1, sizeof(buf) - 1, f);
        fclose(f);
        buf[n] = 0;
        s = buf;
    }
    std::string entry = "\"" + key + "\": " + std::to_string(v);
    size_t p = s.find("\"" + key + "\"");
    if (p != std::string::npos) {
        size_t e = s.find_first_of(",}", p);
        s = s.substr(0, p) + entry + s.substr(e);
    } else if (s.find('}') != std::string::npos && s.find(':') != std::string::npos) {
        s = s.substr(0, s.rfind('}')) + ", " + entry + "}";
    } else {
        s = "{" + entry + "}";
    }
    if (FILE* f = fopen(path.c_str(), "w")) {
        fputs(s.c_str(), f);
        fputs("\n", f);
        fclose(f);
    }
}

// Street levels of a heads-up game for the batched path: river games (one
// level, the single street) and turn+river games from a flop (the turn
// template, then every river line as its own template).
int hu_levels(const mtt_hu_game& g, const HuTree& tr, bool& two, LevelHost& H0, LevelHost& H1) {
    two = g.n_board == 3 && g.root_chance == 1 && g.last_street == 2;
    const bool one = g.n_board == 5;
    MTT_ARG(two || one, "batched mode supports river games (5-card board) and turn+river games from a flop");
    int rc = MTT_OK;
    if (two) {
        const Template& T0 = tr.templates[0];
        rc = add_template(T0, H0, 0);
        if (rc) return rc;
        for (int l = 0; l < T0.n_leaf; ++l) {
            rc = add_template(tr.templates[T0.leaf_next[l]], H1, 0);
            if (rc) return rc;
        }
    } else {
        rc = add_template(tr.templates[0], H1, 0);
    }
    return rc;
}

// The generated-sweep description of the levels (pointers into H0 / H1).
std::vector<JitLevel> jit_levels(const LevelHost& H0, const LevelHost& H1, bool two) {
    auto desc = [](const LevelHost& H, int level, int np) {
        JitLevel J;
        J.level = level;
        J.np = np;
        J.ntmpl = (int)H.tnode0.size() - 1;
        J.nodes = &H.nodes;
        J.tnode0 = &H.tnode0;
        J.tev0 = &H.tev0;
        J.events = &H.events;
        J.slot0 = H.slot0;
        J.slot0.push_back(H.nslot);
        J.term0 = H.term0;
        J.term0.push_back(H.nterm);
        J.col0 = H.col0;
        J.col0.push_back(H.ncol);
        return J;
    };
    std::vector<JitLevel> lv;
    if (two) lv.push_back(desc(H0, 0, NPT));
    lv.push_back(desc(H1, 1, NPR));
    return lv;
}

}  // namespace

int river_gemm_autotune(BatchedState& B, hipStream_t st);

// River showdown by the sign x blocker GEMM (autotuned form, or
// MTT_RIVER_GEMM=<shape code>), by the exact strength-order sweep
// (MTT_RIVER_SD=sweep) or by prefix scans (MTT_RIVER_SD=scan); a forced
// GEMM form wins.  0 GEMM, 1 sweep, 2 scan.
int river_sd_mode() {
    const char* g = getenv("MTT_RIVER_GEMM");
    if (g && *g) return 0;
    const char* m = getenv("MTT_RIVER_SD");
    if (m && strcmp(m, "sweep") == 0) return 1;
    if (m && strcmp(m, "scan") == 0) return 2;
    return 0;
}

int batched_create(SolverImpl& s) {
    const mtt_hu_game& g = s.game;
    const HuTree& tr = s.tree;
    LevelHost H0, H1;
    bool two = false;
    int rc = hu_levels(g, tr, two, H0, H1);
    if (rc) return rc;
    auto st = new (std::nothrow) BatchedState;
    if (!st) { set_error("out of host memory"); return MTT_ERR_OOM; }
    s.batched = st;
    BatchedState& B = *st;
    B.two_level = two;
    memset(&B.D, 0, sizeof(B.D));
    B.D.two_level = two;
    B.D.pot0 = g.pot;
    DeviceArena& A = B.arena;
    // river template l is turn leaf l
    std::vector<int> leaf_tmpl, tmpl_leaf;
    if (two)
        for (int l = 0; l < tr.templates[0].n_leaf; ++l) {
            leaf_tmpl.push_back(l);
            tmpl_leaf.push_back(l);
        }
    // ---------------------------------------------------------- boards --
    const int nb1 = two ? 49 * 48 : 1;
    B.nb1 = nb1;
    // rmap padding is 0 (a valid turn combo): the generated sweeps read it on every lane
    std::vector<uint16_t> rank1((size_t)nb1 * NPR, 0), rmap((size_t)nb1 * NPR, 0);
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
        return upload_level(A, s.stream, H, L, np, nboards, cc, nlive, false, BWD_NT);
    };
    auto pick = [](int d) { return d + 1 <= 4 ? 4 : 8; };
    B.dm[0] = pick(H0.maxdepth);
    B.dm[1] = pick(H1.maxdepth);
    rc = up_level(H1, B.D.L[1], NPR, nb1, cc1, nlive1);
    if (rc) return rc;
    LevelDev& L1 = B.D.L[1];
    MTT_ARG(L1.ncolp <= GCOLS, "river showdown columns %d exceed %d", L1.ncol, GCOLS);
    MTT_ARG(L1.maxslot <= 64, "a river template has %d action slots (the batched path holds 64)", L1.maxslot);
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
        B.sd_mode = river_sd_mode();
        rc = A.alloc(nbc, &B.D.C);
        if (rc) return rc;
        if (B.sd_mode == 2) {
            // FP32 showdown columns and the per-board scan tables
            rc = A.alloc(nbc, &B.D.Bf);
            if (rc) return rc;
            MTT_HIP(hipMemsetAsync(B.D.Bf, 0, nbc * 4, s.stream));
            std::vector<uint32_t> ht((size_t)nb1 * 3 * NPR), lt((size_t)nb1 * SC_LTAB_WORDS);
            for (int bi = 0; bi < nb1; ++bi)
                MTT_ARG(sc_tables(&cc1[(size_t)bi * NPR * 2], &rank1[(size_t)bi * NPR], nlive1[bi],
                                  &ht[(size_t)bi * 3 * NPR], &lt[(size_t)bi * SC_LTAB_WORDS]),
                        "river board %d: not 1081 combos in strength order", bi);
            rc = A.upload(ht, &B.d_schtab);
            rc = rc ? rc : A.upload(lt, &B.d_scltab);
            if (rc) return rc;
        } else if (B.sd_mode == 1) {
            // FP32 showdown columns, the per-board sweep programs and the column -> (terminal, player) map
            rc = A.alloc(nbc, &B.D.Bf);
            if (rc) return rc;
            MTT_HIP(hipMemsetAsync(B.D.Bf, 0, nbc * 4, s.stream));
            std::vector<uint32_t> prog, hdr((size_t)nb1 * 4);
            for (int bi = 0; bi < nb1; ++bi)
                MTT_ARG(sd_program(&cc1[(size_t)bi * NPR * 2], &rank1[(size_t)bi * NPR], nlive1[bi], &prog,
                                   &hdr[(size_t)bi * 4]),
                        "river board %d: combos not in strength order", bi);
            prog.resize(prog.size() + (size_t)SD_PAD_RECORDS * SD_REC, 0u);
            std::vector<int> ci(std::max(L1.ncol, 1), 0);
            for (const DNode& d : H1.nodes)
                if (d.type == DN_SD) {
                    ci[d.col] = d.sidx;
                    ci[d.col + 1] = (int)((uint32_t)d.sidx | 0x80000000u);
                }
            rc = A.upload(prog, &B.d_sdprog);
            rc = rc ? rc : A.upload(hdr, &B.d_sdhdr);
            rc = rc ? rc : A.upload(ci, &B.d_colinfo);
            if (rc) return rc;
        } else {
            rc = rc ? rc : A.alloc(nbc, &B.D.Bh);
            rc = rc ? rc : A.alloc(nbc, &B.D.Bl);
            if (rc) return rc;
            rc = A.alloc((size_t)nb1 * NPR * NPR, &B.d_S8);
            if (rc) return rc;
            hipLaunchKernelGGL(k_bt_s8, dim3(NPR, nb1), dim3(256), 0, s.stream, B.D, B.d_S8);
            rc = A.alloc((size_t)nb1 * GNCH * NPR * 2, &B.d_S2);
            if (rc) return rc;
            hipLaunchKernelGGL(k_bt_s2, dim3(NPR, nb1), dim3(GNCH * 2), 0, s.stream, B.D, B.d_S2);
            MTT_HIP(hipGetLastError());
            MTT_HIP(hipMemsetAsync(B.D.Bh, 0, nbc * 2, s.stream));
            MTT_HIP(hipMemsetAsync(B.D.Bl, 0, nbc * 2, s.stream));
        }
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
        MTT_ARG(L0.maxslot <= 64, "the turn template has %d action slots (the batched path holds 64)", L0.maxslot);
        B.D.nleaf = (int)leaf_tmpl.size();
        int *d_lt, *d_tl;
        rc = rc ? rc : A.upload(leaf_tmpl, &d_lt);
        rc = rc ? rc : A.upload(tmpl_leaf, &d_tl);
        rc = rc ? rc : A.alloc((size_t)49 * B.D.nleaf * 2 * NPT, &B.D.leafreach);
        rc = rc ? rc : A.alloc((size_t)49 * B.D.nleaf * 2 * NPT, &B.D.leafval);
        rc = rc ? rc : A.alloc((size_t)nb1 * L1.ntmpl * 2 * NPR, &B.D.rootval);
        rc = rc ? rc : A.alloc((size_t)49 * 2 * NPT, &B.D.rootval0);
        rc = rc ? rc : A.alloc((size_t)49 * L0.ncolp * NPT, &B.D.B0);
        rc = rc ? rc : A.alloc((size_t)49 * L0.ncolp * NPT, &B.D.C0);
        float* d_E;
        rc = rc ? rc : A.alloc((size_t)49 * NPT * NPT, &d_E);
        if (rc) return rc;
        B.D.leaf_tmpl = d_lt; B.D.tmpl_leaf = d_tl; B.D.Eturn = d_E;
        MTT_HIP(hipMemsetAsync(B.D.leafreach, 0, (size_t)49 * B.D.nleaf * 2 * NPT * 4, s.stream));
        MTT_HIP(hipMemsetAsync(B.D.leafval, 0, (size_t)49 * B.D.nleaf * 2 * NPT * 4, s.stream));
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
    // ------------------------------------------------ generated sweeps --
    if (bt_jit_enabled()) {
        const std::vector<JitLevel> lv = jit_levels(H0, H1, two);
        JitConfig jc = bt_jit_config();
        jc.sd_fp32 = B.sd_mode ? 1 : 0;
        rc = bt_jit_build(lv.data(), (int)lv.size(), jc, &B.jit);
        if (rc) return rc;
    }
    if (!B.sd_mode) {
        rc = river_gemm_autotune(B, s.stream);
        if (rc) return rc;
    }
    MTT_HIP(hipStreamSynchronize(s.stream));
    return MTT_OK;
}

// River showdown GEMM in one of its forms, by shape code: nw * 10 + pd for
// the column-major form (k_bt_gemm5: NW column tiles per workgroup, PD K
// chunks prefetched), 100 for the hero-major form (k_bt_gemm6), 101 for
// the hero-major form reading S at two bits per pair (k_bt_gemm6s), 102
// for the latter with the reach fetched two chunks ahead (k_bt_gemm6d), and
// 700 + GH for the register-streamed single-wave form with GH hero tiles
// per wave (k_bt_gemm7, GH = 5, 6, 7 at two waves per SIMD; 12 at one wave per
// SIMD, its accumulators in AGPRs).
int river_gemm(const BatchedState& B, int shape, hipStream_t st) {
    const int ntiles = B.D.L[1].ncolp / 32;
    if (shape >= 700 && shape < 800) {
        const int gh = shape - 700, nhg = (NPR / 32 + gh - 1) / gh;
        // a last column tile with at most 16 live columns is computed as 16-column products
        const int ntail = ntiles > 1 && B.D.L[1].ncol - 32 * (ntiles - 1) <= 16 ? 1 : 0;
        const int nmain = ntiles - ntail;
        const int nlog = B.nb1 * (nmain * nhg + ntail * ((NPR / 32 + G7T_GH - 1) / G7T_GH));
        const dim3 g(((nlog + 7) / 8) * 8);
        switch (gh) {
        case 5: hipLaunchKernelGGL(k_bt_gemm7<5>, g, dim3(64), 0, st, B.D, B.d_S2, nlog, nmain, ntail); break;
        case 6: hipLaunchKernelGGL(k_bt_gemm7<6>, g, dim3(64), 0, st, B.D, B.d_S2, nlog, nmain, ntail); break;
        case 7: hipLaunchKernelGGL(k_bt_gemm7<7>, g, dim3(64), 0, st, B.D, B.d_S2, nlog, nmain, ntail); break;
        case 12: hipLaunchKernelGGL((k_bt_gemm7<12, 1>), g, dim3(64), 0, st, B.D, B.d_S2, nlog, nmain, ntail); break;
        default: set_error("river GEMM shape %d not available", shape); return MTT_ERR_ARG;
        }
        MTT_HIP(hipGetLastError());
        return MTT_OK;
    }
    if (shape == 102) {
        const int nlog = B.nb1 * G6_HB;
        const dim3 g(((nlog + 7) / 8) * 8);
        switch ((ntiles + 1) / 2) {
        case 1: hipLaunchKernelGGL(k_bt_gemm6d<1>, g, dim3(512), 0, st, B.D, B.d_S2, nlog); break;
        case 2: hipLaunchKernelGGL(k_bt_gemm6d<2>, g, dim3(512), 0, st, B.D, B.d_S2, nlog); break;
        case 3: hipLaunchKernelGGL(k_bt_gemm6d<3>, g, dim3(512), 0, st, B.D, B.d_S2, nlog); break;
        case 4: hipLaunchKernelGGL(k_bt_gemm6d<4>, g, dim3(512), 0, st, B.D, B.d_S2, nlog); break;
        default: hipLaunchKernelGGL(k_bt_gemm6d<5>, g, dim3(512), 0, st, B.D, B.d_S2, nlog); break;
        }
        MTT_HIP(hipGetLastError());
        return MTT_OK;
    }
    if (shape == 101) {
        const int nlog = B.nb1 * G6_HB;
        const dim3 g(((nlog + 7) / 8) * 8);
        switch ((ntiles + 1) / 2) {
        case 1: hipLaunchKernelGGL(k_bt_gemm6s<1>, g, dim3(512), 0, st, B.D, B.d_S2, nlog); break;
        case 2: hipLaunchKernelGGL(k_bt_gemm6s<2>, g, dim3(512), 0, st, B.D, B.d_S2, nlog); break;
        case 3: hipLaunchKernelGGL(k_bt_gemm6s<3>, g, dim3(512), 0, st, B.D, B.d_S2, nlog); break;
        case 4: hipLaunchKernelGGL(k_bt_gemm6s<4>, g, dim3(512), 0, st, B.D, B.d_S2, nlog); break;
        default: hipLaunchKernelGGL(k_bt_gemm6s<5>, g, dim3(512), 0, st, B.D, B.d_S2, nlog); break;
        }
        MTT_HIP(hipGetLastError());
        return MTT_OK;
    }
    if (shape == 100) {
        const int nlog = B.nb1 * G6_HB;
        const dim3 g(((nlog + 7) / 8) * 8);
        switch ((ntiles + 1) / 2) {
        case 1: hipLaunchKernelGGL(k_bt_gemm6<1>, g, dim3(512), 0, st, B.D, B.d_S8, nlog); break;
        case 2: hipLaunchKernelGGL(k_bt_gemm6<2>, g, dim3(512), 0, st, B.D, B.d_S8, nlog); break;
        case 3: hipLaunchKernelGGL(k_bt_gemm6<3>, g, dim3(512), 0, st, B.D, B.d_S8, nlog); break;
        case 4: hipLaunchKernelGGL(k_bt_gemm6<4>, g, dim3(512), 0, st, B.D, B.d_S8, nlog); break;
        default: hipLaunchKernelGGL(k_bt_gemm6<5>, g, dim3(512), 0, st, B.D, B.d_S8, nlog); break;
        }
        MTT_HIP(hipGetLastError());
        return MTT_OK;
    }
    const int nw = shape / 10, pd = shape % 10;
    MTT_ARG(nw >= 1 && nw <= 9 && pd >= 2 && pd <= 3, "river GEMM shape %d not available", shape);
    const int ncg = (ntiles + nw - 1) / nw;
    const int nlog = B.nb1 * GHB * ncg;
    const dim3 g(((nlog + 7) / 8) * 8);
#define MTT_RG(NW_)                                                                                               \
    case NW_:                                                                                                     \
        if (pd == 2) hipLaunchKernelGGL((k_bt_gemm5<NW_, 2>), g, dim3(NW_ * 64), 0, st, B.D, B.d_S8, nlog, ncg);  \
        else hipLaunchKernelGGL((k_bt_gemm5<NW_, 3>), g, dim3(NW_ * 64), 0, st, B.D, B.d_S8, nlog, ncg);          \
        break;
    switch (nw) { MTT_RG(1) MTT_RG(2) MTT_RG(3) MTT_RG(4) MTT_RG(5) MTT_RG(6) MTT_RG(7) MTT_RG(8) MTT_RG(9) }
#undef MTT_RG
    MTT_HIP(hipGetLastError());
    return MTT_OK;
}

// Autotuner of the river GEMM: the register-streamed form (5, 6 or 7 hero
// tiles per wave), the hero-major forms, and the column-major form with
// all of the board's column tiles, or a half or a third of them, per
// workgroup, prefetching 2 or 3 K chunks; each timed on the real operands,
// the winner kept in the tuning file per column-tile count.
// MTT_RIVER_GEMM=<shape code> forces one.
int river_gemm_autotune(BatchedState& B, hipStream_t st) {
    const int ntiles = B.D.L[1].ncolp / 32;
    auto valid = [&](int v) {
        return (v >= 100 && v <= 102) || v == 705 || v == 706 || v == 707 || v == 712 ||
               (v / 10 >= 1 && v / 10 <= std::min(ntiles, 9) && v % 10 >= 2 && v % 10 <= 3);
    };
    if (const char* e = getenv("MTT_RIVER_GEMM")) {
        const int v = atoi(e);
        MTT_ARG(valid(v), "MTT_RIVER_GEMM=%s is not a river GEMM form for %d column tiles", e, ntiles);
        B.gemm_shape = v;
        return MTT_OK;
    }
    const std::string key = "river_gemm_shape/tiles" + std::to_string(ntiles);
    const int v = tune_lookup(key);
    if (valid(v)) { B.gemm_shape = v; return MTT_OK; }
    std::vector<int> cand{707, 712, 706, 705, 102, 101, 100};
    for (int d = 1; d <= 3; ++d) {
        const int nw = (ntiles + d - 1) / d;
        for (int pd = 2; pd <= 3; ++pd)
            if (std::find(cand.begin(), cand.end(), nw * 10 + pd) == cand.end()) cand.push_back(nw * 10 + pd);
    }
    hipEvent_t a, b;
    MTT_HIP(hipEventCreate(&a));
    MTT_HIP(hipEventCreate(&b));
    float best = 1e30f;
    int rc = MTT_OK;
    // the clocks ramp up under load: a few untimed launches first, so the first candidate is not penalised
    for (int r = 0; r < 4 && !rc; ++r) rc = river_gemm(B, cand[0], st);
    for (int shape : cand) {
        rc = rc ? rc : river_gemm(B, shape, st);  // warm
        if (rc) break;
        (void)hipEventRecord(a, st);
        for (int r = 0; r < 3 && !rc; ++r) rc = river_gemm(B, shape, st);
        if (rc) break;
        (void)hipEventRecord(b, st);
        (void)hipEventSynchronize(b);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, a, b);
        if (ms < best) { best = ms; B.gemm_shape = shape; }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc) return rc;
    tune_store(key, B.gemm_shape);
    return MTT_OK;
}

// One generated sweep over the grid.
int jit_launch(hipFunction_t f, dim3 grid, int nt, size_t shmem, hipStream_t st, void** args) {
    MTT_HIP(hipModuleLaunchKernel(f, grid.x, 1, 1, nt, 1, 1, (unsigned)shmem, st, args, nullptr));
    return MTT_OK;
}

int batched_run(SolverImpl& s, int iters) {
    BatchedState& B = *s.batched;
    const LevelDev& L1 = B.D.L[1];
    int rc = MTT_OK;
    for (int it = 0; it < iters && rc == MTT_OK; ++it) {
        int t = s.iters_done + 1;
        float wp, wn, ws;
        iteration_weights(s.cfg, t, wp, wn, ws);
        const float* ranges = B.d_ranges;
        if (B.two_level) {
            const LevelDev& L0 = B.D.L[0];
            const size_t sh = B.jit.active ? fwd_lds(L0.maxterm, B.jit.cfg) : (size_t)L0.maxterm * 104 * 8;
            const dim3 g(49 * L0.ntmpl * (B.jit.active ? fwd_groups(0, B.jit.cfg) : 2));
            void* args[] = {&B.D, &ranges, &ws};
            MTT_TIMED(s, "turn_forward",
                      if (B.jit.active) rc = jit_launch(B.jit.fwd[0], g, fwd_threads(0, B.jit.cfg), sh, s.stream, args);
                      else if (B.dm[0] == 4) hipLaunchKernelGGL((k_bt_forward<0, 4, SWEEP_NT>), g, dim3(SWEEP_NT), sh, s.stream, B.D, ranges, ws, nullptr);
                      else hipLaunchKernelGGL((k_bt_forward<0, 8, SWEEP_NT>), g, dim3(SWEEP_NT), sh, s.stream, B.D, ranges, ws, nullptr));
        }
        {
            const size_t sh = B.jit.active ? fwd_lds(L1.maxterm, B.jit.cfg) : (size_t)L1.maxterm * 104 * 8;
            const dim3 g(B.nb1 * L1.ntmpl * (B.jit.active ? fwd_groups(1, B.jit.cfg) : 2));
            void* args[] = {&B.D, &ranges, &ws};
            MTT_TIMED(s, "river_forward",
                      if (B.jit.active) rc = jit_launch(B.jit.fwd[1], g, fwd_threads(1, B.jit.cfg), sh, s.stream, args);
                      else if (B.dm[1] == 4) hipLaunchKernelGGL((k_bt_forward<1, 4, SWEEP_NT>), g, dim3(SWEEP_NT), sh, s.stream, B.D, ranges, ws, nullptr);
                      else hipLaunchKernelGGL((k_bt_forward<1, 8, SWEEP_NT>), g, dim3(SWEEP_NT), sh, s.stream, B.D, ranges, ws, nullptr));
        }
        if (B.sd_mode == 2) {
            MTT_TIMED(s, "river_showdown", sc_launch(B.D, B.d_schtab, B.d_scltab, B.nb1, B.D.L[1].ncol, s.stream));
        } else if (B.sd_mode == 1) {
            MTT_TIMED(s, "river_showdown",
                      sd_launch(B.D, B.d_sdprog, reinterpret_cast<const uint4*>(B.d_sdhdr), B.d_colinfo, B.nb1,
                                B.D.L[1].ncol, s.stream));
        } else {
            MTT_TIMED(s, "river_showdown", rc = river_gemm(B, B.gemm_shape, s.stream));
        }
        {
            const int jnt = B.jit.cfg.bwd_nt, jnh = bwd_groups(1, B.jit.cfg);
            const size_t sh = B.jit.active ? bwd_lds(L1, jnt) : bwd_lds(L1);
            const dim3 g(B.nb1 * L1.ntmpl * (B.jit.active ? jnh : BWD_NH));
            void* args[] = {&B.D, &ranges, &wp, &wn};
            MTT_TIMED(s, "river_backward",
                      if (B.jit.active) rc = jit_launch(B.jit.bwd[1], g, jnt, sh, s.stream, args);
                      else if (B.dm[1] == 4) hipLaunchKernelGGL((k_bt_backward<1, 4, BWD_NT>), g, dim3(BWD_NT), sh, s.stream, B.D, ranges, wp, wn, nullptr);
                      else hipLaunchKernelGGL((k_bt_backward<1, 8, BWD_NT>), g, dim3(BWD_NT), sh, s.stream, B.D, ranges, wp, wn, nullptr));
        }
        if (B.two_level) {
            const LevelDev& L0 = B.D.L[0];
            MTT_TIMED(s, "turn_leafval",
                      hipLaunchKernelGGL(k_bt_leafval, dim3(NPT / 256 + 1, 49 * B.D.nleaf), dim3(256), 0, s.stream, B.D));
            MTT_TIMED(s, "turn_showdown",
                      hipLaunchKernelGGL(k_bt_turn_sd<16>, dim3(NPT / 64, 49), dim3(256), 0, s.stream, B.D,
                                         L0.ncol));
            const int jnt = B.jit.cfg.bwd_nt, jnh = bwd_groups(0, B.jit.cfg);
            const size_t sh = B.jit.active ? bwd_lds(L0, jnt) : bwd_lds(L0);
            const dim3 g(49 * L0.ntmpl * (B.jit.active ? jnh : BWD_NH));
            void* args[] = {&B.D, &ranges, &wp, &wn};
            MTT_TIMED(s, "turn_backward",
                      if (B.jit.active) rc = jit_launch(B.jit.bwd[0], g, jnt, sh, s.stream, args);
                      else if (B.dm[0] == 4) hipLaunchKernelGGL((k_bt_backward<0, 4, BWD_NT>), g, dim3(BWD_NT), sh, s.stream, B.D, ranges, wp, wn, nullptr);
                      else hipLaunchKernelGGL((k_bt_backward<0, 8, BWD_NT>), g, dim3(BWD_NT), sh, s.stream, B.D, ranges, wp, wn, nullptr));
        }
        if (rc) return rc;
        MTT_HIP(hipGetLastError());
        s.iters_done = t;
    }
    return rc;
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

int batched_jit_info(const SolverImpl& s, int* active, double* compile_s) {
    const bool on = s.batched && s.batched->jit.active;
    if (active) *active = on ? 1 : 0;
    if (compile_s) *compile_s = on ? s.batched->jit.compile_s : 0.0;
    return MTT_OK;
}

// The last iteration's river showdown columns and values of one board, with
// the board's combo order (for inspection and tests).
int batched_river_showdown(SolverImpl& s, int board, float* reach, float* values, uint8_t* cards, uint16_t* grank,
                           int* nlive, int* ncol) {
    MTT_ARG(s.batched, "river showdown: not a batched handle");
    BatchedState& B = *s.batched;
    const LevelDev& L1 = B.D.L[1];
    MTT_ARG(board >= 0 && board < B.nb1, "river showdown: board %d outside [0, %d)", board, B.nb1);
    MTT_HIP(hipStreamSynchronize(s.stream));
    int nl = 0;
    MTT_HIP(hipMemcpy(&nl, L1.nlive + board, sizeof(int), hipMemcpyDeviceToHost));
    if (nlive) *nlive = nl;
    if (ncol) *ncol = L1.ncol;
    const size_t n = (size_t)L1.ncol * NPR, o = (size_t)board * L1.ncolp * NPR;
    if (values) MTT_HIP(hipMemcpy(values, B.D.C + o, n * 4, hipMemcpyDeviceToHost));
    if (reach) {
        if (B.D.Bf) {
            MTT_HIP(hipMemcpy(reach, B.D.Bf + o, n * 4, hipMemcpyDeviceToHost));
        } else {
            std::vector<uint16_t> hi(n), lo(n);
            MTT_HIP(hipMemcpy(hi.data(), B.D.Bh + o, n * 2, hipMemcpyDeviceToHost));
            MTT_HIP(hipMemcpy(lo.data(), B.D.Bl + o, n * 2, hipMemcpyDeviceToHost));
            for (size_t x = 0; x < n; ++x) {
                uint32_t uh = (uint32_t)hi[x] << 16, ul = (uint32_t)lo[x] << 16;
                float fh, fl;
                memcpy(&fh, &uh, 4);
                memcpy(&fl, &ul, 4);
                reach[x] = fh + fl;
            }
        }
    }
    if (cards) MTT_HIP(hipMemcpy(cards, L1.cc + (size_t)board * NPR * 2, NPR * 2, hipMemcpyDeviceToHost));
    if (grank) MTT_HIP(hipMemcpy(grank, B.D.grank + (size_t)board * NPR, NPR * 2, hipMemcpyDeviceToHost));
    return MTT_OK;
}

// Generated sweep source of a game's street templates (no device needed).
int hu_sweep_source(const mtt_hu_game& g, std::string* src) {
    HuTree tr;
    int rc = build_hu_tree(g, tr, false);
    if (rc) return rc;
    LevelHost H0, H1;
    bool two = false;
    rc = hu_levels(g, tr, two, H0, H1);
    if (rc) return rc;
    const std::vector<JitLevel> lv = jit_levels(H0, H1, two);
    JitConfig jc = bt_jit_config();
    jc.sd_fp32 = river_sd_mode() ? 1 : 0;
    *src = bt_jit_source(lv.data(), (int)lv.size(), jc);
    return MTT_OK;
}


// ===================================================== flop instances ===
// Many independent heads-up flop subgames (each its own board, stacks, pot
// and bet menu) solved together: one private template per instance in the
// street sweeps, the showdown over the 990 turn/river runouts as a batched
// GEMM of exact W - L counts against the reach columns, and iterations
// replayed from a captured hipGraph.
namespace {

constexpr int FNCH = NPF / 32;  // K chunks of 32 villains

// W - L of every (hero, villain) pair from exhaustive counts (alive-1216
// layout) as int16 (|W - L| <= 990; blocked and padded pairs 0), layout
// [board][chunk][hero][32] like the sign bytes of the river path.  The GEMM
// splits each value exactly into BF16 parts when it stages S.
__global__ void k_flop_s16(const int32_t* __restrict__ W, const int32_t* __restrict__ T, const uint8_t* __restrict__ cc,
                           int b0, int16_t* __restrict__ S16) {
    const int i = blockIdx.x, bl = blockIdx.y, b = b0 + bl;
    const uint8_t a1 = cc[((size_t)b * NPF + i) * 2], a2 = cc[((size_t)b * NPF + i) * 2 + 1];
    const int32_t* w = W + ((size_t)bl * NPF + i) * NPF;
    const int32_t* t = T + ((size_t)bl * NPF + i) * NPF;
    for (int j = threadIdx.x; j < NPF; j += blockDim.x) {
        int v = 0;
        if (i < kAliveFlop && j < kAliveFlop) {
            const uint8_t b1 = cc[((size_t)b * NPF + j) * 2], b2 = cc[((size_t)b * NPF + j) * 2 + 1];
            if (a1 != b1 && a1 != b2 && a2 != b1 && a2 != b2) v = 2 * w[j] + t[j] - 990;
        }
        S16[(((size_t)b * FNCH + (j >> 5)) * NPF + i) * 32 + (j & 31)] = (int16_t)v;
    }
}

// Eight int16 W - L values (one uint4) split exactly into BF16 parts:
// v = 16 * (v >> 4) + (v & 15), |16 * (v >> 4)| <= 992 (six significant
// bits) and v & 15 in [0, 15].  Element order is kept.
__device__ __forceinline__ unsigned s16_hi2(unsigned u) {  // 16 * (v >> 4) of both halves, as bf16 pair
    const float h0 = (float)(((int)(u << 16) >> 20) * 16), h1 = (float)(((int)u >> 20) * 16);
    return __builtin_amdgcn_perm(__float_as_uint(h1), __float_as_uint(h0), 0x07060302u);
}
__device__ __forceinline__ unsigned s16_lo2(unsigned u) {  // v & 15 of both halves, as bf16 pair
    const float l0 = (float)(u & 15u), l1 = (float)((u >> 16) & 15u);
    return __builtin_amdgcn_perm(__float_as_uint(l1), __float_as_uint(l0), 0x07060302u);
}
__device__ __forceinline__ void s16_split(const uint4& w, uint4& hi, uint4& lo) {
    const unsigned a = w.x, b = w.y, c = w.z, d = w.w;
    hi = make_uint4(s16_hi2(a), s16_hi2(b), s16_hi2(c), s16_hi2(d));
    lo = make_uint4(s16_lo2(a), s16_lo2(b), s16_lo2(c), s16_lo2(d));
}

// Flop showdown GEMM: C[b][col][i] = sum_j (Sh + Sl)[i][j] * (Bh + Bl)[b][col][j]
// in four BF16 products per K step, S = W - L read as int16 and split into
// Sh + Sl while it is staged.  HT 32-hero tiles per workgroup, one 32-column
// tile per wave (waves past the board's columns skip the math); S chunks
// staged in LDS (double-buffered), reach fragments streamed one chunk ahead;
// XCD-grouped numbering as in the river GEMM.
// S staging: piece x < npieces of a chunk is uint4 x of the int16 array,
// x = hero * 4 + (villains 8q .. 8q+7).
template <int NT, int PIECES, int SPT>
__device__ __forceinline__ void flop_load_s(uint4 (&sv)[SPT], const uint4* __restrict__ s16, int c, int npieces) {
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
        const int x = threadIdx.x + NT * q;
        const bool ok = x < PIECES && x < npieces;
        sv[q] = s16[(size_t)c * (NPF * 4) + (ok ? x : 0)];  // every path writes sv (keeps it in registers)
    }
}
template <int NT, int PIECES, int SPT, int HT>
__device__ __forceinline__ void flop_store_s(const uint4 (&sv)[SPT], uint4 (&dst)[2][HT][2][2][32], int npieces) {
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
        const int x = threadIdx.x + NT * q;
        if (x < PIECES && x < npieces) {
            const int hero = x >> 2, qq = x & 3;
            uint4 hi, lo;
            s16_split(sv[q], hi, lo);
            dst[0][hero >> 5][qq >> 1][qq & 1][hero & 31] = hi;
            dst[1][hero >> 5][qq >> 1][qq & 1][hero & 31] = lo;
        }
    }
}

template <int NW, int HT>
__global__ __launch_bounds__((NW < 4 ? 4 : NW) * 64) void k_bt_gemm_flop(BatchedDev D, const int16_t* __restrict__ S16,
                                                                         const int* __restrict__ bncol, int nlogical) {
    constexpr int NT = (NW < 4 ? 4 : NW) * 64;  // at least four waves share the S staging
    constexpr int HB = (NPF + 32 * HT - 1) / (32 * HT);
    constexpr int PIECES = HT * 128;            // uint4 pieces of the int16 S array per chunk
    constexpr int SPT = (PIECES + NT - 1) / NT;
    __shared__ uint4 Ssh[2][2][HT][2][2][32];   // [buf][hi/lo][hero tile][h][k-step][hero]
    const int per = (nlogical + 7) >> 3;
    const int logical = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (logical >= nlogical) return;
    const int b = logical / HB, i0 = (logical % HB) * (32 * HT);
    const int nht = min(HT, (NPF - i0) >> 5);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, half = lane >> 5, l32 = lane & 31;
    const int ncolp = NW * 32;
    const bool wave_on = wave < NW && wave * 32 < bncol[b];
    const int col = (wave < NW ? wave : 0) * 32 + l32;
    const uint4* ah_p = reinterpret_cast<const uint4*>(D.Bh + ((size_t)b * ncolp + col) * NPF + 16 * half);
    const uint4* al_p = reinterpret_cast<const uint4*>(D.Bl + ((size_t)b * ncolp + col) * NPF + 16 * half);
    const uint4* sp = reinterpret_cast<const uint4*>(S16 + ((size_t)b * FNCH * NPF + i0) * 32);
    const int npieces = nht * 128;
    f32x16_t acc[HT];
#pragma unroll
    for (int t = 0; t < HT; ++t)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[t][r] = 0.0f;
    uint4 sv[SPT];
    uint4 a_cur[4], a_nxt[4];
    flop_load_s<NT, PIECES, SPT>(sv, sp, 0, npieces);
    a_cur[0] = ah_p[0]; a_cur[1] = ah_p[1]; a_cur[2] = al_p[0]; a_cur[3] = al_p[1];
    flop_store_s<NT, PIECES, SPT, HT>(sv, Ssh[0], npieces);
    __syncthreads();
    int buf = 0;
    for (int c = 0; c < FNCH; ++c) {
        const bool more = c + 1 < FNCH;
        if (more) {
            flop_load_s<NT, PIECES, SPT>(sv, sp, c + 1, npieces);
            const int o = (c + 1) * 4;
            a_nxt[0] = ah_p[o]; a_nxt[1] = ah_p[o + 1]; a_nxt[2] = al_p[o]; a_nxt[3] = al_p[o + 1];
        }
        if (wave_on) {
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) {
                bf16x8_t ah, al;
                __builtin_memcpy(&ah, &a_cur[ks], 16);
                __builtin_memcpy(&al, &a_cur[2 + ks], 16);
#pragma unroll
                for (int t = 0; t < HT; ++t) {
                    if (t < nht) {
                        bf16x8_t bh, bl;
                        __builtin_memcpy(&bh, &Ssh[buf][0][t][half][ks][l32], 16);
                        __builtin_memcpy(&bl, &Ssh[buf][1][t][half][ks][l32], 16);
                        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc[t], 0, 0, 0);
                        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc[t], 0, 0, 0);
                        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc[t], 0, 0, 0);
                        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bl, acc[t], 0, 0, 0);
                    }
                }
            }
        }
        if (more) flop_store_s<NT, PIECES, SPT, HT>(sv, Ssh[buf ^ 1], npieces);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) a_cur[q] = a_nxt[q];
        buf ^= 1;
    }
    if (!wave_on) return;
    float* Cb = D.C + (size_t)b * ncolp * NPF;
#pragma unroll
    for (int t = 0; t < HT; ++t) {
        if (t < nht) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int cc = wave * 32 + (r >> 2) * 8 + half * 4 + (r & 3);
                Cb[(size_t)cc * NPF + i0 + 32 * t + l32] = acc[t][r];
            }
        }
    }
}

// Register-streamed flop showdown GEMM: one single-wave workgroup per
// (instance, group of HT hero tiles), no LDS and no barrier.  The W - L rows
// are the stream that matters (3 MB per instance, read once): per 32-villain
// chunk a lane loads its hero row's 16 villains (two 16-byte loads per hero
// tile) and its column's reach in both planes, three chunks deep in register
// sets used in rotation; each W - L piece is split into BF16 parts in
// registers.  The instance's column tiles are taken one after the other
// (most instances have one).  Per accumulator the order of the four
// products per K step is that of k_bt_gemm_flop, so the results are
// identical to it.
template <int HT>
__device__ __forceinline__ void gemm_flop7_wave(const BatchedDev& D, const int16_t* __restrict__ S16,
                                                const int* __restrict__ bncol, int b, int hg) {
    constexpr int NTH = NPF / 32;                 // 38 hero tiles
    const int lane = threadIdx.x & 63, half = lane >> 5, l32 = lane & 31;
    const int ht0 = hg * HT;
    const int nht = __builtin_amdgcn_readfirstlane(min(HT, NTH - ht0));    // HT, or fewer in the last group
    constexpr bool kOneShort = NTH % HT == 0 || NTH % HT == HT - 1;    // at most one tile absent: guard only the last
    const int nct = __builtin_amdgcn_readfirstlane((bncol[b] + 31) >> 5);
    const int ncolp = D.L[2].ncolp;
    // W - L of hero tile t: uint4 [b][chunk][hero][4], this lane's half = pieces 2h, 2h + 1
    const uint4* sp = reinterpret_cast<const uint4*>(S16) + (size_t)b * FNCH * NPF * 4 + 2 * half;
    int hrow[HT];
#pragma unroll
    for (int t = 0; t < HT; ++t) hrow[t] = (min(ht0 + t, NTH - 1) * 32 + l32) * 4;
    for (int ct = 0; ct < nct; ++ct) {
        const size_t row = ((size_t)b * ncolp + ct * 32 + l32) * NPF;
        const uint4* rh = reinterpret_cast<const uint4*>(D.Bh + row) + 2 * half;
        const uint4* rl = reinterpret_cast<const uint4*>(D.Bl + row) + 2 * half;
        f32x16_t acc[HT];
#pragma unroll
        for (int t = 0; t < HT; ++t)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[t][r] = 0.0f;
        uint4 ah[2], al[2], as[HT][2], bh[2], bl[2], bs[HT][2], ch[2], cl[2], cs[HT][2];
#define GF7_LOAD(C_, H_, L_, S_)                                                                  \
        do {                                                                                      \
            const int c_ = (C_);                                                                  \
            _Pragma("unroll") for (int k = 0; k < 2; ++k) {                                       \
                H_[k] = rh[4 * c_ + k];                                                           \
                L_[k] = rl[4 * c_ + k];                                                           \
            }                                                                                     \
            _Pragma("unroll") for (int t = 0; t < HT; ++t)                                        \
                _Pragma("unroll") for (int k = 0; k < 2; ++k)                                     \
                    S_[t][k] = sp[(size_t)c_ * NPF * 4 + hrow[t] + k];                            \
        } while (0)
#define GF7_MMA(H_, L_, S_)                                                                       \
        do {                                                                                      \
            _Pragma("unroll") for (int ks = 0; ks < 2; ++ks) {                                    \
                bf16x8_t fh, fl, sh[HT], sl[HT];                                                  \
                __builtin_memcpy(&fh, &H_[ks], 16);                                               \
                __builtin_memcpy(&fl, &L_[ks], 16);                                               \
                _Pragma("unroll") for (int t = 0; t < HT; ++t) {                                  \
                    uint4 hi, lo;                                                                 \
                    s16_split(S_[t][ks], hi, lo);                                                 \
                    __builtin_memcpy(&sh[t], &hi, 16);                                            \
                    __builtin_memcpy(&sl[t], &lo, 16);                                            \
                }                                                                                 \
                _Pragma("unroll") for (int t = 0; t < HT; ++t)                                    \
                    if ((kOneShort && t < HT - 1) || t < nht) acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fh, sh[t], acc[t], 0, 0, 0); \
                _Pragma("unroll") for (int t = 0; t < HT; ++t)                                    \
                    if ((kOneShort && t < HT - 1) || t < nht) acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fl, sh[t], acc[t], 0, 0, 0); \
                _Pragma("unroll") for (int t = 0; t < HT; ++t)                                    \
                    if ((kOneShort && t < HT - 1) || t < nht) acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fh, sl[t], acc[t], 0, 0, 0); \
                _Pragma("unroll") for (int t = 0; t < HT; ++t)                                    \
                    if ((kOneShort && t < HT - 1) || t < nht) acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fl, sl[t], acc[t], 0, 0, 0); \
            }                                                                                     \
        } while (0)
        constexpr int LAST = FNCH - 1;
        GF7_LOAD(0, ah, al, as);
        GF7_LOAD(1, bh, bl, bs);
        int c = 0;
        for (; c + 2 < FNCH; c += 3) {
            GF7_LOAD(c + 2, ch, cl, cs);
            __builtin_amdgcn_s_setprio(1);
            GF7_MMA(ah, al, as);
            __builtin_amdgcn_s_setprio(0);
            GF7_LOAD(min(c + 3, LAST), ah, al, as);
            __builtin_amdgcn_s_setprio(1);
            GF7_MMA(bh, bl, bs);
            __builtin_amdgcn_s_setprio(0);
            GF7_LOAD(min(c + 4, LAST), bh, bl, bs);
            __builtin_amdgcn_s_setprio(1);
            GF7_MMA(ch, cl, cs);
            __builtin_amdgcn_s_setprio(0);
        }
        if (c < FNCH) GF7_MMA(ah, al, as);
        if (c + 1 < FNCH) GF7_MMA(bh, bl, bs);
#undef GF7_LOAD
#undef GF7_MMA
        float* Cb = D.C + ((size_t)b * ncolp + ct * 32) * NPF;
#pragma unroll
        for (int t = 0; t < HT; ++t) {
            if ((kOneShort && t < HT - 1) || t < nht) {
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int cc = (r >> 2) * 8 + half * 4 + (r & 3);
                    Cb[(size_t)cc * NPF + (ht0 + t) * 32 + l32] = acc[t][r];
                }
            }
        }
    }
}

template <int HT>
__global__ __launch_bounds__(64) void k_bt_gemm_flop7(BatchedDev D, const int16_t* __restrict__ S16,
                                                      const int* __restrict__ bncol, int nlogical) {
    constexpr int NHG = (NPF / 32 + HT - 1) / HT;
    const int per = (nlogical + 7) >> 3;
    const int logical = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (logical >= nlogical) return;
    gemm_flop7_wave<HT>(D, S16, bncol, logical / NHG, logical % NHG);
}

constexpr int FLOP_NT = 640;   // sweep workgroup size for flop instances: two cover the 1216 padded combos

// Next iteration's weights (w_pos, w_neg, w_strat) from the precomputed
// table: part of the captured iteration, so replays advance by themselves.
// With the persistent iteration (pq non-null) it also rewinds the work queue
// and advances the epoch that the completion counters are compared with.
__global__ void k_next_weights(int* __restrict__ t, const float* __restrict__ wtab, float* __restrict__ wcur,
                               int* __restrict__ pq) {
    if (threadIdx.x == 0) {
        const int n = *t + 1;
        *t = n;
        wcur[0] = wtab[3 * (size_t)(n - 1)];
        wcur[1] = wtab[3 * (size_t)(n - 1) + 1];
        wcur[2] = wtab[3 * (size_t)(n - 1) + 2];
        if (pq) {
            pq[0] = 0;
            pq[1] += 1;
        }
    }
}

// ------------------------------------------- persistent flop iteration --
// One launch per iteration: workgroups of PF_NT threads take work items
// from a queue until it is empty.  Items (phase << 28 | instance << 8 | sub):
//   phase 0, forward half `sub` of an instance (640 combos in three passes);
//   phase 1, showdown GEMM of hero groups sub * 4 .. sub * 4 + 3 (one per
//            wave, PF_HT hero tiles each), after both forward halves;
//   phase 2, backward slice `sub` (256 combos), after every GEMM item.
// Four waves per workgroup, one per SIMD, so each may hold up to 512
// registers (the deepest backward needs about 250 architectural ones).
// The host orders the queue by blocks of instances, largest first, as
// forward(k), GEMM(k - 1), backward(k - 2), so the phases of different
// instances overlap.  An item depends only on items earlier in the queue,
// which running workgroups hold (a workgroup takes its next item only after
// finishing the current one), so every wait ends.  Completion counters grow
// by one per finished item over the whole solve and are compared with
// epoch x items-per-instance.  The bodies are those of the phase kernels,
// so the results are identical to theirs.
constexpr int PF_NT = 256;
constexpr int PF_HT = 4;
constexpr int PF_NWV = PF_NT / 64;
constexpr int PF_NHG = (NPF / 32 + PF_HT - 1) / PF_HT;
constexpr int PF_NFWD = 2;
constexpr int PF_NGEMM = (PF_NHG + PF_NWV - 1) / PF_NWV;
constexpr int PF_NBWD = (NPF + PF_NT - 1) / PF_NT;

struct FlopQueue {
    int* q;                    // [0] next item, [1] epoch, [2] a wait ran out of its bound
    int* fdone;                // [instance] forward items finished
    int* gdone;                // [instance] GEMM items finished
    const uint32_t* items;
    int nitems;
};

// Completion of an item: the barrier makes every wave's stores complete (to
// L2), then one lane writes the L2 back for the other XCDs (release fence at
// device scope) and increments the instance's counter.  One fence per item:
// a device-scope fence on this GPU writes back or invalidates a whole XCD's
// L2, so it must not run per wave or per poll.
__device__ __forceinline__ void flop_signal(int* cnt) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_fetch_add(cnt, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Wait until an instance's counter reaches target: one lane polls with
// relaxed device-scope loads (coherent across XCDs, no cache maintenance)
// and a bounded number of sleeps; past the bound (several seconds) it raises
// the error flag and the item runs anyway, so the kernel always ends.  Then
// one acquire fence (the CU's L1 and the XCD's L2 are shared by the
// workgroup's waves) and the barrier.
__device__ __forceinline__ void flop_wait(const int* cnt, int target, int* err) {
    if (threadIdx.x == 0) {
        int spins = 0;
        while (__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (++spins > (1 << 22)) {
                __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(8);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

template <int DM, int NRV>
__global__ __launch_bounds__(PF_NT) void k_flop_persistent(BatchedDev D, const float* __restrict__ ranges,
                                                           const float* __restrict__ wcur,
                                                           const int16_t* __restrict__ S16,
                                                           const int* __restrict__ bncol, FlopQueue Q) {
    extern __shared__ unsigned long long dyn[];   // forward bins or backward card sums + showdown values
    __shared__ EvRec sev[MAXE];
    __shared__ int s_item;
    const int ep = Q.q[1];
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(Q.q, 1);
        __syncthreads();
        const int it = s_item;
        __syncthreads();
        if (it >= Q.nitems) break;
        const uint32_t w = Q.items[it];
        const int phase = (int)(w >> 28), inst = (int)((w >> 8) & 0xFFFFFu), sub = (int)(w & 0xFFu);
        // the argument block read through an opaque zero offset per item: otherwise its uniform loads are
        // hoisted out of the item loop and stay live across all three bodies (scalar, then vector spills)
        int z;
        asm volatile("s_mov_b32 %0, 0" : "=s"(z));
        const BatchedDev& Dz = *reinterpret_cast<const BatchedDev*>(reinterpret_cast<const char*>(&D) + z);
        if (phase == 0) {
            bt_forward_body<2, DM, PF_NT, NRV, (FLOP_NT + PF_NT - 1) / PF_NT>(Dz, ranges, 0.0f, wcur, inst, sub, sev, dyn);
            flop_signal(Q.fdone + inst);
        } else if (phase == 1) {
            flop_wait(Q.fdone + inst, ep * PF_NFWD, Q.q + 2);
            const int hg = sub * PF_NWV + (int)(threadIdx.x >> 6);
            if (hg < PF_NHG) gemm_flop7_wave<PF_HT>(Dz, S16, bncol, inst, hg);
            flop_signal(Q.gdone + inst);
        } else {
            flop_wait(Q.gdone + inst, ep * PF_NGEMM, Q.q + 2);
            bt_backward_body<2, DM, PF_NT, NRV>(Dz, ranges, 0.0f, 0.0f, wcur, inst, sub, sev, dyn);
        }
    }
}

}  // namespace

struct BatchImpl {
    std::vector<mtt_hu_game> games;
    std::vector<float> ranges_h;     // [instance][2][1326], as given (best responses rebuild the games)
    mtt_solver_cfg cfg;
    std::vector<HuTree> trees;
    std::vector<int> ncol;           // showdown columns per instance
    int n = 0, dm = 4, ht = 4;
    int ntb = 640;                   // backward workgroup size (640, or 320 for wide, deep or LDS-heavy trees)
    int iters_done = 0, tmax = 0;
    hipStream_t stream = nullptr;
    KernelTimer timer;
    DeviceArena arena;
    BatchedDev D{};            // value-initialised: pointers a mode does not use stay null
    float* d_ranges = nullptr;
    int16_t* d_S16 = nullptr;        // W - L, [instance][chunk][hero][32]
    int* d_bncol = nullptr;
    int* d_t = nullptr;
    float *d_wtab = nullptr, *d_wcur = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool use_graph = true;
    double gemm_ms[5] = {0, 0, 0, 0, 0};   // autotuner times of the flop GEMM forms 2, 4, 12, 13, 14
    // iteration form: 0 = forward, GEMM and backward launches; 1 = one persistent launch; -1 = to be tuned
    int iter_mode = -1;
    bool pf_ok = false;              // the persistent form fits (LDS)
    size_t pf_lds = 0;
    int pf_grid = 0;
    int* d_pq = nullptr;             // queue head, epoch, timeout flag
    int* d_fdone = nullptr;
    int* d_gdone = nullptr;
    uint32_t* d_items = nullptr;
    int nitems = 0;
    double mode_ms[2] = {0, 0};      // tuned per-iteration times of the two forms
    int graph_mode = -1;             // the form the captured graph holds
};

namespace {

// Flop GEMM by tuning code: 2 or 4 = k_bt_gemm_flop with that many hero
// tiles per workgroup, 10 + HT = the register-streamed k_bt_gemm_flop7<HT>.
int batch_launch_gemm(BatchImpl& B, int ht) {
    const LevelDev& L = B.D.L[2];
    if (ht > 10) {
        const int h = ht - 10, nlog = B.n * ((NPF / 32 + h - 1) / h);
        const dim3 g(((nlog + 7) / 8) * 8);
        switch (h) {
        case 2: hipLaunchKernelGGL(k_bt_gemm_flop7<2>, g, dim3(64), 0, B.stream, B.D, B.d_S16, B.d_bncol, nlog); break;
        case 3: hipLaunchKernelGGL(k_bt_gemm_flop7<3>, g, dim3(64), 0, B.stream, B.D, B.d_S16, B.d_bncol, nlog); break;
        case 4: hipLaunchKernelGGL(k_bt_gemm_flop7<4>, g, dim3(64), 0, B.stream, B.D, B.d_S16, B.d_bncol, nlog); break;
        default: set_error("flop GEMM tuning code %d not available", ht); return MTT_ERR_ARG;
        }
        MTT_HIP(hipGetLastError());
        return MTT_OK;
    }
    const int hb = ht == 4 ? (NPF + 127) / 128 : (NPF + 63) / 64;
    const int nlog = B.n * hb;
    const dim3 g(((nlog + 7) / 8) * 8);
    const int nw = L.ncolp / 32;
#define MTT_FG(NW_)                                                                                                   \
    case NW_:                                                                                                         \
        if (ht == 4) hipLaunchKernelGGL((k_bt_gemm_flop<NW_, 4>), g, dim3((NW_ < 4 ? 4 : NW_) * 64), 0, B.stream, B.D,  \
                                        B.d_S16, B.d_bncol, nlog);                                                   \
        else hipLaunchKernelGGL((k_bt_gemm_flop<NW_, 2>), g, dim3((NW_ < 4 ? 4 : NW_) * 64), 0, B.stream, B.D,       \
                                B.d_S16, B.d_bncol, nlog);                                                           \
        break;
    switch (nw) {
        MTT_FG(1) MTT_FG(2) MTT_FG(3) MTT_FG(4) MTT_FG(5) MTT_FG(6) MTT_FG(7) MTT_FG(8) MTT_FG(9)
        default: set_error("flop batch: %d showdown column tiles exceed 9", nw); return MTT_ERR_ARG;
    }
#undef MTT_FG
    MTT_HIP(hipGetLastError());
    return MTT_OK;
}

int batch_iteration(BatchImpl& B) {
    const LevelDev& L = B.D.L[2];
    // a third register vector for more than 64 slots; the backward runs 320-thread groups (more registers per
    // thread) for those and for deep trees
    const bool wide = L.maxslot > 64;
    const int ntb = B.ntb;
    const size_t shf = (size_t)L.maxterm * 104 * 8;
    const size_t shb = bwd_lds(L, ntb);
    const dim3 gf(B.n * 2), gb(B.n * ((NPF + ntb - 1) / ntb));  // backward: ceil(NPF / ntb) workgroups per instance
    if (B.iter_mode == 1) {
        hipLaunchKernelGGL(k_next_weights, dim3(1), dim3(64), 0, B.stream, B.d_t, B.d_wtab, B.d_wcur, B.d_pq);
        const FlopQueue Q{B.d_pq, B.d_fdone, B.d_gdone, B.d_items, B.nitems};
#define MTT_PF(DM_, NRV_)                                                                                           \
    hipLaunchKernelGGL((k_flop_persistent<DM_, NRV_>), dim3(B.pf_grid), dim3(PF_NT), B.pf_lds, B.stream, B.D,       \
                       B.d_ranges, B.d_wcur, B.d_S16, B.d_bncol, Q)
        MTT_TIMED(B, "flop_persistent",
                  if (B.dm == 4) { if (wide) MTT_PF(4, 3); else MTT_PF(4, 2); }
                  else { if (wide) MTT_PF(8, 3); else MTT_PF(8, 2); });
#undef MTT_PF
        MTT_HIP(hipGetLastError());
        return MTT_OK;
    }
    hipLaunchKernelGGL(k_next_weights, dim3(1), dim3(64), 0, B.stream, B.d_t, B.d_wtab, B.d_wcur, (int*)nullptr);
#define MTT_FWD(DM_, NRV_) \
    hipLaunchKernelGGL((k_bt_forward<2, DM_, FLOP_NT, NRV_>), gf, dim3(FLOP_NT), shf, B.stream, B.D, B.d_ranges, 0.0f, B.d_wcur)
#define MTT_BWD(DM_, NT_, NRV_) \
    hipLaunchKernelGGL((k_bt_backward<2, DM_, NT_, NRV_>), gb, dim3(NT_), shb, B.stream, B.D, B.d_ranges, 0.0f, 0.0f, B.d_wcur)
    MTT_TIMED(B, "flop_forward",
              if (B.dm == 4) { if (wide) MTT_FWD(4, 3); else MTT_FWD(4, 2); }
              else { if (wide) MTT_FWD(8, 3); else MTT_FWD(8, 2); });
    int rc = MTT_OK;
    MTT_TIMED(B, "flop_showdown_gemm", rc = batch_launch_gemm(B, B.ht));
    if (rc) return rc;
    MTT_TIMED(B, "flop_backward",
              if (B.dm == 4) {
                  if (wide) MTT_BWD(4, FLOP_NT / 2, 3);
                  else if (ntb == FLOP_NT) MTT_BWD(4, FLOP_NT, 2);
                  else MTT_BWD(4, FLOP_NT / 2, 2);
              } else { if (wide) MTT_BWD(8, FLOP_NT / 2, 3); else MTT_BWD(8, FLOP_NT / 2, 2); });
#undef MTT_FWD
#undef MTT_BWD
    MTT_HIP(hipGetLastError());
    return MTT_OK;
}

// Weight table for iterations [1, tmax] on the device.
int batch_weights(BatchImpl& B, int tmax) {
    std::vector<float> w((size_t)3 * tmax);
    for (int t = 1; t <= tmax; ++t) iteration_weights(B.cfg, t, w[3 * (t - 1)], w[3 * (t - 1) + 1], w[3 * (t - 1) + 2]);
    float* d = nullptr;
    int rc = B.arena.upload(w, &d);
    if (rc) return rc;
    B.d_wtab = d;
    B.tmax = tmax;
    return MTT_OK;
}

// Autotuner: the flop GEMM's form and hero-tile count (k_bt_gemm_flop with 2
// or 4 tiles of 32 per workgroup, k_bt_gemm_flop7 with 2, 3 or 4 per wave)
// is timed on the real operands once per column-tile count and batch-size
// bucket, and the winner is kept in the tuning file (see tune_path).
int batch_autotune(BatchImpl& B) {
    static const int kCand[5] = {2, 4, 12, 13, 14};
    auto valid = [](int v) { return v == 2 || v == 4 || v == 12 || v == 13 || v == 14; };
    if (const char* e = getenv("MTT_FLOP_GEMM")) {
        const int v = atoi(e);
        if (valid(v)) { B.ht = v; return MTT_OK; }
    }
    // the best form depends on how many workgroups fill the GPU: tuned per power-of-two batch size
    int bucket = 1;
    while (bucket < B.n) bucket <<= 1;
    const std::string key = "flop_gemm_form/nw" + std::to_string(B.D.L[2].ncolp / 32) + "/n" + std::to_string(bucket);
    int v = tune_lookup(key);
    if (valid(v)) { B.ht = v; return MTT_OK; }
    hipEvent_t a, b;
    MTT_HIP(hipEventCreate(&a));
    MTT_HIP(hipEventCreate(&b));
    int best = 0;
    for (int r = 0; r < 4; ++r) {   // the clocks ramp up under load: untimed launches first
        int rc = batch_launch_gemm(B, kCand[0]);
        if (rc) return rc;
    }
    for (int k = 0; k < 5; ++k) {
        int rc = batch_launch_gemm(B, kCand[k]);  // warm
        if (rc) return rc;
        MTT_HIP(hipEventRecord(a, B.stream));
        for (int r = 0; r < 3; ++r) batch_launch_gemm(B, kCand[k]);
        MTT_HIP(hipEventRecord(b, B.stream));
        MTT_HIP(hipEventSynchronize(b));
        float ms = 0;
        MTT_HIP(hipEventElapsedTime(&ms, a, b));
        B.gemm_ms[k] = ms / 3.0;
        if (B.gemm_ms[k] < B.gemm_ms[best]) best = k;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    B.ht = kCand[best];
    tune_store(key, B.ht);
    return MTT_OK;
}

}  // namespace

int batch_persistent_setup(BatchImpl& B);

int batch_create(const mtt_hu_game* games, int n, const float* ranges, const mtt_solver_cfg* cfg, BatchImpl** out) {
    *out = nullptr;
    auto h = new (std::nothrow) BatchImpl;
    if (!h) { set_error("mtt_batch_create: out of host memory"); return MTT_ERR_OOM; }
    std::unique_ptr<BatchImpl> own(h);
    BatchImpl& B = *h;
    B.n = n;
    B.games.assign(games, games + n);
    B.ranges_h.assign(ranges, ranges + (size_t)n * 2 * 1326);
    for (int k = 0; k < n; ++k)
        for (int p = 0; p < 2; ++p) {
            double tot = 0.0;
            for (int i = 0; i < 1326; ++i) {
                const float v = ranges[((size_t)k * 2 + p) * 1326 + i];
                MTT_ARG(std::isfinite(v) && v >= 0, "mtt_batch_create: instance %d range %d has a negative or "
                        "non-finite weight", k, p);
                tot += v;
            }
            MTT_ARG(tot < kMaxRangeTotal, "mtt_batch_create: instance %d range %d totals %g, at least %g", k, p, tot,
                    kMaxRangeTotal);
        }
    if (cfg) B.cfg = *cfg;
    else { memset(&B.cfg, 0, sizeof(B.cfg)); B.cfg.scheme = MTT_SCHEME_CFR_PLUS; }
    memset(&B.D, 0, sizeof(B.D));
    B.trees.resize(n);
    LevelHost H;
    int maxdepth = 0;
    for (int k = 0; k < n; ++k) {
        const mtt_hu_game& g = games[k];
        MTT_ARG(g.n_board == 3 && !g.root_chance && g.last_street == 0,
                "mtt_batch_create: instance %d is not a flop subgame (3-card board, flop betting, showdown over the "
                "turn and river)", k);
        int rc = build_hu_tree(g, B.trees[k], true);
        if (rc) return rc;
        MTT_ARG(B.trees[k].templates.size() == 1, "mtt_batch_create: instance %d has more than one street", k);
        rc = add_template(B.trees[k].templates[0], H, 0);
        if (rc) return rc;
        B.ncol.push_back(2 * B.trees[k].templates[0].n_showdown);
        maxdepth = std::max(maxdepth, H.maxdepth);
    }
    B.dm = maxdepth + 1 <= 4 ? 4 : 8;
    MTT_ARG(maxdepth < 8, "mtt_batch_create: a tree deeper than 8 decisions");
    // boards: alive combos of each flop in the reference order
    std::vector<uint8_t> cc((size_t)n * NPF * 2, 0);
    std::vector<int> nlive(n, kAliveFlop);
    std::vector<float> rr((size_t)n * 2 * NPF, 0.0f), pot0(n);
    std::vector<int> btmpl(n);
    std::vector<uint8_t> flops((size_t)n * 3);
    for (int k = 0; k < n; ++k) {
        const mtt_hu_game& g = games[k];
        const uint64_t fm = (1ull << g.board[0]) | (1ull << g.board[1]) | (1ull << g.board[2]);
        int i = 0;
        for (int a = 0; a < 52; ++a)
            for (int c = a + 1; c < 52; ++c) {
                if (((1ull << a) | (1ull << c)) & fm) continue;
                cc[((size_t)k * NPF + i) * 2] = (uint8_t)a;
                cc[((size_t)k * NPF + i) * 2 + 1] = (uint8_t)c;
                const int gi = combo_index(a, c);
                rr[((size_t)k * 2 + 0) * NPF + i] = ranges[((size_t)k * 2 + 0) * 1326 + gi];
                rr[((size_t)k * 2 + 1) * NPF + i] = ranges[((size_t)k * 2 + 1) * 1326 + gi];
                i++;
            }
        pot0[k] = g.pot;
        btmpl[k] = k;
        for (int q = 0; q < 3; ++q) flops[(size_t)k * 3 + q] = g.board[q];
    }
    MTT_HIP(hipStreamCreate(&B.stream));
    DeviceArena& A = B.arena;
    LevelDev& L = B.D.L[2];
    // the backward must fit at the smaller workgroup size; it runs at FLOP_NT when the trees are neither wide
    // (more than 64 slots: a third register vector) nor deep (8-level reach vectors) and their LDS still fits
    int rc = upload_level(A, B.stream, H, L, NPF, n, cc, nlive, true, FLOP_NT / 2);
    if (rc) return rc;
    B.ntb = (L.maxslot > 64 || B.dm == 8 || bwd_lds(L, FLOP_NT) > 152 * 1024) ? FLOP_NT / 2 : FLOP_NT;
    int* d_bt;
    float* d_pot0;
    rc = rc ? rc : A.upload(btmpl, &d_bt);
    rc = rc ? rc : A.upload(pot0, &d_pot0);
    rc = rc ? rc : A.upload(rr, &B.d_ranges);
    rc = rc ? rc : A.upload(B.ncol, &B.d_bncol);
    if (rc) return rc;
    L.btmpl = d_bt;
    L.bpot0 = d_pot0;
    MTT_ARG(L.ncolp <= GCOLS, "mtt_batch_create: %d showdown columns exceed %d", L.ncol, GCOLS);
    const size_t nbc = (size_t)n * L.ncolp * NPF;
    rc = rc ? rc : A.alloc(nbc, &B.D.Bh);
    rc = rc ? rc : A.alloc(nbc, &B.D.Bl);
    rc = rc ? rc : A.alloc(nbc, &B.D.C);
    rc = rc ? rc : A.alloc((size_t)n * 2 * NPF, &B.D.rootval0);
    rc = rc ? rc : A.alloc((size_t)n * FNCH * NPF * 32, &B.d_S16);
    rc = rc ? rc : A.alloc(1, &B.d_t);
    rc = rc ? rc : A.alloc(3, &B.d_wcur);
    if (rc) return rc;
    MTT_HIP(hipMemsetAsync(B.D.Bh, 0, nbc * 2, B.stream));
    MTT_HIP(hipMemsetAsync(B.D.Bl, 0, nbc * 2, B.stream));
    MTT_HIP(hipMemsetAsync(B.D.C, 0, nbc * 4, B.stream));
    MTT_HIP(hipMemsetAsync(B.d_t, 0, sizeof(int), B.stream));
    // W - L of every instance, from the exhaustive counts, 16 flops at a time
    {
        const int CH = 16;
        DeviceArena tmp;
        int32_t *dW, *dT;
        rc = tmp.alloc((size_t)CH * NPF * NPF, &dW);
        rc = rc ? rc : tmp.alloc((size_t)CH * NPF * NPF, &dT);
        if (rc) return rc;
        for (int k0 = 0; k0 < n; k0 += CH) {
            const int m = std::min(CH, n - k0);
            rc = mtt_flop_counts(flops.data() + (size_t)k0 * 3, m, dW, dT, MTT_LAYOUT_ALIVE1216, B.stream);
            if (rc) return rc;
            hipLaunchKernelGGL(k_flop_s16, dim3(NPF, m), dim3(256), 0, B.stream, dW, dT, L.cc, k0, B.d_S16);
            MTT_HIP(hipGetLastError());
        }
        MTT_HIP(hipStreamSynchronize(B.stream));
    }
    rc = batch_weights(B, 4096);
    if (rc) return rc;
    rc = batch_autotune(B);
    if (rc) return rc;
    rc = batch_persistent_setup(B);
    if (rc) return rc;
    MTT_HIP(hipStreamSynchronize(B.stream));
    *out = own.release();
    return MTT_OK;
}

namespace {

void batch_drop_graph(BatchImpl& B) {
    if (B.exec) { (void)hipGraphExecDestroy(B.exec); B.exec = nullptr; }
    if (B.graph) { (void)hipGraphDestroy(B.graph); B.graph = nullptr; }
}

// Iterations in the current form: replayed from the captured graph, or
// launched one by one (graph replay off, or per-kernel timing on).
int batch_iters(BatchImpl& B, int iters) {
    if (!B.use_graph || B.timer.enabled) {
        for (int it = 0; it < iters; ++it) {
            int rc = batch_iteration(B);
            if (rc) return rc;
        }
    } else {
        if (B.exec && B.graph_mode != B.iter_mode) batch_drop_graph(B);
        if (!B.exec) {
            MTT_HIP(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal));
            int rc = batch_iteration(B);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(B.stream, &g);
            if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
            if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
            B.graph = g;
            MTT_HIP(hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0));
            B.graph_mode = B.iter_mode;
        }
        for (int it = 0; it < iters; ++it) MTT_HIP(hipGraphLaunch(B.exec, B.stream));
    }
    B.iters_done += iters;
    return MTT_OK;
}

// The iteration form (three phase launches, or the persistent launch): from
// MTT_FLOP_PERSISTENT, the tuning file, or timed on the solve itself -- two
// replayed iterations of each form after one that captures its graph (both
// forms give identical states, so the tuning iterations are ordinary
// iterations of the solve).  Returns the iterations it ran.
int batch_pick_mode(BatchImpl& B, int iters, int* ran) {
    *ran = 0;
    if (const char* e = getenv("MTT_FLOP_PERSISTENT")) {
        if (atoi(e) && !B.pf_ok) {
            set_error("mtt_batch_run: the persistent iteration is switched off in this build");
            return MTT_ERR_ARG;
        }
        B.iter_mode = atoi(e) ? 1 : 0;
        return MTT_OK;
    }
    if (!B.pf_ok) { B.iter_mode = 0; return MTT_OK; }
    int bucket = 1;
    while (bucket < B.n) bucket <<= 1;
    const std::string key = "flop_iter_mode/nw" + std::to_string(B.D.L[2].ncolp / 32) + "/n" + std::to_string(bucket);
    const int v = tune_lookup(key);
    if (v == 0 || v == 1) { B.iter_mode = v; return MTT_OK; }
    if (iters < 6 || !B.use_graph || B.timer.enabled) return MTT_OK;   // not yet: this call runs the phase launches
    hipEvent_t a, b;
    MTT_HIP(hipEventCreate(&a));
    MTT_HIP(hipEventCreate(&b));
    int rc = MTT_OK;
    for (int m = 0; m < 2 && !rc; ++m) {
        B.iter_mode = m;
        batch_drop_graph(B);
        rc = batch_iters(B, 1);
        if (rc) break;
        (void)hipEventRecord(a, B.stream);
        rc = batch_iters(B, 2);
        if (rc) break;
        (void)hipEventRecord(b, B.stream);
        (void)hipEventSynchronize(b);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, a, b);
        B.mode_ms[m] = ms / 2.0;
        *ran += 3;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc) return rc;
    const int best = B.mode_ms[1] < B.mode_ms[0] ? 1 : 0;
    if (best != B.iter_mode) batch_drop_graph(B);
    B.iter_mode = best;
    tune_store(key, best);
    return MTT_OK;
}

}  // namespace

// Work items, completion counters and launch shape of the persistent form.
// The form is switched off: its first GPU runs did not finish within three
// minutes, twice (the second after the polling lane stopped invalidating the
// L2 on every poll), while a CPU simulation of the same queue order and
// counters completes for every workgroup count; the cause is not known, so
// the three phase launches are used and MTT_FLOP_PERSISTENT=1 is refused.
constexpr bool kPersistentFlopIteration = false;

int batch_persistent_setup(BatchImpl& B) {
    B.pf_ok = false;
    if (!kPersistentFlopIteration) return MTT_OK;
    const LevelDev& L = B.D.L[2];
    B.pf_lds = std::max((size_t)L.maxterm * 104 * 8, bwd_lds(L, PF_NT));
    B.pf_ok = B.pf_lds + sizeof(EvRec) * MAXE + 64 <= 160 * 1024;
    if (!B.pf_ok) return MTT_OK;
    // instances by tree size, largest first, in blocks: forward(k), GEMM(k - 1), backward(k - 2)
    std::vector<int> ord(B.n);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return B.trees[x].flat.size() > B.trees[y].flat.size(); });
    constexpr int BK = 64;
    const int nblk = (B.n + BK - 1) / BK;
    std::vector<uint32_t> items;
    auto emit = [&](int phase, int blk, int nsub) {
        if (blk < 0 || blk >= nblk) return;
        for (int i = blk * BK; i < std::min(B.n, (blk + 1) * BK); ++i)
            for (int sub = 0; sub < nsub; ++sub)
                items.push_back((uint32_t)phase << 28 | (uint32_t)ord[i] << 8 | (uint32_t)sub);
    };
    for (int st = 0; st < nblk + 2; ++st) {
        emit(0, st, PF_NFWD);
        emit(1, st - 1, PF_NGEMM);
        emit(2, st - 2, PF_NBWD);
    }
    B.nitems = (int)items.size();
    int rc = B.arena.upload(items, &B.d_items);
    rc = rc ? rc : B.arena.alloc(4, &B.d_pq);
    rc = rc ? rc : B.arena.alloc((size_t)B.n, &B.d_fdone);
    rc = rc ? rc : B.arena.alloc((size_t)B.n, &B.d_gdone);
    if (rc) return rc;
    MTT_HIP(hipMemsetAsync(B.d_pq, 0, 4 * sizeof(int), B.stream));
    MTT_HIP(hipMemsetAsync(B.d_fdone, 0, (size_t)B.n * sizeof(int), B.stream));
    MTT_HIP(hipMemsetAsync(B.d_gdone, 0, (size_t)B.n * sizeof(int), B.stream));
    // as many workgroups as stay resident together (any count would be correct)
    const bool wide = L.maxslot > 64;
    int per_cu = 0, dev = 0;
    const void* fn = B.dm == 4 ? (wide ? reinterpret_cast<const void*>(&k_flop_persistent<4, 3>)
                                       : reinterpret_cast<const void*>(&k_flop_persistent<4, 2>))
                               : (wide ? reinterpret_cast<const void*>(&k_flop_persistent<8, 3>)
                                       : reinterpret_cast<const void*>(&k_flop_persistent<8, 2>));
    MTT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, PF_NT, B.pf_lds));
    MTT_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    MTT_HIP(hipGetDeviceProperties(&prop, dev));
    B.pf_grid = std::max(1, std::min(B.nitems, std::max(per_cu, 1) * prop.multiProcessorCount));
    return MTT_OK;
}

int batch_run(BatchImpl& B, int iters) {
    if (B.iters_done + iters > B.tmax) {
        int rc = batch_weights(B, std::max(2 * B.tmax, B.iters_done + iters));
        if (rc) return rc;
        batch_drop_graph(B);
    }
    int ran = 0;
    if (B.iter_mode < 0) {
        int rc = batch_pick_mode(B, iters, &ran);
        if (rc) return rc;
    }
    const bool tuning_pending = B.iter_mode < 0;
    if (tuning_pending) B.iter_mode = 0;
    int rc = batch_iters(B, iters - ran);
    if (tuning_pending) B.iter_mode = -1;   // the graph of the phase launches stays until a call can tune
    if (rc) return rc;
    MTT_HIP(hipStreamSynchronize(B.stream));
    if (B.pf_ok) {
        int flag = 0;
        MTT_HIP(hipMemcpy(&flag, B.d_pq + 2, sizeof(int), hipMemcpyDeviceToHost));
        if (flag) {
            set_error("mtt_batch_run: a persistent iteration's dependency wait exceeded its bound");
            return MTT_ERR_HIP;
        }
    }
    return MTT_OK;
}

// Raw regrets / strategy sums of one instance in its own flat layout
// [infoset][MTT_MAX_ACTIONS][1216].
int batch_state(BatchImpl& B, int k, float* R, float* S, size_t nf) {
    const HuTree& tr = B.trees[k];
    const size_t need = (size_t)tr.n_infosets * MTT_MAX_ACTIONS * NPF;
    MTT_ARG(nf >= need, "mtt_batch_get_state: buffer of %zu floats, need %zu", nf, need);
    const LevelDev& L = B.D.L[2];
    const size_t per = (size_t)L.nslot * NPF;
    std::vector<float> hR(per), hS(per);
    MTT_HIP(hipStreamSynchronize(B.stream));
    MTT_HIP(hipMemcpy(hR.data(), L.R + (size_t)k * per, per * 4, hipMemcpyDeviceToHost));
    MTT_HIP(hipMemcpy(hS.data(), L.S + (size_t)k * per, per * 4, hipMemcpyDeviceToHost));
    if (R) memset(R, 0, need * 4);
    if (S) memset(S, 0, need * 4);
    const Template& T = tr.templates[0];
    for (size_t id = 0; id < tr.flat.size(); ++id) {
        const mtt_flat_node& x = tr.flat[id];
        if (x.type != MTT_NODE_DECISION) continue;
        const int slot = T.nodes[tr.flat_tnode[id]].slot;
        for (int a = 0; a < x.num_actions; ++a)
            for (int i = 0; i < kAliveFlop; ++i) {
                const size_t o = ((size_t)x.infoset_offset * MTT_MAX_ACTIONS + a) * NPF + i;
                if (R) R[o] = hR[(size_t)(slot + a) * NPF + i];
                if (S) S[o] = hS[(size_t)(slot + a) * NPF + i];
            }
    }
    return MTT_OK;
}

// Average strategy of instance k in the reference layout: strategy sums
// normalised over each decision's actions, uniform where they sum to 0.
int batch_strategy(BatchImpl& B, int k, float* out, size_t nf) {
    const HuTree& tr = B.trees[k];
    const size_t need = (size_t)tr.n_infosets * MTT_MAX_ACTIONS * NPF;
    MTT_ARG(nf >= need, "mtt_batch_get_strategy: buffer of %zu floats, need %zu", nf, need);
    std::vector<float> S(need);
    int rc = batch_state(B, k, nullptr, S.data(), need);
    if (rc) return rc;
    memset(out, 0, need * 4);
    for (const mtt_flat_node& x : tr.flat) {
        if (x.type != MTT_NODE_DECISION) continue;
        const size_t base = (size_t)x.infoset_offset * MTT_MAX_ACTIONS * NPF;
        for (int i = 0; i < NPF; ++i) {
            float sum = 0.0f;
            for (int a = 0; a < x.num_actions; ++a) sum += S[base + (size_t)a * NPF + i];
            for (int a = 0; a < x.num_actions; ++a)
                out[base + (size_t)a * NPF + i] = sum > 0.0f ? S[base + (size_t)a * NPF + i] / sum : 1.0f / x.num_actions;
        }
    }
    return MTT_OK;
}

void batch_destroy(BatchImpl* B) {
    if (!B) return;
    (void)hipStreamSynchronize(B->stream);
    if (B->exec) (void)hipGraphExecDestroy(B->exec);
    if (B->graph) (void)hipGraphDestroy(B->graph);
    B->arena.release();
    (void)hipStreamDestroy(B->stream);
    delete B;
}

}  // namespace mtt

struct mtt_batch {
    mtt::BatchImpl* impl;
};

extern "C" {

int mtt_river_scan_tables(const uint8_t* cards, const uint16_t* grank, int n, uint32_t* htab, uint32_t* ltab) {
    using namespace mtt;
    MTT_ARG(cards && grank && htab && ltab, "mtt_river_scan_tables: null argument");
    MTT_ARG(bt::sc_tables(cards, grank, n, htab, ltab),
            "mtt_river_scan_tables: %d combos (need the 1081 of a river board), bad cards or unsorted ranks", n);
    return MTT_OK;
}

int mtt_river_sweep_program(const uint8_t* cards, const uint16_t* grank, int n, uint32_t* out, size_t cap,
                            size_t* len) {
    using namespace mtt;
    MTT_ARG(cards && grank && len, "mtt_river_sweep_program: null argument");
    std::vector<uint32_t> prog(bt::SD_REC, 0u);  // record 0: the board header
    uint32_t hdr[4];
    MTT_ARG(bt::sd_program(cards, grank, n, &prog, hdr),
            "mtt_river_sweep_program: %d combos, bad cards or unsorted ranks", n);
    for (int x = 0; x < 4; ++x) prog[x] = hdr[x];
    *len = prog.size();
    if (out) memcpy(out, prog.data(), std::min(cap, prog.size()) * 4);
    return MTT_OK;
}

int mtt_hu_sweep_source(const mtt_hu_game* game, char* buf, size_t cap, size_t* len) {
    using namespace mtt;
    MTT_ARG(game, "mtt_hu_sweep_source: null game");
    std::string src;
    int rc = hu_sweep_source(*game, &src);
    if (rc) return rc;
    if (len) *len = src.size();
    if (buf && cap) {
        const size_t n = std::min(cap - 1, src.size());
        memcpy(buf, src.data(), n);
        buf[n] = 0;
    }
    return MTT_OK;
}

int mtt_hu_sweep_compile(const mtt_hu_game* game, size_t* code_bytes, double* seconds) {
    using namespace mtt;
    MTT_ARG(game, "mtt_hu_sweep_compile: null game");
    std::string src;
    int rc = hu_sweep_source(*game, &src);
    if (rc) return rc;
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<char> code;
    rc = bt_jit_compile(src, &code);
    if (rc) return rc;
    if (code_bytes) *code_bytes = code.size();
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return MTT_OK;
}

int mtt_batch_create(const mtt_hu_game* games, int n, const float* ranges, const mtt_solver_cfg* cfg, mtt_batch** out) {
    (void)hipGetLastError();   // an error left by an earlier call is not this call's
    using namespace mtt;
    MTT_ARG(games && ranges && out, "mtt_batch_create: null pointer");
    MTT_ARG(n >= 1 && n <= 65535, "mtt_batch_create: instance count %d out of 1..65535", n);
    *out = nullptr;
    BatchImpl* b = nullptr;
    int rc = batch_create(games, n, ranges, cfg, &b);
    if (rc) return rc;
    auto h = new (std::nothrow) mtt_batch;
    if (!h) { batch_destroy(b); set_error("mtt_batch_create: out of host memory"); return MTT_ERR_OOM; }
    h->impl = b;
    *out = h;
    return MTT_OK;
}

int mtt_batch_run(mtt_batch* h, int iters) {
    using namespace mtt;
    MTT_ARG(h && iters >= 0, "mtt_batch_run: bad arguments");
    return batch_run(*h->impl, iters);
}

int mtt_batch_iterations(const mtt_batch* h) { return h ? h->impl->iters_done : -1; }

int mtt_batch_tree_info(mtt_batch* h, int instance, int* n_nodes, int* n_infosets) {
    using namespace mtt;
    MTT_ARG(h && n_nodes && n_infosets && instance >= 0 && instance < h->impl->n, "mtt_batch_tree_info: bad arguments");
    *n_nodes = (int)h->impl->trees[instance].flat.size();
    *n_infosets = h->impl->trees[instance].n_infosets;
    return MTT_OK;
}

int mtt_batch_get_state(mtt_batch* h, int instance, float* R_out, float* S_out, size_t n_floats) {
    using namespace mtt;
    MTT_ARG(h && instance >= 0 && instance < h->impl->n, "mtt_batch_get_state: bad arguments");
    return batch_state(*h->impl, instance, R_out, S_out, n_floats);
}

int mtt_batch_get_strategy(mtt_batch* h, int instance, float* out, size_t n_floats) {
    using namespace mtt;
    MTT_ARG(h && out && instance >= 0 && instance < h->impl->n, "mtt_batch_get_strategy: bad arguments");
    return batch_strategy(*h->impl, instance, out, n_floats);
}

int mtt_batch_exploitability_fp64(mtt_batch* h, int instance, double* out) {
    using namespace mtt;
    MTT_ARG(h && out && instance >= 0 && instance < h->impl->n, "mtt_batch_exploitability_fp64: bad arguments");
    BatchImpl& B = *h->impl;
    const HuTree& tr = B.trees[instance];
    const size_t need = (size_t)tr.n_infosets * MTT_MAX_ACTIONS * NPF;
    std::vector<float> st(need);
    int rc = batch_strategy(B, instance, st.data(), need);
    if (rc) return rc;
    const float* r = B.ranges_h.data() + (size_t)instance * 2 * 1326;
    return solver_exploitability_of(B.games[instance], r, r + 1326, st.data(), need, tr.n_infosets, out);
}

int mtt_batch_set_graph(mtt_batch* h, int on) {
    using namespace mtt;
    MTT_ARG(h, "mtt_batch_set_graph: null handle");
    h->impl->use_graph = on != 0;
    return MTT_OK;
}

int mtt_batch_tuning(mtt_batch* h, int* hero_tiles, double* gemm_ms) {
    using namespace mtt;
    MTT_ARG(h && hero_tiles, "mtt_batch_tuning: bad arguments");
    *hero_tiles = h->impl->ht;
    if (gemm_ms)
        for (int k = 0; k < 5; ++k) gemm_ms[k] = h->impl->gemm_ms[k];
    return MTT_OK;
}

int mtt_batch_iter_mode(mtt_batch* h, int* mode, double* ms) {
    using namespace mtt;
    MTT_ARG(h && mode, "mtt_batch_iter_mode: bad arguments");
    *mode = h->impl->iter_mode;
    if (ms) {
        ms[0] = h->impl->mode_ms[0];
        ms[1] = h->impl->mode_ms[1];
    }
    return MTT_OK;
}

int mtt_batch_timing(mtt_batch* h, int enable) {
    using namespace mtt;
    MTT_ARG(h, "mtt_batch_timing: null handle");
    MTT_HIP(hipStreamSynchronize(h->impl->stream));
    h->impl->timer.collect();
    h->impl->timer.enabled = enable != 0;
    h->impl->timer.names.clear();
    h->impl->timer.ms.clear();
    h->impl->timer.count.clear();
    return MTT_OK;
}

int mtt_batch_kernel_times(mtt_batch* h, int idx, char* name, int name_cap, double* total_ms, int64_t* launches) {
    using namespace mtt;
    MTT_ARG(h && name && total_ms && launches && name_cap > 0, "mtt_batch_kernel_times: bad arguments");
    MTT_HIP(hipStreamSynchronize(h->impl->stream));
    h->impl->timer.collect();
    if (idx < 0 || idx >= (int)h->impl->timer.names.size()) return -1;
    snprintf(name, name_cap, "%s", h->impl->timer.names[idx].c_str());
    *total_ms = h->impl->timer.ms[idx];
    *launches = h->impl->timer.count[idx];
    return MTT_OK;
}

int mtt_batch_destroy(mtt_batch* h) {
    if (!h) return MTT_OK;
    mtt::batch_destroy(h->impl);
    delete h;
    return MTT_OK;
}

}  // extern "C"
