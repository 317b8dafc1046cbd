// Generator and hipRTC build of the per-template sweeps (see bt_jit.h).
//
// Code shape (per level, one forward and one backward kernel; the template of
// a workgroup is a wave-uniform switch):
//   forward   regrets of the template's slots into named registers; for each
//             node in preorder: its reach from the parent's reach and
//             strategy, regret matching and the strategy-sum atomics
//             (decisions), the showdown reach columns and the card-sum LDS
//             atomics (terminals), the leaf reaches (turn level);
//   backward  card sums, totals and showdown values staged in LDS as in the
//             interpreted kernel; the DFS walk recomputes reaches top-down,
//             takes terminal values from the card sums and showdown values,
//             and at each decision's exit forms the node value and applies
//             R <- (R + r) * (R + r > 0 ? w+ : w-).
// Each statement mirrors the interpreted kernels' arithmetic in the same
// order (sums of positive regrets left to right, node value as a chain of
// fmaf over the actions, parent value accumulation in action order).
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include <hip/hiprtc.h>

#include "bt_jit.h"
#include "mtt_common.h"

namespace mtt {

namespace {

// The shared device header's text (generated from bt_dev.h at build time).
#include "bt_dev_src.inc"

using bt::DNode;

struct Src {
    std::string s;
    void operator()(const char* fmt, ...) __attribute__((format(printf, 2, 3))) {
        char buf[1024];
        va_list ap;
        va_start(ap, fmt);
        int n = vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        if (n >= (int)sizeof buf) n = (int)sizeof buf - 1;
        s.append(buf, (size_t)n);
        s.push_back('\n');
    }
};

// Exact float literal (hexadecimal significand).
std::string flit(float f) {
    char b[64];
    snprintf(b, sizeof b, "%af", (double)f);
    return b;
}

// 1/n exactly as the interpreted kernels' uniform_share.
float uniform_of(int n) { return n == 2 ? 0.5f : (n == 3 ? 1.0f / 3.0f : (n == 4 ? 0.25f : 1.0f)); }

struct Tmpl {
    const JitLevel& J;
    int t, n0, e0, e1, s0, nsl, t0, nt, c0, ncolt;
    Tmpl(const JitLevel& j, int tt) : J(j), t(tt) {
        n0 = (*J.tnode0)[t];
        e0 = (*J.tev0)[t];
        e1 = (*J.tev0)[t + 1];
        s0 = J.slot0[t];
        nsl = J.slot0[t + 1] - s0;
        t0 = J.term0[t];
        nt = J.term0[t + 1] - t0;
        c0 = J.col0[t];
        ncolt = J.col0[t + 1] - c0;
    }
    const DNode& node(int local) const { return (*J.nodes)[n0 + local]; }
};

// Reach of node n (template-local) from its parent p (-1 at the root):
// defines x0_n and x1_n.
void emit_reach(Src& o, const Tmpl& T, int n, int p) {
    const DNode& nd = T.node(n);
    if (p < 0) {
        o("const float x0_%d = root0, x1_%d = root1;", n, n);
        return;
    }
    const int pp = T.node(p).player;
    if (pp == 0)
        o("const float x0_%d = x0_%d * s%d_%d, x1_%d = x1_%d;", n, p, p, nd.act, n, p);
    else
        o("const float x0_%d = x0_%d, x1_%d = x1_%d * s%d_%d;", n, p, n, p, p, nd.act);
}

// Regret matching at decision n: defines s<n>_<a>.
void emit_strategy(Src& o, const Tmpl& T, int n) {
    const DNode& nd = T.node(n);
    const int ls = nd.slot - T.s0, A = nd.nact;
    std::string decl = "float", sum;
    for (int a = 0; a < A; ++a) {
        decl += (a ? ", s" : " s") + std::to_string(n) + "_" + std::to_string(a);
        sum += (a ? " + q" : "q") + std::to_string(a);
    }
    o("%s;", decl.c_str());
    o("{");
    for (int a = 0; a < A; ++a) o("const float q%d = fmaxf(R%d, 0.0f);", a, ls + a);
    o("const float sum = %s;", sum.c_str());
    o("const bool pz = sum > 1e-7f;");
    o("const float inv = pz ? __builtin_amdgcn_rcpf(sum) : 0.0f;");
    for (int a = 0; a < A; ++a) o("s%d_%d = pz ? q%d * inv : %s;", n, a, a, flit(uniform_of(A)).c_str());
    o("}");
}

// Card-sum bins of terminal n (forward).
void emit_bins(Src& o, const Tmpl& T, int n) {
    const DNode& nd = T.node(n);
    o("if (live) {");
    o("unsigned long long* bn = binl + %d;", (nd.sidx - T.t0) * 104);
    o("const unsigned long long u0 = fx48_of(x0_%d);", n);
    o("const unsigned long long u1 = fx48_of(x1_%d);", n);
    o("if (u0) { atomicAdd(&bn[c1], u0); atomicAdd(&bn[c2], u0); }");
    o("if (u1) { atomicAdd(&bn[52 + c1], u1); atomicAdd(&bn[52 + c2], u1); }");
    o("}");
}

void gen_forward_case(Src& o, const JitLevel& J, bool sd_fp32, int nbins, int slices, int smask, bool xpre, int t) {
    Tmpl T(J, t);
    xpre = xpre && slices > 1 && T.nsl > 0;
    o("case %d: {", t);
    o("fwd_bins_clear<NT>(bins, %d);", T.nt * nbins);
    if (nbins == 2) o("unsigned long long* const binl = bins + (threadIdx.x & 1) * %d;", T.nt * 104);
    else o("unsigned long long* const binl = bins;");
    if (slices > 1) {
        // one workgroup walks the combos in slices, adding every slice into the same bins
        o("__syncthreads();");
        if (xpre) {
            // xpre: the first slice's regrets here, each later slice's loaded during the walk before it
            std::string decl;
            for (int j = 0; j < T.nsl; ++j) decl += (j ? ", Rn" : "Rn") + std::to_string(j);
            o("float %s;", decl.c_str());
            o("{");
            o("const float* __restrict__ Rp = L.R + ((size_t)b * L.nslot + %d) * NP + min((int)threadIdx.x, NP - 1);",
              T.s0);
            for (int j = 0; j < T.nsl; ++j) o("Rn%d = Rp[%d];", j, j * J.np);
            o("}");
        }
        o("#pragma unroll 1");
        o("for (int pass = 0; pass < %d; ++pass) {", slices);
        o("if (__builtin_amdgcn_readfirstlane(pass * NT + ((int)threadIdx.x & ~63)) >= NP) continue;");
        o("const int k = pass * NT + (int)threadIdx.x;");
        o("const bool inb = k < NP, live = k < nl;");
        o("const int kc = min(k, NP - 1);");
        o("(void)inb;");
    }
    o("const size_t rb = ((size_t)b * L.nslot + %d) * NP + k;", T.s0);
    o("const float* __restrict__ Rl = L.R + ((size_t)b * L.nslot + %d) * NP + kc;", T.s0);
    o("float* __restrict__ Sg = L.S + rb;");
    o("const int c1r = L.cc[((size_t)b * NP + kc) * 2], c2r = L.cc[((size_t)b * NP + kc) * 2 + 1];");
    o("const int ri = root_index(D, LEVEL, b, kc);");
    if (xpre)
        for (int j = 0; j < T.nsl; ++j) o("const float Rv%d = Rn%d;", j, j);
    else
        for (int j = 0; j < T.nsl; ++j) o("const float Rv%d = Rl[%d];", j, j * J.np);
    // the strategy sums are updated by plain loads and stores: each (slot, combo) has one writer per
    // iteration, and a float atomic per entry cost the river forward ~22% of its time.  Either they are
    // read with the regrets, or (smask) each decision reads its own through a bounded buffer descriptor
    // once the acting player's reach is known, with lanes of zero reach sent out of range: those loads
    // return 0 without touching memory, and those lanes store nothing
    if (smask) {
        o("const __amdgpu_buffer_rsrc_t rsS = __builtin_amdgcn_make_buffer_rsrc("
          "(void*)(L.S + ((size_t)b * L.nslot + %d) * NP), (short)0, %d, 0x00020000);",
          T.s0, T.nsl * J.np * 4);
    } else {
        o("const float* __restrict__ Sl = L.S + ((size_t)b * L.nslot + %d) * NP + kc;", T.s0);
        for (int j = 0; j < T.nsl; ++j) o("const float Sv%d = Sl[%d];", j, j * J.np);
    }
    o("const float rr0 = root_reach_at(D, ranges, LEVEL, b, tm, NP, ri, 0);");
    o("const float rr1 = root_reach_at(D, ranges, LEVEL, b, tm, NP, ri, 1);");
    if (xpre) {
        // the next slice's regrets, behind this slice's loads and pinned there (as in the backward);
        // unconditional, the last slice's reading the clamped last combo
        o("__builtin_amdgcn_sched_barrier(0);");
        o("{");
        o("const float* __restrict__ Rp = L.R + ((size_t)b * L.nslot + %d) * NP + min(k + NT, NP - 1);", T.s0);
        for (int j = 0; j < T.nsl; ++j) o("Rn%d = Rp[%d];", j, j * J.np);
        o("}");
        o("__builtin_amdgcn_sched_barrier(0);");
    }
    o("const float root0 = live ? rr0 : 0.0f, root1 = live ? rr1 : 0.0f;");
    o("const int c1 = live ? c1r : 0, c2 = live ? c2r : 0;");
    for (int j = 0; j < T.nsl; ++j) o("const float R%d = live ? Rv%d : 0.0f;", j, j);
    if (slices == 1) o("__syncthreads();");
    // smask: strategy-sum updates of decisions deferred in groups of smask, each group's loads issued
    // together and then its stores (a load cannot pass an earlier store to the same descriptor), all
    // branch-free: the loads and the stores of a zero-reach lane (not-live lanes carry zero reach) both
    // fall outside the descriptor, so the loads return 0 and the stores are dropped
    std::vector<int> pend;
    auto flush_sums = [&]() {
        if (pend.empty()) return;
        o("{");
        o("#pragma clang fp contract(off)");
        for (int n : pend) {
            o("const float wr%d = ws * x%d_%d;", n, T.node(n).player, n);
            o("const uint32_t so%d = wr%d != 0.0f ? (uint32_t)kc * 4u : 0x80000000u;", n, n);
        }
        for (int n : pend) {
            const int ls = T.node(n).slot - T.s0;
            for (int a = 0; a < T.node(n).nact; ++a)
                o("const float Sv%d = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsS, so%d, "
                  "%d, 0));", ls + a, n, (ls + a) * J.np * 4);
        }
        for (int n : pend) {
            const int ls = T.node(n).slot - T.s0;
            for (int a = 0; a < T.node(n).nact; ++a)
                o("__builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, Sv%d + wr%d * s%d_%d), rsS, "
                  "so%d, %d, 0);", ls + a, n, n, a, n, (ls + a) * J.np * 4);
        }
        o("}");
        pend.clear();
    };
    int stack[16];
    for (int e = T.e0; e < T.e1; ++e) {
        const int ev = (*J.events)[e], n = ev >> 2, kind = ev & 3;
        if (kind == bt::EV_LEAVE) continue;
        const DNode& nd = T.node(n);
        const int d = nd.depth;
        const int p = d ? stack[d - 1] : -1;
        emit_reach(o, T, n, p);
        if (kind == bt::EV_ENTER) {
            stack[d] = n;
            emit_strategy(o, T, n);
            const int ls = nd.slot - T.s0;
            // a zero reach adds +0 to every strategy sum (all sums are >= +0), so its stores are skipped; the
            // product is rounded before the sum, as an atomic add of it would (no contraction into an FMA)
            if (smask) {
                pend.push_back(n);
                if ((int)pend.size() >= smask) flush_sums();
                continue;
            } else {
                o("if (live) {");
                o("const float wr = ws * x%d_%d;", nd.player, n);
                o("if (wr != 0.0f) {");
            }
            o("#pragma clang fp contract(off)");
            for (int a = 0; a < nd.nact; ++a) o("Sg[%d] = Sv%d + wr * s%d_%d;", (ls + a) * J.np, ls + a, n, a);
            o("}");
            o("}");
            continue;
        }
        if (nd.type == bt::DN_LEAF) {
            o("if (inb) {");
            o("float* lr = D.leafreach + ((size_t)b * D.nleaf + %d) * 2 * NPT + k;", nd.leaf);
            o("lr[0] = live ? x0_%d : 0.0f;", n);
            o("lr[NPT] = live ? x1_%d : 0.0f;", n);
            o("}");
            continue;
        }
        if (nd.type == bt::DN_SD) {
            o("if (inb) {");
            o("const float o0 = live ? x1_%d : 0.0f, o1 = live ? x0_%d : 0.0f;", n, n);
            if (J.level == 1) {
                o("const size_t cb = ((size_t)b * L.ncolp + %d) * NPR + k;", nd.col);
                if (sd_fp32) {
                    o("D.Bf[cb] = o0;");
                    o("D.Bf[cb + NPR] = o1;");
                } else {
                    o("store_split(D.Bh, D.Bl, cb, o0);");
                    o("store_split(D.Bh, D.Bl, cb + NPR, o1);");
                }
            } else {
                o("const size_t cb = ((size_t)b * L.ncolp + %d) * NPT + k;", nd.col);
                o("D.B0[cb] = o0;");
                o("D.B0[cb + NPT] = o1;");
            }
            o("}");
        }
        emit_bins(o, T, n);
    }
    flush_sums();
    if (slices > 1) o("}");  // pass loop; the sums of all slices go to the first half of the partials
    const char* half = slices > 1 ? "0" : "half";
    if (nbins == 2) o("fwd_bins_flush2<NT>(L, bins, b, %d, %s, %d);", T.t0, half, T.nt);
    else o("fwd_bins_flush<NT>(L, bins, b, %d, %s, %d);", T.t0, half, T.nt);
    o("break;");
    o("}");
}

// Value of node n for both players is in v0_<n>, v1_<n>; hand it to the
// parent p (or to the template root's outputs).
void emit_up(Src& o, const Tmpl& T, int n, int p) {
    if (p < 0) {
        o("out0 = v0_%d;", n);
        o("out1 = v1_%d;", n);
        return;
    }
    const int pp = T.node(p).player;
    const int a = T.node(n).act;
    o("cv%d_%d = v%d_%d;", p, a, pp, n);
    o("op%d += v%d_%d;", p, 1 - pp, n);
}

void gen_backward_case(Src& o, const JitLevel& J, const JitConfig& cfg, int passes, int t) {
    Tmpl T(J, t);
    // rpre > 0: a decision's regrets are loaded at the entry of the decision rpre places before it in
    // walk order (the first rpre in the prologue) instead of all in the prologue, so only the regrets of
    // the open decisions and of the next few stay live
    const int rpre = passes > 1 ? cfg.bwd_rpre : 0;
    // xpre: the next slice's regrets are loaded into Rn* while the current slice walks
    const bool xpre = passes > 1 && rpre == 0 && cfg.bwd_xpre;
    std::vector<int> dec;
    for (int e = T.e0; e < T.e1; ++e)
        if (((*J.events)[e] & 3) == bt::EV_ENTER) dec.push_back((*J.events)[e] >> 2);
    auto emit_rload = [&](int n) {
        const int ls = T.node(n).slot - T.s0;
        // unconditional loads from the clamped combo index, dead lanes zeroed after (no branch per load)
        for (int a = 0; a < T.node(n).nact; ++a) o("const float Rv%d = Rl[%d];", ls + a, (ls + a) * J.np);
        for (int a = 0; a < T.node(n).nact; ++a) o("const float R%d = live ? Rv%d : 0.0f;", ls + a, ls + a);
    };
    int nenter = 0;
    o("case %d: {", t);
    if (passes > 1) {
        // card sums and totals once; then each slice of the combos in turn (the showdown values a thread stages
        // in LDS are its own, so the passes need no barrier)
        if (T.nt > 0) o("bwd_sums_n<NTB, %d, %d>(L, ssum, b, %d);", T.nt, fwd_sliced(J.level, cfg) ? 1 : 2, T.t0);
        o("bwd_totals<NTB>(ssum, %d);", T.nt);
        if (xpre) {
            // the first slice's regrets (clamped combo index: every lane loads from a valid address)
            std::string decl;
            for (int j = 0; j < T.nsl; ++j) decl += (j ? ", Rn" : "Rn") + std::to_string(j);
            if (T.nsl > 0) o("float %s;", decl.c_str());
            o("{");
            o("const float* __restrict__ Rp = L.R + ((size_t)b * L.nslot + %d) * NP + min(tid, NP - 1);", T.s0);
            for (int j = 0; j < T.nsl; ++j) o("Rn%d = Rp[%d];", j, j * J.np);
            o("}");
        }
        o("#pragma unroll 1");  // one walk's registers at a time
        o("for (int pass = 0; pass < %d; ++pass) {", passes);
        // an offset the compiler cannot see is zero: the card-sum reads are not hoisted out of the pass loop
        // (where they would all stay live across the walk)
        o("int soff = 0;");
        o("asm volatile(\"\" : \"+v\"(soff));");
        o("const unsigned long long* ssv = ssum + soff;");
        // branch-free prologue (clamped combo index, dead lanes zeroed after the loads): the cards and the
        // root index first, then the regrets and the showdown values, then the root reaches, so every load
        // of the slice is in flight before the first wait
        // a wave whose combos all lie past the padded count has nothing to walk (no barrier follows)
        o("if (__builtin_amdgcn_readfirstlane(pass * NTB + (tid & ~63)) >= NP) continue;");
        o("const int k = pass * NTB + tid;");
        o("const bool live = k < nl;");
        o("const int kc = min(k, NP - 1);");
        o("const int c1r = L.cc[((size_t)b * NP + kc) * 2], c2r = L.cc[((size_t)b * NP + kc) * 2 + 1];");
        o("const int ri = root_index(D, LEVEL, b, kc);");
        o("const float* __restrict__ Rl = L.R + ((size_t)b * L.nslot + %d) * NP + kc;", T.s0);
        if (xpre)
            for (int j = 0; j < T.nsl; ++j) o("const float Rv%d = Rn%d;", j, j);
        else if (rpre == 0)
            for (int j = 0; j < T.nsl; ++j) o("const float Rv%d = Rl[%d];", j, j * J.np);
        else
            for (int i = 0; i < std::min(rpre, (int)dec.size()); ++i) emit_rload(dec[i]);
        if (T.ncolt > 0) o("bwd_cs_n<LEVEL, NTB, %d>(D, L, Cs, b, kc, live, %d);", T.ncolt, T.c0);
        o("const float rr0 = root_reach_at(D, ranges, LEVEL, b, tm, NP, ri, 0);");
        o("const float rr1 = root_reach_at(D, ranges, LEVEL, b, tm, NP, ri, 1);");
        if (xpre && T.nsl > 0) {
            // the next slice's regrets, in flight during this slice's walk; the scheduling barriers keep them
            // behind this slice's loads (whose waits would otherwise cover them) and keep the compiler from
            // sinking them towards their use in the next pass.  Unconditional (a branch around them would
            // make the waits after the join count them): in the last pass every lane reads the clamped last
            // combo, one cache line per slot, and the values go unused
            o("__builtin_amdgcn_sched_barrier(0);");
            o("{");
            o("const float* __restrict__ Rp = L.R + ((size_t)b * L.nslot + %d) * NP + min(k + NTB, NP - 1);", T.s0);
            for (int j = 0; j < T.nsl; ++j) o("Rn%d = Rp[%d];", j, j * J.np);
            o("}");
            o("__builtin_amdgcn_sched_barrier(0);");
        }
        o("const int c1 = live ? c1r : 0, c2 = live ? c2r : 0;");
        o("const float root0 = live ? rr0 : 0.0f, root1 = live ? rr1 : 0.0f;");
        o("float* __restrict__ Rg = L.R + ((size_t)b * L.nslot + %d) * NP + k;", T.s0);
        if (rpre == 0)
            for (int j = 0; j < T.nsl; ++j) o("const float R%d = live ? Rv%d : 0.0f;", j, j);
        o("float out0 = 0.0f, out1 = 0.0f;");
    } else {
        o("bwd_stage<LEVEL, NTB>(D, L, ssum, Cs, b, k, live, %d, %d, %d, %d);", T.nt, T.t0, T.c0, T.ncolt);
        o("float* __restrict__ Rg = L.R + ((size_t)b * L.nslot + %d) * NP + k;", T.s0);
        for (int j = 0; j < T.nsl; ++j) o("const float R%d = live ? Rg[%d] : 0.0f;", j, j * J.np);
        o("bwd_totals<NTB>(ssum, %d);", T.nt);
        o("const unsigned long long* ssv = ssum;");
    }
    int stack[16];
    for (int e = T.e0; e < T.e1; ++e) {
        const int ev = (*J.events)[e], n = ev >> 2, kind = ev & 3;
        const DNode& nd = T.node(n);
        const int d = nd.depth;
        const int p = d ? stack[d - 1] : -1;
        if (kind == bt::EV_ENTER) {
            stack[d] = n;
            if (rpre > 0 && nenter + rpre < (int)dec.size()) emit_rload(dec[nenter + rpre]);
            ++nenter;
            emit_reach(o, T, n, p);
            emit_strategy(o, T, n);
            std::string cv;
            for (int a = 0; a < nd.nact; ++a) cv += (a ? ", cv" : "cv") + std::to_string(n) + "_" + std::to_string(a);
            o("float %s;", cv.c_str());
            o("float op%d = 0.0f;", n);
            continue;
        }
        if (kind == bt::EV_TERM) {
            emit_reach(o, T, n, p);
            o("float v0_%d, v1_%d;", n, n);
            o("{");
            if (nd.type == bt::DN_FOLD) {
                const int w = nd.player;
                const int off = (nd.sidx - T.t0) * 106;
                const float cw = w == 0 ? nd.c0 : nd.c1, cl = w == 0 ? nd.c1 : nd.c0;
                o("const float vw = %s * mass_u(ssv + %d, c1, c2, x%d_%d);", flit(nd.pot - cw).c_str(),
                  off + (1 - w) * 53, 1 - w, n);
                o("const float vl = %s * mass_u(ssv + %d, c1, c2, x%d_%d);", flit(-cl).c_str(), off + w * 53, w, n);
                o("v0_%d = live ? %s : 0.0f;", n, w == 0 ? "vw" : "vl");
                o("v1_%d = live ? %s : 0.0f;", n, w == 0 ? "vl" : "vw");
            } else if (nd.type == bt::DN_SD) {
                const int off = (nd.sidx - T.t0) * 106;
                const int lc = nd.col - T.c0;
                o("const float m0 = mass_u(ssv + %d, c1, c2, x1_%d), m1 = mass_u(ssv + %d, c1, c2, x0_%d);",
                  off + 53, n, off, n);
                o("const float q0 = Cs[%d * NTB], q1 = Cs[%d * NTB];", lc, lc + 1);
                if (J.level == 1) {
                    const std::string hp = flit(0.5f * nd.pot);
                    o("v0_%d = live ? %s * q0 + hp0 * m0 : 0.0f;", n, hp.c_str());
                    o("v1_%d = live ? %s * q1 + hp0 * m1 : 0.0f;", n, hp.c_str());
                } else {
                    o("v0_%d = live ? %s * q0 - %s * m0 : 0.0f;", n, flit(nd.pot).c_str(), flit(nd.c0).c_str());
                    o("v1_%d = live ? %s * q1 - %s * m1 : 0.0f;", n, flit(nd.pot).c_str(), flit(nd.c1).c_str());
                }
            } else {  // leaf (turn level): mean of the river-line roots
                o("const float* lv = D.leafval + ((size_t)b * D.nleaf + %d) * 2 * NPT + k;", nd.leaf);
                o("v0_%d = live ? lv[0] : 0.0f;", n);
                o("v1_%d = live ? lv[NPT] : 0.0f;", n);
            }
            o("}");
            emit_up(o, T, n, p);
            continue;
        }
        // EV_LEAVE of decision n
        const int pl = nd.player, ls = nd.slot - T.s0;
        o("float v0_%d, v1_%d;", n, n);
        o("{");
        o("float nv = 0.0f;");
        for (int a = 0; a < nd.nact; ++a) o("nv = fmaf(s%d_%d, cv%d_%d, nv);", n, a, n, a);
        // a regret whose update leaves its bits unchanged is not stored again
        o("if (live) {");
        for (int a = 0; a < nd.nact; ++a) {
            o("{ const float q = R%d + (cv%d_%d - nv), nr = q * (q > 0.0f ? wp : wn);", ls + a, n, a);
            o("if (__float_as_uint(nr) != __float_as_uint(R%d)) Rg[%d] = nr; }", ls + a, (ls + a) * J.np);
        }
        o("}");
        o("v%d_%d = nv;", pl, n);
        o("v%d_%d = op%d;", 1 - pl, n, n);
        o("}");
        emit_up(o, T, n, p);
    }
    if (passes > 1) {
        o("bwd_root_out<LEVEL>(D, L, b, tm, k, NP, live, out0, out1);");
        o("}");
    }
    o("break;");
    o("}");
}

std::string waves_attr(int w) {
    return w > 0 ? " __attribute__((amdgpu_waves_per_eu(" + std::to_string(w) + ")))" : std::string();
}

void gen_level(Src& o, const JitLevel& J, const JitConfig& cfg) {
    const int L = J.level;
    // ---- forward
    o("extern \"C\" __global__ __launch_bounds__(%d)%s void mtt_jit_fwd%d(mtt::bt::BatchedDev D, "
      "const float* __restrict__ ranges, float ws) {", fwd_threads(J.level, cfg), waves_attr(cfg.fwd_waves).c_str(), L);
    o("using namespace mtt::bt;");
    o("constexpr int LEVEL = %d, NT = %d, NP = %d;", L, fwd_threads(J.level, cfg), J.np);
    o("const LevelDev& L = D.L[LEVEL];");
    o("extern __shared__ unsigned long long bins[];");
    const int fsl = fwd_sliced(J.level, cfg) ? (J.np + cfg.fwd_nt - 1) / cfg.fwd_nt : 1;
    if (fsl > 1) {
        o("const int inst = blockIdx.x;");
        o("const int b = inst / %d, tm = inst %% %d;", J.ntmpl, J.ntmpl);
        o("const int nl = ldu(L.nlive + b);");
    } else {
        o("const int inst = blockIdx.x >> 1, half = blockIdx.x & 1;");
        o("const int b = inst / %d, tm = inst %% %d;", J.ntmpl, J.ntmpl);
        o("const int k = half * NT + (int)threadIdx.x;");
        o("const bool inb = k < NP;");
        o("const bool live = k < ldu(L.nlive + b);");
        // branch-free prologue: clamped combo index, dead lanes zeroed after the loads (each case requests the
        // cards and the root index, then its regrets, then the root reaches)
        o("const int kc = min(k, NP - 1);");
        o("(void)inb;");
    }
    o("switch (tm) {");
    for (int t = 0; t < J.ntmpl; ++t) gen_forward_case(o, J, cfg.sd_fp32 != 0, cfg.fwd_bins, fsl, J.level == 1 ? cfg.fwd_smask : 0,
                                                    J.level == 1 && cfg.fwd_xpre != 0, t);
    o("default: break;");
    o("}");
    o("}");
    // ---- backward
    o("extern \"C\" __global__ __launch_bounds__(%d)%s void mtt_jit_bwd%d(mtt::bt::BatchedDev D, "
      "const float* __restrict__ ranges, float wp, float wn) {", cfg.bwd_nt, waves_attr(cfg.bwd_waves).c_str(), L);
    o("using namespace mtt::bt;");
    o("constexpr int LEVEL = %d, NTB = %d, NP = %d;", L, cfg.bwd_nt, J.np);
    o("const LevelDev& L = D.L[LEVEL];");
    o("extern __shared__ unsigned long long ssum[];");
    const int passes = bwd_passes(J.level, cfg);
    if (passes > 1) {
        o("const int inst = blockIdx.x;");
        o("const int b = inst / %d, tm = inst %% %d;", J.ntmpl, J.ntmpl);
        o("const int tid = threadIdx.x;");
        o("const int nl = ldu(L.nlive + b);");
        o("float* Cs = reinterpret_cast<float*>(ssum + (size_t)L.maxterm * 106) + tid;");
        o("const float hp0 = 0.5f * D.pot0;");
        o("(void)hp0;");
        o("switch (tm) {");
        for (int t = 0; t < J.ntmpl; ++t) gen_backward_case(o, J, cfg, passes, t);
        o("default: break;");
        o("}");
        o("}");
        return;
    }
    o("constexpr int NH = (NPT + NTB - 1) / NTB;");
    o("const int inst = blockIdx.x / NH, half = blockIdx.x %% NH;");
    o("const int b = inst / %d, tm = inst %% %d;", J.ntmpl, J.ntmpl);
    o("const int tid = threadIdx.x;");
    o("const int k = half * NTB + tid;");
    o("const bool live = k < ldu(L.nlive + b);");
    o("float* Cs = reinterpret_cast<float*>(ssum + (size_t)L.maxterm * 106) + tid;");
    o("int c1 = 0, c2 = 0;");
    o("float root0 = 0.0f, root1 = 0.0f;");
    o("if (live) {");
    o("c1 = L.cc[((size_t)b * NP + k) * 2];");
    o("c2 = L.cc[((size_t)b * NP + k) * 2 + 1];");
    o("root0 = root_reach(D, ranges, LEVEL, b, tm, NP, k, 0);");
    o("root1 = root_reach(D, ranges, LEVEL, b, tm, NP, k, 1);");
    o("}");
    o("const float hp0 = 0.5f * D.pot0;");
    o("(void)hp0;");
    o("float out0 = 0.0f, out1 = 0.0f;");
    o("switch (tm) {");
    for (int t = 0; t < J.ntmpl; ++t) gen_backward_case(o, J, cfg, passes, t);
    o("default: break;");
    o("}");
    o("bwd_root_out<LEVEL>(D, L, b, tm, k, NP, live, out0, out1);");
    o("}");
}

struct CacheEntry {
    hipModule_t mod = nullptr;
    hipFunction_t fwd[2] = {nullptr, nullptr}, bwd[2] = {nullptr, nullptr};
};

std::mutex g_mu;
std::map<std::pair<int, std::string>, CacheEntry> g_cache;

}  // namespace

bool bt_jit_enabled() {
    const char* e = getenv("MTT_JIT");
    return !(e && e[0] == '0' && e[1] == 0);
}

JitConfig bt_jit_config() {
    JitConfig c;
    auto env = [](const char* k, int lo, int hi, int def) {
        const char* e = getenv(k);
        if (!e || !*e) return def;
        char* end = nullptr;
        const long v = strtol(e, &end, 10);
        return (end && *end == 0 && v >= lo && v <= hi) ? (int)v : def;
    };
    c.bwd_nt = env("MTT_JIT_BWD_NT", 192, 576, c.bwd_nt);
    if (c.bwd_nt % 64) c.bwd_nt = JitConfig().bwd_nt;
    c.bwd_waves = env("MTT_JIT_BWD_WAVES", 0, 8, c.bwd_waves);
    c.fwd_waves = env("MTT_JIT_FWD_WAVES", 0, 8, c.fwd_waves);
    c.bwd_passes = env("MTT_JIT_BWD_PASSES", 1, 2, c.bwd_passes);
    c.fwd_bins = env("MTT_JIT_FWD_BINS", 1, 2, c.fwd_bins);
    c.fwd_nt = env("MTT_JIT_FWD_NT", 192, 576, c.fwd_nt);
    c.fwd_smask = env("MTT_JIT_FWD_SMASK", 0, 64, c.fwd_smask);
    c.bwd_rpre = env("MTT_JIT_BWD_RPRE", 0, 64, c.bwd_rpre);
    c.bwd_xpre = env("MTT_JIT_BWD_XPRE", 0, 1, c.bwd_xpre);
    c.fwd_xpre = env("MTT_JIT_FWD_XPRE", 0, 1, c.fwd_xpre);
    if (c.fwd_nt % 64) c.fwd_nt = JitConfig().fwd_nt;
    return c;
}

std::string bt_jit_source(const JitLevel* levels, int n, const JitConfig& cfg) {
    Src o;
    o("// generated per-template sweeps");
    o("%s", "");
    std::string s = kBtDevSrc;
    const size_t po = s.find("#pragma once");  // the header is the main file here
    if (po != std::string::npos) s.erase(po, 12);
    for (int i = 0; i < n; ++i) gen_level(o, levels[i], cfg);
    return s + "\n" + o.s;
}

namespace {

int check_levels(const JitLevel* levels, int n) {
    for (int i = 0; i < n; ++i) {
        const JitLevel& J = levels[i];
        MTT_ARG(J.level == 0 || J.level == 1, "generated sweeps: level %d not supported", J.level);
        MTT_ARG(J.ntmpl >= 1 && (int)J.slot0.size() == J.ntmpl + 1 && (int)J.term0.size() == J.ntmpl + 1 &&
                    (int)J.col0.size() == J.ntmpl + 1,
                "generated sweeps: level %d tables inconsistent", J.level);
    }
    return MTT_OK;
}

}  // namespace

int bt_jit_compile(const std::string& src, std::vector<char>* code) {
    hiprtcProgram prog = nullptr;
    hiprtcResult r = hiprtcCreateProgram(&prog, src.c_str(), "mtt_sweeps.hip", 0, nullptr, nullptr);
    if (r != HIPRTC_SUCCESS) {
        set_error("hiprtcCreateProgram: %s", hiprtcGetErrorString(r));
        return MTT_ERR_HIP;
    }
    // the library's device code is built for gfx950 with packed-FP32 VALU forms disabled; so is this
    const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-DMTT_JIT=1",
                          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    r = hiprtcCompileProgram(prog, (int)(sizeof opts / sizeof opts[0]), opts);
    if (r != HIPRTC_SUCCESS) {
        size_t ls = 0;
        hiprtcGetProgramLogSize(prog, &ls);
        std::string log(ls + 1, '\0');
        if (ls) hiprtcGetProgramLog(prog, &log[0]);
        set_error("hipRTC compile of the generated sweeps failed (%s; MTT_JIT=0 selects the interpreted sweeps): "
                  "%.900s", hiprtcGetErrorString(r), log.c_str());
        hiprtcDestroyProgram(&prog);
        return MTT_ERR_HIP;
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    code->resize(cs);
    hiprtcGetCode(prog, code->data());
    hiprtcDestroyProgram(&prog);
    return MTT_OK;
}

int bt_jit_build(const JitLevel* levels, int n, const JitConfig& cfg, JitSweeps* out) {
    *out = JitSweeps();
    int rc = check_levels(levels, n);
    if (rc) return rc;
    MTT_ARG(cfg.bwd_nt >= 192 && cfg.bwd_nt <= 576 && cfg.bwd_nt % 64 == 0,
            "generated sweeps: backward workgroup %d not a multiple of 64 in 192..576", cfg.bwd_nt);
    MTT_ARG(cfg.bwd_passes == 1 || cfg.bwd_passes == 2, "generated sweeps: backward passes %d not 1 or 2", cfg.bwd_passes);
    MTT_ARG(cfg.fwd_bins == 1 || cfg.fwd_bins == 2, "generated sweeps: forward bin copies %d not 1 or 2", cfg.fwd_bins);
    MTT_ARG(cfg.fwd_nt >= 192 && cfg.fwd_nt <= 576 && cfg.fwd_nt % 64 == 0,
            "generated sweeps: forward workgroup %d not a multiple of 64 in 192..576", cfg.fwd_nt);
    const std::string src = bt_jit_source(levels, n, cfg);
    int dev = 0;
    MTT_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_mu);
    auto key = std::make_pair(dev, src);
    auto it = g_cache.find(key);
    if (it == g_cache.end()) {
        const auto t0 = std::chrono::steady_clock::now();
        std::vector<char> code;
        rc = bt_jit_compile(src, &code);
        if (rc) return rc;
        CacheEntry ce;
        MTT_HIP(hipModuleLoadData(&ce.mod, code.data()));
        for (int i = 0; i < n; ++i) {
            const int L = levels[i].level;
            char name[32];
            snprintf(name, sizeof name, "mtt_jit_fwd%d", L);
            MTT_HIP(hipModuleGetFunction(&ce.fwd[L], ce.mod, name));
            snprintf(name, sizeof name, "mtt_jit_bwd%d", L);
            MTT_HIP(hipModuleGetFunction(&ce.bwd[L], ce.mod, name));
        }
        it = g_cache.emplace(key, ce).first;
        out->compile_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    for (int L = 0; L < 2; ++L) {
        out->fwd[L] = it->second.fwd[L];
        out->bwd[L] = it->second.bwd[L];
    }
    out->cfg = cfg;
    out->active = true;
    return MTT_OK;
}

}  // namespace mtt
