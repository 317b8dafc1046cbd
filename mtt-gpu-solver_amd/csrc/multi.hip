// 4-way single-street engine (multiway_4way_engine.cu), the fused DCFR step
// (preflop_grid_generator.cu:387-410) and Malmuth-Harville ICM in FP64.
//
// Layouts keep the reference's: per-action state [infoset][3][1216],
// per-node state [node][4][1216]; batched solves add a leading flop index.
// The multiway showdown keeps the reference's structure (reaches in LDS,
// strength-sorted exclusive prefix scans) and adds what it lacks: each
// opponent's reach is taken without the combos that share a card with the
// hero (per-card prefix lists: total - card1 - card2 + own), and equal
// strengths split the pot.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "solver.h"

namespace mtt {

namespace {

constexpr int P4 = 4;
constexpr int A4 = MTT_4WAY_MAX_ACTIONS;
constexpr int NP = kPadFlop;     // 1216
constexpr int NA = kAliveFlop;   // 1176
constexpr int TT = 256;          // terminal kernel threads
constexpr int HPT = (NP + TT - 1) / TT;  // heroes per thread (5)

// ------------------------------------------------------------ forward ----
// Regret matching and reach propagation for the decision nodes of one
// depth; blockIdx.z = flop (state strides per flop).
__global__ __launch_bounds__(256) void k4_forward(const int* __restrict__ idx, int n, const mtt_flat_node4* __restrict__ nodes,
                                                  const float* __restrict__ R, float* __restrict__ sigma,
                                                  float* __restrict__ reach, size_t info_stride, size_t node_stride) {
    const int j = blockIdx.y;
    if (j >= n) return;
    const int id = idx[j];
    const mtt_flat_node4 nd = nodes[id];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (nd.type != MTT_NODE_DECISION || i >= NP) return;
    const size_t fi = (size_t)blockIdx.z * info_stride, fn = (size_t)blockIdx.z * node_stride;
    const float* Rk = R + fi + (size_t)nd.infoset_offset * A4 * NP + i;
    float* Sg = sigma + fi + (size_t)nd.infoset_offset * A4 * NP + i;
    float pos[A4], sum = 0.0f;
#pragma unroll
    for (int a = 0; a < A4; ++a) {
        pos[a] = (a < nd.num_actions && i < NA) ? fmaxf(Rk[(size_t)a * NP], 0.0f) : 0.0f;
        sum += pos[a];
    }
    float sg[A4];
#pragma unroll
    for (int a = 0; a < A4; ++a) {
        sg[a] = (a < nd.num_actions && i < NA) ? (sum > 1e-7f ? pos[a] / sum : 1.0f / (float)nd.num_actions) : 0.0f;
        Sg[(size_t)a * NP] = sg[a];
    }
    float pr[P4];
    const float* rp = reach + fn + (size_t)id * P4 * NP + i;
#pragma unroll
    for (int p = 0; p < P4; ++p) pr[p] = rp[(size_t)p * NP];
    for (int a = 0; a < nd.num_actions; ++a) {
        float* rc = reach + fn + (size_t)(nd.children_offset + a) * P4 * NP + i;
#pragma unroll
        for (int p = 0; p < P4; ++p) rc[(size_t)p * NP] = p == nd.player ? pr[p] * sg[a] : pr[p];
    }
}

// ----------------------------------------------------------- backward ----
// Node values and the DCFR-generalised regret / strategy-sum update for
// the decision nodes of one depth (CFR+ is w_pos = 1, w_neg = 0).
__global__ __launch_bounds__(256) void k4_backward(const int* __restrict__ idx, int n, const mtt_flat_node4* __restrict__ nodes,
                                                   const float* __restrict__ sigma, const float* __restrict__ reach,
                                                   float* __restrict__ cfv, float* __restrict__ R, float* __restrict__ S,
                                                   float wp, float wn, float ws, size_t info_stride, size_t node_stride) {
    const int j = blockIdx.y;
    if (j >= n) return;
    const int id = idx[j];
    const mtt_flat_node4 nd = nodes[id];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (nd.type != MTT_NODE_DECISION || i >= NP) return;
    const size_t fi = (size_t)blockIdx.z * info_stride, fn = (size_t)blockIdx.z * node_stride;
    const int ap = nd.player;
    const size_t ib = fi + (size_t)nd.infoset_offset * A4 * NP + i;
    float av[A4], node_v = 0.0f;
#pragma unroll
    for (int a = 0; a < A4; ++a) {
        av[a] = 0.0f;
        if (a < nd.num_actions) {
            av[a] = cfv[fn + ((size_t)(nd.children_offset + a) * P4 + ap) * NP + i];
            node_v += sigma[ib + (size_t)a * NP] * av[a];
        }
    }
    cfv[fn + ((size_t)id * P4 + ap) * NP + i] = node_v;
#pragma unroll
    for (int p = 0; p < P4; ++p) {
        if (p == ap) continue;
        float so = 0.0f;
        for (int a = 0; a < nd.num_actions; ++a) so += cfv[fn + ((size_t)(nd.children_offset + a) * P4 + p) * NP + i];
        cfv[fn + ((size_t)id * P4 + p) * NP + i] = so;
    }
    if (i < NA) {
        const float rp = reach[fn + ((size_t)id * P4 + ap) * NP + i];
        for (int a = 0; a < nd.num_actions; ++a) {
            const size_t k = ib + (size_t)a * NP;
            const float q = R[k] + (av[a] - node_v);
            R[k] = q * (q > 0.0f ? wp : wn);
            S[k] += ws * rp * sigma[k];
        }
    }
}

// ---------------------------------------------------------- terminals ----
// One workgroup per (terminal, flop).  Reaches of the four players are
// staged in LDS; for each opponent o the exclusive prefix over strength
// positions (block scan) and over each card's list (one thread per card)
// give, per hero, o's reach strictly weaker (F), weaker or tied (G) and in
// total (W), all without combos that share a hero card.  The prefixes are
// exact integers (reach in 2^48 fixed point), so the outcome masses F,
// G - F and W - G are exact before their one rounding to FP32; the payoff
// tables and their contraction are FP32.
__device__ __forceinline__ long long wave_incl_scan(long long v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const long long t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// r * 2^48 truncated to an integer (0 <= r < 2^15), branch-free: the
// integer and fraction parts of r * 2^16 (both exact in FP32) give the high
// and low words.
__device__ __forceinline__ long long fx48_of(float r) {
    const float t = r * 65536.0f;
    const uint32_t hi = (uint32_t)t;
    const uint32_t lo = (uint32_t)((t - (float)hi) * 4294967296.0f);
    return (long long)(((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ float fx48_to_float(long long v) { return ldexpf((float)v, -48); }

__device__ __forceinline__ float pay_net(float net, float bf) { return net >= 0.0f ? net : bf * net; }

// Showdown payoff of player p (still in) when its active opponents, in seat
// order, are weaker (digit 0), tied (1) or stronger (2) than p: outcome
// s = d0 + 3 d1 + 9 d2.  The pot is split into layers at the distinct
// commitments of the players still in (all seats' chips up to each level);
// p takes each layer it is eligible for unless an eligible opponent is
// stronger, shared with the eligible tied ones.  Net chips, scaled by the
// bubble factor when negative.
__device__ float k4_showdown_pay(const mtt_flat_node4& nd, int p, int s, float bpp) {
    int opp[3] = {0, 0, 0}, na = 0;
    for (int o = 0; o < P4; ++o)
        if (o != p && (nd.active_mask & (1 << o))) opp[na++] = o;
    int dig[3];
    for (int t = 0; t < 3; ++t) { dig[t] = s % 3; s /= 3; }
    // distinct commitments of the players still in, ascending
    float lv[P4];
    int k = 0;
    for (int q = 0; q < P4; ++q) {
        if (!(nd.active_mask & (1 << q))) continue;
        const float c = nd.committed[q];
        int at = k;
        bool dup = false;
        for (int u = 0; u < k; ++u) {
            if (fabsf(lv[u] - c) <= 1e-6f) dup = true;
            else if (lv[u] > c && at == k) at = u;
        }
        if (dup) continue;
        for (int u = k; u > at; --u) lv[u] = lv[u - 1];
        lv[at] = c;
        k++;
    }
    float won = 0.0f, prev = 0.0f;
    for (int l = 0; l < k; ++l) {
        const float top = lv[l];
        float chips = 0.0f;
        for (int q = 0; q < P4; ++q) chips += fminf(fmaxf(nd.committed[q] - prev, 0.0f), top - prev);
        prev = top;
        if (nd.committed[p] < top - 1e-6f) continue;      // p is not in this layer
        bool lost = false;
        int tied = 0;
        for (int t = 0; t < na; ++t) {
            if (nd.committed[opp[t]] < top - 1e-6f) continue;
            lost |= dig[t] == 2;
            tied += dig[t] == 1;
        }
        if (!lost) won += chips / (float)(tied + 1);
    }
    return pay_net(won - nd.committed[p], bpp);
}

__global__ __launch_bounds__(TT) void k4_terminals(const int* __restrict__ terms, int nterm,
                                                   const mtt_flat_node4* __restrict__ nodes,
                                                   const float* __restrict__ reach, const mtt_board4* __restrict__ boards,
                                                   const float* __restrict__ bubble, float* __restrict__ cfv,
                                                   size_t node_stride) {
    __shared__ long long rx[P4][NP];   // reach in 2^48 fixed point
    __shared__ long long P[NA + 1];
    __shared__ long long Q[52][49];
    __shared__ long long wsum[TT / 64];
    __shared__ uint16_t pc[NP];
    __shared__ uint16_t cl[52][48];
    __shared__ uint8_t ncard[52];
    __shared__ float pay[P4][27];
    const int t = blockIdx.x;
    if (t >= nterm) return;
    const int f = blockIdx.y;
    const int id = terms[t];
    const mtt_flat_node4 nd = nodes[id];
    const mtt_board4& B = boards[f];
    const size_t fn = (size_t)f * node_stride;
    const int tid = threadIdx.x;
    for (int e = tid; e < P4 * NP; e += TT) {
        const int p = e / NP, i = e % NP;
        rx[p][i] = i < NA ? fx48_of(reach[fn + ((size_t)id * P4 + p) * NP + i]) : 0;
    }
    for (int e = tid; e < NP; e += TT) pc[e] = B.pos_combo[e];
    for (int e = tid; e < 52 * 48; e += TT) cl[e / 48][e % 48] = B.card_list[e / 48][e % 48];
    if (tid < 52) ncard[tid] = B.n_card[tid];
    if (nd.type == MTT_NODE_SHOWDOWN && tid < P4 * 27) {
        const int p = tid / 27;
        pay[p][tid % 27] = (nd.active_mask & (1 << p)) ? k4_showdown_pay(nd, p, tid % 27, bubble[p * P4 + p]) : 0.0f;
    }
    // hero data, packed: cards | lt counts, tie group, le counts
    uint32_t hcl[HPT], hgr[HPT], hle[HPT];
#pragma unroll
    for (int h = 0; h < HPT; ++h) {
        const int i = tid + TT * h;
        const bool ok = i < NA;
        hcl[h] = ok ? (uint32_t)B.c1[i] | ((uint32_t)B.c2[i] << 8) | ((uint32_t)B.card_lt[i][0] << 16) |
                          ((uint32_t)B.card_lt[i][1] << 24)
                    : 0u;
        hgr[h] = ok ? (uint32_t)B.grp_lo[i] | ((uint32_t)B.grp_hi[i] << 16) : 0u;
        hle[h] = ok ? (uint32_t)B.card_le[i][0] | ((uint32_t)B.card_le[i][1] << 8) : 0u;
    }
    // per opponent and hero, FP32: weaker (F), tied (G - F) and stronger (W - G) reach for an opponent
    // still in a showdown; otherwise the total W in the first slot (all a fold or an inactive seat needs)
    float MF[P4][HPT], MT[P4][HPT], MS[P4][HPT];
    __syncthreads();
    const bool fold = nd.type == MTT_NODE_FOLD;
#pragma unroll
    for (int o = 0; o < P4; ++o) {   // unrolled: the per-opponent arrays stay in registers
        const bool need_fg = !fold && (nd.active_mask & (1 << o));
        // exclusive prefix over positions: thread owns positions 5*tid..5*tid+4
        long long loc[HPT], s = 0;
#pragma unroll
        for (int h = 0; h < HPT; ++h) {
            const int q = tid * HPT + h;
            loc[h] = q < NA ? rx[o][pc[q]] : 0;
            s += loc[h];
        }
        const long long inc = wave_incl_scan(s);
        if ((tid & 63) == 63) wsum[tid >> 6] = inc;
        // per-card lists
        if (tid < 52) {
            long long acc = 0;
            Q[tid][0] = 0;
            for (int k = 0; k < ncard[tid]; ++k) {
                acc += rx[o][cl[tid][k]];
                Q[tid][k + 1] = acc;
            }
        }
        __syncthreads();
        long long base = inc - s;
        for (int w = 0; w < (tid >> 6); ++w) base += wsum[w];
#pragma unroll
        for (int h = 0; h < HPT; ++h) {
            const int q = tid * HPT + h;
            if (q <= NA) P[q] = base;
            base += loc[h];
        }
        __syncthreads();
        const long long tot = P[NA];
#pragma unroll
        for (int h = 0; h < HPT; ++h) {
            const int i = tid + TT * h;
            if (i >= NA) { MF[o][h] = MT[o][h] = MS[o][h] = 0.0f; continue; }
            const long long ri = rx[o][i];
            const int a = hcl[h] & 0xFF, b = (hcl[h] >> 8) & 0xFF;
            const long long Wv = tot - Q[a][ncard[a]] - Q[b][ncard[b]] + ri;
            if (need_fg) {
                const long long Fv = P[hgr[h] & 0xFFFF] - Q[a][(hcl[h] >> 16) & 0xFF] - Q[b][hcl[h] >> 24];
                const long long Gv = P[hgr[h] >> 16] - Q[a][hle[h] & 0xFF] - Q[b][(hle[h] >> 8) & 0xFF] + ri;
                MF[o][h] = fx48_to_float(Fv);
                MT[o][h] = fx48_to_float(Gv - Fv);
                MS[o][h] = fx48_to_float(Wv - Gv);
            } else {
                MF[o][h] = fx48_to_float(Wv);
                MT[o][h] = MS[o][h] = 0.0f;
            }
        }
        __syncthreads();
    }
    int winner = 0;
    if (fold)
        for (int p = P4 - 1; p >= 0; --p)
            if (nd.active_mask & (1 << p)) winner = p;
#pragma unroll
    for (int p = 0; p < P4; ++p) {
        const float cp = nd.committed[p];
        const float bpp = bubble[p * P4 + p];
#pragma unroll
        for (int h = 0; h < HPT; ++h) {
            const int i = tid + TT * h;
            if (i >= NP) continue;
            float v = 0.0f;
            if (i < NA) {
                if (fold || !(nd.active_mask & (1 << p))) {
                    float opp = 1.0f;
#pragma unroll
                    for (int o = 0; o < P4; ++o) {
                        if (o == p) continue;
                        // a fold, or a seat not in the showdown: W of every other seat (an opponent still in
                        // the showdown holds its three outcome masses, which sum to W)
                        const bool parts = !fold && (nd.active_mask & (1 << o));
                        opp *= parts ? (MF[o][h] + MT[o][h]) + MS[o][h] : MF[o][h];
                    }
                    const float pw = fold ? (p == winner ? nd.pot_size - cp : -bubble[p * P4 + winner] * cp) : -bpp * cp;
                    v = pw * opp;
                } else {
                    // outcome masses per active opponent (weaker, tied, stronger),
                    // contracted with p's payoff table; absent digits are "weaker" with mass 1
                    float inact = 1.0f, m[3][3];
                    int na = 0;
#pragma unroll
                    for (int o = 0; o < P4; ++o) {
                        if (o == p) continue;
                        if (nd.active_mask & (1 << o)) {
                            m[na][0] = MF[o][h];
                            m[na][1] = MT[o][h];
                            m[na][2] = MS[o][h];
                            na++;
                        } else {
                            inact *= MF[o][h];                 // an inactive seat's slot holds W
                        }
                    }
                    for (int u = na; u < 3; ++u) { m[u][0] = 1.0f; m[u][1] = 0.0f; m[u][2] = 0.0f; }
                    float acc = 0.0f;
#pragma unroll
                    for (int d2 = 0; d2 < 3; ++d2) {
                        float a1 = 0.0f;
#pragma unroll
                        for (int d1 = 0; d1 < 3; ++d1) {
                            const float* row = &pay[p][3 * d1 + 9 * d2];
                            a1 += m[1][d1] * (m[0][0] * row[0] + m[0][1] * row[1] + m[0][2] * row[2]);
                        }
                        acc += m[2][d2] * a1;
                    }
                    v = acc * inact;
                }
            }
            cfv[fn + ((size_t)id * P4 + p) * NP + i] = v;
        }
    }
}

// ---------------------------------------------------------- DCFR step ----
__global__ void k_dcfr_step(int na, float* __restrict__ R, float* __restrict__ sg, float* __restrict__ S,
                            const float* __restrict__ cfv, const float* __restrict__ reach, float wp, float ws, float wn,
                            int n, int stride) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float node_v = 0.0f;
    for (int a = 0; a < na; ++a) node_v += sg[(size_t)a * stride + i] * cfv[(size_t)a * stride + i];
    float nr[8], pos_sum = 0.0f;
    for (int a = 0; a < na; ++a) {
        const size_t k = (size_t)a * stride + i;
        const float q = R[k] + (cfv[k] - node_v);
        nr[a] = q * (q > 0.0f ? wp : wn);
        R[k] = nr[a];
        S[k] += ws * reach[i] * sg[k];
        pos_sum += fmaxf(nr[a], 0.0f);
    }
    for (int a = 0; a < na; ++a)
        sg[(size_t)a * stride + i] = pos_sum > 1e-7f ? fmaxf(nr[a], 0.0f) / pos_sum : 1.0f / (float)na;
}

// ---------------------------------------------------------------- ICM ----
// One workgroup per stack vector.  Q[m] = probability that the players in
// subset m finish first (in some order): Q[m] = sum_{i in m} Q[m\i] s_i /
// T(m\i), layer by layer (pull form: deterministic).  Then each thread owns
// one player and sums its place probabilities over the subsets in order.
constexpr int ICM_MAXN = 11;  // Q and T tables: 2 x 2^11 doubles of LDS
__global__ __launch_bounds__(256) void k_icm(const double* __restrict__ stacks, int n, const double* __restrict__ pay,
                                             int k, double* __restrict__ eq) {
    __shared__ double Q[1 << ICM_MAXN];
    __shared__ double Tm[1 << ICM_MAXN];
    __shared__ double st[ICM_MAXN];
    const int b = blockIdx.x;
    const int ns = 1 << n;
    if (threadIdx.x < n) st[threadIdx.x] = stacks[(size_t)b * n + threadIdx.x];
    __syncthreads();
    double tot = 0.0;
    for (int i = 0; i < n; ++i) tot += st[i];
    for (int m = threadIdx.x; m < ns; m += blockDim.x) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) if (m & (1 << i)) s += st[i];
        Tm[m] = tot - s;
        Q[m] = m == 0 ? 1.0 : 0.0;
    }
    __syncthreads();
    const int kk = k < n ? k : n;
    for (int L = 1; L < kk; ++L) {
        for (int m = threadIdx.x; m < ns; m += blockDim.x) {
            if (__builtin_popcount((unsigned)m) != L) continue;
            double q = 0.0;
            for (int i = 0; i < n; ++i) {
                if (!(m & (1 << i)) || st[i] <= 0.0) continue;
                const int pm = m & ~(1 << i);
                if (Tm[pm] > 0.0) q += Q[pm] * st[i] / Tm[pm];
            }
            Q[m] = q;
        }
        __syncthreads();
    }
    if (threadIdx.x < n) {
        const int i = threadIdx.x;
        double e = 0.0;
        for (int m = 0; m < ns; ++m) {
            const int placed = __builtin_popcount((unsigned)m);
            if (placed >= kk || (m & (1 << i)) || Q[m] == 0.0) continue;
            if (Tm[m] > 0.0) {
                if (st[i] > 0.0) e += Q[m] * st[i] / Tm[m] * pay[placed];
            } else {  // only zero stacks left: equal shares of the remaining places
                double rest = 0.0;
                for (int q = placed; q < kk; ++q) rest += pay[q];
                e += Q[m] * rest / (double)(n - placed);
            }
        }
        eq[(size_t)b * n + i] = e;
    }
}

}  // namespace

// ------------------------------------------------------------ boards ----
// Host build of the flop tables (the same definitions as the CPU oracle's
// or_4way_board): combos in reference order, strength = score of the five
// cards, stable ascending order, tie groups, per-card lists.
void build_board4(const uint8_t flop[3], mtt_board4& B) {
    memset(&B, 0, sizeof(B));
    uint8_t deck[49];
    int nd = 0;
    for (int c = 0; c < 52; ++c)
        if (c != flop[0] && c != flop[1] && c != flop[2]) deck[nd++] = (uint8_t)c;
    std::vector<uint32_t> score(NA);
    const uint64_t fm = card_bits16(flop[0]) | card_bits16(flop[1]) | card_bits16(flop[2]);
    int k = 0;
    for (int i = 0; i < 49; ++i)
        for (int j = i + 1; j < 49; ++j) {
            B.c1[k] = deck[i];
            B.c2[k] = deck[j];
            score[k] = score_packed(fm | card_bits16(deck[i]) | card_bits16(deck[j]));
            k++;
        }
    std::vector<int> order(NA);
    for (int i = 0; i < NA; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return score[x] < score[y]; });
    for (int q = 0; q < NA; ++q) {
        B.pos_combo[q] = (uint16_t)order[q];
        B.rank_pos[order[q]] = (uint16_t)q;
    }
    for (int q = NA; q < NP; ++q) { B.pos_combo[q] = (uint16_t)q; B.rank_pos[q] = (uint16_t)q; }
    for (int q = 0; q < NA;) {
        int e = q;
        while (e < NA && score[order[e]] == score[order[q]]) e++;
        for (int t = q; t < e; ++t) { B.grp_lo[order[t]] = (uint16_t)q; B.grp_hi[order[t]] = (uint16_t)e; }
        q = e;
    }
    for (int q = 0; q < NA; ++q) {
        const int i = order[q];
        B.card_list[B.c1[i]][B.n_card[B.c1[i]]++] = (uint16_t)i;
        B.card_list[B.c2[i]][B.n_card[B.c2[i]]++] = (uint16_t)i;
    }
    for (int i = 0; i < NA; ++i) {
        const int cs[2] = {B.c1[i], B.c2[i]};
        for (int h = 0; h < 2; ++h) {
            int lt = 0, le = 0;
            for (int t = 0; t < B.n_card[cs[h]]; ++t) {
                const int j = B.card_list[cs[h]][t];
                lt += B.rank_pos[j] < B.grp_lo[i];
                le += B.rank_pos[j] < B.grp_hi[i];
            }
            B.card_lt[i][h] = (uint8_t)lt;
            B.card_le[i][h] = (uint8_t)le;
        }
    }
}

int build_4way_tree(const mtt_4way_game& g, std::vector<mtt_flat_node4>& out, int& n_info);  // multi_tree.cpp

// --------------------------------------------------------- multi solve ----
struct MultiImpl {
    mtt_4way_game game;
    mtt_solver_cfg cfg;
    int nf = 0;
    std::vector<mtt_flat_node4> nodes;
    int n_info = 0;
    std::vector<std::vector<int>> depth_lists;
    std::vector<int> terms;
    hipStream_t stream = nullptr;
    KernelTimer timer;
    DeviceArena arena;
    mtt_flat_node4* d_nodes = nullptr;
    std::vector<int*> d_depth;
    int* d_terms = nullptr;
    mtt_board4* d_boards = nullptr;
    float* d_bf = nullptr;
    float *d_R = nullptr, *d_S = nullptr, *d_sig = nullptr, *d_reach = nullptr, *d_cfv = nullptr;
    size_t info_stride = 0, node_stride = 0;
    int iters_done = 0;
};

int launch_4way_forward(const int* d_idx, int n, const mtt_flat_node4* d_nodes, const float* R, float* sig, float* reach,
                        int nf, size_t is, size_t ns, hipStream_t st) {
    if (n <= 0) return MTT_OK;
    MTT_ARG(n <= 65535 && nf <= 65535, "4-way sweep: %d nodes x %d flops exceeds the grid", n, nf);
    hipLaunchKernelGGL(k4_forward, dim3((NP + 255) / 256, n, nf), dim3(256), 0, st, d_idx, n, d_nodes, R, sig, reach, is, ns);
    MTT_HIP(hipGetLastError());
    return MTT_OK;
}
int launch_4way_backward(const int* d_idx, int n, const mtt_flat_node4* d_nodes, const float* sig, const float* reach,
                         float* cfv, float* R, float* S, float wp, float wn, float ws, int nf, size_t is, size_t ns,
                         hipStream_t st) {
    if (n <= 0) return MTT_OK;
    MTT_ARG(n <= 65535 && nf <= 65535, "4-way sweep: %d nodes x %d flops exceeds the grid", n, nf);
    hipLaunchKernelGGL(k4_backward, dim3((NP + 255) / 256, n, nf), dim3(256), 0, st, d_idx, n, d_nodes, sig, reach, cfv, R, S,
                       wp, wn, ws, is, ns);
    MTT_HIP(hipGetLastError());
    return MTT_OK;
}
int launch_4way_terminals(const int* d_terms, int n, const mtt_flat_node4* d_nodes, const float* reach,
                          const mtt_board4* boards, const float* bf, float* cfv, int nf, size_t ns, hipStream_t st) {
    if (n <= 0) return MTT_OK;
    MTT_ARG(nf <= 65535, "4-way terminals: %d flops exceeds the grid", nf);
    hipLaunchKernelGGL(k4_terminals, dim3(n, nf), dim3(TT), 0, st, d_terms, n, d_nodes, reach, boards, bf, cfv, ns);
    MTT_HIP(hipGetLastError());
    return MTT_OK;
}

}  // namespace mtt

struct mtt_multi {
    mtt::MultiImpl impl;
};

using namespace mtt;

extern "C" {

int mtt_board4_build(const uint8_t flop[3], mtt_board4* out) {
    MTT_ARG(flop && out, "mtt_board4_build: null pointer");
    MTT_ARG(flop[0] < 52 && flop[1] < 52 && flop[2] < 52 && flop[0] != flop[1] && flop[0] != flop[2] &&
                flop[1] != flop[2], "mtt_board4_build: bad flop");
    build_board4(flop, *out);
    return MTT_OK;
}

int mtt_4way_forward_sweep(const int* d_depth_node_indices, int num_nodes_at_depth, const mtt_flat_node4* d_nodes,
                           const float* d_cumulative_regrets, float* d_current_strategy, float* d_reach_probs,
                           hipStream_t stream) {
    MTT_ARG(num_nodes_at_depth >= 0, "mtt_4way_forward_sweep: negative node count");
    if (!num_nodes_at_depth) return MTT_OK;
    MTT_ARG(d_depth_node_indices && d_nodes && d_cumulative_regrets && d_current_strategy && d_reach_probs,
            "mtt_4way_forward_sweep: null pointer");
    return launch_4way_forward(d_depth_node_indices, num_nodes_at_depth, d_nodes, d_cumulative_regrets,
                               d_current_strategy, d_reach_probs, 1, 0, 0, stream);
}

int mtt_4way_evaluate_terminals_board(const int* d_terminal_node_indices, int num_terminals,
                                      const mtt_flat_node4* d_nodes, const float* d_reach_probs,
                                      const mtt_board4* d_board, const float* d_bubble_factors, float* d_cfvs,
                                      hipStream_t stream) {
    MTT_ARG(num_terminals >= 0, "mtt_4way_evaluate_terminals_board: negative count");
    if (!num_terminals) return MTT_OK;
    MTT_ARG(d_terminal_node_indices && d_nodes && d_reach_probs && d_board && d_bubble_factors && d_cfvs,
            "mtt_4way_evaluate_terminals_board: null pointer");
    return launch_4way_terminals(d_terminal_node_indices, num_terminals, d_nodes, d_reach_probs, d_board,
                                 d_bubble_factors, d_cfvs, 1, 0, stream);
}

int mtt_4way_evaluate_terminals(const int* d_terminal_node_indices, int num_terminals, const mtt_flat_node4* d_nodes,
                                const float* d_reach_probs, const uint16_t* d_combo_rank_pos,
                                const float* d_bubble_factors, float* d_cfvs, const uint8_t* flop,
                                hipStream_t stream) {
    MTT_ARG(num_terminals >= 0, "mtt_4way_evaluate_terminals: negative count");
    if (!num_terminals) return MTT_OK;
    MTT_ARG(d_terminal_node_indices && d_nodes && d_reach_probs && d_combo_rank_pos && d_bubble_factors && d_cfvs &&
                flop, "mtt_4way_evaluate_terminals: null pointer");
    std::unique_ptr<mtt_board4> B(new (std::nothrow) mtt_board4);
    MTT_ARG(B, "mtt_4way_evaluate_terminals: out of host memory");
    int rc = mtt_board4_build(flop, B.get());
    if (rc) return rc;
    // The caller's strength order must agree with the flop's: ascending
    // positions, combos of equal strength (one tie group) adjacent.
    uint16_t pos[NP];
    MTT_HIP(hipMemcpyAsync(pos, d_combo_rank_pos, sizeof(pos), hipMemcpyDeviceToHost, stream));
    MTT_HIP(hipStreamSynchronize(stream));
    std::vector<uint8_t> seen(NA, 0);
    for (int i = 0; i < NA; ++i) {
        MTT_ARG(pos[i] < NA && !seen[pos[i]], "mtt_4way_evaluate_terminals: combo_rank_pos is not a permutation of "
                "0..%d (combo %d)", NA - 1, i);
        seen[pos[i]] = 1;
        MTT_ARG(pos[i] >= B->grp_lo[i] && pos[i] < B->grp_hi[i],
                "mtt_4way_evaluate_terminals: combo_rank_pos of combo %d (%d) is not the flop's strength order "
                "(positions %d..%d)", i, (int)pos[i], (int)B->grp_lo[i], (int)B->grp_hi[i] - 1);
    }
    mtt_board4* d_board = nullptr;
    MTT_HIP(hipMalloc(&d_board, sizeof(mtt_board4)));
    hipError_t e = hipMemcpyAsync(d_board, B.get(), sizeof(mtt_board4), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess)
        rc = launch_4way_terminals(d_terminal_node_indices, num_terminals, d_nodes, d_reach_probs, d_board,
                                   d_bubble_factors, d_cfvs, 1, 0, stream);
    hipError_t e2 = hipStreamSynchronize(stream);   // the board buffer is freed below
    (void)hipFree(d_board);
    if (e != hipSuccess) return hip_fail(e, "mtt_4way_evaluate_terminals: board upload", __FILE__, __LINE__);
    if (e2 != hipSuccess) return hip_fail(e2, "mtt_4way_evaluate_terminals", __FILE__, __LINE__);
    return rc;
}

int mtt_4way_backward_cfr_plus(const int* d_depth_node_indices, int num_nodes_at_depth, const mtt_flat_node4* d_nodes,
                               const float* d_current_strategy, const float* d_reach_probs, float* d_cfvs,
                               float* d_cumulative_regrets, float* d_cumulative_strategy_sums, float iter_weight,
                               hipStream_t stream) {
    MTT_ARG(num_nodes_at_depth >= 0, "mtt_4way_backward_cfr_plus: negative node count");
    if (!num_nodes_at_depth) return MTT_OK;
    MTT_ARG(d_depth_node_indices && d_nodes && d_current_strategy && d_reach_probs && d_cfvs && d_cumulative_regrets &&
                d_cumulative_strategy_sums, "mtt_4way_backward_cfr_plus: null pointer");
    return launch_4way_backward(d_depth_node_indices, num_nodes_at_depth, d_nodes, d_current_strategy, d_reach_probs,
                                d_cfvs, d_cumulative_regrets, d_cumulative_strategy_sums, 1.0f, 0.0f, iter_weight, 1, 0,
                                0, stream);
}

int mtt_4way_backward_dcfr(const int* d_depth_node_indices, int num_nodes_at_depth, const mtt_flat_node4* d_nodes,
                           const float* d_current_strategy, const float* d_reach_probs, float* d_cfvs,
                           float* d_cumulative_regrets, float* d_cumulative_strategy_sums, float w_pos, float w_neg,
                           float w_strat, hipStream_t stream) {
    MTT_ARG(num_nodes_at_depth >= 0, "mtt_4way_backward_dcfr: negative node count");
    if (!num_nodes_at_depth) return MTT_OK;
    MTT_ARG(d_depth_node_indices && d_nodes && d_current_strategy && d_reach_probs && d_cfvs && d_cumulative_regrets &&
                d_cumulative_strategy_sums, "mtt_4way_backward_dcfr: null pointer");
    return launch_4way_backward(d_depth_node_indices, num_nodes_at_depth, d_nodes, d_current_strategy, d_reach_probs,
                                d_cfvs, d_cumulative_regrets, d_cumulative_strategy_sums, w_pos, w_neg, w_strat, 1, 0, 0,
                                stream);
}

int mtt_dcfr_step(int num_actions, float* d_regrets, float* d_current_strat, float* d_strat_sums,
                  const float* d_action_cfvs, const float* d_reach, float w_pos, float w_strat, float w_neg, int n,
                  int stride, hipStream_t stream) {
    MTT_ARG(num_actions >= 1 && num_actions <= 8, "mtt_dcfr_step: num_actions %d out of 1..8", num_actions);
    MTT_ARG(n >= 0 && stride >= n, "mtt_dcfr_step: need 0 <= n <= stride");
    if (!n) return MTT_OK;
    MTT_ARG(d_regrets && d_current_strat && d_strat_sums && d_action_cfvs && d_reach, "mtt_dcfr_step: null pointer");
    hipLaunchKernelGGL(k_dcfr_step, dim3((n + 255) / 256), dim3(256), 0, stream, num_actions, d_regrets,
                       d_current_strat, d_strat_sums, d_action_cfvs, d_reach, w_pos, w_strat, w_neg, n, stride);
    MTT_HIP(hipGetLastError());
    return MTT_OK;
}

int mtt_icm_mh_fp64_batch(const double* d_stacks, int count, int n, const double* d_payouts, int k, double* d_equity,
                          hipStream_t stream) {
    MTT_ARG(n >= 1 && n <= ICM_MAXN, "ICM: player count %d out of 1..%d", n, ICM_MAXN);
    MTT_ARG(k >= 1 && count >= 0, "ICM: bad payout count or batch");
    if (!count) return MTT_OK;
    MTT_ARG(d_stacks && d_payouts && d_equity, "ICM: null pointer");
    hipLaunchKernelGGL(k_icm, dim3(count), dim3(256), 0, stream, d_stacks, n, d_payouts, k, d_equity);
    MTT_HIP(hipGetLastError());
    return MTT_OK;
}

static int icm_host_batch(const std::vector<double>& stacks, int count, int n, const double* payouts, int k,
                          std::vector<double>& eq) {
    for (double s : stacks) MTT_ARG(s >= 0.0 && std::isfinite(s), "ICM: stacks must be finite and >= 0");
    for (int q = 0; q < k; ++q) MTT_ARG(std::isfinite(payouts[q]), "ICM: payouts must be finite");
    double *ds = nullptr, *dp = nullptr, *de = nullptr;
    auto cleanup = [&]() { (void)hipFree(ds); (void)hipFree(dp); (void)hipFree(de); };
    if (hipMalloc(&ds, stacks.size() * sizeof(double)) != hipSuccess || hipMalloc(&dp, (size_t)k * sizeof(double)) != hipSuccess ||
        hipMalloc(&de, (size_t)count * n * sizeof(double)) != hipSuccess) {
        cleanup();
        set_error("ICM: device allocation failed");
        return MTT_ERR_OOM;
    }
    int rc = MTT_OK;
    if (hipMemcpy(ds, stacks.data(), stacks.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dp, payouts, (size_t)k * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        set_error("ICM: copy to device failed");
        rc = MTT_ERR_HIP;
    }
    if (!rc) rc = mtt_icm_mh_fp64_batch(ds, count, n, dp, k, de, nullptr);
    eq.assign((size_t)count * n, 0.0);
    if (!rc && hipMemcpy(eq.data(), de, eq.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
        set_error("ICM: copy to host failed");
        rc = MTT_ERR_HIP;
    }
    cleanup();
    return rc;
}

int mtt_icm_mh_fp64(const double* stacks, int n, const double* payouts, int k, double* out_equity) {
    MTT_ARG(stacks && payouts && out_equity, "mtt_icm_mh_fp64: null pointer");
    MTT_ARG(n >= 1 && n <= ICM_MAXN && k >= 1, "mtt_icm_mh_fp64: n %d (1..%d), k %d", n, ICM_MAXN, k);
    std::vector<double> st(stacks, stacks + n), eq;
    int rc = icm_host_batch(st, 1, n, payouts, k, eq);
    if (rc) return rc;
    memcpy(out_equity, eq.data(), (size_t)n * sizeof(double));
    return MTT_OK;
}

int mtt_icm_bubble_factors_fp64(const double* stacks, int n, const double* payouts, int k, const int* in_pot, int m,
                                double* out_bf) {
    MTT_ARG(stacks && payouts && in_pot && out_bf, "mtt_icm_bubble_factors_fp64: null pointer");
    MTT_ARG(n >= 2 && n <= ICM_MAXN && k >= 1 && m >= 1 && m <= n, "mtt_icm_bubble_factors_fp64: bad sizes");
    for (int a = 0; a < m; ++a) {
        MTT_ARG(in_pot[a] >= 0 && in_pot[a] < n, "mtt_icm_bubble_factors_fp64: player index out of range");
        for (int b = 0; b < a; ++b) MTT_ARG(in_pot[a] != in_pot[b], "mtt_icm_bubble_factors_fp64: repeated player");
    }
    // scenario 0: base; then per ordered pair (a, b): p loses e to q, p wins e from q
    std::vector<double> st((size_t)(1 + 2 * m * m) * n);
    for (int i = 0; i < n; ++i) st[i] = stacks[i];
    int count = 1;
    for (int a = 0; a < m; ++a)
        for (int b = 0; b < m; ++b) {
            if (a == b) continue;
            const int p = in_pot[a], q = in_pot[b];
            const double e = std::min(stacks[p], stacks[q]);
            for (int w = 0; w < 2; ++w) {
                double* s = st.data() + (size_t)count * n;
                for (int i = 0; i < n; ++i) s[i] = stacks[i];
                s[p] += w ? e : -e;
                s[q] += w ? -e : e;
                count++;
            }
        }
    st.resize((size_t)count * n);
    std::vector<double> eq;
    int rc = icm_host_batch(st, count, n, payouts, k, eq);
    if (rc) return rc;
    int c = 1;
    for (int a = 0; a < m; ++a) {
        double sum = 0.0;
        for (int b = 0; b < m; ++b) {
            if (a == b) continue;
            const int p = in_pot[a];
            const double lose = eq[p] - eq[(size_t)c * n + p];
            const double gain = eq[(size_t)(c + 1) * n + p] - eq[p];
            c += 2;
            out_bf[a * m + b] = gain > 0.0 ? lose / gain : 1.0;
            sum += out_bf[a * m + b];
        }
        out_bf[a * m + a] = m > 1 ? sum / (double)(m - 1) : 1.0;
    }
    return MTT_OK;
}

int mtt_multi_create(const mtt_4way_game* game, int n_flops, const uint8_t* flops, const float* ranges,
                     const float* bubble_factors, const mtt_solver_cfg* cfg, mtt_multi** out) {
    (void)hipGetLastError();   // an error left by an earlier call is not this call's
    MTT_ARG(game && flops && ranges && bubble_factors && out, "mtt_multi_create: null pointer");
    MTT_ARG(n_flops >= 1 && n_flops <= 65535, "mtt_multi_create: n_flops %d out of 1..65535", n_flops);
    for (int i = 0; i < 16; ++i)
        MTT_ARG(std::isfinite(bubble_factors[i]) && bubble_factors[i] > 0.0f,
                "mtt_multi_create: bubble factor %d must be finite and > 0", i);
    for (size_t f = 0; f < (size_t)n_flops; ++f)
        for (int p = 0; p < P4; ++p)
            for (int i = 0; i < kAliveFlop; ++i) {
                const float v = ranges[(f * P4 + p) * NP + i];
                MTT_ARG(std::isfinite(v) && v >= 0.0f,
                        "mtt_multi_create: flop %zu seat %d combo %d: weight must be finite and >= 0", f, p, i);
            }
    *out = nullptr;
    auto h = new (std::nothrow) mtt_multi;
    if (!h) { set_error("mtt_multi_create: out of host memory"); return MTT_ERR_OOM; }
    MultiImpl& s = h->impl;
    s.game = *game;
    if (cfg) s.cfg = *cfg;
    else { memset(&s.cfg, 0, sizeof(s.cfg)); s.cfg.scheme = MTT_SCHEME_CFR_PLUS; }
    s.nf = n_flops;
    int rc = build_4way_tree(*game, s.nodes, s.n_info);
    if (rc) { delete h; return rc; }
    // depth lists and terminals
    std::vector<int> depth(s.nodes.size(), 0);
    for (size_t n = 0; n < s.nodes.size(); ++n)
        if (s.nodes[n].type == MTT_NODE_DECISION)
            for (int a = 0; a < s.nodes[n].num_actions; ++a) depth[s.nodes[n].children_offset + a] = depth[n] + 1;
    for (size_t n = 0; n < s.nodes.size(); ++n) {
        if (s.nodes[n].type == MTT_NODE_DECISION) {
            if ((int)s.depth_lists.size() <= depth[n]) s.depth_lists.resize(depth[n] + 1);
            s.depth_lists[depth[n]].push_back((int)n);
        } else {
            s.terms.push_back((int)n);
        }
    }
    std::vector<mtt_board4> boards(n_flops);
    for (int f = 0; f < n_flops; ++f) {
        const uint8_t* fl = flops + 3 * f;
        if (!(fl[0] < 52 && fl[1] < 52 && fl[2] < 52 && fl[0] != fl[1] && fl[0] != fl[2] && fl[1] != fl[2])) {
            delete h;
            set_error("mtt_multi_create: flop %d is not three distinct cards", f);
            return MTT_ERR_ARG;
        }
        build_board4(fl, boards[f]);
    }
    if (hipStreamCreate(&s.stream) != hipSuccess) { delete h; set_error("mtt_multi_create: stream"); return MTT_ERR_HIP; }
    DeviceArena& A = s.arena;
    s.info_stride = (size_t)s.n_info * A4 * NP;
    s.node_stride = s.nodes.size() * P4 * NP;
    rc = A.upload(s.nodes, &s.d_nodes);
    for (auto& L : s.depth_lists) {
        int* d = nullptr;
        rc = rc ? rc : A.upload(L, &d);
        s.d_depth.push_back(d);
    }
    rc = rc ? rc : A.upload(s.terms, &s.d_terms);
    rc = rc ? rc : A.upload(boards, &s.d_boards);
    std::vector<float> bf(bubble_factors, bubble_factors + 16);
    rc = rc ? rc : A.upload(bf, &s.d_bf);
    rc = rc ? rc : A.alloc(s.info_stride * n_flops, &s.d_R);
    rc = rc ? rc : A.alloc(s.info_stride * n_flops, &s.d_S);
    rc = rc ? rc : A.alloc(s.info_stride * n_flops, &s.d_sig);
    rc = rc ? rc : A.alloc(s.node_stride * n_flops, &s.d_reach);
    rc = rc ? rc : A.alloc(s.node_stride * n_flops, &s.d_cfv);
    if (rc) { (void)hipStreamDestroy(s.stream); delete h; return rc; }
    hipError_t e = hipMemsetAsync(s.d_R, 0, s.info_stride * n_flops * 4, s.stream);
    if (e == hipSuccess) e = hipMemsetAsync(s.d_S, 0, s.info_stride * n_flops * 4, s.stream);
    if (e == hipSuccess) e = hipMemsetAsync(s.d_sig, 0, s.info_stride * n_flops * 4, s.stream);
    if (e == hipSuccess) e = hipMemsetAsync(s.d_reach, 0, s.node_stride * n_flops * 4, s.stream);
    if (e == hipSuccess) e = hipMemsetAsync(s.d_cfv, 0, s.node_stride * n_flops * 4, s.stream);
    // root reaches: ranges [flop][4][1216] into node 0
    for (int f = 0; f < n_flops && e == hipSuccess; ++f)
        e = hipMemcpyAsync(s.d_reach + (size_t)f * s.node_stride, ranges + (size_t)f * P4 * NP, (size_t)P4 * NP * 4,
                           hipMemcpyHostToDevice, s.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s.stream);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(s.stream);
        delete h;
        return hip_fail(e, "mtt_multi_create", __FILE__, __LINE__);
    }
    *out = h;
    return MTT_OK;
}

int mtt_multi_run(mtt_multi* h, int iters) {
    MTT_ARG(h && iters >= 0, "mtt_multi_run: bad arguments");
    MultiImpl& s = h->impl;
    const int D = (int)s.depth_lists.size();
    for (int it = 0; it < iters; ++it) {
        const int t = s.iters_done + 1;
        float wp, wn, ws;
        iteration_weights(s.cfg, t, wp, wn, ws);
        int rc = MTT_OK;
        MTT_TIMED(s, "4way_forward",
                  for (int d = 0; d < D && !rc; ++d)
                      rc = launch_4way_forward(s.d_depth[d], (int)s.depth_lists[d].size(), s.d_nodes, s.d_R, s.d_sig,
                                               s.d_reach, s.nf, s.info_stride, s.node_stride, s.stream));
        if (rc) return rc;
        MTT_TIMED(s, "4way_terminals",
                  rc = launch_4way_terminals(s.d_terms, (int)s.terms.size(), s.d_nodes, s.d_reach, s.d_boards, s.d_bf,
                                             s.d_cfv, s.nf, s.node_stride, s.stream));
        if (rc) return rc;
        MTT_TIMED(s, "4way_backward",
                  for (int d = D - 1; d >= 0 && !rc; --d)
                      rc = launch_4way_backward(s.d_depth[d], (int)s.depth_lists[d].size(), s.d_nodes, s.d_sig,
                                                s.d_reach, s.d_cfv, s.d_R, s.d_S, wp, wn, ws, s.nf, s.info_stride,
                                                s.node_stride, s.stream));
        if (rc) return rc;
        s.iters_done = t;
    }
    MTT_HIP(hipStreamSynchronize(s.stream));
    return MTT_OK;
}

int mtt_multi_get_state(mtt_multi* h, int flop, float* R_out, float* S_out, float* root_cfv_out) {
    MTT_ARG(h && flop >= 0 && flop < h->impl.nf, "mtt_multi_get_state: bad arguments");
    MultiImpl& s = h->impl;
    MTT_HIP(hipStreamSynchronize(s.stream));
    if (R_out) MTT_HIP(hipMemcpy(R_out, s.d_R + (size_t)flop * s.info_stride, s.info_stride * 4, hipMemcpyDeviceToHost));
    if (S_out) MTT_HIP(hipMemcpy(S_out, s.d_S + (size_t)flop * s.info_stride, s.info_stride * 4, hipMemcpyDeviceToHost));
    if (root_cfv_out)
        MTT_HIP(hipMemcpy(root_cfv_out, s.d_cfv + (size_t)flop * s.node_stride, (size_t)P4 * NP * 4, hipMemcpyDeviceToHost));
    return MTT_OK;
}

int mtt_multi_timing(mtt_multi* h, int enable) {
    MTT_ARG(h, "mtt_multi_timing: null handle");
    MTT_HIP(hipStreamSynchronize(h->impl.stream));
    h->impl.timer.collect();
    h->impl.timer.enabled = enable != 0;
    h->impl.timer.names.clear();
    h->impl.timer.ms.clear();
    h->impl.timer.count.clear();
    return MTT_OK;
}

int mtt_multi_kernel_times(mtt_multi* h, int idx, char* name, int name_cap, double* total_ms, int64_t* launches) {
    MTT_ARG(h && name && total_ms && launches && name_cap > 0, "mtt_multi_kernel_times: bad arguments");
    MTT_HIP(hipStreamSynchronize(h->impl.stream));
    h->impl.timer.collect();
    if (idx < 0 || idx >= (int)h->impl.timer.names.size()) return -1;
    snprintf(name, name_cap, "%s", h->impl.timer.names[idx].c_str());
    *total_ms = h->impl.timer.ms[idx];
    *launches = h->impl.timer.count[idx];
    return MTT_OK;
}

int mtt_multi_destroy(mtt_multi* h) {
    if (!h) return MTT_OK;
    (void)hipStreamSynchronize(h->impl.stream);
    h->impl.arena.release();
    (void)hipStreamDestroy(h->impl.stream);
    delete h;
    return MTT_OK;
}

}  // extern "C"
