// Preflop pieces of the reference's preflop program
// (src/preflop/preflop_grid_generator.cu) on CDNA4 (gfx950):
//   - the Monte-Carlo all-in equity matrix over the 169 hand classes
//     (kernel_gen_preflop_matrix, :95-127);
//   - class equity matrices of a flop set (kernel_gen_100_flop_matrices +
//     kernel_reduce_combo_to_class_matrices, :130-223), over all 1,326
//     combos;
//   - the preflop DCFR co-solver loop (:228-421, :588-614).
//
// The Monte-Carlo boards repeat cards (a board card is drawn from the deck
// without removing the cards drawn before it), and the reference's
// evaluator counts a repeated card twice.  The suit-mask evaluator of the
// solver path merges repeated cards, so this file scores boards with a
// multiset evaluator that keeps the reference's counting: rank counts as
// 4-bit fields of one 64-bit word, suit counts as 8-bit fields.
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "mtt_common.h"
#include "solver.h"

namespace mtt {

namespace {

// Representative hand of a class (class_to_cards, preflop_grid_generator.cu:
// 79-92): pairs in suits 0/1, suited hands both suit 0, offsuit suits 0/1.
// The rank-block walk advances while cls >= base + r1 (a block of high rank
// r1 holds r1 classes); the reference's test base + (r1 - 1) <= cls
// advances one class early, which gives the last class of each block two
// cards of one rank and never ends for 32s / 32o.
MTT_HD void class_cards(int cls, int& h1, int& h2) {
    if (cls < 13) {
        const int r = 12 - cls;
        h1 = 4 * r;
        h2 = 4 * r + 1;
        return;
    }
    int base = cls < 91 ? 13 : 91, r1 = 12;
    while (cls >= base + r1) {
        base += r1;
        r1--;
    }
    const int r2 = r1 - 1 - (cls - base);
    h1 = 4 * r1;
    h2 = 4 * r2 + (cls < 91 ? 0 : 1);
}

// Score of a 7-card multiset with the reference's rules (mtt_ev_dumper.cu:
// 53-83): the first suit (0..3) holding five or more cards is a flush (its
// rank mask may have fewer than five ranks when cards repeat: only the ranks
// present are placed); rank counts of exactly 4, 3 and 2 make quads (the
// lowest such rank), trips (highest, second highest) and pairs; other
// counts are ignored; the high card is the rank mask.
__device__ uint32_t eval7_multiset(const int* c) {
    uint64_t rc = 0;         // 4-bit count per rank
    uint32_t sc = 0;         // 8-bit count per suit
    uint64_t sm = 0;         // 16-bit rank mask per suit
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        rc += 1ull << (4 * (c[k] >> 2));
        sc += 1u << (8 * (c[k] & 3));
        sm |= card_bits16(c[k]);
    }
    for (int s = 0; s < 4; ++s) {
        if (((sc >> (8 * s)) & 0xFFu) < 5) continue;
        const uint32_t m = (uint32_t)(sm >> (16 * s)) & 0x1FFFu;
        const int sf = straight_top(m);
        if (sf >= 0) return (8u << 24) | (uint32_t)sf;
        uint32_t score = 5u << 24, rest = m;
        for (int p = 0; p < 5 && rest; ++p) {
            const int r = hibit(rest);
            score |= (uint32_t)r << (4 * (4 - p));
            rest &= ~(1u << r);
        }
        return score;
    }
    int q = -1, t1 = -1, t2 = -1, p1 = -1, p2 = -1;
    uint32_t rm = 0;
    for (int r = 12; r >= 0; --r) {
        const int n = (int)((rc >> (4 * r)) & 15u);
        if (n) rm |= 1u << r;
        if (n == 4) q = r;
        else if (n == 3) { if (t1 < 0) t1 = r; else if (t2 < 0) t2 = r; }
        else if (n == 2) { if (p1 < 0) p1 = r; else if (p2 < 0) p2 = r; }
    }
    if (q >= 0) return (7u << 24) | ((uint32_t)q << 4);
    if (t1 >= 0 && (t2 >= 0 || p1 >= 0)) return (6u << 24) | ((uint32_t)t1 << 4) | (uint32_t)(t2 >= 0 ? t2 : p1);
    const int st = straight_top(rm);
    if (st >= 0) return (4u << 24) | (uint32_t)st;
    if (t1 >= 0) return (3u << 24) | ((uint32_t)t1 << 8);
    if (p1 >= 0 && p2 >= 0) return (2u << 24) | ((uint32_t)p1 << 8) | ((uint32_t)p2 << 4);
    if (p1 >= 0) return (1u << 24) | ((uint32_t)p1 << 12);
    return rm;
}

// The reference's LCG, rng <- rng * 1664525 + 1013904223 (mod 2^32), advanced
// n steps at once: the n-fold composition of the affine map, by squaring.
__device__ __forceinline__ uint32_t lcg_jump(uint32_t s, uint32_t n) {
    uint32_t A = 1u, C = 0u, a = 1664525u, c = 1013904223u;
    while (n) {
        if (n & 1u) {
            A = a * A;
            C = a * C + c;
        }
        c = a * c + c;
        a = a * a;
        n >>= 1;
    }
    return A * s + C;
}

// Position of the r-th (0-based) set bit of m (r < popcount(m)): a binary
// search by popcounts over halves, quarters, .. of the word, branch-free.
__device__ __forceinline__ int nth_set_bit(uint64_t m, int r) {
    const int lo = __builtin_popcount((uint32_t)m);
    const bool up = r >= lo;
    uint32_t x = up ? (uint32_t)(m >> 32) : (uint32_t)m;
    int pos = up ? 32 : 0;
    r -= up ? lo : 0;
#pragma unroll
    for (int w = 16; w >= 1; w >>= 1) {
        const int c = __builtin_popcount(x & ((1u << w) - 1u));
        const bool hi = r >= c;
        r -= hi ? c : 0;
        x = hi ? x >> w : x;
        pos += hi ? w : 0;
    }
    return pos;
}

// kMcSplit lanes per (hero class, villain class) cell, each taking a
// contiguous share of the cell's samples: the reference's sample stream
// (32-bit LCG seeded with cell * 2654435761 + 0x9E3779B9, five steps per
// sample; board card k = deck[k + (rng >> 16) % (48 - k)]) entered at the
// share's first sample by a jump; the lanes' win and tie counts are summed
// exactly, so the counts are the single-stream ones.
constexpr int kMcSplit = 16;
__global__ __launch_bounds__(256) void k_preflop_matrix(int samples, float* __restrict__ eq,
                                                        int32_t* __restrict__ wins, int32_t* __restrict__ ties) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int cell = tid / kMcSplit, part = tid % kMcSplit;
    const bool valid = cell < 169 * 169;
    const int h = valid ? cell / 169 : 0, v = valid ? cell % 169 : 1;
    int w = 0, t = 0;
    if (valid && h != v) {
        int h1, h2, v1, v2;
        class_cards(h, h1, h2);
        class_cards(v, v1, v2);
        const int ov1 = v1, ov2 = v2;
        if (h1 == ov1 || h2 == ov1) v1 = (ov1 >> 2) * 4 + 2;     // collision rule, :103-106
        if (h1 == ov2 || h2 == ov2) v2 = (ov2 >> 2) * 4 + 3;
        const uint64_t dead = (1ull << h1) | (1ull << h2) | (1ull << v1) | (1ull << v2);
        // deck[i] = i-th smallest card outside `dead`: the 48 live cards as a mask, the i-th one found
        // by popcount steps
        const uint64_t live = ~dead & ((1ull << 52) - 1);
        const int per = (samples + kMcSplit - 1) / kMcSplit;
        const int s0 = min(samples, part * per), s1 = min(samples, s0 + per);
        uint32_t rng = lcg_jump((uint32_t)cell * 2654435761u + 0x9E3779B9u, 5u * (uint32_t)s0);
        int hh[7] = {h1, h2, 0, 0, 0, 0, 0}, vv[7] = {v1, v2, 0, 0, 0, 0, 0};
        for (int s = s0; s < s1; ++s) {
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                rng = rng * 1664525u + 1013904223u;
                const int pick = k + (int)((rng >> 16) % (uint32_t)(48 - k));
                const int card = nth_set_bit(live, pick);
                hh[2 + k] = card;
                vv[2 + k] = card;
            }
            const uint32_t sh = eval7_multiset(hh), sv = eval7_multiset(vv);
            w += sh > sv;
            t += sh == sv;
        }
    }
    // the cell's kMcSplit lanes are consecutive lanes of one wave
#pragma unroll
    for (int o = kMcSplit / 2; o > 0; o >>= 1) {
        w += __shfl_xor(w, o, 64);
        t += __shfl_xor(t, o, 64);
    }
    if (!valid || part != 0) return;
    if (h == v) {
        eq[cell] = 0.5f;
        if (wins) wins[cell] = 0;
        if (ties) ties[cell] = 0;
        return;
    }
    eq[cell] = ((float)w + 0.5f * (float)t) / (float)samples;
    if (wins) wins[cell] = w;
    if (ties) ties[cell] = t;
}

}  // namespace

namespace {

// ------------------------------------------------ class equity matrices --
// Tables: the cards of the 1,326 combos (lexicographic), the members of each
// class in that order, and the 192 (turn, river) index pairs into a hero's
// 47-card deck that the reference samples (turn 0, 2, .., 46; river r1 + 1,
// r1 + 4, .. < 47; preflop_grid_generator.cu:164-167).
constexpr int kPfCombos = 1326;
constexpr int kPfLd = 1344;          // code-table pitch: 1,326 combos padded to 21 x 64
constexpr int kPfRows = 1176;        // (turn, river) pairs of the 49 cards off a flop
constexpr int kPfPts = 192;          // sampled runouts per hero
constexpr int kPfEqThreads = 704;    // 672 villain pairs (2 combos per lane), 11 waves
constexpr int kPfBatch = 64;         // flops per code-table batch (203 MB of codes)
constexpr int kPfMaxFlops = 4096;

constexpr int class_of(int a, int b) {
    const int r1 = a >> 2, s1 = a & 3, r2 = b >> 2, s2 = b & 3;
    const int hi = r1 > r2 ? r1 : r2, lo = r1 < r2 ? r1 : r2;
    if (r1 == r2) return 12 - hi;
    int base = s1 == s2 ? 13 : 91;
    for (int r = 12; r > hi; --r) base += r;
    return base + (hi - 1 - lo);
}

struct PfTables {
    uint8_t c1[kPfCombos], c2[kPfCombos];
    uint16_t members[169][12];
    uint8_t n_members[169];
    uint8_t r1[kPfPts], r2[kPfPts];
};
constexpr PfTables make_pf_tables() {
    PfTables t{};
    int g = 0;
    for (int a = 0; a < 52; ++a)
        for (int b = a + 1; b < 52; ++b, ++g) {
            t.c1[g] = (uint8_t)a;
            t.c2[g] = (uint8_t)b;
            const int c = class_of(a, b);
            t.members[c][t.n_members[c]++] = (uint16_t)g;
        }
    int n = 0;
    for (int x = 0; x < 47; x += 2)
        for (int y = x + 1; y < 47; y += 3) { t.r1[n] = (uint8_t)x; t.r2[n] = (uint8_t)y; n++; }
    return t;
}
__constant__ PfTables c_pf = make_pf_tables();
static_assert(make_pf_tables().r1[kPfPts - 1] == 44, "192 sampled runouts, the last one (44, 45)");

struct PfFlops {                    // sorted flop cards by value
    uint8_t c[3 * kPfBatch];
};

// Deck position p (0..48) of the cards off a flop f0 < f1 < f2 -> card.
__device__ __forceinline__ int pf_card49(int p, int f0, int f1, int f2) {
    p += p >= f0;
    p += p >= f1;
    p += p >= f2;
    return p;
}

// Row of the (x < y) pair of deck positions among the 1,176 runouts.
__device__ __forceinline__ int pf_row(int x, int y) { return (97 * x - x * x) / 2 + (y - x - 1); }

// codes[f][b][g] = code16 of combo g with flop f and runout b, 0xFFFF when
// the combo holds a board card (and for the padding 1326..1343).  Two combos
// per thread and dword.
__global__ __launch_bounds__(256) void k_pf_codes(PfFlops fl, uint32_t* __restrict__ codes) {
    const int b = blockIdx.x, f = blockIdx.y;
    const int f0 = fl.c[3 * f], f1 = fl.c[3 * f + 1], f2 = fl.c[3 * f + 2];
    int x = 0;
    while (pf_row(x + 1, x + 2) <= b) ++x;
    const int y = b - pf_row(x, x + 1) + x + 1;
    const int t = pf_card49(x, f0, f1, f2), r = pf_card49(y, f0, f1, f2);
    const uint64_t dead = (1ull << f0) | (1ull << f1) | (1ull << f2) | (1ull << t) | (1ull << r);
    const uint64_t board = card_bits16(f0) | card_bits16(f1) | card_bits16(f2) | card_bits16(t) | card_bits16(r);
    uint32_t* row = codes + ((size_t)f * kPfRows + b) * (kPfLd / 2);
    for (int q = threadIdx.x; q < kPfLd / 2; q += blockDim.x) {
        uint32_t v[2] = {0xFFFFu, 0xFFFFu};
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int g = 2 * q + k;
            if (g >= kPfCombos) continue;
            const int a = c_pf.c1[g], c = c_pf.c2[g];
            if (dead & ((1ull << a) | (1ull << c))) continue;
            v[k] = code16(score_packed(board | card_bits16(a) | card_bits16(c)));
        }
        row[q] = v[0] | (v[1] << 16);
    }
}

typedef unsigned short pf_u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ pf_u16x2 pf_bits(uint32_t v) { return __builtin_bit_cast(pf_u16x2, v); }

// One workgroup per (hero class H, flop): for each member h of H that avoids
// the flop, the sampled runouts of h's deck are placed in LDS (row offset,
// hero code + 1 in both 16-bit halves); each lane then counts two villain
// combos at once in 16-bit lanes:
//     x = min(sat_sub(hero + 1, villain), 2)   2 win, 1 tie, 0 loss or dead
//     s1 += x, s2 += x * x, d += villain >> 15  (dead codes are 0xFFFF)
// W = (s2 - s1) / 2, T = 2 s1 - s2, live runouts n = 192 - d, and the pair's
// equity is (2W + T/2) / (2n), the reference's entry (a tie counts a
// quarter).  The villain classes then add the entries of their members that
// avoid the flop and the hero, hero members outer and villain members inner,
// in FP32: M[f][v][H] = sum / count, 0.5 when no pair is valid.
// Flop-to-XCD mapping: workgroup i runs on XCD i % 8 and XCD k takes the
// flops k, k + 8, .., so each XCD's L2 holds the code table of one flop
// (3.2 MB) at a time.
__global__ __launch_bounds__(kPfEqThreads) void k_pf_class_eq(PfFlops fl, int n_flops,
                                                              const uint32_t* __restrict__ codes,
                                                              float* __restrict__ M) {
    __shared__ float eqv[kPfLd];
    __shared__ uint2 pts[kPfPts];
    const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
    const int f = xcd + 8 * (slot / 169), H = slot % 169;
    if (f >= n_flops) return;
    const int tid = threadIdx.x;
    const int f0 = fl.c[3 * f], f1 = fl.c[3 * f + 1], f2 = fl.c[3 * f + 2];
    const uint64_t fm = (1ull << f0) | (1ull << f1) | (1ull << f2);
    const uint32_t* tab = codes + (size_t)f * kPfRows * (kPfLd / 2);
    const unsigned short* tab16 = reinterpret_cast<const unsigned short*>(tab);
    const pf_u16x2 two = {2, 2};
    const pf_u16x2 fifteen = {15, 15};
    float sum = 0.0f;
    int cnt = 0;
    const int nh = c_pf.n_members[H];
    for (int i = 0; i < nh; ++i) {
        const int hg = c_pf.members[H][i];
        const int h1 = c_pf.c1[hg], h2 = c_pf.c2[hg];
        const uint64_t hm = (1ull << h1) | (1ull << h2);
        if (hm & fm) continue;
        // deck positions of the hero cards among the 49 cards off the flop
        const int hx = h1 - (f0 < h1) - (f1 < h1) - (f2 < h1);
        const int hy = h2 - (f0 < h2) - (f1 < h2) - (f2 < h2);
        if (tid < kPfPts) {
            int p1 = c_pf.r1[tid], p2 = c_pf.r2[tid];
            p1 += p1 >= hx;
            p1 += p1 >= hy;
            p2 += p2 >= hx;
            p2 += p2 >= hy;
            const uint32_t off = (uint32_t)pf_row(p1, p2) * (kPfLd / 2);
            const uint32_t hs1 = (uint32_t)tab16[2 * (size_t)off + hg] + 1u;
            pts[tid] = make_uint2(off, hs1 | (hs1 << 16));
        }
        __syncthreads();
        if (tid < kPfLd / 2) {
            // buffer descriptor over this flop's code table (every load lies inside it)
            const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                const_cast<uint32_t*>(tab), (short)0, kPfRows * kPfLd * (int)sizeof(uint16_t), 0x00020000);
            pf_u16x2 s1 = {0, 0}, s2 = {0, 0}, d = {0, 0};
#pragma unroll 8
            for (int p = 0; p < kPfPts; ++p) {
                const uint2 e = pts[p];
                // row offset in the scalar offset, the lane's dword in the vector offset
                const pf_u16x2 v = pf_bits(__builtin_amdgcn_raw_buffer_load_b32(
                    rs, 4 * tid, (int)(4 * __builtin_amdgcn_readfirstlane(e.x)), 0));
                const pf_u16x2 x = __builtin_elementwise_min(__builtin_elementwise_sub_sat(pf_bits(e.y), v), two);
                s1 += x;
                s2 += x * x;
                d += v >> fifteen;
            }
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int a = k ? s1.y : s1.x, b = k ? s2.y : s2.x, n = kPfPts - (k ? d.y : d.x);
                const int w = (b - a) >> 1, t = 2 * a - b;
                eqv[2 * tid + k] = n > 0 ? ((float)(2 * w) + 0.5f * (float)t) / (float)(2 * n) : 0.0f;
            }
        }
        __syncthreads();
        if (tid < 169) {
            const int nv = c_pf.n_members[tid];
            for (int j = 0; j < nv; ++j) {
                const int g = c_pf.members[tid][j];
                if ((fm | hm) & ((1ull << c_pf.c1[g]) | (1ull << c_pf.c2[g]))) continue;
                sum += eqv[g];
                cnt++;
            }
        }
        __syncthreads();
    }
    if (tid < 169) M[((size_t)f * 169 + tid) * 169 + H] = cnt ? sum / (float)cnt : 0.5f;
}

// ----------------------------------------------------- co-solver loop ----
// State of the closed tree, one block of floats (classes padded to 192):
//   R, sigma, S  [4 nodes][3 actions][192]  (button root 3 actions, big
//                blind defence 3, button facing the 3-bet jam 2, big blind
//                facing the open jam 2)
//   reach        [2][192]  root reach x call probability into the flop:
//                button (min-raise), big blind (call)
//   ev           [2][192]  flop net EV per class: button, big blind
//   root         [192]     class combo count / 1326
constexpr int kPfC = 169, kPfCP = 192;
constexpr int kPfNode = 3 * kPfCP;
constexpr int kPfAllin4 = (kPfC * kPfC + 3) / 4;   // all-in matrix as float4 (28,564 floats)
constexpr int kPfTreeThreads = 1024;
struct PfState {
    float R[4 * kPfNode], sg[4 * kPfNode], S[4 * kPfNode];
    float reach[2 * kPfCP], ev[2 * kPfCP], root[kPfCP];
};

__global__ __launch_bounds__(256) void k_pf_init(PfState* __restrict__ st) {
    const int c = threadIdx.x;
    if (c >= kPfCP) return;
    const float root = c < 13 ? 6.0f / 1326.0f : (c < 91 ? 4.0f / 1326.0f : (c < kPfC ? 12.0f / 1326.0f : 0.0f));
    for (int n = 0; n < 4; ++n)
        for (int a = 0; a < 3; ++a) {
            const int k = n * kPfNode + a * kPfCP + c;
            const int na = n < 2 ? 3 : 2;
            st->R[k] = 0.0f;
            st->S[k] = 0.0f;
            st->sg[k] = a < na ? 1.0f / (float)na : 0.0f;
        }
    st->root[c] = root;
    st->ev[c] = 0.0f;
    st->ev[kPfCP + c] = 0.0f;
    st->reach[c] = root * (1.0f / 3.0f);
    st->reach[kPfCP + c] = root * (1.0f / 3.0f);
}

// MT[f][o][c] = 1 - M[f][c][o]: the big blind's side of the co-solve reads
// the flop matrix by rows, stored once as columns so both sides read
// consecutive classes.  Also copies the all-in matrix into a 16-byte aligned,
// zero-padded buffer for the tree kernel's vector loads.
__global__ __launch_bounds__(256) void k_pf_flip(const float* __restrict__ M, float* __restrict__ MT, int n_flops,
                                                 const float* __restrict__ allin, float* __restrict__ allin_pad) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 4 * (size_t)kPfAllin4) allin_pad[i] = i < (size_t)kPfC * kPfC ? allin[i] : 0.0f;
    if (i >= (size_t)n_flops * kPfC * kPfC) return;
    const size_t f = i / (kPfC * kPfC);
    const int r = (int)(i % (kPfC * kPfC)), o = r / kPfC, c = r % kPfC;
    MT[i] = 1.0f - M[f * kPfC * kPfC + (size_t)c * kPfC + o];
}

// kernel_cosolve_100_board_matrices (preflop_grid_generator.cu:228-285):
// for each flop and class, the class's mean equity against the opponent's
// reach into the flop, realised at 1.06 (button) / 0.94 (big blind) of the
// flop pot, minus the 2.1bb investment (a loss scaled by the bubble
// factor).  Three one-wave workgroups per flop (classes 0-63, 64-127,
// 128-168), so a flop's 228 KB of matrix reads spread over three CUs.
// net[side][c][f], flop index fastest (row pitch `pitch`), so the tree
// kernel reads a class's flops with vector loads.
__global__ __launch_bounds__(64) void k_pf_cosolve(const float* __restrict__ M, const float* __restrict__ MT,
                                                   const PfState* __restrict__ st, float* __restrict__ net,
                                                   int pitch, float pot_flop, float bf) {
#pragma clang fp contract(off)
    const int f = blockIdx.x, c = blockIdx.y * 64 + threadIdx.x;
    if (c >= kPfC) return;
    const float* E = M + (size_t)f * kPfC * kPfC;
    const float* ET = MT + (size_t)f * kPfC * kPfC;
    const float* rbc = st->reach;
    const float* rbb = st->reach + kPfCP;
    // both sides in one fully unrolled loop: 338 independent loads in flight,
    // four chains in class order
    float s1 = 0.0f, e0 = 0.0f, s0 = 0.0f, e1 = 0.0f;
#pragma unroll
    for (int o = 0; o < kPfC; ++o) {
        s1 += rbb[o];
        e0 = fmaf(rbb[o], E[o * kPfC + c], e0);
        s0 += rbc[o];
        e1 = fmaf(rbc[o], ET[o * kPfC + c], e1);
    }
    if (s1 > 1e-7f) e0 /= s1;
    if (s0 > 1e-7f) e1 /= s0;
    const float sh_btn = e0 * 1.06f * pot_flop, sh_bb = e1 * 0.94f * pot_flop;
    net[(size_t)c * pitch + f] = sh_btn >= 2.1f ? sh_btn - 2.1f : -((2.1f - sh_btn) * bf);
    net[((size_t)kPfCP + c) * pitch + f] = sh_bb >= 2.1f ? sh_bb - 2.1f : -((2.1f - sh_bb) * bf);
}

// Discounted regret update and regret matching of one node for one class
// (kernel_dcfr_step, preflop_grid_generator.cu:387-410).
__device__ __forceinline__ void pf_dcfr(int na, float* __restrict__ R, float* __restrict__ sg,
                                        float* __restrict__ S, const float* cfv, const float* sg_old, float reach,
                                        float wp, float ws, int c) {
#pragma clang fp contract(off)
    float node_v = 0.0f;
    for (int a = 0; a < na; ++a) node_v = fmaf(sg_old[a], cfv[a], node_v);
    float pos = 0.0f, nx[3] = {0.0f, 0.0f, 0.0f};
    for (int a = 0; a < na; ++a) {
        const int k = a * kPfCP + c;
        const float q = R[k] + (cfv[a] - node_v);
        const float nr = wp * (q > 0.0f ? q : 0.0f);
        R[k] = nr;
        pos += nr;
        nx[a] = nr;
        S[k] = fmaf(ws * reach, sg_old[a], S[k]);
    }
    for (int a = 0; a < na; ++a) sg[a * kPfCP + c] = pos > 1e-7f ? nx[a] / pos : 1.0f / (float)na;
}

// One iteration's closed tree (kernel_build_closed_preflop_cfvs,
// preflop_grid_generator.cu:290-385) and DCFR steps of the four nodes, in
// one workgroup:
//   0. the all-in matrix, every class's strategies and the root reach are
//      staged in LDS; meanwhile each class thread sums its subset-weighted
//      flop net EVs (flops in order, 16-byte loads of its net row);
//   1. the opponent weights of the tree values (root reach x strategy, per
//      opponent class) are formed once, and the nine opponent masses that
//      do not depend on the hero class (e.g. the 3-bet jamming mass) are
//      summed once, one lane per chain, in class order;
//   2. each class thread forms its four all-in equities (rows and columns of
//      the all-in matrix against those weights), its tree values and its
//      DCFR updates, and writes the next reaches into the flop
//      (kernel_update_dynamic_reaches, :412-421).
// Every sum keeps the reference's class order and operation shapes.
__global__ __launch_bounds__(kPfTreeThreads) void k_pf_tree(PfState* __restrict__ st, const float* __restrict__ net,
                                                            int pitch, const float* __restrict__ w, int n_flops,
                                                            const float4* __restrict__ allin4, float wp, float ws,
                                                            float stack_bb, float bf, float dead_pot) {
#pragma clang fp contract(off)
    __shared__ float4 A4[kPfAllin4];
    __shared__ float4 sg4[kPfNode];            // 4 nodes x 3 actions x 192 floats
    __shared__ float root[kPfCP];
    __shared__ float4 V[kPfC];                 // per opponent: p12, u, p02, q3 (below)
    __shared__ float RO[kPfC];                 // per opponent: root x BTN min-raise
    __shared__ float mass[9];
    __shared__ float one;
    float* A = reinterpret_cast<float*>(A4);
    float* sg = reinterpret_cast<float*>(sg4);
    const int tid = threadIdx.x;
    const int c = tid;
    // 0. staging, and the flop sums of the class threads
#pragma unroll
    for (int k = 0; k < (kPfAllin4 + kPfTreeThreads - 1) / kPfTreeThreads; ++k) {
        const int i = tid + k * kPfTreeThreads;
        if (i < kPfAllin4) A4[i] = allin4[i];
    }
    if (tid < kPfNode) sg4[tid] = reinterpret_cast<const float4*>(st->sg)[tid];
    if (tid < kPfCP) root[tid] = st->root[tid];
    if (tid == 0) one = 1.0f;
    float ev_btn = 0.0f, ev_bb = 0.0f;
    if (c < kPfC) {
        const float4* nb = reinterpret_cast<const float4*>(net + (size_t)c * pitch);
        const float4* nbb = reinterpret_cast<const float4*>(net + ((size_t)kPfCP + c) * pitch);
        float acc_btn = 0.0f, acc_bb = 0.0f, tw = 0.0f;
        const int n4 = n_flops >> 2;
#pragma unroll 8
        for (int q = 0; q < n4; ++q) {
            const float4 x = nb[q], y = nbb[q];
            const float4 wq = reinterpret_cast<const float4*>(w)[q];
            tw += wq.x; acc_btn = fmaf(wq.x, x.x, acc_btn); acc_bb = fmaf(wq.x, y.x, acc_bb);
            tw += wq.y; acc_btn = fmaf(wq.y, x.y, acc_btn); acc_bb = fmaf(wq.y, y.y, acc_bb);
            tw += wq.z; acc_btn = fmaf(wq.z, x.z, acc_btn); acc_bb = fmaf(wq.z, y.z, acc_bb);
            tw += wq.w; acc_btn = fmaf(wq.w, x.w, acc_btn); acc_bb = fmaf(wq.w, y.w, acc_bb);
        }
        for (int f = 4 * n4; f < n_flops; ++f) {
            const float wf = w[f];
            tw += wf;
            acc_btn = fmaf(wf, net[(size_t)c * pitch + f], acc_btn);
            acc_bb = fmaf(wf, net[((size_t)kPfCP + c) * pitch + f], acc_bb);
        }
        if (tw > 0.0f) {
            ev_btn = acc_btn / tw;
            ev_bb = acc_bb / tw;
            st->ev[c] = ev_btn;
            st->ev[kPfCP + c] = ev_bb;
        } else {
            ev_btn = st->ev[c];
            ev_bb = st->ev[kPfCP + c];
        }
    }
    __syncthreads();
    const float* s0 = sg;                 // button root
    const float* s1 = sg + kPfNode;       // big blind defence
    const float* s2 = sg + 2 * kPfNode;   // button facing the 3-bet jam
    const float* s3 = sg + 3 * kPfNode;   // big blind facing the open jam
    // 1. per opponent o: p12 = root x BB 3-bet jam, u = (root x BTN min-raise) x
    //    BTN call of the jam, p02 = root x BTN open jam, q3 = root x BB call of
    //    the open jam (each product rounded once, as the reference forms it)
    if (tid < kPfC) {
        const int o = tid;
        const float ro = root[o] * s0[kPfCP + o];
        RO[o] = ro;
        V[o] = make_float4(root[o] * s1[2 * kPfCP + o], ro * s2[kPfCP + o], root[o] * s0[2 * kPfCP + o],
                           root[o] * s3[kPfCP + o]);
    }
    __syncthreads();
    if (tid < 9) {
        // lane k: one opponent mass sum_o a[o] * b[o] in class order, as an FMA
        // chain (a plain sum is the chain with b = 1: an FMA by 1 rounds as the
        // addition does)
        const float* Vf = reinterpret_cast<const float*>(V);
        const float* pa;
        const float* pb;
        int sa = 1, sb = 1;
        switch (tid) {
            case 0: pa = Vf; sa = 4; pb = &one; sb = 0; break;                 // BB 3-bet jam mass
            case 1: pa = RO; pb = s2; break;                                   // min-raise, then fold to the jam
            case 2: pa = RO; pb = s2 + kPfCP; break;                           // min-raise, then call the jam
            case 3: pa = Vf + 2; sa = 4; pb = &one; sb = 0; break;             // BTN open-jam mass
            case 4: pa = root; pb = s1; break;                                 // BB folds to the min-raise
            case 5: pa = root; pb = s1 + kPfCP; break;                         // BB calls
            case 6: pa = root; pb = s1 + 2 * kPfCP; break;                     // BB 3-bet jams
            case 7: pa = root; pb = s3; break;                                 // BB folds to the open jam
            default: pa = root; pb = s3 + kPfCP; break;                        // BB calls the open jam
        }
        float m = 0.0f;
#pragma unroll 13
        for (int o = 0; o < kPfC; ++o) m = fmaf(pa[o * sa], pb[o * sb], m);
        mass[tid] = m;
    }
    __syncthreads();
    if (c >= kPfC) return;
    const float m3 = mass[0], bf3 = mass[1], bc3 = mass[2], mj = mass[3];
    const float bff = mass[4], bfc = mass[5], bfj = mass[6], fj = mass[7], cj = mass[8];
    // 2. the class's all-in equities against those opponent weights
    float eq_call = 0.0f, eq_jam = 0.0f, eq_oj = 0.0f, eq_b = 0.0f;
#pragma unroll 13
    for (int o = 0; o < kPfC; ++o) {
        const float4 v = V[o];
        const float row = A[c * kPfC + o], col = 1.0f - A[o * kPfC + c];
        eq_call = fmaf(v.x, row, eq_call);
        eq_jam = fmaf(v.y, col, eq_jam);
        eq_oj = fmaf(v.z, col, eq_oj);
        eq_b = fmaf(v.w, row, eq_b);
    }
    const float jam_pot = 2.0f * stack_bb + dead_pot;
    float cfv[4][3];
    // button facing the 3-bet jam
    cfv[2][0] = -2.1f * bf;
    if (m3 > 1e-7f) eq_call /= m3;
    cfv[2][1] = eq_call * (jam_pot - stack_bb) - (1.0f - eq_call) * bf * stack_bb;
    cfv[2][2] = 0.0f;
    const float v3b = s2[c] * cfv[2][0] + s2[kPfCP + c] * cfv[2][1];
    // big blind facing the min-raise
    cfv[1][0] = -1.0f * bf;
    cfv[1][1] = ev_bb;
    if (bc3 > 1e-7f) eq_jam /= bc3;
    const float tot3 = bf3 + bc3 + 1e-9f;
    const float win_unc = dead_pot + 2.1f;
    const float ev_sh = eq_jam * (jam_pot - stack_bb) - (1.0f - eq_jam) * bf * stack_bb;
    cfv[1][2] = (bf3 * win_unc + bc3 * ev_sh) / tot3;
    // big blind facing the open jam
    cfv[3][0] = -1.0f * bf;
    if (mj > 1e-7f) eq_oj /= mj;
    cfv[3][1] = eq_oj * (jam_pot - stack_bb) - (1.0f - eq_oj) * bf * stack_bb;
    cfv[3][2] = 0.0f;
    // button root
    cfv[0][0] = 0.0f;
    const float tbb = bff + bfc + bfj + 1e-9f;
    cfv[0][1] = (bff * dead_pot + bfc * ev_btn + bfj * v3b) / tbb;
    const float toj = fj + cj + 1e-9f;
    if (cj > 1e-7f) eq_b /= cj;
    cfv[0][2] = (fj * dead_pot + cj * (eq_b * (jam_pot - stack_bb) - (1.0f - eq_b) * bf * stack_bb)) / toj;
    // DCFR on the four nodes (own class only: the LDS strategies stay the
    // old ones every other class read above)
    const float rc = root[c];
    const int na[4] = {3, 3, 2, 2};
    for (int n = 0; n < 4; ++n) {
        const float old[3] = {sg[n * kPfNode + c], sg[n * kPfNode + kPfCP + c], sg[n * kPfNode + 2 * kPfCP + c]};
        pf_dcfr(na[n], st->R + n * kPfNode, st->sg + n * kPfNode, st->S + n * kPfNode, cfv[n], old, rc, wp, ws, c);
    }
    st->reach[c] = rc * st->sg[kPfCP + c];
    st->reach[kPfCP + c] = rc * st->sg[kPfNode + kPfCP + c];
}

__global__ __launch_bounds__(256) void k_pf_sums(const PfState* __restrict__ st, float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 12 * kPfC) return;
    const int nc = i / kPfC, c = i % kPfC;
    out[i] = st->S[nc * kPfCP + c];
}

}  // namespace

}  // namespace mtt

using namespace mtt;

namespace {

std::mutex g_pf_timer_mu;
KernelTimer g_pf_timer;

struct PfMem {
    void* p = nullptr;
    ~PfMem() {
        if (p) (void)hipFree(p);
    }
};

bool pf_flop_ok(const uint8_t* f) {
    return f[0] < 52 && f[1] < 52 && f[2] < 52 && f[0] != f[1] && f[0] != f[2] && f[1] != f[2];
}

// Launch with the kernel timer's marks around it when timing is on.
template <typename F>
void pf_timed(const char* name, hipStream_t stream, F&& launch) {
    if (!g_pf_timer.enabled) {
        launch();
        return;
    }
    KernelTimer::Mark m;
    g_pf_timer.begin(name, stream, m);
    launch();
    g_pf_timer.end(m, stream);
}

int pf_class_matrices(const uint8_t* flops, int n_flops, float* d_M, hipStream_t stream) {
    const int nb = n_flops < kPfBatch ? n_flops : kPfBatch;
    PfMem codes;
    MTT_HIP(hipMalloc(&codes.p, (size_t)nb * kPfRows * kPfLd * sizeof(uint16_t)));
    std::lock_guard<std::mutex> lock(g_pf_timer_mu);
    for (int f0 = 0; f0 < n_flops; f0 += kPfBatch) {
        const int cnt = n_flops - f0 < kPfBatch ? n_flops - f0 : kPfBatch;
        PfFlops fb{};
        for (int f = 0; f < cnt; ++f) {
            uint8_t c[3] = {flops[3 * (f0 + f)], flops[3 * (f0 + f) + 1], flops[3 * (f0 + f) + 2]};
            if (c[0] > c[1]) std::swap(c[0], c[1]);
            if (c[1] > c[2]) std::swap(c[1], c[2]);
            if (c[0] > c[1]) std::swap(c[0], c[1]);
            memcpy(fb.c + 3 * f, c, 3);
        }
        uint32_t* cw = static_cast<uint32_t*>(codes.p);
        pf_timed("pf_codes", stream, [&] {
            hipLaunchKernelGGL(k_pf_codes, dim3(kPfRows, cnt), dim3(256), 0, stream, fb, cw);
        });
        const int cnt8 = (cnt + 7) & ~7;
        float* out = d_M + (size_t)f0 * kPfC * kPfC;
        pf_timed("pf_class_eq", stream, [&] {
            hipLaunchKernelGGL(k_pf_class_eq, dim3(169 * cnt8), dim3(kPfEqThreads), 0, stream, fb, cnt, cw, out);
        });
        MTT_HIP(hipGetLastError());
    }
    MTT_HIP(hipStreamSynchronize(stream));
    return MTT_OK;
}

int pf_solve(const float* d_M, const float* weights, int n_flops, const float* d_allin, int iters, float stack_bb,
             float bf, float dead_pot, float* d_sums, hipStream_t stream) {
    const size_t mat = (size_t)n_flops * kPfC * kPfC;
    const size_t bytes_state = (sizeof(PfState) + 255) & ~(size_t)255;
    const size_t bytes_mt = (mat * sizeof(float) + 255) & ~(size_t)255;
    const int pitch = (n_flops + 3) & ~3;   // net row pitch: 16-byte aligned rows
    const size_t bytes_net = ((size_t)2 * kPfCP * pitch * sizeof(float) + 255) & ~(size_t)255;
    const size_t bytes_allin = ((size_t)kPfAllin4 * sizeof(float4) + 255) & ~(size_t)255;
    PfMem ws;
    MTT_HIP(hipMalloc(&ws.p, bytes_state + bytes_mt + bytes_net + bytes_allin + (size_t)pitch * sizeof(float)));
    char* base = static_cast<char*>(ws.p);
    PfState* st = reinterpret_cast<PfState*>(base);
    float* MT = reinterpret_cast<float*>(base + bytes_state);
    float* net = reinterpret_cast<float*>(base + bytes_state + bytes_mt);
    float4* allin4 = reinterpret_cast<float4*>(base + bytes_state + bytes_mt + bytes_net);
    float* w = reinterpret_cast<float*>(base + bytes_state + bytes_mt + bytes_net + bytes_allin);
    MTT_HIP(hipMemcpyAsync(w, weights, (size_t)n_flops * sizeof(float), hipMemcpyHostToDevice, stream));
    std::lock_guard<std::mutex> lock(g_pf_timer_mu);
    hipLaunchKernelGGL(k_pf_init, dim3(1), dim3(256), 0, stream, st);
    const size_t flip_n = mat > 4 * (size_t)kPfAllin4 ? mat : 4 * (size_t)kPfAllin4;
    hipLaunchKernelGGL(k_pf_flip, dim3((unsigned)((flip_n + 255) / 256)), dim3(256), 0, stream, d_M, MT, n_flops,
                       d_allin, reinterpret_cast<float*>(allin4));
    const float pot_flop = dead_pot + 3.1f;
    for (int t = 1; t <= iters; ++t) {
        // the discount weights on the host (libm powf), as the reference's
        // host loop computes them (preflop_grid_generator.cu:589-591)
        const float tf = (float)t;
        const float t15 = powf(tf, 1.5f);
        const float wp = t15 / (t15 + 1.0f), wsg = tf * tf;
        pf_timed("pf_cosolve", stream, [&] {
            hipLaunchKernelGGL(k_pf_cosolve, dim3(n_flops, 3), dim3(64), 0, stream, d_M, MT, st, net, pitch,
                               pot_flop, bf);
        });
        pf_timed("pf_tree", stream, [&] {
            hipLaunchKernelGGL(k_pf_tree, dim3(1), dim3(kPfTreeThreads), 0, stream, st, net, pitch, w, n_flops,
                               allin4, wp, wsg, stack_bb, bf, dead_pot);
        });
    }
    hipLaunchKernelGGL(k_pf_sums, dim3((12 * kPfC + 255) / 256), dim3(256), 0, stream, st, d_sums);
    MTT_HIP(hipGetLastError());
    MTT_HIP(hipStreamSynchronize(stream));
    return MTT_OK;
}

}  // namespace

extern "C" {

int mtt_gen_preflop_matrix(float* d_eq, int32_t* d_wins, int32_t* d_ties, int samples, hipStream_t stream) {
    MTT_ARG(d_eq, "mtt_gen_preflop_matrix: null output");
    MTT_ARG(samples > 0, "mtt_gen_preflop_matrix: samples %d must be positive", samples);
    hipLaunchKernelGGL(k_preflop_matrix, dim3((169 * 169 * kMcSplit + 255) / 256), dim3(256), 0, stream, samples,
                       d_eq, d_wins, d_ties);
    MTT_HIP(hipGetLastError());
    return MTT_OK;
}

int mtt_class_cards(int cls, uint8_t* c1, uint8_t* c2) {
    MTT_ARG(cls >= 0 && cls < 169 && c1 && c2, "mtt_class_cards: bad arguments");
    int a, b;
    class_cards(cls, a, b);
    *c1 = (uint8_t)a;
    *c2 = (uint8_t)b;
    return MTT_OK;
}

int mtt_flop_class_matrices(const uint8_t* flops, int n_flops, float* d_class_matrices, hipStream_t stream) {
    MTT_ARG(n_flops >= 0 && n_flops <= kPfMaxFlops, "mtt_flop_class_matrices: n_flops %d out of 0..%d", n_flops,
            kPfMaxFlops);
    if (!n_flops) return MTT_OK;
    MTT_ARG(flops && d_class_matrices, "mtt_flop_class_matrices: null pointer");
    for (int f = 0; f < n_flops; ++f)
        MTT_ARG(pf_flop_ok(flops + 3 * f), "mtt_flop_class_matrices: flop %d is not three distinct cards < 52", f);
    return pf_class_matrices(flops, n_flops, d_class_matrices, stream);
}

int mtt_preflop_solve(const float* d_class_matrices, const float* flop_weights, int n_flops, const float* d_allin,
                      int iters, float stack_bb, float bubble_factor, float dead_pot, float* d_strategy_sums,
                      hipStream_t stream) {
    MTT_ARG(n_flops >= 1 && n_flops <= kPfMaxFlops, "mtt_preflop_solve: n_flops %d out of 1..%d", n_flops,
            kPfMaxFlops);
    MTT_ARG(iters >= 0 && iters <= 1000000, "mtt_preflop_solve: iters %d out of 0..1000000", iters);
    MTT_ARG(d_class_matrices && flop_weights && d_allin && d_strategy_sums, "mtt_preflop_solve: null pointer");
    MTT_ARG(std::isfinite(stack_bb) && stack_bb > 0.0f && std::isfinite(bubble_factor) && bubble_factor > 0.0f &&
                std::isfinite(dead_pot) && dead_pot >= 0.0f,
            "mtt_preflop_solve: need stack_bb > 0, bubble_factor > 0, dead_pot >= 0");
    for (int f = 0; f < n_flops; ++f)
        MTT_ARG(std::isfinite(flop_weights[f]) && flop_weights[f] >= 0.0f,
                "mtt_preflop_solve: flop weight %d must be finite and >= 0", f);
    return pf_solve(d_class_matrices, flop_weights, n_flops, d_allin, iters, stack_bb, bubble_factor, dead_pot,
                    d_strategy_sums, stream);
}

int mtt_preflop_cosolve(const mtt_subset_flop* subset, int n_flops, const float* d_allin, int iters, float stack_bb,
                        float bubble_factor, float dead_pot, float* d_strategy_sums, hipStream_t stream) {
    MTT_ARG(n_flops >= 1 && n_flops <= kPfMaxFlops, "mtt_preflop_cosolve: n_flops %d out of 1..%d", n_flops,
            kPfMaxFlops);
    MTT_ARG(subset && d_allin && d_strategy_sums, "mtt_preflop_cosolve: null pointer");
    std::vector<uint8_t> cards(3 * (size_t)n_flops);
    std::vector<float> w(n_flops);
    for (int f = 0; f < n_flops; ++f) {
        cards[3 * f] = subset[f].c1;
        cards[3 * f + 1] = subset[f].c2;
        cards[3 * f + 2] = subset[f].c3;
        w[f] = subset[f].weight;
    }
    PfMem M;
    MTT_HIP(hipMalloc(&M.p, (size_t)n_flops * kPfC * kPfC * sizeof(float)));
    int rc = mtt_flop_class_matrices(cards.data(), n_flops, static_cast<float*>(M.p), stream);
    if (rc != MTT_OK) return rc;
    return mtt_preflop_solve(static_cast<const float*>(M.p), w.data(), n_flops, d_allin, iters, stack_bb,
                             bubble_factor, dead_pot, d_strategy_sums, stream);
}

int mtt_preflop_timing(int enable) {
    std::lock_guard<std::mutex> lock(g_pf_timer_mu);
    g_pf_timer.collect();
    g_pf_timer.enabled = enable != 0;
    g_pf_timer.names.clear();
    g_pf_timer.ms.clear();
    g_pf_timer.count.clear();
    return MTT_OK;
}

int mtt_preflop_kernel_times(int idx, char* name, int name_cap, double* total_ms, int64_t* launches) {
    MTT_ARG(name && total_ms && launches && name_cap > 0, "mtt_preflop_kernel_times: bad arguments");
    std::lock_guard<std::mutex> lock(g_pf_timer_mu);
    g_pf_timer.collect();
    if (idx < 0 || idx >= (int)g_pf_timer.names.size()) return -1;
    snprintf(name, name_cap, "%s", g_pf_timer.names[idx].c_str());
    *total_ms = g_pf_timer.ms[idx];
    *launches = g_pf_timer.count[idx];
    return MTT_OK;
}

}  // extern "C"
