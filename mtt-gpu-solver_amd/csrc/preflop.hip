// Preflop pieces of the reference's preflop program
// (src/preflop/preflop_grid_generator.cu) on CDNA4 (gfx950):
//   - the Monte-Carlo all-in equity matrix over the 169 hand classes
//     (kernel_gen_preflop_matrix, :95-127).
//
// The Monte-Carlo boards repeat cards (a board card is drawn from the deck
// without removing the cards drawn before it), and the reference's
// evaluator counts a repeated card twice.  The suit-mask evaluator of the
// solver path merges repeated cards, so this file scores boards with a
// multiset evaluator that keeps the reference's counting: rank counts as
// 4-bit fields of one 64-bit word, suit counts as 8-bit fields.
#include "mtt_common.h"

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

// One thread per (hero class, villain class) cell: the reference's sample
// stream (32-bit LCG rng = rng * 1664525 + 1013904223 seeded with
// cell * 2654435761 + 0x9E3779B9; board card k = deck[k + (rng >> 16) %
// (48 - k)]), counting wins and ties of the hero.
__global__ __launch_bounds__(256) void k_preflop_matrix(int samples, float* __restrict__ eq,
                                                        int32_t* __restrict__ wins, int32_t* __restrict__ ties) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= 169 * 169) return;
    const int h = cell / 169, v = cell % 169;
    if (h == v) {
        eq[cell] = 0.5f;
        if (wins) wins[cell] = 0;
        if (ties) ties[cell] = 0;
        return;
    }
    int h1, h2, v1, v2;
    class_cards(h, h1, h2);
    class_cards(v, v1, v2);
    const int ov1 = v1, ov2 = v2;
    if (h1 == ov1 || h2 == ov1) v1 = (ov1 >> 2) * 4 + 2;     // collision rule, :103-106
    if (h1 == ov2 || h2 == ov2) v2 = (ov2 >> 2) * 4 + 3;
    const uint64_t dead = (1ull << h1) | (1ull << h2) | (1ull << v1) | (1ull << v2);
    // deck[i] = i-th smallest card outside `dead`: the 48 live cards as a mask,
    // the i-th one found by popcount steps
    const uint64_t live = ~dead & ((1ull << 52) - 1);
    uint32_t rng = (uint32_t)cell * 2654435761u + 0x9E3779B9u;
    int hh[7] = {h1, h2, 0, 0, 0, 0, 0}, vv[7] = {v1, v2, 0, 0, 0, 0, 0};
    int w = 0, t = 0;
    for (int s = 0; s < samples; ++s) {
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            rng = rng * 1664525u + 1013904223u;
            int pick = k + (int)((rng >> 16) % (uint32_t)(48 - k));
            uint64_t m = live;
            for (int i = 0; i < pick; ++i) m &= m - 1;        // drop the pick lowest live cards
            const int card = __builtin_ctzll(m);
            hh[2 + k] = card;
            vv[2 + k] = card;
        }
        const uint32_t sh = eval7_multiset(hh), sv = eval7_multiset(vv);
        w += sh > sv;
        t += sh == sv;
    }
    eq[cell] = ((float)w + 0.5f * (float)t) / (float)samples;
    if (wins) wins[cell] = w;
    if (ties) ties[cell] = t;
}

}  // namespace

}  // namespace mtt

using namespace mtt;

extern "C" {

int mtt_gen_preflop_matrix(float* d_eq, int32_t* d_wins, int32_t* d_ties, int samples, hipStream_t stream) {
    MTT_ARG(d_eq, "mtt_gen_preflop_matrix: null output");
    MTT_ARG(samples > 0, "mtt_gen_preflop_matrix: samples %d must be positive", samples);
    hipLaunchKernelGGL(k_preflop_matrix, dim3((169 * 169 + 255) / 256), dim3(256), 0, stream, samples, d_eq, d_wins,
                       d_ties);
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

}  // extern "C"
