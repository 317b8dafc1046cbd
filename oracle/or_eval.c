/*
 * CPU oracle, part 1: cards, the 7-card evaluator, canonical flops, the EV dump
 * and the exhaustive equity counts.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP solver:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and never as the thing measured or shipped.
 *
 * Every function restates the reference's algorithm from its source text
 * (file:line given per function, relative to the reference repository root).
 * Floating point that the reference computes on the device with FMA
 * contraction is written with explicit fmaf(); the file is compiled with
 * -ffp-contract=off so that nothing else is contracted.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- cards -- */

/* Card c = rank*4 + suit; ranks 0='2'..12='A' (scripts/optimize_subset.py:59-61). */
static inline int rank_of(int c) { return c >> 2; }
static inline int suit_of(int c) { return c & 3; }

/* Highest straight in a 13-bit rank mask: top rank 12..4, wheel = 3, none = -1
 * (mtt_ev_dumper.cu:57-61). */
static int straight_top(uint32_t m) {
    for (int r = 12; r >= 4; --r) {
        uint32_t need = 0x1Fu << (r - 4);
        if ((m & need) == need) return r;
    }
    if ((m & 0x100Fu) == 0x100Fu) return 3;
    return -1;
}

/* 7-card score with the reference's total preorder (mtt_ev_dumper.cu:53-83):
 * flush first (straight flush 8<<24|top, flush 5<<24|five rank nibbles), then
 * quads 7<<24|q<<4 (no kicker), full house 6<<24|t<<4|(second trips or pair),
 * straight 4<<24|top, trips 3<<24|t<<8 (no kickers), two pair
 * 2<<24|p1<<8|p2<<4 (no kicker), pair 1<<24|p<<12 (no kickers), and high card
 * as the raw 13-bit mask of all seven ranks. */
uint32_t or_eval7(const uint8_t c[7]) {
    uint32_t suit_cnt[4] = {0, 0, 0, 0}, suit_mask[4] = {0, 0, 0, 0};
    uint32_t rank_cnt[13] = {0}, rank_mask = 0;
    for (int k = 0; k < 7; ++k) {
        int r = rank_of(c[k]), s = suit_of(c[k]);
        suit_cnt[s]++;
        suit_mask[s] |= 1u << r;
        rank_cnt[r]++;
        rank_mask |= 1u << r;
    }
    for (int s = 0; s < 4; ++s) {
        if (suit_cnt[s] < 5) continue;
        int sf = straight_top(suit_mask[s]);
        if (sf >= 0) return (8u << 24) | (uint32_t)sf;
        uint32_t score = 5u << 24;
        int placed = 0;
        for (int r = 12; r >= 0 && placed < 5; --r)
            if (suit_mask[s] & (1u << r)) { score |= (uint32_t)r << (4 * (4 - placed)); placed++; }
        return score;
    }
    int quad = -1, trip1 = -1, trip2 = -1, pair1 = -1, pair2 = -1;
    for (int r = 12; r >= 0; --r) {
        if (rank_cnt[r] == 4) quad = r;
        else if (rank_cnt[r] == 3) { if (trip1 < 0) trip1 = r; else if (trip2 < 0) trip2 = r; }
        else if (rank_cnt[r] == 2) { if (pair1 < 0) pair1 = r; else if (pair2 < 0) pair2 = r; }
    }
    if (quad >= 0) return (7u << 24) | ((uint32_t)quad << 4);
    if (trip1 >= 0 && (trip2 >= 0 || pair1 >= 0))
        return (6u << 24) | ((uint32_t)trip1 << 4) | (uint32_t)(trip2 >= 0 ? trip2 : pair1);
    int st = straight_top(rank_mask);
    if (st >= 0) return (4u << 24) | (uint32_t)st;
    if (trip1 >= 0) return (3u << 24) | ((uint32_t)trip1 << 8);
    if (pair1 >= 0 && pair2 >= 0) return (2u << 24) | ((uint32_t)pair1 << 8) | ((uint32_t)pair2 << 4);
    if (pair1 >= 0) return (1u << 24) | ((uint32_t)pair1 << 12);
    return rank_mask;
}

void or_eval7_batch(const uint8_t* cards, int64_t n, uint32_t* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) out[i] = or_eval7(cards + 7 * i);
}

/* Category histogram over all C(52,7) hands (category = score >> 24, with the
 * high-card category 0 holding raw rank masks below 1<<24). */
void or_category_counts(uint64_t out[9]) {
    uint64_t acc[9] = {0};
#pragma omp parallel
    {
        uint64_t loc[9] = {0};
#pragma omp for schedule(dynamic, 1)
        for (int a = 0; a < 52; ++a) {
            uint8_t h[7];
            h[0] = (uint8_t)a;
            for (int b = a + 1; b < 52; ++b) { h[1] = (uint8_t)b;
            for (int c = b + 1; c < 52; ++c) { h[2] = (uint8_t)c;
            for (int d = c + 1; d < 52; ++d) { h[3] = (uint8_t)d;
            for (int e = d + 1; e < 52; ++e) { h[4] = (uint8_t)e;
            for (int f = e + 1; f < 52; ++f) { h[5] = (uint8_t)f;
            for (int g = f + 1; g < 52; ++g) { h[6] = (uint8_t)g;
                loc[or_eval7(h) >> 24]++;
            } } } } } }
        }
#pragma omp critical
        for (int k = 0; k < 9; ++k) acc[k] += loc[k];
    }
    for (int k = 0; k < 9; ++k) out[k] = acc[k];
}

/* ------------------------------------------------------ classes / flops -- */

/* 169-class index: pairs 0..12 (AA=0), suited 13..90, offsuit 91..168
 * (mtt_ev_dumper.cu:26-32). */
int or_canonical_class(int c1, int c2) {
    int r1 = rank_of(c1), s1 = suit_of(c1), r2 = rank_of(c2), s2 = suit_of(c2);
    int hi = r1 > r2 ? r1 : r2, lo = r1 < r2 ? r1 : r2;
    if (r1 == r2) return 12 - hi;
    int base = (s1 == s2) ? 13 : 91;
    for (int r = 12; r > hi; --r) base += r;
    return base + (hi - 1 - lo);
}

/* The 1,755 canonical flops with orbit weights, in the reference's order:
 * 13 trips, 312 paired, 286x5 unpaired (mtt_ev_dumper.cu:34-51).
 * out: 4 bytes per flop {c1, c2, c3, orbit_weight}. Returns the count. */
int or_canonical_flops(uint8_t* out) {
    int n = 0;
#define MC(r, s) ((uint8_t)((r) * 4 + (s)))
#define PUT(a, b, c, w) do { out[4*n] = (a); out[4*n+1] = (b); out[4*n+2] = (c); out[4*n+3] = (uint8_t)(w); n++; } while (0)
    for (int r = 12; r >= 0; --r) PUT(MC(r, 0), MC(r, 1), MC(r, 2), 4);
    for (int rp = 12; rp >= 0; --rp)
        for (int rk = 12; rk >= 0; --rk) {
            if (rp == rk) continue;
            PUT(MC(rp, 0), MC(rp, 1), MC(rk, 0), 12);
            PUT(MC(rp, 0), MC(rp, 1), MC(rk, 2), 12);
        }
    for (int r1 = 12; r1 >= 2; --r1)
        for (int r2 = r1 - 1; r2 >= 1; --r2)
            for (int r3 = r2 - 1; r3 >= 0; --r3) {
                PUT(MC(r1, 0), MC(r2, 0), MC(r3, 0), 4);
                PUT(MC(r1, 0), MC(r2, 0), MC(r3, 1), 12);
                PUT(MC(r1, 0), MC(r2, 1), MC(r3, 0), 12);
                PUT(MC(r1, 1), MC(r2, 0), MC(r3, 0), 12);
                PUT(MC(r1, 0), MC(r2, 1), MC(r3, 2), 24);
            }
#undef PUT
#undef MC
    return n;
}

/* ------------------------------------------------------------ EV dump -- */

/* Lexicographic (c1<c2) combos that avoid `dead`; returns the count
 * (mtt_ev_dumper.cu:147-152). */
static int alive_combos(uint64_t dead, uint8_t* c1s, uint8_t* c2s) {
    int n = 0;
    for (int a = 0; a < 52; ++a)
        for (int b = a + 1; b < 52; ++b) {
            uint64_t m = (1ull << a) | (1ull << b);
            if (m & dead) continue;
            c1s[n] = (uint8_t)a; c2s[n] = (uint8_t)b; n++;
        }
    return n;
}

/* Per-combo subsampled EV of one flop (mtt_ev_dumper.cu:85-116): villains
 * j = 0,4,8,..; turn deck index 0,2,..,46; river index r1+1, r1+4, .. < 47;
 * eq = (2W+T)/(2n); ev0 += r1[j]*(2eq-1), ev1 += r0[j]*(2(1-eq)-1), each a
 * device fused multiply-add.  Scores come from a per-runout table rather
 * than two evaluator calls per pair; the values are the same. */
static void ev_one_flop(const uint8_t flop[3], float* ev0, float* ev1) {
    uint64_t fmask = (1ull << flop[0]) | (1ull << flop[1]) | (1ull << flop[2]);
    uint8_t deck[49]; int nd = 0;
    for (int c = 0; c < 52; ++c) if (!(fmask & (1ull << c))) deck[nd++] = (uint8_t)c;
    uint8_t c1s[1326], c2s[1326];
    int na = alive_combos(fmask, c1s, c2s); /* 1176 */
    /* score table over the (r1, r2) grid actually visited */
    static const int NT = 47;
    uint32_t* sc = (uint32_t*)malloc(sizeof(uint32_t) * NT * NT * na);
    for (int r1 = 0; r1 < 47; r1 += 2)
        for (int r2 = r1 + 1; r2 < 47; r2 += 3)
            for (int h = 0; h < na; ++h) {
                uint8_t hand[7] = {flop[0], flop[1], flop[2], c1s[h], c2s[h], deck[r1], deck[r2]};
                sc[((size_t)r1 * NT + r2) * na + h] = or_eval7(hand);
            }
    const float reach = 1.0f / 1176.0f;
    for (int i = 0; i < na; ++i) {
        uint64_t hm = (1ull << c1s[i]) | (1ull << c2s[i]);
        float t0 = 0.0f, t1 = 0.0f;
        for (int j = 0; j < na; j += 4) {
            uint64_t vm = (1ull << c1s[j]) | (1ull << c2s[j]);
            if (hm & vm) continue;
            uint64_t dead = hm | vm;
            int wins = 0, ties = 0, total = 0;
            for (int r1 = 0; r1 < 47; r1 += 2) {
                if (dead & (1ull << deck[r1])) continue;
                for (int r2 = r1 + 1; r2 < 47; r2 += 3) {
                    if (dead & (1ull << deck[r2])) continue;
                    const uint32_t* row = sc + ((size_t)r1 * NT + r2) * na;
                    uint32_t sh = row[i], sv = row[j];
                    if (sh > sv) wins += 2; else if (sh == sv) ties += 1;
                    total += 2;
                }
            }
            if (total > 0) {
                float eq = (float)(wins + ties) / (float)total;
                t0 = fmaf(reach, fmaf(eq, 2.0f, -1.0f), t0);
                t1 = fmaf(reach, fmaf(1.0f - eq, 2.0f, -1.0f), t1);
            }
        }
        ev0[i] = t0; ev1[i] = t1;
    }
    free(sc);
}

/* EV-dump tensor float32[flop][4 scen][2 player][169 class] for flops
 * [first, first+count) (mtt_ev_dumper.cu:141-168).  Host accumulation in
 * combo order, plain multiply then add.  out has count*1352 floats and is
 * indexed from flop `first`. */
void or_ev_dump(int first, int count, float* out) {
    uint8_t flops[1755 * 4];
    or_canonical_flops(flops);
    memset(out, 0, sizeof(float) * (size_t)count * 1352);
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < count; ++k) {
        int f = first + k;
        const uint8_t* fl = flops + 4 * f;
        float ev0[1176], ev1[1176];
        ev_one_flop(fl, ev0, ev1);
        uint64_t fmask = (1ull << fl[0]) | (1ull << fl[1]) | (1ull << fl[2]);
        uint8_t c1s[1326], c2s[1326];
        int na = alive_combos(fmask, c1s, c2s);
        for (int s = 0; s < 4; ++s) {
            float* base = out + ((size_t)k * 4 + s) * 2 * 169;
            float mult = 1.0f + 0.1f * (float)s;
            for (int c = 0; c < na; ++c) {
                int cls = or_canonical_class(c1s[c], c2s[c]);
                float a = ev0[c] * mult, b = ev1[c] * mult;
                base[cls] = base[cls] + a;
                base[169 + cls] = base[169 + cls] + b;
            }
        }
    }
}

/* ----------------------------------------------- exhaustive equity counts */

int or_combo_index(int a, int b) {
    if (a > b) { int t = a; a = b; b = t; }
    /* lexicographic rank of (a,b), a<b, among the 1326 pairs */
    return a * 51 - a * (a - 1) / 2 + (b - a - 1);
}

/* Exhaustive win/tie counts for one flop over all C(45,2)=990 runouts that
 * avoid both hands (config-2 builder; the evaluator is mtt_ev_dumper.cu:53-83
 * and the runout loop follows honest_multistreet_solver.cu:71-82 extended to
 * two cards).  W, T: int32 [1326][1326] global-combo layout; W[i][j] counts
 * runouts where i beats j; blocked pairs (sharing a card with each other or
 * the flop) stay 0. */
void or_flop_counts(const uint8_t flop[3], int32_t* W, int32_t* T) {
    uint64_t fmask = (1ull << flop[0]) | (1ull << flop[1]) | (1ull << flop[2]);
    uint8_t deck[49]; int nd = 0;
    for (int c = 0; c < 52; ++c) if (!(fmask & (1ull << c))) deck[nd++] = (uint8_t)c;
    int nb = nd * (nd - 1) / 2; /* 1176 runouts */
    uint8_t gc1[1326], gc2[1326];
    alive_combos(0, gc1, gc2);
    /* scores[b][g] for every runout b and global combo g (0 if conflicting) */
    uint32_t* sc = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)nb * 1326);
    uint64_t* bm = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)nb);
    int b = 0;
    for (int x = 0; x < nd; ++x)
        for (int y = x + 1; y < nd; ++y) { bm[b] = (1ull << deck[x]) | (1ull << deck[y]); b++; }
#pragma omp parallel for schedule(static)
    for (int bb = 0; bb < nb; ++bb) {
        int x = 0, y = 0, k = 0;
        for (x = 0; x < nd; ++x) { int row = nd - 1 - x; if (bb < k + row) { y = x + 1 + (bb - k); break; } k += row; }
        for (int g = 0; g < 1326; ++g) {
            uint64_t gm = (1ull << gc1[g]) | (1ull << gc2[g]);
            if (gm & (fmask | bm[bb])) { sc[(size_t)bb * 1326 + g] = 0; continue; }
            uint8_t hand[7] = {flop[0], flop[1], flop[2], deck[x], deck[y], gc1[g], gc2[g]};
            sc[(size_t)bb * 1326 + g] = or_eval7(hand);
        }
    }
    memset(W, 0, sizeof(int32_t) * 1326 * 1326);
    memset(T, 0, sizeof(int32_t) * 1326 * 1326);
#pragma omp parallel for schedule(dynamic, 8)
    for (int i = 0; i < 1326; ++i) {
        uint64_t im = (1ull << gc1[i]) | (1ull << gc2[i]);
        if (im & fmask) continue;
        int32_t* wrow = W + (size_t)i * 1326;
        int32_t* trow = T + (size_t)i * 1326;
        for (int bb = 0; bb < nb; ++bb) {
            if (im & bm[bb]) continue;
            const uint32_t* row = sc + (size_t)bb * 1326;
            uint32_t si = row[i];
            for (int j = 0; j < 1326; ++j) {
                uint32_t sj = row[j];
                if (sj == 0) continue; /* j conflicts with flop or runout */
                uint64_t jm = (1ull << gc1[j]) | (1ull << gc2[j]);
                if (im & jm) continue;
                wrow[j] += (si > sj);
                trow[j] += (si == sj);
            }
        }
    }
    free(sc); free(bm);
}

/* River "equity" matrices of the reference layout for one 4-card board
 * (turn or river street), transposed villain-major E[j][h] over the
 * flop-alive 1216-padded combo list (honest_multistreet_solver.cu:56-86):
 * E = (W + 0.5 T)/n over all rivers that avoid both hands, n = their count.
 * `board4` = flop + turn; returns E for every (j, h), zero when blocked. */
void or_turn_matrix(const uint8_t flop[3], int turn_card, float* E /* [1216][1216] */) {
    uint64_t fmask = (1ull << flop[0]) | (1ull << flop[1]) | (1ull << flop[2]);
    uint8_t deck[49]; int nd = 0;
    for (int c = 0; c < 52; ++c) if (!(fmask & (1ull << c))) deck[nd++] = (uint8_t)c;
    uint8_t c1s[1326], c2s[1326];
    int na = alive_combos(fmask, c1s, c2s);
    uint64_t tbit = 1ull << turn_card;
    /* scores per river card and combo */
    uint32_t* sc = (uint32_t*)calloc((size_t)52 * na, sizeof(uint32_t));
    for (int r = 0; r < nd; ++r) {
        int rc = deck[r];
        if (rc == turn_card) continue;
        for (int h = 0; h < na; ++h) {
            uint8_t hand[7] = {flop[0], flop[1], flop[2], (uint8_t)turn_card, c1s[h], c2s[h], (uint8_t)rc};
            sc[(size_t)rc * na + h] = or_eval7(hand);
        }
    }
    memset(E, 0, sizeof(float) * 1216 * 1216);
#pragma omp parallel for schedule(dynamic, 8)
    for (int h = 0; h < na; ++h) {
        uint64_t hm = (1ull << c1s[h]) | (1ull << c2s[h]);
        if (hm & tbit) continue;
        for (int j = 0; j < na; ++j) {
            uint64_t vm = (1ull << c1s[j]) | (1ull << c2s[j]);
            if ((hm & vm) || (vm & tbit)) continue;
            uint64_t dead = hm | vm | tbit;
            int w = 0, t = 0, n = 0;
            for (int r = 0; r < nd; ++r) {
                int rc = deck[r];
                if (dead & (1ull << rc)) continue;
                uint32_t a = sc[(size_t)rc * na + h], v = sc[(size_t)rc * na + j];
                if (a > v) w++; else if (a == v) t++;
                n++;
            }
            E[(size_t)j * 1216 + h] = n > 0 ? ((float)w + 0.5f * (float)t) / (float)n : 0.0f;
        }
    }
    free(sc);
}

/* One entry of kernel_gen_100_flop_matrices (preflop_grid_generator.cu:
 * 130-183): the hero's deck is the 47 cards off the flop and the hero's
 * hand; turn index 0,2,..,46, river index r1+1 step 3; runouts that hold a
 * villain card are skipped; wins += 2, ties += 1, total += 2 per runout and
 * eq = (wins + 0.5*ties) / total (so a tie counts a quarter).  Returns 0
 * for blocked pairs. */
float or_flop_matrix_entry(const uint8_t flop[3], int hc1, int hc2, int vc1, int vc2) {
    const uint64_t fm = (1ull << flop[0]) | (1ull << flop[1]) | (1ull << flop[2]);
    const uint64_t hm = (1ull << hc1) | (1ull << hc2), vm = (1ull << vc1) | (1ull << vc2);
    if ((hm & vm) || (vm & fm) || (hm & fm)) return 0.0f;
    uint8_t deck[47];
    int d = 0;
    for (int c = 0; c < 52; ++c)
        if (!((hm | fm) & (1ull << c))) deck[d++] = (uint8_t)c;
    int wins = 0, ties = 0, total = 0;
    for (int r1 = 0; r1 < d; r1 += 2) {
        if (vm & (1ull << deck[r1])) continue;
        for (int r2 = r1 + 1; r2 < d; r2 += 3) {
            if (vm & (1ull << deck[r2])) continue;
            uint8_t h7[7] = {flop[0], flop[1], flop[2], (uint8_t)hc1, (uint8_t)hc2, deck[r1], deck[r2]};
            uint8_t v7[7] = {flop[0], flop[1], flop[2], (uint8_t)vc1, (uint8_t)vc2, deck[r1], deck[r2]};
            uint32_t sh = or_eval7(h7), sv = or_eval7(v7);
            if (sh > sv) wins += 2;
            else if (sh == sv) ties += 1;
            total += 2;
        }
    }
    return total > 0 ? ((float)wins + 0.5f * (float)ties) / (float)total : 0.0f;
}
