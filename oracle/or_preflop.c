/*
 * CPU oracle, part 4: the preflop pieces of the reference's preflop program
 * (src/preflop/preflop_grid_generator.cu).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP solver:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and never as the thing measured or shipped.
 *
 * Restated (paths relative to the reference repository root):
 *   or_class_to_cards    class_to_cards, preflop_grid_generator.cu:79-92: the
 *       representative hand of a 169-class (pairs suits 0/1, suited both
 *       suit 0, offsuit suits 0/1), with the rank-block walk corrected: the
 *       reference advances while base + (r1 - 1) <= cls, one block too
 *       early, so the last class of every block (A2s, K2s, ..., A2o, ...)
 *       gets both cards of rank r1 - 1 (the same card when suited) and 32s /
 *       32o never leave the loop; here it advances while cls >= base + r1.
 *   or_preflop_matrix    kernel_gen_preflop_matrix, :95-127: for every class
 *       pair, the reference's Monte-Carlo all-in equity, sample for sample:
 *       villain cards that collide with the hero's moved to suits 2 / 3; the
 *       48 cards left in ascending order; a 32-bit LCG
 *       rng = rng * 1664525 + 1013904223 seeded with
 *       (h * 169 + v) * 2654435761 + 0x9E3779B9; board card k taken as
 *       deck[k + (rng >> 16) % (48 - k)] with no removal (so a board can
 *       repeat a card, and the evaluator then counts it twice: or_eval7 is
 *       the reference's histogram evaluator, which does); eq = (wins +
 *       0.5 ties) / samples; the diagonal is 0.5.
 *   or_preflop_exact     the same board-sampling process enumerated: every
 *       sequence of the five picks (k + 0..47-k), equally likely, giving the
 *       expectation and variance of one sample's score (1 win, 1/2 tie).
 * Pins: the representative hands map back to their classes through
 * or_canonical_class (pinned by closed forms in the tests of or_eval.c);
 * the Monte-Carlo estimate over many samples lies within a few standard
 * errors of or_preflop_exact, which fixes the sampling distribution (deck,
 * collision rule, pick rule, evaluator).  The particular sample stream (LCG
 * constants, seed, shift) follows the reference's source literally; no
 * output of the reference pins it ("parity unpinned" for the exact stream,
 * pinned in distribution).
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

void or_class_to_cards(int cls, uint8_t* h1, uint8_t* h2) {
    if (cls < 13) {
        int r = 12 - cls;
        *h1 = (uint8_t)(r * 4);
        *h2 = (uint8_t)(r * 4 + 1);
        return;
    }
    int base = cls < 91 ? 13 : 91, r1 = 12;
    while (cls >= base + r1) { base += r1; r1--; }   /* block of hi rank r1: r1 classes */
    int r2 = r1 - 1 - (cls - base);
    *h1 = (uint8_t)(r1 * 4);
    *h2 = (uint8_t)(r2 * 4 + (cls < 91 ? 0 : 1));
}

/* Hands and the 48-card deck of one class matchup (collision rule of
 * preflop_grid_generator.cu:103-106). */
static int matchup(int h, int v, uint8_t hh[2], uint8_t vv[2], uint8_t deck[48]) {
    or_class_to_cards(h, &hh[0], &hh[1]);
    or_class_to_cards(v, &vv[0], &vv[1]);
    uint8_t v1 = vv[0], v2 = vv[1];
    if (hh[0] == v1 || hh[1] == v1) v1 = (uint8_t)((v1 / 4) * 4 + 2);
    if (hh[0] == vv[1] || hh[1] == vv[1]) v2 = (uint8_t)((v2 / 4) * 4 + 3);
    vv[0] = v1;
    vv[1] = v2;
    uint64_t dead = (1ull << hh[0]) | (1ull << hh[1]) | (1ull << vv[0]) | (1ull << vv[1]);
    int n = 0;
    for (int c = 0; c < 52; ++c)
        if (!(dead & (1ull << c))) deck[n++] = (uint8_t)c;
    return n;
}

/* One off-diagonal cell: wins and ties of the hero class over `samples`
 * sampled boards. */
void or_preflop_cell(int h, int v, int samples, int32_t* wins, int32_t* ties) {
    uint8_t hh[2], vv[2], deck[48];
    int nd = matchup(h, v, hh, vv, deck);
    uint32_t rng = (uint32_t)(h * 169 + v) * 2654435761u + 0x9E3779B9u;
    int w = 0, t = 0;
    uint8_t h7[7] = {hh[0], hh[1], 0, 0, 0, 0, 0}, v7[7] = {vv[0], vv[1], 0, 0, 0, 0, 0};
    for (int s = 0; s < samples; ++s) {
        for (int k = 0; k < 5; ++k) {
            rng = rng * 1664525u + 1013904223u;
            int pick = k + (int)((rng >> 16) % (uint32_t)(nd - k));
            h7[2 + k] = v7[2 + k] = deck[pick];
        }
        uint32_t sh = or_eval7(h7), sv = or_eval7(v7);
        if (sh > sv) w++;
        else if (sh == sv) t++;
    }
    *wins = w;
    *ties = t;
}

void or_preflop_matrix(int samples, int32_t* wins, int32_t* ties, float* eq) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int h = 0; h < 169; ++h)
        for (int v = 0; v < 169; ++v) {
            const int cell = h * 169 + v;
            if (h == v) {
                wins[cell] = 0;
                ties[cell] = 0;
                eq[cell] = 0.5f;
                continue;
            }
            or_preflop_cell(h, v, samples, &wins[cell], &ties[cell]);
            eq[cell] = ((float)wins[cell] + 0.5f * (float)ties[cell]) / (float)samples;
        }
}

void or_preflop_exact(int h, int v, double* mean, double* var) {
    uint8_t hh[2], vv[2], deck[48];
    int nd = matchup(h, v, hh, vv, deck);
    double sw = 0.0, sw2 = 0.0, n = 0.0;
#pragma omp parallel for reduction(+ : sw, sw2, n) schedule(dynamic, 1)
    for (int p0 = 0; p0 < nd; ++p0) {
        uint8_t h7[7] = {hh[0], hh[1], deck[p0], 0, 0, 0, 0}, v7[7] = {vv[0], vv[1], deck[p0], 0, 0, 0, 0};
        for (int p1 = 1; p1 < nd; ++p1) {
            h7[3] = v7[3] = deck[p1];
            for (int p2 = 2; p2 < nd; ++p2) {
                h7[4] = v7[4] = deck[p2];
                for (int p3 = 3; p3 < nd; ++p3) {
                    h7[5] = v7[5] = deck[p3];
                    for (int p4 = 4; p4 < nd; ++p4) {
                        h7[6] = v7[6] = deck[p4];
                        uint32_t sh = or_eval7(h7), sv = or_eval7(v7);
                        double x = sh > sv ? 1.0 : (sh == sv ? 0.5 : 0.0);
                        sw += x;
                        sw2 += x * x;
                        n += 1.0;
                    }
                }
            }
        }
    }
    *mean = sw / n;
    *var = sw2 / n - (sw / n) * (sw / n);
}
