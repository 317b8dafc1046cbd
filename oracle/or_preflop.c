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
 *   or_flop_class_matrix  kernel_gen_100_flop_matrices +
 *       kernel_reduce_combo_to_class_matrices, :130-223: per flop, the mean
 *       of or_flop_matrix_entry over the member pairs of each (hero class,
 *       villain class) that avoid the flop and each other, over all 1,326
 *       combos (the reference keeps 1,176, filters eq > 0 and reads unwritten
 *       entries: its :486, :214, :146).
 *   or_preflop_solve     the DCFR loop, :588-614 over :228-421 (co-solve,
 *       closed tree, DCFR step, reach update), FP32 with the reference's
 *       operation shapes and explicit fused multiply-adds.
 * Pins: the representative hands map back to their classes through
 * or_canonical_class (pinned by closed forms in the tests of or_eval.c);
 * the Monte-Carlo estimate over many samples lies within a few standard
 * errors of or_preflop_exact, which fixes the sampling distribution (deck,
 * collision rule, pick rule, evaluator).  The particular sample stream (LCG
 * constants, seed, shift) follows the reference's source literally; no
 * output of the reference pins it ("parity unpinned" for the exact stream,
 * pinned in distribution).  or_flop_class_matrix is pinned entry by entry
 * by an independent brute force over class members of or_flop_matrix_entry
 * (itself pinned by the quarter-tie and blocked-pair cases) and by the
 * no-valid-pair rows of a three-ace flop.  or_preflop_solve: the strategy
 * sums are pinned by closed forms (sums over actions = root reach x sum t^2
 * at every node; on coin-flip data the big blind folds to an open jam from
 * the second iteration on) and the orientation of both equity matrices by a
 * game whose class index orders strength (strong classes call jams, the
 * weakest folds first).  The closed-tree payoff formulas themselves (the
 * 1.06 / 0.94 realisation, the 2.1bb investment, the pot terms) are the
 * reference's heuristics restated from its source; no output of the
 * reference exists to pin them ("parity unpinned" for those formulas).
 */
#include "oracle.h"

#include <math.h>
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

/* --------------------------------------------------- preflop co-solver -- */
/* Members of each of the 169 classes: lexicographic (c1 < c2) combos of the
 * whole deck (all 1,326: the reference keeps only the first 1,176 of them,
 * preflop_grid_generator.cu:486, which drops every combo whose first card is
 * 35 or higher). */
static uint8_t g_mem1[169][12], g_mem2[169][12];
static int g_nmem[169];
static int g_mem_ready;

static void class_members_init(void) {
    if (g_mem_ready) return;
#pragma omp critical(or_class_members)
    {
        if (!g_mem_ready) {
            memset(g_nmem, 0, sizeof(g_nmem));
            for (int a = 0; a < 52; ++a)
                for (int b = a + 1; b < 52; ++b) {
                    int c = or_canonical_class(a, b);
                    g_mem1[c][g_nmem[c]] = (uint8_t)a;
                    g_mem2[c][g_nmem[c]] = (uint8_t)b;
                    g_nmem[c]++;
                }
            g_mem_ready = 1;
        }
    }
}

/* kernel_reduce_combo_to_class_matrices, preflop_grid_generator.cu:189-223,
 * over the flop equities of kernel_gen_100_flop_matrices (:130-183, here
 * or_flop_matrix_entry): M[v * 169 + h] = mean equity of hero class h
 * against villain class v over the member pairs that avoid the flop and
 * each other (hero members outer, villain members inner, FP32 sum), 0.5
 * when there is none.  The reference instead averages the entries > 0 of a
 * matrix that holds only 1,176 combos and leaves heroes holding a flop card
 * unwritten (its :146, :214); a pair with equity exactly 0 is a valid pair
 * here. */
void or_flop_class_matrix(const uint8_t flop[3], float* M) {
    class_members_init();
    const uint64_t fm = (1ull << flop[0]) | (1ull << flop[1]) | (1ull << flop[2]);
#pragma omp parallel for schedule(dynamic, 1)
    for (int h = 0; h < 169; ++h)
        for (int v = 0; v < 169; ++v) {
            float sum = 0.0f;
            int cnt = 0;
            for (int i = 0; i < g_nmem[h]; ++i) {
                const uint64_t hm = (1ull << g_mem1[h][i]) | (1ull << g_mem2[h][i]);
                if (hm & fm) continue;
                for (int j = 0; j < g_nmem[v]; ++j) {
                    const uint64_t vm = (1ull << g_mem1[v][j]) | (1ull << g_mem2[v][j]);
                    if ((vm & fm) || (vm & hm)) continue;
                    sum += or_flop_matrix_entry(flop, g_mem1[h][i], g_mem2[h][i], g_mem1[v][j], g_mem2[v][j]);
                    cnt++;
                }
            }
            M[v * 169 + h] = cnt ? sum / (float)cnt : 0.5f;
        }
}

/* The reference's preflop DCFR loop (preflop_grid_generator.cu:588-614) over
 * its closed tree, FP32:
 *   reaches into the flop (kernel_update_dynamic_reaches, :412-421);
 *   flop co-solve (kernel_cosolve_100_board_matrices, :228-285): per class,
 *     the subset-weighted mean over flops of the net EV of taking the class
 *     to the flop against the opponent's calling range, with realisation
 *     factors 1.06 (button) / 0.94 (big blind), the flop pot dead_pot + 3.1
 *     and the 2.1bb investment, losses scaled by the bubble factor;
 *   closed tree values (kernel_build_closed_preflop_cfvs, :290-385):
 *     button root fold / min-raise / jam, big blind fold / call / 3-bet jam,
 *     button fold / call facing the jam, big blind fold / call facing an
 *     open jam, all-ins by the Monte-Carlo all-in matrix;
 *   DCFR (kernel_dcfr_step, :387-410) on the four nodes, w_pos = t^1.5 /
 *     (t^1.5 + 1), w_strat = t^2, negatives floored, strategy sums weighted
 *     by the root reach.
 * Root reach: class combo count / 1326 (6 pairs, 4 suited, 12 offsuit).
 * Outputs the strategy sums of the four nodes, sums[node][3][169]
 * (nodes: button root, big blind defence, button facing the 3-bet jam,
 * big blind facing the open jam; 2-action nodes leave row 2 zero). */
static void dcfr_node(int na, float* R, float* sg, float* S, const float* cfv, const float* reach, float wp,
                      float ws) {
    for (int c = 0; c < 169; ++c) {
        float node_v = 0.0f;
        for (int a = 0; a < na; ++a) node_v = fmaf(sg[a * 169 + c], cfv[a * 169 + c], node_v);
        float pos = 0.0f, nx[3] = {0, 0, 0};
        for (int a = 0; a < na; ++a) {
            const int k = a * 169 + c;
            const float q = R[k] + (cfv[k] - node_v);
            const float nr = wp * (q > 0.0f ? q : 0.0f);
            R[k] = nr;
            pos += nr;
            nx[a] = nr;
            S[k] = fmaf(ws * reach[c], sg[k], S[k]);
        }
        for (int a = 0; a < na; ++a) sg[a * 169 + c] = pos > 1e-7f ? nx[a] / pos : 1.0f / (float)na;
    }
}

void or_preflop_solve(int n_flops, const float* flop_w, const float* M, const float* allin, int iters,
                      float stack_bb, float bf, float dead_pot, float* sums_out) {
    enum { C = 169 };
    float R[4][3 * C], sg[4][3 * C], S[4][3 * C], cfv[4][3 * C];
    float root[C], rbc[C], rbb[C], ev_btn[C], ev_bb[C];
    memset(R, 0, sizeof(R));
    memset(S, 0, sizeof(S));
    memset(cfv, 0, sizeof(cfv));
    memset(ev_btn, 0, sizeof(ev_btn));
    memset(ev_bb, 0, sizeof(ev_bb));
    const int na[4] = {3, 3, 2, 2};
    for (int n = 0; n < 4; ++n)
        for (int k = 0; k < 3 * C; ++k) sg[n][k] = k < na[n] * C ? 1.0f / (float)na[n] : 0.0f;
    for (int c = 0; c < C; ++c) root[c] = (c < 13 ? 6.0f : (c < 91 ? 4.0f : 12.0f)) / 1326.0f;
    const float pot_flop = dead_pot + 3.1f;
    const float jam_pot = 2.0f * stack_bb + dead_pot;
    for (int t = 1; t <= iters; ++t) {
        const float tf = (float)t;
        const float t15 = powf(tf, 1.5f);
        const float wp = t15 / (t15 + 1.0f), ws = tf * tf;
        for (int c = 0; c < C; ++c) {
            rbc[c] = root[c] * sg[0][1 * C + c];
            rbb[c] = root[c] * sg[1][1 * C + c];
        }
        /* flop co-solve */
        for (int c = 0; c < C; ++c) {
            float acc_btn = 0.0f, acc_bb = 0.0f, tw = 0.0f;
            for (int f = 0; f < n_flops; ++f) {
                const float w = flop_w[f];
                tw += w;
                const float* E = M + (size_t)f * C * C;
                float s1 = 0.0f, e0 = 0.0f;
                for (int o = 0; o < C; ++o) { s1 += rbb[o]; e0 = fmaf(rbb[o], E[o * C + c], e0); }
                if (s1 > 1e-7f) e0 /= s1;
                float s0 = 0.0f, e1 = 0.0f;
                for (int o = 0; o < C; ++o) { s0 += rbc[o]; e1 = fmaf(rbc[o], 1.0f - E[c * C + o], e1); }
                if (s0 > 1e-7f) e1 /= s0;
                const float sh_btn = e0 * 1.06f * pot_flop, sh_bb = e1 * 0.94f * pot_flop;
                const float net_btn = sh_btn >= 2.1f ? sh_btn - 2.1f : -((2.1f - sh_btn) * bf);
                const float net_bb = sh_bb >= 2.1f ? sh_bb - 2.1f : -((2.1f - sh_bb) * bf);
                acc_btn = fmaf(w, net_btn, acc_btn);
                acc_bb = fmaf(w, net_bb, acc_bb);
            }
            if (tw > 0.0f) { ev_btn[c] = acc_btn / tw; ev_bb[c] = acc_bb / tw; }
        }
        /* closed tree values */
        for (int c = 0; c < C; ++c) {
            /* node 2: button facing the 3-bet jam */
            cfv[2][0 * C + c] = -2.1f * bf;
            float m3 = 0.0f, eq_call = 0.0f;
            for (int o = 0; o < C; ++o) {
                const float r = root[o] * sg[1][2 * C + o];
                m3 += r;
                eq_call = fmaf(r, allin[c * C + o], eq_call);
            }
            if (m3 > 1e-7f) eq_call /= m3;
            cfv[2][1 * C + c] = eq_call * (jam_pot - stack_bb) - (1.0f - eq_call) * bf * stack_bb;
            const float v3b = sg[2][0 * C + c] * cfv[2][0 * C + c] + sg[2][1 * C + c] * cfv[2][1 * C + c];
            /* node 1: big blind facing the min-raise */
            cfv[1][0 * C + c] = -1.0f * bf;
            cfv[1][1 * C + c] = ev_bb[c];
            float bf3 = 0.0f, bc3 = 0.0f, eq_jam = 0.0f;
            for (int o = 0; o < C; ++o) {
                const float ro = root[o] * sg[0][1 * C + o];
                bf3 = fmaf(ro, sg[2][0 * C + o], bf3);
                bc3 = fmaf(ro, sg[2][1 * C + o], bc3);
                eq_jam = fmaf(ro * sg[2][1 * C + o], 1.0f - allin[o * C + c], eq_jam);
            }
            if (bc3 > 1e-7f) eq_jam /= bc3;
            const float tot3 = bf3 + bc3 + 1e-9f;
            const float win_unc = dead_pot + 2.1f;
            const float ev_sh = eq_jam * (jam_pot - stack_bb) - (1.0f - eq_jam) * bf * stack_bb;
            cfv[1][2 * C + c] = (bf3 * win_unc + bc3 * ev_sh) / tot3;
            /* node 3: big blind facing the open jam */
            cfv[3][0 * C + c] = -1.0f * bf;
            float mj = 0.0f, eq_oj = 0.0f;
            for (int o = 0; o < C; ++o) {
                const float r = root[o] * sg[0][2 * C + o];
                mj += r;
                eq_oj = fmaf(r, 1.0f - allin[o * C + c], eq_oj);
            }
            if (mj > 1e-7f) eq_oj /= mj;
            cfv[3][1 * C + c] = eq_oj * (jam_pot - stack_bb) - (1.0f - eq_oj) * bf * stack_bb;
            /* node 0: button root */
            cfv[0][0 * C + c] = 0.0f;
            float bff = 0.0f, bfc = 0.0f, bfj = 0.0f;
            for (int o = 0; o < C; ++o) {
                bff = fmaf(root[o], sg[1][0 * C + o], bff);
                bfc = fmaf(root[o], sg[1][1 * C + o], bfc);
                bfj = fmaf(root[o], sg[1][2 * C + o], bfj);
            }
            const float tbb = bff + bfc + bfj + 1e-9f;
            cfv[0][1 * C + c] = (bff * dead_pot + bfc * ev_btn[c] + bfj * v3b) / tbb;
            float fj = 0.0f, cj = 0.0f, eq_b = 0.0f;
            for (int o = 0; o < C; ++o) {
                fj = fmaf(root[o], sg[3][0 * C + o], fj);
                cj = fmaf(root[o], sg[3][1 * C + o], cj);
                eq_b = fmaf(root[o] * sg[3][1 * C + o], allin[c * C + o], eq_b);
            }
            const float toj = fj + cj + 1e-9f;
            if (cj > 1e-7f) eq_b /= cj;
            cfv[0][2 * C + c] =
                (fj * dead_pot + cj * (eq_b * (jam_pot - stack_bb) - (1.0f - eq_b) * bf * stack_bb)) / toj;
        }
        dcfr_node(3, R[0], sg[0], S[0], cfv[0], root, wp, ws);
        dcfr_node(3, R[1], sg[1], S[1], cfv[1], root, wp, ws);
        dcfr_node(2, R[2], sg[2], S[2], cfv[2], root, wp, ws);
        dcfr_node(2, R[3], sg[3], S[3], cfv[3], root, wp, ws);
    }
    memcpy(sums_out, S, sizeof(S));
}
