/*
 * CPU oracle, part 3: the 4-way single-street engine, Malmuth-Harville ICM
 * and the fused DCFR step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP solver:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and never as the thing measured or shipped.
 *
 * Restated from the reference (paths relative to its repository root):
 *   or_4way_forward / or_4way_backward   multiway_4way_engine.cu:41-83, 219-257
 *   or_4way_terminals                    multiway_4way_engine.cu:85-217, with
 *       card removal against the hero's hand, split pots for ties and side
 *       pots for unequal commitments (the reference has none of these); the
 *       strength-sorted exclusive prefix scan is kept, extended with
 *       per-card prefix lists for the card removal.
 *   or_dcfr_step                         preflop_grid_generator.cu:387-410,
 *       generalised to R <- (R + r) * (R + r > 0 ? w_pos : w_neg); w_neg = 0
 *       is the reference's floor-then-scale form.
 *   or_icm_*                             absent from the reference: Malmuth-
 *       Harville finish probabilities (subset DP) and a brute-force
 *       enumeration of finishing orders that pins it.
 * Pins: or_4way_showdown_brute (enumeration of opponent hands) for the
 * terminal values; or_icm_brute plus closed forms for the ICM; the scalar
 * closed form of the DCFR recurrence for or_dcfr_step.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NP4 1216
#define NA4 1176
#define P4 4
#define A4 3

/* -------------------------------------------------------- evaluator -- */
/* The reference's score on any number of cards (3..7): the same histogram
 * logic as or_eval7 (mtt_ev_dumper.cu:53-83); the 4-way showdown ranks
 * flop hands (board + 2 hole cards = 5 cards) with it. */
uint32_t or_evaln(const uint8_t* c, int n) {
    uint32_t suit_cnt[4] = {0, 0, 0, 0}, suit_mask[4] = {0, 0, 0, 0};
    uint32_t rank_cnt[13] = {0}, rank_mask = 0;
    for (int k = 0; k < n; ++k) {
        int r = c[k] >> 2, s = c[k] & 3;
        suit_cnt[s]++;
        suit_mask[s] |= 1u << r;
        rank_cnt[r]++;
        rank_mask |= 1u << r;
    }
    for (int s = 0; s < 4; ++s) {
        if (suit_cnt[s] < 5) continue;
        int sf = -1;
        for (int r = 12; r >= 4 && sf < 0; --r) {
            uint32_t need = 0x1Fu << (r - 4);
            if ((suit_mask[s] & need) == need) sf = r;
        }
        if (sf < 0 && (suit_mask[s] & 0x100Fu) == 0x100Fu) sf = 3;
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
    int st = -1;
    for (int r = 12; r >= 4 && st < 0; --r) {
        uint32_t need = 0x1Fu << (r - 4);
        if ((rank_mask & need) == need) st = r;
    }
    if (st < 0 && (rank_mask & 0x100Fu) == 0x100Fu) st = 3;
    if (st >= 0) return (4u << 24) | (uint32_t)st;
    if (trip1 >= 0) return (3u << 24) | ((uint32_t)trip1 << 8);
    if (pair1 >= 0 && pair2 >= 0) return (2u << 24) | ((uint32_t)pair1 << 8) | ((uint32_t)pair2 << 4);
    if (pair1 >= 0) return (1u << 24) | ((uint32_t)pair1 << 12);
    return rank_mask;
}

/* ------------------------------------------------------ board tables -- */
/* Flop combos in the reference order (mtt_ev_dumper.cu:145-152): the 49
 * deck cards ascending, pairs (i<j).  Strength = or_evaln of the 5 cards;
 * positions ascend in strength (ties by combo index); groups of equal score. */
void or_4way_board(const uint8_t flop[3], or_board4* B) {
    memset(B, 0, sizeof(*B));
    uint8_t deck[49];
    int nd = 0;
    for (int c = 0; c < 52; ++c)
        if (c != flop[0] && c != flop[1] && c != flop[2]) deck[nd++] = (uint8_t)c;
    int k = 0;
    uint32_t score[NA4];
    for (int i = 0; i < 49; ++i)
        for (int j = i + 1; j < 49; ++j) {
            B->c1[k] = deck[i];
            B->c2[k] = deck[j];
            uint8_t h[5] = {flop[0], flop[1], flop[2], deck[i], deck[j]};
            score[k] = or_evaln(h, 5);
            k++;
        }
    int order[NA4];
    for (int i = 0; i < NA4; ++i) order[i] = i;
    for (int i = 1; i < NA4; ++i) { /* stable insertion sort by score */
        int x = order[i], j = i - 1;
        while (j >= 0 && score[order[j]] > score[x]) { order[j + 1] = order[j]; j--; }
        order[j + 1] = x;
    }
    for (int q = 0; q < NA4; ++q) {
        B->pos_combo[q] = (uint16_t)order[q];
        B->rank_pos[order[q]] = (uint16_t)q;
    }
    for (int q = NA4; q < NP4; ++q) { B->pos_combo[q] = (uint16_t)q; B->rank_pos[q] = (uint16_t)q; }
    for (int q = 0; q < NA4;) {
        int e = q;
        while (e < NA4 && score[order[e]] == score[order[q]]) e++;
        for (int t = q; t < e; ++t) { B->grp_lo[order[t]] = (uint16_t)q; B->grp_hi[order[t]] = (uint16_t)e; }
        q = e;
    }
    /* per-card lists in ascending position */
    for (int q = 0; q < NA4; ++q) {
        int i = order[q];
        int a = B->c1[i], b = B->c2[i];
        B->card_list[a][B->n_card[a]++] = (uint16_t)i;
        B->card_list[b][B->n_card[b]++] = (uint16_t)i;
    }
    for (int i = 0; i < NA4; ++i) {
        int cs[2] = {B->c1[i], B->c2[i]};
        for (int h = 0; h < 2; ++h) {
            int c = cs[h], lt = 0, le = 0;
            for (int t = 0; t < B->n_card[c]; ++t) {
                int j = B->card_list[c][t];
                if (B->rank_pos[j] < B->grp_lo[i]) lt++;
                if (B->rank_pos[j] < B->grp_hi[i]) le++;
            }
            B->card_lt[i][h] = (uint8_t)lt;
            B->card_le[i][h] = (uint8_t)le;
        }
    }
}

/* ------------------------------------------------------ terminal values -- */
static double pay_net(double net, double bf) { return net >= 0.0 ? net : bf * net; }

/* Pot layers of a showdown: the distinct commitments L1 < L2 < ... of the
 * players still in; layer l holds every seat's chips between L(l-1) and Ll
 * (folded seats' included) and is contested by the players still in whose
 * commitment reaches Ll.  One layer when everyone in matched. */
static int pot_layers(const or_node4* nd, double* amt, int* elig) {
    double lv[P4];
    int k = 0;
    for (int p = 0; p < P4; ++p) {
        if (!(nd->active_mask & (1 << p))) continue;
        double c = nd->committed[p];
        int dup = 0;
        for (int t = 0; t < k; ++t) if (fabs(lv[t] - c) <= 1e-6) dup = 1;
        if (!dup) lv[k++] = c;
    }
    for (int a = 1; a < k; ++a)             /* insertion sort, ascending */
        for (int b = a; b > 0 && lv[b] < lv[b - 1]; --b) { double t = lv[b]; lv[b] = lv[b - 1]; lv[b - 1] = t; }
    double prev = 0.0;
    for (int l = 0; l < k; ++l) {
        amt[l] = 0.0;
        elig[l] = 0;
        for (int q = 0; q < P4; ++q) {
            double x = (double)nd->committed[q] - prev, cap = lv[l] - prev;
            amt[l] += x <= 0.0 ? 0.0 : (x < cap ? x : cap);
            if ((nd->active_mask & (1 << q)) && nd->committed[q] >= lv[l] - 1e-6) elig[l] |= 1 << q;
        }
        prev = lv[l];
    }
    return k;
}

/* Payoff of player p (still in) at a showdown when each active opponent
 * opp[t] is weaker (d_t = 0), tied (1) or stronger (2) than p, outcome
 * s = sum d_t 3^t: per layer p is eligible for, p wins it when no eligible
 * opponent is stronger, shared with the eligible tied ones; the net result
 * is scaled by the bubble factor when negative. */
static double showdown_pay(const or_node4* nd, int p, const int* opp, int na, int s, double bpp) {
    double amt[P4];
    int elig[P4];
    int nl = pot_layers(nd, amt, elig);
    int d[3] = {0, 0, 0};
    for (int t = 0; t < na; ++t) { d[t] = s % 3; s /= 3; }
    double won = 0.0;
    for (int l = 0; l < nl; ++l) {
        if (!(elig[l] & (1 << p))) continue;
        int lost = 0, tied = 0;
        for (int t = 0; t < na; ++t) {
            if (!(elig[l] & (1 << opp[t]))) continue;
            if (d[t] == 2) lost = 1;
            if (d[t] == 1) tied++;
        }
        if (!lost) won += amt[l] / (double)(tied + 1);
    }
    return pay_net(won - (double)nd->committed[p], bpp);
}

/* Values of one terminal for all players and combos, in double.
 * reach[p*NP4 + i]; out[p*NP4 + i]. */
void or_4way_terminal(const or_board4* B, const or_node4* nd, const float* reach, const double* bf, double* out) {
    static double F[P4][NA4], G[P4][NA4], W[P4][NA4];
    for (int o = 0; o < P4; ++o) {
        const float* r = reach + (size_t)o * NP4;
        double tot = 0.0, card[52], P[NA4 + 1], Q[52][49];
        memset(card, 0, sizeof(card));
        for (int j = 0; j < NA4; ++j) { tot += r[j]; card[B->c1[j]] += r[j]; card[B->c2[j]] += r[j]; }
        P[0] = 0.0;
        for (int q = 0; q < NA4; ++q) P[q + 1] = P[q] + r[B->pos_combo[q]];
        for (int c = 0; c < 52; ++c) {
            Q[c][0] = 0.0;
            for (int t = 0; t < B->n_card[c]; ++t) Q[c][t + 1] = Q[c][t] + r[B->card_list[c][t]];
        }
        for (int i = 0; i < NA4; ++i) {
            int a = B->c1[i], b = B->c2[i];
            F[o][i] = P[B->grp_lo[i]] - Q[a][B->card_lt[i][0]] - Q[b][B->card_lt[i][1]];
            G[o][i] = P[B->grp_hi[i]] - Q[a][B->card_le[i][0]] - Q[b][B->card_le[i][1]] + r[i];
            W[o][i] = tot - card[a] - card[b] + r[i];
        }
    }
    int winner = -1;
    if (nd->type == MTT_NODE_FOLD)
        for (int p = 0; p < P4; ++p)
            if (nd->active_mask & (1 << p)) { winner = p; break; }
    for (int p = 0; p < P4; ++p) {
        int opp_act[P4], na = 0;
        for (int o = 0; o < P4; ++o)
            if (o != p && (nd->active_mask & (1 << o))) opp_act[na++] = o;
        double pay[27];
        const int ns = na == 3 ? 27 : (na == 2 ? 9 : 3);
        if (nd->type == MTT_NODE_SHOWDOWN && (nd->active_mask & (1 << p)))
            for (int s = 0; s < ns; ++s) pay[s] = showdown_pay(nd, p, opp_act, na, s, bf[p * P4 + p]);
        for (int i = 0; i < NP4; ++i) {
            double v = 0.0;
            if (i < NA4) {
                const double cp = nd->committed[p];
                if (nd->type == MTT_NODE_FOLD) {
                    double opp = 1.0;
                    for (int o = 0; o < P4; ++o) if (o != p) opp *= W[o][i];
                    const double py = p == winner ? nd->pot_size - cp : -bf[p * P4 + winner] * cp;
                    v = py * opp;
                } else if (!(nd->active_mask & (1 << p))) {
                    double opp = 1.0;
                    for (int o = 0; o < P4; ++o) if (o != p) opp *= W[o][i];
                    v = -bf[p * P4 + p] * cp * opp;
                } else {
                    double inact = 1.0;
                    for (int o = 0; o < P4; ++o)
                        if (o != p && !(nd->active_mask & (1 << o))) inact *= W[o][i];
                    /* outcome masses per active opponent: weaker, tied, stronger */
                    double m[3][3];
                    for (int t = 0; t < na; ++t) {
                        int o = opp_act[t];
                        m[t][0] = F[o][i];
                        m[t][1] = G[o][i] - F[o][i];
                        m[t][2] = W[o][i] - G[o][i];
                    }
                    double acc = 0.0;
                    for (int s = 0; s < ns; ++s) {
                        double pr = 1.0;
                        int x = s;
                        for (int t = 0; t < na; ++t) { pr *= m[t][x % 3]; x /= 3; }
                        acc += pr * pay[s];
                    }
                    v = acc * inact;
                }
            }
            out[(size_t)p * NP4 + i] = v;
        }
    }
}

/* Brute force of the same values: enumerates every tuple of opponent hands
 * with nonzero reach that avoids the hero's cards (opponents independent). */
void or_4way_terminal_brute(const or_board4* B, const or_node4* nd, const float* reach, const double* bf,
                            const uint32_t* score, double* out) {
    int nz[P4][NA4], nn[P4];
    for (int o = 0; o < P4; ++o) {
        nn[o] = 0;
        for (int j = 0; j < NA4; ++j) if (reach[(size_t)o * NP4 + j] != 0.0f) nz[o][nn[o]++] = j;
    }
    int winner = -1;
    if (nd->type == MTT_NODE_FOLD)
        for (int p = 0; p < P4; ++p)
            if (nd->active_mask & (1 << p)) { winner = p; break; }
    for (int p = 0; p < P4; ++p) {
        int opp[3], k = 0;
        for (int o = 0; o < P4; ++o) if (o != p) opp[k++] = o;
        for (int i = 0; i < NP4; ++i) {
            double v = 0.0;
            if (i < NA4) {
                const uint64_t hm = (1ull << B->c1[i]) | (1ull << B->c2[i]);
                const double cp = nd->committed[p];
                for (int x = 0; x < nn[opp[0]]; ++x) {
                    int j0 = nz[opp[0]][x];
                    if (((1ull << B->c1[j0]) | (1ull << B->c2[j0])) & hm) continue;
                    double r0 = reach[(size_t)opp[0] * NP4 + j0];
                    for (int y = 0; y < nn[opp[1]]; ++y) {
                        int j1 = nz[opp[1]][y];
                        if (((1ull << B->c1[j1]) | (1ull << B->c2[j1])) & hm) continue;
                        double r1 = reach[(size_t)opp[1] * NP4 + j1];
                        for (int z = 0; z < nn[opp[2]]; ++z) {
                            int j2 = nz[opp[2]][z];
                            if (((1ull << B->c1[j2]) | (1ull << B->c2[j2])) & hm) continue;
                            double w = r0 * r1 * reach[(size_t)opp[2] * NP4 + j2];
                            double pay;
                            if (nd->type == MTT_NODE_FOLD) {
                                pay = p == winner ? nd->pot_size - cp : -bf[p * P4 + winner] * cp;
                            } else if (!(nd->active_mask & (1 << p))) {
                                pay = -bf[p * P4 + p] * cp;
                            } else {
                                /* chips layer by layer: each pot layer goes to the strongest of the
                                 * players still in whose commitment reaches it, split on ties */
                                int js[3] = {j0, j1, j2};
                                uint32_t sc[P4];
                                sc[p] = score[i];
                                for (int t = 0; t < 3; ++t) sc[opp[t]] = score[js[t]];
                                double lv[P4];
                                int nl = 0;
                                for (int q = 0; q < P4; ++q) {
                                    if (!(nd->active_mask & (1 << q))) continue;
                                    int seen = 0;
                                    for (int u = 0; u < nl; ++u) if (fabs(lv[u] - nd->committed[q]) <= 1e-6) seen = 1;
                                    if (!seen) lv[nl++] = nd->committed[q];
                                }
                                double won = 0.0, prev = 0.0;
                                for (int round = 0; round < nl; ++round) {
                                    double lo = 1e300;             /* next level above prev */
                                    for (int u = 0; u < nl; ++u) if (lv[u] > prev + 1e-6 && lv[u] < lo) lo = lv[u];
                                    double chips = 0.0;
                                    for (int q = 0; q < P4; ++q) {
                                        double c = nd->committed[q];
                                        double hi = c < lo ? c : lo;
                                        if (hi > prev) chips += hi - prev;
                                    }
                                    uint32_t best = 0;
                                    int any = 0, nwin = 0, hero_in = 0;
                                    for (int q = 0; q < P4; ++q)
                                        if ((nd->active_mask & (1 << q)) && nd->committed[q] >= lo - 1e-6) {
                                            if (!any || sc[q] > best) best = sc[q];
                                            any = 1;
                                        }
                                    for (int q = 0; q < P4; ++q)
                                        if ((nd->active_mask & (1 << q)) && nd->committed[q] >= lo - 1e-6 && sc[q] == best) {
                                            nwin++;
                                            if (q == p) hero_in = 1;
                                        }
                                    if (hero_in) won += chips / (double)nwin;
                                    prev = lo;
                                }
                                pay = pay_net(won - cp, bf[p * P4 + p]);
                            }
                            v += w * pay;
                        }
                    }
                }
            }
            out[(size_t)p * NP4 + i] = v;
        }
    }
}

/* ------------------------------------------------------------ DCFR step -- */
/* preflop_grid_generator.cu:387-410 generalised (see header). */
void or_dcfr_step(int num_actions, int n, int stride, float* R, float* sigma, float* S, const float* cfv,
                  const float* reach, float w_pos, float w_neg, float w_strat) {
    for (int i = 0; i < n; ++i) {
        float node_v = 0.0f;
        for (int a = 0; a < num_actions; ++a) node_v += sigma[a * stride + i] * cfv[a * stride + i];
        float pos_sum = 0.0f, nr[8];
        for (int a = 0; a < num_actions; ++a) {
            const int k = a * stride + i;
            const float q = R[k] + (cfv[k] - node_v);
            nr[a] = q * (q > 0.0f ? w_pos : w_neg);
            R[k] = nr[a];
            S[k] += w_strat * reach[i] * sigma[k];
            pos_sum += nr[a] > 0.0f ? nr[a] : 0.0f;
        }
        for (int a = 0; a < num_actions; ++a) {
            const float pr = nr[a] > 0.0f ? nr[a] : 0.0f;
            sigma[a * stride + i] = pos_sum > 1e-7f ? pr / pos_sum : 1.0f / (float)num_actions;
        }
    }
}

/* ----------------------------------------------------------------- ICM -- */
/* Malmuth-Harville: the next finisher among the remaining players is drawn
 * with probability proportional to stack.  Subset DP over the players who
 * already finished (in any order).  Players with a zero stack finish after
 * every positive stack; places left when only zero stacks remain are paid
 * to them in equal shares. */
void or_icm(const double* stacks, int n, const double* payouts, int k, double* eq) {
    const int ns = 1 << n;
    double* Q = (double*)calloc((size_t)ns, sizeof(double));
    double* T = (double*)calloc((size_t)ns, sizeof(double));
    double tot = 0.0;
    for (int i = 0; i < n; ++i) tot += stacks[i];
    for (int m = 0; m < ns; ++m) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) if (m & (1 << i)) s += stacks[i];
        T[m] = tot - s;  /* stack still in play */
    }
    for (int i = 0; i < n; ++i) eq[i] = 0.0;
    Q[0] = 1.0;
    for (int m = 0; m < ns; ++m) {
        if (Q[m] == 0.0) continue;
        int placed = __builtin_popcount((unsigned)m);
        if (placed >= k) continue;
        if (T[m] > 0.0) {
            for (int i = 0; i < n; ++i) {
                if ((m & (1 << i)) || stacks[i] <= 0.0) continue;
                const double pr = Q[m] * stacks[i] / T[m];
                eq[i] += pr * payouts[placed];
                Q[m | (1 << i)] += pr;
            }
        } else {  /* only zero stacks left: share the remaining places */
            int left = n - placed;
            double rest = 0.0;
            for (int q = placed; q < k; ++q) rest += payouts[q];
            for (int i = 0; i < n; ++i)
                if (!(m & (1 << i))) eq[i] += Q[m] * rest / (double)left;
        }
    }
    free(Q);
    free(T);
}

/* Brute force over every finishing order (n <= 9): the order's probability
 * is the product of stack shares, zero stacks last. */
static void perm_rec(const double* st, int n, const double* pay, int k, int* used, int depth, double pr, double rem,
                     int* order, double* eq) {
    int pos_left = 0;
    for (int i = 0; i < n; ++i) if (!used[i] && st[i] > 0.0) pos_left++;
    if (depth == n || pos_left == 0) {
        /* remaining players (zero stacks) share the remaining places */
        int left = n - depth;
        double rest = 0.0;
        for (int q = depth; q < k; ++q) rest += pay[q];
        for (int d = 0; d < depth && d < k; ++d) eq[order[d]] += pr * pay[d];
        if (left > 0)
            for (int i = 0; i < n; ++i) if (!used[i]) eq[i] += pr * rest / (double)left;
        return;
    }
    for (int i = 0; i < n; ++i) {
        if (used[i] || st[i] <= 0.0) continue;
        used[i] = 1;
        order[depth] = i;
        perm_rec(st, n, pay, k, used, depth + 1, pr * st[i] / rem, rem - st[i], order, eq);
        used[i] = 0;
    }
}

void or_icm_brute(const double* stacks, int n, const double* payouts, int k, double* eq) {
    int used[16] = {0}, order[16];
    double tot = 0.0;
    for (int i = 0; i < n; ++i) { eq[i] = 0.0; tot += stacks[i]; }
    double pay[16] = {0};
    for (int q = 0; q < k && q < 16; ++q) pay[q] = payouts[q];
    perm_rec(stacks, n, pay, k < n ? k : n, used, 0, 1.0, tot, order, eq);
}

/* Bubble factors of the players in the pot: for p != q, with the all-in
 * amount e = min(S_p, S_q), bf[p][q] = (EQ_p - EQ_p(p loses e to q)) /
 * (EQ_p(p wins e from q) - EQ_p).  The diagonal bf[p][p] is the mean of
 * bf[p][q] over the other players in the pot (the multiway showdown loss
 * has no single opponent). */
void or_bubble_factors(const double* stacks, int n, const double* payouts, int k, const int* in_pot, int m, double* bf) {
    double base[16], eq[16], st[16];
    or_icm(stacks, n, payouts, k, base);
    for (int a = 0; a < m; ++a) {
        double sum = 0.0;
        for (int b = 0; b < m; ++b) {
            if (a == b) continue;
            const int p = in_pot[a], q = in_pot[b];
            const double e = stacks[p] < stacks[q] ? stacks[p] : stacks[q];
            memcpy(st, stacks, sizeof(double) * (size_t)n);
            st[p] -= e; st[q] += e;
            or_icm(st, n, payouts, k, eq);
            const double lose = base[p] - eq[p];
            memcpy(st, stacks, sizeof(double) * (size_t)n);
            st[p] += e; st[q] -= e;
            or_icm(st, n, payouts, k, eq);
            const double gain = eq[p] - base[p];
            bf[a * m + b] = gain > 0.0 ? lose / gain : 1.0;
            sum += bf[a * m + b];
        }
        bf[a * m + a] = m > 1 ? sum / (double)(m - 1) : 1.0;
    }
}

/* ------------------------------------------------------- 4-way CFR loop -- */
/* One flop, FP32 state in the reference layouts:
 * R/S/sigma [infoset][A4][NP4], reach/cfv [node][P4][NP4].  Forward by
 * depth (regret matching, multiway_4way_engine.cu:41-83), terminals, then
 * backward by depth with the DCFR-generalised update (:219-257). */
struct or_4way_s {
    or_board4 B;
    const or_node4* nodes;
    int n_nodes, n_info;
    int* depth;          /* node depth */
    int max_depth;
    float *R, *S, *sig, *reach, *cfv;
    double bf[16];
    double* tmp;
};

or_4way* or_4way_create(const or_node4* nodes, int n_nodes, int n_infosets, const uint8_t flop[3], const float* ranges,
                        const double* bf) {
    or_4way* h = (or_4way*)calloc(1, sizeof(or_4way));
    or_4way_board(flop, &h->B);
    h->nodes = nodes;
    h->n_nodes = n_nodes;
    h->n_info = n_infosets;
    h->depth = (int*)calloc((size_t)n_nodes, sizeof(int));
    for (int n = 0; n < n_nodes; ++n)
        if (nodes[n].type == MTT_NODE_DECISION)
            for (int a = 0; a < nodes[n].num_actions; ++a) h->depth[nodes[n].children_offset + a] = h->depth[n] + 1;
    for (int n = 0; n < n_nodes; ++n) if (h->depth[n] > h->max_depth) h->max_depth = h->depth[n];
    size_t ni = (size_t)n_infosets * A4 * NP4, nn = (size_t)n_nodes * P4 * NP4;
    h->R = (float*)calloc(ni, sizeof(float));
    h->S = (float*)calloc(ni, sizeof(float));
    h->sig = (float*)calloc(ni, sizeof(float));
    h->reach = (float*)calloc(nn, sizeof(float));
    h->cfv = (float*)calloc(nn, sizeof(float));
    h->tmp = (double*)calloc((size_t)P4 * NP4, sizeof(double));
    for (int p = 0; p < P4; ++p)
        for (int i = 0; i < NA4; ++i) h->reach[(size_t)p * NP4 + i] = ranges[(size_t)p * NP4 + i];
    memcpy(h->bf, bf, sizeof(h->bf));
    return h;
}

void or_4way_destroy(or_4way* h) {
    if (!h) return;
    free(h->depth); free(h->R); free(h->S); free(h->sig); free(h->reach); free(h->cfv); free(h->tmp);
    free(h);
}

void or_4way_iterate(or_4way* h, int iters, int t0, float alpha, float beta, float gamma, int scheme) {
    for (int it = 0; it < iters; ++it) {
        const int t = t0 + it;
        float wp, wn, ws;
        if (scheme == 0) { wp = 1.0f; wn = 0.0f; ws = (float)t; }   /* CFR+ linear averaging */
        else {                                                       /* DCFR(alpha, beta, gamma) */
            const double ta = pow((double)t, alpha), tb = beta < -1e30f ? 0.0 : pow((double)t, beta);
            wp = (float)(ta / (ta + 1.0));
            wn = beta < -1e30f ? 0.0f : (float)(tb / (tb + 1.0));
            ws = (float)pow((double)t, gamma);
        }
        /* forward */
        for (int d = 0; d <= h->max_depth; ++d)
            for (int n = 0; n < h->n_nodes; ++n) {
                const or_node4* x = &h->nodes[n];
                if (h->depth[n] != d || x->type != MTT_NODE_DECISION) continue;
                for (int i = 0; i < NP4; ++i) {
                    float pos[A4], sum = 0.0f;
                    for (int a = 0; a < A4; ++a) {
                        pos[a] = 0.0f;
                        if (a < x->num_actions && i < NA4) {
                            float r = h->R[((size_t)x->infoset_offset * A4 + a) * NP4 + i];
                            pos[a] = r > 0.0f ? r : 0.0f;
                            sum += pos[a];
                        }
                    }
                    for (int a = 0; a < A4; ++a) {
                        float s = 0.0f;
                        if (a < x->num_actions && i < NA4) s = sum > 1e-7f ? pos[a] / sum : 1.0f / (float)x->num_actions;
                        h->sig[((size_t)x->infoset_offset * A4 + a) * NP4 + i] = s;
                    }
                    for (int a = 0; a < x->num_actions; ++a) {
                        const int c = x->children_offset + a;
                        for (int p = 0; p < P4; ++p) {
                            const float pr = h->reach[((size_t)n * P4 + p) * NP4 + i];
                            h->reach[((size_t)c * P4 + p) * NP4 + i] =
                                p == x->player ? pr * h->sig[((size_t)x->infoset_offset * A4 + a) * NP4 + i] : pr;
                        }
                    }
                }
            }
        /* terminals */
        for (int n = 0; n < h->n_nodes; ++n) {
            const or_node4* x = &h->nodes[n];
            if (x->type == MTT_NODE_DECISION) continue;
            or_4way_terminal(&h->B, x, h->reach + (size_t)n * P4 * NP4, h->bf, h->tmp);
            for (size_t q = 0; q < (size_t)P4 * NP4; ++q) h->cfv[(size_t)n * P4 * NP4 + q] = (float)h->tmp[q];
        }
        /* backward */
        for (int d = h->max_depth; d >= 0; --d)
            for (int n = 0; n < h->n_nodes; ++n) {
                const or_node4* x = &h->nodes[n];
                if (h->depth[n] != d || x->type != MTT_NODE_DECISION) continue;
                const int ap = x->player;
                for (int i = 0; i < NP4; ++i) {
                    float av[A4], node_v = 0.0f;
                    for (int a = 0; a < A4; ++a) {
                        av[a] = 0.0f;
                        if (a < x->num_actions) {
                            av[a] = h->cfv[((size_t)(x->children_offset + a) * P4 + ap) * NP4 + i];
                            node_v += h->sig[((size_t)x->infoset_offset * A4 + a) * NP4 + i] * av[a];
                        }
                    }
                    h->cfv[((size_t)n * P4 + ap) * NP4 + i] = node_v;
                    for (int p = 0; p < P4; ++p) {
                        if (p == ap) continue;
                        float so = 0.0f;
                        for (int a = 0; a < x->num_actions; ++a) so += h->cfv[((size_t)(x->children_offset + a) * P4 + p) * NP4 + i];
                        h->cfv[((size_t)n * P4 + p) * NP4 + i] = so;
                    }
                    if (i < NA4) {
                        const float rp = h->reach[((size_t)n * P4 + ap) * NP4 + i];
                        for (int a = 0; a < x->num_actions; ++a) {
                            const size_t k = ((size_t)x->infoset_offset * A4 + a) * NP4 + i;
                            const float q = h->R[k] + (av[a] - node_v);
                            h->R[k] = q * (q > 0.0f ? wp : wn);
                            h->S[k] += ws * rp * h->sig[k];
                        }
                    }
                }
            }
    }
}

void or_4way_state(const or_4way* h, float* R, float* S, float* root_cfv) {
    size_t ni = (size_t)h->n_info * A4 * NP4;
    if (R) memcpy(R, h->R, ni * sizeof(float));
    if (S) memcpy(S, h->S, ni * sizeof(float));
    if (root_cfv) memcpy(root_cfv, h->cfv, (size_t)P4 * NP4 * sizeof(float));
}
