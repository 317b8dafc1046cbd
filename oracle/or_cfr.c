/*
 * CPU oracle, part 2: heads-up CFR+/DCFR over the reference FlatNode tree,
 * FP64 best response and exploitability.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  The per-iteration recursion is
 * in or_cfr_impl.h, compiled for double and for float.  Showdowns are
 * evaluated with the oracle's own evaluator (or_eval7): river boards by a
 * strength-sorted scan with per-card removal (same values as the dense
 * E-matrix product of honest_multistreet_solver.cu:181-187, and checked
 * against that dense form in the tests), turn boards as the mean over the
 * 44 rivers, flop boards with the exhaustive counts of or_flop_counts.
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* ------------------------------------------------------------ problem -- */
typedef struct {
    int n;             /* live combos on this board */
    uint32_t* score;   /* [N] score of each root-alive combo (0 if dead) */
    int32_t* order;    /* live combos sorted by ascending score */
} or_board;

typedef struct {
    int32_t* W;
    int32_t* T;
} or_flopcounts;

struct or_hu_s {
    const mtt_flat_node* nodes;
    int n_nodes, n_infosets, A;
    uint8_t board[5];
    int n_board;
    uint64_t root_mask;
    int deck_n;
    uint8_t deck[52];
    int N;                 /* root-alive combos */
    uint8_t c1[1326], c2[1326];
    uint64_t cmask[1326];
    int gidx[1326];
    double* range[2];      /* [N] */
    or_board* boards[50 * 50];
    or_flopcounts* fc;
    int max_depth;
    /* solver state */
    int prec;              /* 0 double, 1 float */
    void* R;
    void* S;
    int iters_done;
};
typedef struct or_hu_s or_hu;

static int cmp_u64(const void* a, const void* b) {
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : (x > y);
}

static const or_board* or_get_board(const or_hu* hc, const uint8_t* dealt, int nd) {
    or_hu* h = (or_hu*)hc;
    int key = (nd >= 1 ? dealt[0] + 1 : 0) * 50 + (nd >= 2 ? dealt[1] + 1 : 0);
    or_board* b = h->boards[key];
    if (b) return b;
    uint8_t cards[5];
    int nb = 0;
    for (int k = 0; k < h->n_board; ++k) cards[nb++] = h->board[k];
    for (int k = 0; k < nd; ++k) cards[nb++] = h->deck[dealt[k]];
    uint64_t bm = 0;
    for (int k = 0; k < 5; ++k) bm |= 1ull << cards[k];
    b = (or_board*)calloc(1, sizeof(or_board));
    b->score = (uint32_t*)calloc(h->N, sizeof(uint32_t));
    b->order = (int32_t*)malloc(sizeof(int32_t) * h->N);
    for (int i = 0; i < h->N; ++i) {
        if (h->cmask[i] & bm) continue;
        uint8_t hand[7] = {cards[0], cards[1], cards[2], cards[3], cards[4], h->c1[i], h->c2[i]};
        b->score[i] = or_eval7(hand);
        b->order[b->n++] = i;
    }
    /* stable sort by score: pack (score, index) and sort as 64-bit keys */
    uint64_t* kv = (uint64_t*)malloc(sizeof(uint64_t) * (b->n ? b->n : 1));
    for (int a = 0; a < b->n; ++a) kv[a] = ((uint64_t)b->score[b->order[a]] << 32) | (uint32_t)b->order[a];
    qsort(kv, b->n, sizeof(uint64_t), cmp_u64);
    for (int a = 0; a < b->n; ++a) b->order[a] = (int32_t)(kv[a] & 0xFFFFFFFFu);
    free(kv);
    h->boards[key] = b;
    return b;
}

static const or_flopcounts* or_get_flopcounts(const or_hu* hc) {
    or_hu* h = (or_hu*)hc;
    if (!h->fc) {
        h->fc = (or_flopcounts*)calloc(1, sizeof(or_flopcounts));
        h->fc->W = (int32_t*)malloc(sizeof(int32_t) * 1326 * 1326);
        h->fc->T = (int32_t*)malloc(sizeof(int32_t) * 1326 * 1326);
        or_flop_counts(h->board, h->fc->W, h->fc->T);
    }
    return h->fc;
}

/* eq[i] = (E r)(i) = sum_j nb(i,j) r_j (1 win, 1/2 tie) on a river board:
 * 0.5 * (mass_i + less_i - greater_i), less/greater by a scan in strength
 * order with per-card running sums (a blocked j holds c1 or c2 of i, and
 * only j = i holds both, which has i's own score). */
#define SORTED_SHOWDOWN(NAME, RT)                                                              \
    static void NAME(const or_hu* h, const or_board* b, const RT* r, uint64_t bmask, double* eq, \
                     const double* mass) {                                                     \
        (void)bmask;                                                                           \
        double* less = (double*)malloc(sizeof(double) * h->N);                                 \
        double run = 0, card[52];                                                              \
        memset(card, 0, sizeof(card));                                                         \
        int a = 0;                                                                             \
        while (a < b->n) {                                                                     \
            int e = a;                                                                         \
            uint32_t s = b->score[b->order[a]];                                                \
            while (e < b->n && b->score[b->order[e]] == s) e++;                                \
            for (int k = a; k < e; ++k) {                                                      \
                int i = b->order[k];                                                           \
                less[i] = run - card[h->c1[i]] - card[h->c2[i]];                               \
            }                                                                                  \
            for (int k = a; k < e; ++k) {                                                      \
                int i = b->order[k];                                                           \
                double x = r[i];                                                               \
                run += x; card[h->c1[i]] += x; card[h->c2[i]] += x;                            \
            }                                                                                  \
            a = e;                                                                             \
        }                                                                                      \
        run = 0;                                                                               \
        memset(card, 0, sizeof(card));                                                         \
        int e2 = b->n;                                                                         \
        while (e2 > 0) {                                                                       \
            int s0 = e2 - 1;                                                                   \
            uint32_t s = b->score[b->order[s0]];                                               \
            while (s0 > 0 && b->score[b->order[s0 - 1]] == s) s0--;                            \
            for (int k = s0; k < e2; ++k) {                                                    \
                int i = b->order[k];                                                           \
                double gr = run - card[h->c1[i]] - card[h->c2[i]];                             \
                eq[i] = 0.5 * (mass[i] + less[i] - gr);                                        \
            }                                                                                  \
            for (int k = s0; k < e2; ++k) {                                                    \
                int i = b->order[k];                                                           \
                double x = r[i];                                                               \
                run += x; card[h->c1[i]] += x; card[h->c2[i]] += x;                            \
            }                                                                                  \
            e2 = s0;                                                                           \
        }                                                                                      \
        free(less);                                                                            \
    }

SORTED_SHOWDOWN(or_sorted_showdown_d, double)
SORTED_SHOWDOWN(or_sorted_showdown_f, float)

/* ------------------------------------------------- instantiate (double) -- */
#define REAL double
#define SUFFIX _d
#define or_sorted_showdown or_sorted_showdown_d
#include "or_cfr_impl.h"
#undef or_sorted_showdown
#undef REAL
#undef SUFFIX
/* -------------------------------------------------- instantiate (float) -- */
#define REAL float
#define SUFFIX _f
#define or_sorted_showdown or_sorted_showdown_f
#include "or_cfr_impl.h"
#undef or_sorted_showdown
#undef REAL
#undef SUFFIX

/* -------------------------------------------------------------- setup -- */
static int tree_depth(const mtt_flat_node* nodes, int node) {
    const mtt_flat_node* x = &nodes[node];
    if (x->type == MTT_NODE_FOLD || x->type == MTT_NODE_SHOWDOWN) return 1;
    int d = 0;
    for (int k = 0; k < x->num_actions; ++k) {
        int c = tree_depth(nodes, x->children_offset + k);
        if (c > d) d = c;
    }
    return d + 1;
}

or_hu* or_hu_create(const mtt_flat_node* nodes, int n_nodes, int n_infosets, const uint8_t* board, int n_board,
                    const float* range0, const float* range1, int precision) {
    or_hu* h = (or_hu*)calloc(1, sizeof(or_hu));
    h->nodes = nodes;
    h->n_nodes = n_nodes;
    h->n_infosets = n_infosets;
    h->A = MTT_MAX_ACTIONS;
    h->n_board = n_board;
    for (int k = 0; k < n_board; ++k) { h->board[k] = board[k]; h->root_mask |= 1ull << board[k]; }
    for (int c = 0; c < 52; ++c)
        if (!(h->root_mask & (1ull << c))) h->deck[h->deck_n++] = (uint8_t)c;
    int n = 0;
    for (int a = 0; a < 52; ++a)
        for (int b = a + 1; b < 52; ++b) {
            uint64_t m = (1ull << a) | (1ull << b);
            if (m & h->root_mask) continue;
            h->c1[n] = (uint8_t)a; h->c2[n] = (uint8_t)b; h->cmask[n] = m;
            h->gidx[n] = or_combo_index(a, b);
            n++;
        }
    h->N = n;
    for (int p = 0; p < 2; ++p) {
        h->range[p] = (double*)malloc(sizeof(double) * n);
        const float* src = p == 0 ? range0 : range1;
        for (int i = 0; i < n; ++i) h->range[p][i] = src[i];
    }
    h->max_depth = tree_depth(nodes, 0) + 2;
    /* every 5-card board reachable below the root, built up front so that
     * the parallel iteration only reads the cache */
    if (n_board == 5) {
        or_get_board(h, NULL, 0);
    } else if (n_board == 4) {
        for (int a = 0; a < h->deck_n; ++a) { uint8_t d[1] = {(uint8_t)a}; or_get_board(h, d, 1); }
    } else {
        for (int a = 0; a < h->deck_n; ++a)
            for (int b = 0; b < h->deck_n; ++b)
                if (a != b) { uint8_t d[2] = {(uint8_t)a, (uint8_t)b}; or_get_board(h, d, 2); }
    }
    h->prec = precision;
    size_t sz = (size_t)n_infosets * h->A * n;
    size_t el = precision == 0 ? sizeof(double) : sizeof(float);
    h->R = calloc(sz ? sz : 1, el);
    h->S = calloc(sz ? sz : 1, el);
    return h;
}

void or_hu_destroy(or_hu* h) {
    if (!h) return;
    for (int k = 0; k < 2500; ++k)
        if (h->boards[k]) { free(h->boards[k]->score); free(h->boards[k]->order); free(h->boards[k]); }
    if (h->fc) { free(h->fc->W); free(h->fc->T); free(h->fc); }
    free(h->range[0]); free(h->range[1]);
    free(h->R); free(h->S);
    free(h);
}

int or_hu_n_combos(const or_hu* h) { return h->N; }

/* scheme 0 = CFR+ (R floored, linear averaging), 1 = DCFR(alpha,beta,gamma)
 * with beta = -INFINITY meaning negatives are floored (the reference form,
 * preflop_grid_generator.cu:402). */
/* OpenMP threads used by the oracle's parallel loops (timing at a stated
 * thread count). */
void or_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

int or_max_threads(void) { return omp_get_max_threads(); }

void or_hu_iterate(or_hu* h, int iters, int scheme, double alpha, double beta, double gamma) {
    if (h->prec == 0) cfr_iterations_d(h, iters, h->iters_done, scheme, alpha, beta, gamma);
    else cfr_iterations_f(h, iters, h->iters_done, scheme, alpha, beta, gamma);
    h->iters_done += iters;
}

/* Average strategy in the reference layout [(info*A + a)*1216 + combo],
 * normalised over the node's actions, uniform where the sum is 0. */
void or_hu_avg_strategy(const or_hu* h, float* out) {
    int N = h->N, A = h->A;
    memset(out, 0, sizeof(float) * (size_t)h->n_infosets * A * 1216);
    for (int nd = 0; nd < h->n_nodes; ++nd) {
        const mtt_flat_node* x = &h->nodes[nd];
        if (x->type != MTT_NODE_DECISION) continue;
        size_t base = (size_t)x->infoset_offset * A;
        for (int i = 0; i < N; ++i) {
            double s = 0, v[8];
            for (int a = 0; a < x->num_actions; ++a) {
                size_t k = (base + a) * N + i;
                v[a] = h->prec == 0 ? ((double*)h->S)[k] : (double)((float*)h->S)[k];
                s += v[a];
            }
            for (int a = 0; a < x->num_actions; ++a)
                out[(base + a) * 1216 + i] = (float)(s > 0 ? v[a] / s : 1.0 / x->num_actions);
        }
    }
}

/* Raw regrets / strategy sums in the reference layout (as float). */
void or_hu_state(const or_hu* h, float* R_out, float* S_out) {
    int N = h->N, A = h->A;
    size_t tot = (size_t)h->n_infosets * A;
    for (size_t k = 0; k < tot; ++k)
        for (int i = 0; i < 1216; ++i) {
            float r = 0, s = 0;
            if (i < N) {
                size_t q = k * N + i;
                r = h->prec == 0 ? (float)((double*)h->R)[q] : ((float*)h->R)[q];
                s = h->prec == 0 ? (float)((double*)h->S)[q] : ((float*)h->S)[q];
            }
            if (R_out) R_out[k * 1216 + i] = r;
            if (S_out) S_out[k * 1216 + i] = s;
        }
}

/* ------------------------------------------------------ best response -- */
typedef struct {
    const or_hu* h;
    const float* strat;   /* reference layout, normalised */
    int br;               /* best-responding player; -1 = evaluate the profile */
    double** reach;       /* [depth][N] opponent reach (and own for eval) */
    double** vals;
    double** acc;
} br_ctx;

static void br_rec(br_ctx* C, int node, int depth, uint8_t* dealt, int nd, const double* r0, const double* r1,
                   double* v0, double* v1) {
    const or_hu* h = C->h;
    const mtt_flat_node* x = &h->nodes[node];
    int N = h->N;
    if (x->type == MTT_NODE_FOLD || x->type == MTT_NODE_SHOWDOWN) {
        terminal_values_d(h, x, dealt, nd, r0, r1, v0, v1);
        return;
    }
    double* cr = C->reach[depth];
    double* cv = C->vals[depth];
    double* acc = C->acc[depth];
    if (x->type == MTT_NODE_CHANCE) {
        memset(acc, 0, sizeof(double) * 2 * N);
        int nch = x->num_actions;
        for (int k = 0; k < nch; ++k) {
            int di = (x->turn_idx >= 0 && k >= x->turn_idx) ? k + 1 : k;
            uint64_t cb = 1ull << h->deck[di];
            for (int i = 0; i < N; ++i) {
                int dead = (h->cmask[i] & cb) != 0;
                cr[i] = dead ? 0 : r0[i];
                cr[N + i] = dead ? 0 : r1[i];
            }
            dealt[nd] = (uint8_t)di;
            br_rec(C, x->children_offset + k, depth + 1, dealt, nd + 1, cr, cr + N, cv, cv + N);
            for (int i = 0; i < N; ++i) {
                if (h->cmask[i] & cb) continue;
                acc[i] += cv[i];
                acc[N + i] += cv[N + i];
            }
        }
        double inv = 1.0 / (double)(nch - 4);
        for (int i = 0; i < N; ++i) { v0[i] = acc[i] * inv; v1[i] = acc[N + i] * inv; }
        return;
    }
    int p = x->player, na = x->num_actions;
    size_t base = (size_t)x->infoset_offset * h->A;
    double* vp = p == 0 ? v0 : v1;
    double* vo = p == 0 ? v1 : v0;
    for (int i = 0; i < N; ++i) { vp[i] = p == C->br ? -INFINITY : 0; vo[i] = 0; }
    for (int a = 0; a < na; ++a) {
        const float* sg = C->strat + (base + a) * 1216;
        for (int i = 0; i < N; ++i) {
            /* the best responder's own reach is not used by its values */
            double s = sg[i];
            cr[i] = (p == 0 && C->br != 0) ? r0[i] * s : r0[i];
            cr[N + i] = (p == 1 && C->br != 1) ? r1[i] * s : r1[i];
        }
        br_rec(C, x->children_offset + a, depth + 1, dealt, nd, cr, cr + N, cv, cv + N);
        const double* cp = p == 0 ? cv : cv + N;
        const double* co = p == 0 ? cv + N : cv;
        for (int i = 0; i < N; ++i) {
            if (p == C->br) { if (cp[i] > vp[i]) vp[i] = cp[i]; }
            else vp[i] += (double)sg[i] * cp[i];
            vo[i] += co[i];
        }
    }
}

/* br_rec at a chance root with its subtrees on OpenMP threads: each child's
 * values land in their own slot and are summed in card order afterwards, so
 * the result does not depend on the thread count.  Same arithmetic as the
 * chance branch of br_rec (children summed over the cards a combo does not
 * hold, scaled by 1 / (children - 4)). */
static void br_root_chance(const or_hu* h, const br_ctx* C0, double* v) {
    const mtt_flat_node* x = &h->nodes[0];
    int N = h->N, nch = x->num_actions;
    double* kid = (double*)malloc(sizeof(double) * 2 * N * (size_t)nch);
#pragma omp parallel
    {
        br_ctx C = *C0;
        C.reach = (double**)malloc(sizeof(double*) * h->max_depth);
        C.vals = (double**)malloc(sizeof(double*) * h->max_depth);
        C.acc = (double**)malloc(sizeof(double*) * h->max_depth);
        for (int d = 0; d < h->max_depth; ++d) {
            C.reach[d] = (double*)malloc(sizeof(double) * 2 * N);
            C.vals[d] = (double*)malloc(sizeof(double) * 2 * N);
            C.acc[d] = (double*)malloc(sizeof(double) * 2 * N);
        }
        uint8_t dealt[4];
#pragma omp for schedule(dynamic, 1)
        for (int k = 0; k < nch; ++k) {
            int di = (x->turn_idx >= 0 && k >= x->turn_idx) ? k + 1 : k;
            uint64_t cb = 1ull << h->deck[di];
            double* cr = C.reach[0];
            for (int i = 0; i < N; ++i) {
                int dead = (h->cmask[i] & cb) != 0;
                cr[i] = dead ? 0 : h->range[0][i];
                cr[N + i] = dead ? 0 : h->range[1][i];
            }
            dealt[0] = (uint8_t)di;
            double* kv = kid + (size_t)k * 2 * N;
            br_rec(&C, x->children_offset + k, 1, dealt, 1, cr, cr + N, kv, kv + N);
        }
        for (int d = 0; d < h->max_depth; ++d) { free(C.reach[d]); free(C.vals[d]); free(C.acc[d]); }
        free(C.reach); free(C.vals); free(C.acc);
    }
    for (int i = 0; i < 2 * N; ++i) v[i] = 0;
    for (int k = 0; k < nch; ++k) {
        int di = (x->turn_idx >= 0 && k >= x->turn_idx) ? k + 1 : k;
        uint64_t cb = 1ull << h->deck[di];
        const double* kv = kid + (size_t)k * 2 * N;
        for (int i = 0; i < N; ++i) {
            if (h->cmask[i] & cb) continue;
            v[i] += kv[i];
            v[N + i] += kv[N + i];
        }
    }
    double inv = 1.0 / (double)(nch - 4);
    for (int i = 0; i < 2 * N; ++i) v[i] *= inv;
    free(kid);
}

/* Values at the root for a profile `strat` (reference layout, normalised):
 * out[0..1] = best-response values BR0(strat1), BR1(strat0);
 * out[2..3] = the profile's own values; out[4] = pair mass M (sum over
 * non-conflicting pairs of range0*range1); out[5] = exploitability in % of
 * the starting pot, 100*(BR0 + BR1 - pot*M) / (2*M*pot). */
void or_hu_exploitability(const or_hu* h, const float* strat, double pot, double* out) {
    int N = h->N;
    br_ctx C;
    C.h = h;
    C.strat = strat;
    C.reach = (double**)malloc(sizeof(double*) * h->max_depth);
    C.vals = (double**)malloc(sizeof(double*) * h->max_depth);
    C.acc = (double**)malloc(sizeof(double*) * h->max_depth);
    for (int d = 0; d < h->max_depth; ++d) {
        C.reach[d] = (double*)malloc(sizeof(double) * 2 * N);
        C.vals[d] = (double*)malloc(sizeof(double) * 2 * N);
        C.acc[d] = (double*)malloc(sizeof(double) * 2 * N);
    }
    double* v = (double*)malloc(sizeof(double) * 2 * N);
    uint8_t dealt[4];
    double res[3][2];
    const mtt_flat_node* root = &h->nodes[0];
    for (int mode = 0; mode < 3; ++mode) {
        C.br = mode == 2 ? -1 : mode;
        if (root->type == MTT_NODE_CHANCE) br_root_chance(h, &C, v);
        else br_rec(&C, 0, 0, dealt, 0, h->range[0], h->range[1], v, v + N);
        for (int p = 0; p < 2; ++p) {
            double s = 0;
            for (int i = 0; i < N; ++i) s += h->range[p][i] * v[p * N + i];
            res[mode][p] = s;
        }
    }
    double M = 0, tot1 = 0, card1[52];
    memset(card1, 0, sizeof(card1));
    for (int j = 0; j < N; ++j) { tot1 += h->range[1][j]; card1[h->c1[j]] += h->range[1][j]; card1[h->c2[j]] += h->range[1][j]; }
    for (int i = 0; i < N; ++i)
        M += h->range[0][i] * (tot1 - card1[h->c1[i]] - card1[h->c2[i]] + h->range[1][i]);
    out[0] = res[0][0];
    out[1] = res[1][1];
    out[2] = res[2][0];
    out[3] = res[2][1];
    out[4] = M;
    out[5] = 100.0 * (out[0] + out[1] - pot * M) / (2.0 * M * pot);
    for (int d = 0; d < h->max_depth; ++d) { free(C.reach[d]); free(C.vals[d]); free(C.acc[d]); }
    free(C.reach); free(C.vals); free(C.acc); free(v);
}

/* Terminal values at one node for given reaches (reference layout, 1216),
 * for tests of the showdown/fold math.  dealt: deck indices of the cards
 * dealt below the root. */
void or_hu_terminal(or_hu* h, int node, const uint8_t* dealt, int nd, const double* r0, const double* r1,
                    double* v0, double* v1) {
    int N = h->N;
    double* a = (double*)malloc(sizeof(double) * 4 * N);
    for (int i = 0; i < N; ++i) { a[i] = r0[i]; a[N + i] = r1[i]; }
    terminal_values_d(h, &h->nodes[node], dealt, nd, a, a + N, a + 2 * N, a + 3 * N);
    for (int i = 0; i < 1216; ++i) {
        v0[i] = i < N ? a[2 * N + i] : 0;
        v1[i] = i < N ? a[3 * N + i] : 0;
    }
    free(a);
}
