/*
 * HU CFR restatement, instantiated for REAL = double and REAL = float by
 * or_cfr.c (TEST INFRASTRUCTURE ONLY; see oracle.h).
 *
 * It follows the reference's level-synchronous sweeps as a depth-first
 * recursion over the same FlatNode tree, which visits every node once per
 * iteration and computes the same quantities:
 *   forward  (honest_multistreet_solver.cu:88-126): regret matching with the
 *            1e-7 uniform threshold, child reach = parent reach x sigma for
 *            the acting player, chance children zero combos holding the card;
 *   terminal (honest_multistreet_solver.cu:128-196) with the corrected
 *            payoffs: fold and showdown values use the full opponent range
 *            with card removal, showdown is zero-sum in chips won;
 *   backward (honest_multistreet_solver.cu:198-249): chance sums children
 *            whose card the combo does not hold, normalised by the exact
 *            count (45 turn / 44 river cards given both hands); decision
 *            value sum(sigma*v) for the actor and sum(v) for the opponent;
 *            regret update R <- (R+r)*(R+r>0 ? wp : wn), strategy sum
 *            S += ws * reach * sigma (preflop_grid_generator.cu:387-410
 *            generalised; CFR+ is wp=1, wn=0, ws=t).
 * Values of combos that conflict with the current board are 0.
 */
#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SUFFIX)

typedef struct {
    int A;        /* action stride */
    int N;        /* root-alive combos */
    REAL* R;      /* regrets [infosets][A][N] */
    REAL* S;      /* strategy sums */
    /* per-depth scratch */
    int max_depth;
    REAL** reach;   /* [depth][2*N] */
    REAL** vals;    /* [depth][A*2*N] */
    REAL** sig;     /* [depth][A*N] */
    REAL** acc;     /* [depth][2*N] */
} FN(cfr_buf);

/* --------------------------------------------------------- terminals -- */
/* (nb r)(i) and, if eq is not NULL, (E r)(i) on the current board, for
 * live heroes; E = 1 win, 1/2 tie averaged over cards still to come. */
static void FN(terminal_sums)(const or_hu* h, const uint8_t* dealt, int nd, const REAL* r,
                              double* eq, double* mass) {
    int N = h->N;
    uint64_t bmask = h->root_mask;
    for (int k = 0; k < nd; ++k) bmask |= 1ull << h->deck[dealt[k]];
    double tot = 0, card[52];
    memset(card, 0, sizeof(card));
    for (int j = 0; j < N; ++j) {
        if (h->cmask[j] & bmask) continue;
        double x = r[j];
        tot += x;
        card[h->c1[j]] += x;
        card[h->c2[j]] += x;
    }
    for (int i = 0; i < N; ++i) {
        if (eq) eq[i] = 0;
        if (h->cmask[i] & bmask) { mass[i] = 0; continue; }
        mass[i] = tot - card[h->c1[i]] - card[h->c2[i]] + (double)r[i];
    }
    if (!eq) return;
    int cards_left = 5 - (h->n_board + nd);
    if (cards_left == 0) {
        const or_board* b = or_get_board(h, dealt, nd);
        or_sorted_showdown(h, b, r, bmask, eq, mass);
    } else if (cards_left == 1) {
        /* average of the river showdowns over the 44 rivers avoiding both hands */
        double* tmp = (double*)malloc(sizeof(double) * N);
        double* tmpm = (double*)malloc(sizeof(double) * N);
        REAL* rr = (REAL*)malloc(sizeof(REAL) * N);
        uint8_t d2[3];
        for (int k = 0; k < nd; ++k) d2[k] = dealt[k];
        for (int di = 0; di < h->deck_n; ++di) {
            int c = h->deck[di];
            if (bmask & (1ull << c)) continue;
            uint64_t cb = 1ull << c;
            for (int j = 0; j < N; ++j) rr[j] = (h->cmask[j] & cb) ? 0 : r[j];
            d2[nd] = (uint8_t)di;
            const or_board* b = or_get_board(h, d2, nd + 1);
            uint64_t bm2 = bmask | cb;
            for (int j = 0; j < N; ++j) { tmp[j] = 0; tmpm[j] = 0; }
            /* mass on the river board (needed by the sorted showdown) */
            double tot2 = 0, card2[52];
            memset(card2, 0, sizeof(card2));
            for (int j = 0; j < N; ++j) {
                if (h->cmask[j] & bm2) continue;
                tot2 += rr[j]; card2[h->c1[j]] += rr[j]; card2[h->c2[j]] += rr[j];
            }
            for (int i = 0; i < N; ++i)
                if (!(h->cmask[i] & bm2)) tmpm[i] = tot2 - card2[h->c1[i]] - card2[h->c2[i]] + (double)rr[i];
            or_sorted_showdown(h, b, rr, bm2, tmp, tmpm);
            for (int i = 0; i < N; ++i)
                if (!(h->cmask[i] & bm2)) eq[i] += tmp[i];
        }
        for (int i = 0; i < N; ++i) eq[i] /= 44.0;
        free(tmp); free(tmpm); free(rr);
    } else {
        /* flop board: dense counts E = (W + T/2)/990 */
        const or_flopcounts* fc = or_get_flopcounts(h);
        for (int i = 0; i < N; ++i) {
            if (h->cmask[i] & bmask) continue;
            double s = 0;
            const int32_t* w = fc->W + (size_t)h->gidx[i] * 1326;
            const int32_t* t = fc->T + (size_t)h->gidx[i] * 1326;
            for (int j = 0; j < N; ++j) {
                if (h->cmask[j] & bmask) continue;
                int gj = h->gidx[j];
                s += ((double)w[gj] + 0.5 * (double)t[gj]) * (double)r[j];
            }
            eq[i] = s / 990.0;
        }
    }
}

static void FN(terminal_values)(const or_hu* h, const mtt_flat_node* nd_, const uint8_t* dealt, int nd,
                                const REAL* r0, const REAL* r1, REAL* v0, REAL* v1) {
    int N = h->N;
    double* eq = (double*)malloc(sizeof(double) * N);
    double* mass = (double*)malloc(sizeof(double) * N);
    double P = nd_->pot_size;
    if (nd_->type == MTT_NODE_FOLD) {
        int w = (nd_->active_mask == 1) ? 0 : 1, l = 1 - w;
        double gain = P - nd_->committed[w], loss = -(double)nd_->committed[l];
        /* winner's value uses the loser's reach and vice versa */
        const REAL* rl = l == 0 ? r0 : r1;
        const REAL* rw = w == 0 ? r0 : r1;
        REAL* vw = w == 0 ? v0 : v1;
        REAL* vl = l == 0 ? v0 : v1;
        FN(terminal_sums)(h, dealt, nd, rl, NULL, mass);
        for (int i = 0; i < N; ++i) vw[i] = (REAL)(gain * mass[i]);
        FN(terminal_sums)(h, dealt, nd, rw, NULL, mass);
        for (int i = 0; i < N; ++i) vl[i] = (REAL)(loss * mass[i]);
    } else {
        for (int p = 0; p < 2; ++p) {
            const REAL* ro = p == 0 ? r1 : r0;
            REAL* vp = p == 0 ? v0 : v1;
            FN(terminal_sums)(h, dealt, nd, ro, eq, mass);
            double c = nd_->committed[p];
            for (int i = 0; i < N; ++i) vp[i] = (REAL)(P * eq[i] - c * mass[i]);
        }
    }
    free(eq); free(mass);
}

/* ------------------------------------------------------- CFR iteration -- */
static void FN(cfr_rec)(or_hu* h, FN(cfr_buf) * B, int node, int depth, uint8_t* dealt, int nd,
                        const REAL* r0, const REAL* r1, REAL* v0, REAL* v1, double wp, double wn, double ws) {
    const mtt_flat_node* x = &h->nodes[node];
    int N = h->N, A = B->A;
    if (x->type == MTT_NODE_FOLD || x->type == MTT_NODE_SHOWDOWN) {
        FN(terminal_values)(h, x, dealt, nd, r0, r1, v0, v1);
        return;
    }
    REAL* cr = B->reach[depth];
    if (x->type == MTT_NODE_CHANCE) {
        REAL* acc = B->acc[depth];
        memset(acc, 0, sizeof(REAL) * 2 * N);
        REAL* cv = B->vals[depth];
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
            FN(cfr_rec)(h, B, x->children_offset + k, depth + 1, dealt, nd + 1, cr, cr + N, cv, cv + N, wp, wn, ws);
            for (int i = 0; i < N; ++i) {
                if (h->cmask[i] & cb) continue;
                acc[i] += cv[i];
                acc[N + i] += cv[N + i];
            }
        }
        REAL inv = (REAL)(1.0 / (double)(nch - 4));
        for (int i = 0; i < N; ++i) { v0[i] = acc[i] * inv; v1[i] = acc[N + i] * inv; }
        return;
    }
    /* decision */
    int p = x->player, na = x->num_actions, info = x->infoset_offset;
    REAL* sg = B->sig[depth];
    REAL* Rn = B->R + (size_t)info * A * N;
    REAL* Sn = B->S + (size_t)info * A * N;
    const REAL* rp = p == 0 ? r0 : r1;
    for (int i = 0; i < N; ++i) {
        REAL pos[8], sum = 0;
        for (int a = 0; a < na; ++a) { REAL v = Rn[(size_t)a * N + i]; pos[a] = v > 0 ? v : 0; sum += pos[a]; }
        for (int a = 0; a < na; ++a) sg[(size_t)a * N + i] = sum > (REAL)1e-7 ? pos[a] / sum : (REAL)1 / (REAL)na;
    }
    REAL* vals = B->vals[depth];
    for (int a = 0; a < na; ++a) {
        for (int i = 0; i < N; ++i) {
            cr[i] = p == 0 ? r0[i] * sg[(size_t)a * N + i] : r0[i];
            cr[N + i] = p == 1 ? r1[i] * sg[(size_t)a * N + i] : r1[i];
        }
        REAL* va = vals + (size_t)a * 2 * N;
        FN(cfr_rec)(h, B, x->children_offset + a, depth + 1, dealt, nd, cr, cr + N, va, va + N, wp, wn, ws);
    }
    REAL* vp = p == 0 ? v0 : v1;
    REAL* vo = p == 0 ? v1 : v0;
    for (int i = 0; i < N; ++i) {
        REAL node_v = 0, opp = 0;
        for (int a = 0; a < na; ++a) {
            const REAL* va = vals + (size_t)a * 2 * N;
            node_v += sg[(size_t)a * N + i] * va[p * N + i];
            opp += va[(1 - p) * N + i];
        }
        vp[i] = node_v;
        vo[i] = opp;
        for (int a = 0; a < na; ++a) {
            size_t k = (size_t)a * N + i;
            REAL q = Rn[k] + ((vals + (size_t)a * 2 * N)[p * N + i] - node_v);
            Rn[k] = q * (REAL)(q > 0 ? wp : wn);
            Sn[k] += (REAL)ws * rp[i] * sg[k];
        }
    }
}

static void FN(buf_init)(FN(cfr_buf) * B, or_hu* h) {
    B->A = h->A; B->N = h->N; B->R = (REAL*)h->R; B->S = (REAL*)h->S;
    B->max_depth = h->max_depth;
    B->reach = (REAL**)malloc(sizeof(REAL*) * B->max_depth);
    B->vals = (REAL**)malloc(sizeof(REAL*) * B->max_depth);
    B->sig = (REAL**)malloc(sizeof(REAL*) * B->max_depth);
    B->acc = (REAL**)malloc(sizeof(REAL*) * B->max_depth);
    for (int d = 0; d < B->max_depth; ++d) {
        B->reach[d] = (REAL*)malloc(sizeof(REAL) * 2 * h->N);
        B->vals[d] = (REAL*)malloc(sizeof(REAL) * h->A * 2 * h->N);
        B->sig[d] = (REAL*)malloc(sizeof(REAL) * h->A * h->N);
        B->acc[d] = (REAL*)malloc(sizeof(REAL) * 2 * h->N);
    }
}

static void FN(buf_free)(FN(cfr_buf) * B) {
    for (int d = 0; d < B->max_depth; ++d) { free(B->reach[d]); free(B->vals[d]); free(B->sig[d]); free(B->acc[d]); }
    free(B->reach); free(B->vals); free(B->sig); free(B->acc);
}

/* Iterations; a chance root is split over OpenMP threads (its subtrees own
 * disjoint regrets and strategy sums, and boards are precomputed). */
static void FN(cfr_iterations)(or_hu* h, int iters, int t0, int scheme, double alpha, double beta, double gamma) {
    int N = h->N;
    const mtt_flat_node* root = &h->nodes[0];
    for (int it = 1; it <= iters; ++it) {
        double t = (double)(t0 + it);
        double wp, wn, ws;
        if (scheme == 0) { wp = 1; wn = 0; ws = t; }                  /* CFR+ linear averaging */
        else {                                                           /* DCFR(alpha, beta, gamma) */
            double ta = pow(t, alpha), tb = pow(t, beta);
            wp = ta / (ta + 1);
            wn = isinf(beta) && beta < 0 ? 0 : tb / (tb + 1);
            ws = pow(t, gamma);
        }
        if (root->type != MTT_NODE_CHANCE) {
            FN(cfr_buf) B;
            FN(buf_init)(&B, h);
            REAL* r = (REAL*)malloc(sizeof(REAL) * 4 * N);
            for (int i = 0; i < N; ++i) { r[i] = (REAL)h->range[0][i]; r[N + i] = (REAL)h->range[1][i]; }
            uint8_t dealt[4];
            FN(cfr_rec)(h, &B, 0, 0, dealt, 0, r, r + N, r + 2 * N, r + 3 * N, wp, wn, ws);
            free(r);
            FN(buf_free)(&B);
            continue;
        }
#pragma omp parallel
        {
            FN(cfr_buf) B;
            FN(buf_init)(&B, h);
            REAL* r = (REAL*)malloc(sizeof(REAL) * 4 * N);
            uint8_t dealt[4];
#pragma omp for schedule(dynamic, 1)
            for (int k = 0; k < root->num_actions; ++k) {
                uint64_t cb = 1ull << h->deck[k];
                for (int i = 0; i < N; ++i) {
                    int dead = (h->cmask[i] & cb) != 0;
                    r[i] = dead ? 0 : (REAL)h->range[0][i];
                    r[N + i] = dead ? 0 : (REAL)h->range[1][i];
                }
                dealt[0] = (uint8_t)k;
                FN(cfr_rec)(h, &B, root->children_offset + k, 1, dealt, 1, r, r + N, r + 2 * N, r + 3 * N, wp, wn, ws);
            }
            free(r);
            FN(buf_free)(&B);
        }
    }
}
