/*
 * CPU oracle for the MI355X CFR+/DCFR solver core.
 *
 * TEST INFRASTRUCTURE ONLY.  liboracle.so restates the reference's algorithms
 * on the CPU so that the HIP kernels can be checked against it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product never links, calls or falls back to it.
 *
 * Parity pins (what checks each restated function; details in the tests):
 *   or_eval7            C(52,7) category counts (closed form), distinct-value
 *                       counts, hand-picked quirk cases, and the committed EV
 *                       dump (mtt_ev_dumper.cu output) within 2e-6.
 *   or_canonical_*      orbit weights sum to C(52,3); class map covers 1326.
 *   or_ev_dump          the reference's own data/mtt_ev_dump_1755.bin slice.
 *   or_flop_counts      W+W^T+T = 990 on unblocked pairs, T symmetric, brute
 *                       force on sampled pairs, suit-permutation invariance.
 *   or_turn_matrix      equals the mean of river sign matrices (brute force).
 *   or_flop_matrix_entry  identical-rank hands that can never differ give
 *                       exactly 1/4 (the reference's quarter-weight tie);
 *                       sampled-grid counts checked against or_eval7.
 *   HU CFR (or_cfr.c)   river clairvoyance toy game (closed form), brute-force
 *                       pair enumeration of showdown/fold values, zero-sum
 *                       accounting, dense vs sorted showdown.
 *   best response       exploitability of toy games (closed form), >= 0.
 *   multiway showdown   O(N^4) brute force vs prefix scan.
 *   DCFR weights        telescoping closed form.
 *   ICM                 2-player closed form, symmetry, sum = prize pool,
 *                       brute-force permutation enumeration.
 */
#ifndef MTT_ORACLE_H
#define MTT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Tree node layout of the reference (honest_multistreet_solver.cu:16-21),
 * restated here byte for byte; river_idx occupies the padding at offset 14. */
#define MTT_MAX_ACTIONS 4
enum { MTT_NODE_DECISION = 0, MTT_NODE_FOLD = 1, MTT_NODE_SHOWDOWN = 2, MTT_NODE_CHANCE = 3 };
typedef struct mtt_flat_node {
    uint8_t type, player, num_actions, active_mask;
    int32_t children_offset, infoset_offset;
    int16_t turn_idx, river_idx;
    float pot_size, committed[2];
} __attribute__((aligned(16))) mtt_flat_node;

uint32_t or_eval7(const uint8_t c[7]);
void or_eval7_batch(const uint8_t* cards, int64_t n, uint32_t* out);
void or_category_counts(uint64_t out[9]);
int or_canonical_class(int c1, int c2);
int or_canonical_flops(uint8_t* out);
void or_ev_dump(int first, int count, float* out);
int or_combo_index(int a, int b);
void or_flop_counts(const uint8_t flop[3], int32_t* W, int32_t* T);
void or_turn_matrix(const uint8_t flop[3], int turn_card, float* E);
float or_flop_matrix_entry(const uint8_t flop[3], int hc1, int hc2, int vc1, int vc2);

/* heads-up CFR (or_cfr.c) */
typedef struct or_hu_s or_hu;
or_hu* or_hu_create(const mtt_flat_node* nodes, int n_nodes, int n_infosets, const uint8_t* board, int n_board,
                    const float* range0, const float* range1, int precision);
void or_hu_destroy(or_hu* h);
int or_hu_n_combos(const or_hu* h);
void or_set_threads(int n);
int or_max_threads(void);
void or_hu_iterate(or_hu* h, int iters, int scheme, double alpha, double beta, double gamma);
void or_hu_avg_strategy(const or_hu* h, float* out);
void or_hu_state(const or_hu* h, float* R_out, float* S_out);
void or_hu_exploitability(const or_hu* h, const float* strat, double pot, double* out);
void or_hu_terminal(or_hu* h, int node, const uint8_t* dealt, int nd, const double* r0, const double* r1,
                    double* v0, double* v1);

/* 4-way engine (or_multi.c).  Node layout of multiway_4way_engine.cu:20-24. */
typedef struct or_node4 {
    uint8_t type, player, num_actions, active_mask;
    int32_t children_offset, infoset_offset;
    float pot_size, committed[4];
} __attribute__((aligned(16))) or_node4;
/* Flop board tables: combos, strength order, tie groups, per-card lists. */
typedef struct or_board4 {
    uint8_t c1[1216], c2[1216];
    uint16_t rank_pos[1216], pos_combo[1216], grp_lo[1216], grp_hi[1216];
    uint16_t card_list[52][48];
    uint8_t n_card[52];
    uint8_t card_lt[1216][2], card_le[1216][2];
} or_board4;
uint32_t or_evaln(const uint8_t* c, int n);
void or_4way_board(const uint8_t flop[3], or_board4* B);
void or_4way_terminal(const or_board4* B, const or_node4* nd, const float* reach, const double* bf, double* out);
void or_4way_terminal_brute(const or_board4* B, const or_node4* nd, const float* reach, const double* bf,
                            const uint32_t* score, double* out);
void or_dcfr_step(int num_actions, int n, int stride, float* R, float* sigma, float* S, const float* cfv,
                  const float* reach, float w_pos, float w_neg, float w_strat);
void or_icm(const double* stacks, int n, const double* payouts, int k, double* eq);
void or_icm_brute(const double* stacks, int n, const double* payouts, int k, double* eq);
void or_bubble_factors(const double* stacks, int n, const double* payouts, int k, const int* in_pot, int m, double* bf);
typedef struct or_4way_s or_4way;
or_4way* or_4way_create(const or_node4* nodes, int n_nodes, int n_infosets, const uint8_t flop[3], const float* ranges,
                        const double* bf);
void or_4way_destroy(or_4way* h);
void or_4way_iterate(or_4way* h, int iters, int t0, float alpha, float beta, float gamma, int scheme);
void or_4way_state(const or_4way* h, float* R, float* S, float* root_cfv);

/* preflop pieces (or_preflop.c) */
void or_class_to_cards(int cls, uint8_t* h1, uint8_t* h2);
void or_preflop_cell(int h, int v, int samples, int32_t* wins, int32_t* ties);
void or_preflop_matrix(int samples, int32_t* wins, int32_t* ties, float* eq);
void or_preflop_exact(int h, int v, double* mean, double* var);
void or_flop_class_matrix(const uint8_t flop[3], float* M);
void or_preflop_solve(int n_flops, const float* flop_w, const float* M, const float* allin, int iters,
                      float stack_bb, float bf, float dead_pot, float* sums_out);

#ifdef __cplusplus
}
#endif
#endif
