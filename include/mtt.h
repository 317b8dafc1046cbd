/*
 * mtt.h - C ABI of the MI355X-native CFR+/DCFR solver core (libmtt.so).
 *
 * The reference (Azerus96/MTT-GPU-Solver) ships CUDA __global__ kernels with
 * raw device pointers and standalone programs; it has no library API.  Each
 * entry point below names the reference kernel or program it replaces
 * (file:line, relative to the reference repository root) and keeps that
 * kernel's argument meaning, order and memory layouts, adding an explicit
 * hipStream_t.  Caller-owned device pointers; every function returns an
 * int status (MTT_OK = 0) instead of exiting the process, and
 * mtt_last_error() gives the message of the last failure on this thread.
 *
 * Layouts kept from the reference:
 *   card        c = rank*4 + suit, rank 0='2'..12='A'
 *   combos      lexicographic (c1<c2) combos avoiding the flop: 1176 alive,
 *               padded to 1216 (mtt_ev_dumper.cu:20-21,149-152)
 *   per-node    state[(node*P + p)*1216 + combo], P = 2 (HU) or 4 (4-way)
 *   per-action  state[(infoset*MTT_MAX_ACTIONS + a)*1216 + combo]
 *   EV dump     float32[flop][4 scenario][2 player][169 class], no header
 */
#ifndef MTT_H
#define MTT_H

#include <stddef.h>
#include <stdint.h>

#ifndef MTT_NO_HIP_TYPES
#include <hip/hip_runtime_api.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

#ifndef MTT_NO_HIP_TYPES
typedef hipStream_t mtt_stream_t;
#else
typedef void* mtt_stream_t;
#endif

/* ------------------------------------------------------------ status -- */
enum {
    MTT_OK = 0,
    MTT_ERR_ARG = 1,     /* bad argument or size */
    MTT_ERR_OOM = 2,     /* device or host allocation failed */
    MTT_ERR_HIP = 3,     /* HIP runtime error */
    MTT_ERR_TREE = 4,    /* invalid tree (offsets, action counts, types) */
    MTT_ERR_STATE = 5,   /* call not valid in the handle's current state */
    MTT_ERR_IO = 6       /* file could not be written */
};
const char* mtt_last_error(void);
const char* mtt_version(void);

/* ------------------------------------------------------------ layouts -- */
#define MTT_N_ALIVE_COMBOS 1176
#define MTT_N_PADDED_COMBOS 1216
#define MTT_N_GLOBAL_COMBOS 1326
#define MTT_N_CANONICAL_FLOPS 1755
#define MTT_N_CLASSES 169
/* Action stride of per-action state.  The reference uses 3
 * (honest_multistreet_solver.cu:12); the bet menus here need 4. */
#define MTT_MAX_ACTIONS 4

/* honest_multistreet_solver.cu:14 / multiway_4way_engine.cu:18 */
enum { MTT_NODE_DECISION = 0, MTT_NODE_FOLD = 1, MTT_NODE_SHOWDOWN = 2, MTT_NODE_CHANCE = 3 };

/* HU tree node, honest_multistreet_solver.cu:16-21 (32 bytes, same offsets).
 * Chance nodes deal from the deck of cards not on the root board, in order:
 * turn_idx < 0 deals every deck card (num_actions children), turn_idx >= 0
 * is the deck index of the turn already dealt, which is skipped.  Showdown
 * nodes keep the reference meaning of turn_idx: the index of their equity
 * matrix.  river_idx (in the reference's padding at offset 14) records the
 * cards dealt below the root on the node's path: 0 none, d1+1 for one deck
 * index d1, (d1+1) + 64*(d2+1) for two; combos holding a dealt card are
 * dead at the node.  Decision, fold and chance nodes set turn_idx to d1 (or
 * -1).  Fold: active_mask has the winner's bit (0b01 = player 0 wins). */
typedef struct mtt_flat_node {
    uint8_t type, player, num_actions, active_mask;
    int32_t children_offset, infoset_offset;
    int16_t turn_idx, river_idx;
    float pot_size, committed[2];
} __attribute__((aligned(16))) mtt_flat_node;

/* 4-way tree node, multiway_4way_engine.cu:20-24 (32 bytes, same offsets). */
typedef struct mtt_flat_node4 {
    uint8_t type, player, num_actions, active_mask;
    int32_t children_offset, infoset_offset;
    float pot_size, committed[4];
} __attribute__((aligned(16))) mtt_flat_node4;

/* mtt_ev_dumper.cu:24 (16 bytes). */
typedef struct mtt_alive_combo {
    uint8_t c1, c2;
    uint64_t mask;
} __attribute__((aligned(8))) mtt_alive_combo;

/* ---------------------------------------------- cards and evaluator -- */
/* fast_7card_score, mtt_ev_dumper.cu:53-83 (host side). */
uint32_t mtt_eval7(const uint8_t cards[7]);
/* Host batch of the same evaluator: n hands of 7 cards. */
int mtt_eval7_host_batch(const uint8_t* cards, int64_t n, uint32_t* out);
/* Device batch: d_cards n*7 bytes, d_out n scores. */
int mtt_eval7_batch(const uint8_t* d_cards, int64_t n, uint32_t* d_out, mtt_stream_t stream);
/* get_canonical_class, mtt_ev_dumper.cu:26-32; -1 (with mtt_last_error)
 * unless c1, c2 are two distinct cards below 52. */
int mtt_canonical_class(int c1, int c2);
/* generate_all_1755_flops, mtt_ev_dumper.cu:34-51: out 1755*4 bytes
 * {c1,c2,c3,orbit_weight}; returns the flop count. */
int mtt_canonical_flops(uint8_t* out);
/* Lexicographic index of combo {a,b} among 1326. */
int mtt_combo_index(int a, int b);

/* ------------------------------------------------- equity builders -- */
enum { MTT_LAYOUT_GLOBAL1326 = 0, MTT_LAYOUT_ALIVE1216 = 1 };
/* Exhaustive win/tie counts for flops (new, config-2 builder; evaluator
 * mtt_ev_dumper.cu:53-83).  For each flop f and non-conflicting pair (i,j):
 * W[i][j] = runouts (of the C(45,2)=990 avoiding both hands) where i beats
 * j, T[i][j] = ties; blocked pairs are 0.  d_W, d_T: int32 n_flops*S*S with
 * S = 1326 (global combo index) or 1216 (flop-alive index).  flops: host
 * array of n_flops*3 cards.  Bit-exact against the reference evaluator. */
int mtt_flop_counts(const uint8_t* flops, int n_flops, int32_t* d_W, int32_t* d_T,
                    int layout, mtt_stream_t stream);
/* No counterpart in the reference.  mtt_flop_counts without any device
 * allocation, for callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc): same outputs,
 * at most 16 flops per call (MTT_ERR_ARG beyond), workspace in a static
 * 48,294,400-byte device array per GPU.  Asynchronous on `stream`; calls that
 * overlap on one GPU race on that workspace. */
int mtt_flop_counts_noalloc(const uint8_t* flops, int n_flops, int32_t* d_W, int32_t* d_T,
                            int layout, mtt_stream_t stream);
/* HIP-event timing of the count builder's kernels (fc_zero, fc_tables,
 * fc_codes, fc_pairs) on the stream they run on: enable (1) or disable
 * (0) and reset; then read entry idx (total ms, launches), -1 past the end. */
int mtt_flop_counts_timing(int enable);
int mtt_flop_counts_kernel_times(int idx, char* name, int name_cap, double* total_ms, int64_t* launches);

/* kernel_evaluate_flop_ev, mtt_ev_dumper.cu:85-116: subsampled per-combo EV
 * of one flop against reach_p1/reach_p0 (1216 floats).  d_combos 1216
 * entries, d_deck 49 cards, d_flop 3 cards, outputs 1216 floats (1176 used). */
int mtt_evaluate_flop_ev(const mtt_alive_combo* d_combos, const uint8_t* d_deck,
                         const uint8_t* d_flop, const float* d_reach_p0,
                         const float* d_reach_p1, float* d_out_ev_p0, float* d_out_ev_p1,
                         mtt_stream_t stream);
/* EV-dump tensor for canonical flops [first, first+count) into host memory
 * out (count*1352 floats), computed on the GPU with the uniform 1/1176
 * reaches and host-identical class/scenario accumulation
 * (mtt_ev_dumper.cu:141-168). */
int mtt_ev_dump_tensor(int first, int count, float* out);
/* bin/ev_dumper [--limit N] (mtt_ev_dumper.cu:118-176): writes the tensor of
 * the first `limit` flops to `path` (the reference always writes
 * "mtt_ev_dump_1755.bin" in the working directory; NULL means that). */
int mtt_ev_dump(int limit, const char* path);

/* kernel_gen_preflop_matrix, preflop_grid_generator.cu:95-127: for every
 * (hero class, villain class) of the 169, the all-in equity of their
 * representative hands over `samples` Monte-Carlo boards (3000 in the
 * reference), with the reference's sample stream and its repeated board
 * cards; d_eq[169*169] = (wins + ties/2) / samples, 0.5 on the diagonal;
 * d_wins / d_ties (may be NULL) receive the integer counts.  Representative
 * hands as mtt_class_cards.  Asynchronous on `stream`. */
int mtt_gen_preflop_matrix(float* d_eq, int32_t* d_wins, int32_t* d_ties, int samples, mtt_stream_t stream);
/* class_to_cards, preflop_grid_generator.cu:79-92: the representative hand
 * of class cls (pairs suits 0/1, suited both suit 0, offsuit suits 0/1),
 * with the rank-block walk corrected (the reference's gives the last class
 * of each rank block two cards of one rank and never ends for 32s/32o). */
int mtt_class_cards(int cls, uint8_t* c1, uint8_t* c2);

/* kernel_gen_49_turn_matrices, honest_multistreet_solver.cu:56-86:
 * d_out[49][1216][1216] FP32, E[t][j][h] = (W + T/2)/n over the n rivers
 * avoiding both hands and the turn; blocked and padded entries 0. */
int mtt_gen_turn_matrices(const mtt_alive_combo* d_combos, const uint8_t* d_deck,
                          const uint8_t* d_flop, float* d_out, mtt_stream_t stream);
/* River equity matrices of the same layout for boards (turn t, river r):
 * d_out[n_boards][1216][1216], E in {0, 1/2, 1}; boards given as host pairs
 * of deck indices. */
int mtt_gen_river_matrices(const mtt_alive_combo* d_combos, const uint8_t* d_deck,
                           const uint8_t* d_flop, const int16_t* boards, int n_boards,
                           float* d_out, mtt_stream_t stream);

/* kernel_gen_100_flop_matrices, preflop_grid_generator.cu:130-183: for
 * each flop f, hero h and villain j among the caller's first 1176 combos,
 * eq = (2W + T/2)/(2n) over the sampled runouts of the hero's 47-card deck
 * (turn index 0,2,..,46; river r1+1, r1+4, .. < 47; runouts holding a
 * villain card skipped), the reference's quarter weight of ties kept.
 * Output transposed: d_flop_matrices[(f*1216 + j)*1216 + h]; blocked
 * villains 0; heroes holding a flop card are not written (as in the
 * reference).  preflop_grid_generator.cu:30 flop record: */
typedef struct mtt_subset_flop {
    uint8_t c1, c2, c3, pad;
    float weight;
} mtt_subset_flop;
int mtt_gen_flop_matrices(const mtt_subset_flop* d_flops, int n_flops, const mtt_alive_combo* d_combos,
                          float* d_flop_matrices, mtt_stream_t stream);

/* ------------------------------------------- preflop DCFR co-solver -- */
/* kernel_gen_100_flop_matrices + kernel_reduce_combo_to_class_matrices,
 * preflop_grid_generator.cu:130-223, over all 1,326 combos: for each flop
 * (host array flops[n_flops][3], any card order) the class equity matrix
 * d_class_matrices[(f*169 + v)*169 + h] = mean, over the member combos of
 * hero class h and villain class v that avoid the flop and each other (hero
 * members outer, villain inner, lexicographic, FP32 sum), of the flop-matrix
 * entry (2W + T/2)/(2n) on the 192 runouts sampled from the hero's deck;
 * 0.5 when no pair is valid.  The reference keeps only the first 1,176
 * combos (dropping every combo whose first card is 35 or higher), averages
 * only entries > 0 and reads unwritten memory for heroes on the flop
 * (:486, :214, :146); here every combo counts and a valid pair with equity
 * 0 counts.  Returns after the stream has finished. */
int mtt_flop_class_matrices(const uint8_t* flops, int n_flops, float* d_class_matrices, mtt_stream_t stream);
/* The preflop DCFR loop, preflop_grid_generator.cu:588-614 over :228-421:
 * `iters` iterations (500 in the reference) of reaches into the flop ->
 * flop co-solve over the class matrices (subset weights flop_weights, host
 * [n_flops]) -> closed four-node tree values with the all-in matrix d_allin
 * [169*169] (hero class row) -> DCFR on the four nodes (w_pos = t^1.5 /
 * (t^1.5 + 1), w_strat = t^2; root reach = class combos / 1326).  Output
 * d_strategy_sums[4][3][169]: nodes button root (fold, min-raise, jam), big
 * blind facing the min-raise (fold, call, 3-bet jam), button facing the
 * 3-bet jam (fold, call), big blind facing the open jam (fold, call); rows of
 * 2-action nodes end with zeros.  The reference's stack 20bb, bubble factor
 * 1.35, dead pot 2.5bb (:427-430).  Returns after the stream has finished. */
int mtt_preflop_solve(const float* d_class_matrices, const float* flop_weights, int n_flops, const float* d_allin,
                      int iters, float stack_bb, float bubble_factor, float dead_pot, float* d_strategy_sums,
                      mtt_stream_t stream);
/* Class matrices of a subset (host mtt_subset_flop records, as read from
 * mtt_subset.bin) followed by mtt_preflop_solve with the subset weights:
 * the reference program's steps 2-4 (preflop_grid_generator.cu:445-614). */
int mtt_preflop_cosolve(const mtt_subset_flop* subset, int n_flops, const float* d_allin, int iters, float stack_bb,
                        float bubble_factor, float dead_pot, float* d_strategy_sums, mtt_stream_t stream);
/* Per-kernel HIP-event timing of the two calls above (kernels pf_codes,
 * pf_class_eq, pf_cosolve, pf_tree); enabling clears the totals. */
int mtt_preflop_timing(int enable);
int mtt_preflop_kernel_times(int idx, char* name, int name_cap, double* total_ms, int64_t* launches);

/* ------------------------------------------------------- HU game trees -- */
/* Bet menu of one street.  The reference has no bet-menu config in code
 * (menus exist only as README.md:10 prose and MAX_ACTIONS); this struct
 * carries the menus the benchmark configurations describe.  Unopened: check,
 * each size (pot fraction) that stays below the stack, all-in (if allin).
 * Facing a bet: fold, call and, while fewer than raise_cap raises were made
 * this street, a raise to raise_mult x the bet (0 = none) and (if allin) an
 * all-in raise.  At most MTT_MAX_ACTIONS actions per node. */
typedef struct mtt_street_menu {
    int n_sizes;
    float sizes[3];
    float raise_mult;
    int raise_cap;
    int allin;          /* 1: offer all-in (unopened, and as a raise) */
} mtt_street_menu;

/* Heads-up subgame.  n_board cards are known at the root; root_chance = 1
 * makes the root a chance node dealing the next card (the turn for a flop
 * board).  Streets are 0 flop, 1 turn, 2 river; betting runs from the first
 * street to last_street, then showdown over any cards still to come.
 * pot = dead money at the root, stack = effective stack behind. */
typedef struct mtt_hu_game {
    uint8_t board[5];
    int n_board;
    int root_chance;
    int last_street;
    float pot, stack;
    mtt_street_menu menu[3];
} mtt_hu_game;

/* Size of the flat tree (honest_multistreet_solver.cu:16-21 nodes; children
 * contiguous from children_offset, decision infosets numbered in order). */
int mtt_hu_tree_size(const mtt_hu_game* game, int* n_nodes, int* n_infosets, int* n_matrices);
int mtt_hu_tree_build(const mtt_hu_game* game, mtt_flat_node* nodes, int capacity, int* n_nodes);

/* ---------------------------------------------------- seeded inputs -- */
/* Project PRNG (SplitMix64): n cards without replacement. */
int mtt_seeded_cards(uint64_t seed, int n, uint8_t* out);
/* Board (n_board cards) then player 0 and player 1 weights over the 1326
 * combos from one seed; dist 0 = Uniform[0,1] on unblocked combos, 1 =
 * Dirichlet(alpha) over all combos with blocked ones zeroed; normalised. */
int mtt_seeded_inputs(uint64_t seed, int n_board, int dist, float alpha, uint8_t* board_out,
                      float* range0_out, float* range1_out);

/* ------------------------------------------- HU tree-sweep kernels -- */
/* Per-depth kernels with the reference's argument meaning and order, for a
 * host loop that drives the sweeps itself; arguments the reference's
 * kernels do not take follow its list.  Device pointers; layouts:
 * regrets/strategy/strategy sums [infoset][action_stride][1216] (the
 * reference's MAX_ACTIONS = 3, honest_multistreet_solver.cu:12; 4 for bet
 * menus with four actions; decisions with more actions than the stride are
 * left untouched), reach/cfv [node][2][1216], matrices [n][1216][1216]
 * (E[j][h] = equity of hero h against villain j, 0 where blocked).
 * n_alive = live combos of the root board (<= 1216).  stream may be NULL
 * (the null stream). */
/* Regret matching + reach propagation for the decision nodes of one depth;
 * replaces kernel_forward_sweep, honest_multistreet_solver.cu:88-126. */
int mtt_hu_forward_sweep(const int* d_depth_node_indices, int num_nodes_at_depth, const mtt_flat_node* d_nodes,
                         const uint8_t* d_deck, const mtt_alive_combo* d_combos, const float* d_cumulative_regrets,
                         float* d_current_strategy, float* d_reach_probs, int n_alive, int action_stride,
                         hipStream_t stream);
/* Showdown and fold values of terminal nodes (card removal, exact chance
 * normalisation); replaces kernel_evaluate_terminals,
 * honest_multistreet_solver.cu:128-196 (same first six arguments). */
int mtt_hu_evaluate_terminals(const int* d_terminal_node_indices, int num_terminals, const mtt_flat_node* d_nodes,
                              const float* d_reach_probs, const float* d_turn_matrices, float* d_cfvs,
                              const uint8_t* d_deck, const mtt_alive_combo* d_combos, int n_alive, hipStream_t stream);
/* CFR+ backward pass for one depth: node values, regrets floored at 0,
 * strategy sums weighted by iter_weight; replaces kernel_backward_sweep,
 * honest_multistreet_solver.cu:198-249. */
int mtt_hu_backward_sweep(const int* d_depth_node_indices, int num_nodes_at_depth, const mtt_flat_node* d_nodes,
                          const uint8_t* d_deck, const mtt_alive_combo* d_combos, const float* d_current_strategy,
                          const float* d_reach_probs, float* d_cfvs, float* d_cumulative_regrets,
                          float* d_cumulative_strategy_sums, float iter_weight, int n_alive, int action_stride,
                          hipStream_t stream);
/* Same pass with DCFR discounting fused: R <- (R + r) * (R + r > 0 ? w_pos :
 * w_neg), S += w_strat * reach * sigma; the reference's DCFR step is
 * kernel_dcfr_step, preflop_grid_generator.cu:387-410. */
int mtt_hu_backward_sweep_dcfr(const int* d_depth_node_indices, int num_nodes_at_depth, const mtt_flat_node* d_nodes,
                               const uint8_t* d_deck, const mtt_alive_combo* d_combos, const float* d_current_strategy,
                               const float* d_reach_probs, float* d_cfvs, float* d_cumulative_regrets,
                               float* d_cumulative_strategy_sums, float w_pos, float w_neg, float w_strat, int n_alive,
                               int action_stride, hipStream_t stream);

/* -------------------------------------------------------- HU solver -- */
/* Solve entry points (new: the reference's HU main is a stub,
 * honest_multistreet_solver.cu:251-254).  The handle owns all device state
 * on the current device; handles are not shared across threads. */
enum { MTT_MODE_PLAIN_FP32 = 0, MTT_MODE_BF16_SPLIT = 1 };
enum { MTT_SCHEME_CFR_PLUS = 0, MTT_SCHEME_DCFR = 1 };
typedef struct mtt_solver_cfg {
    int mode;          /* MTT_MODE_* */
    int scheme;        /* MTT_SCHEME_*: CFR+ (R floored, weights t) or DCFR */
    float alpha, beta, gamma;  /* DCFR discounts; beta <= -1e30 floors negatives */
} mtt_solver_cfg;
typedef struct mtt_solver mtt_solver;

/* ranges: 1326 floats each (global combo order), blocked combos ignored;
 * weights finite and >= 0, each range totalling less than 32768 (card
 * sums are signed 2^48 fixed point), else MTT_ERR_ARG. */
int mtt_solver_create(const mtt_hu_game* game, const float* range0, const float* range1,
                      const mtt_solver_cfg* cfg, mtt_solver** out);
int mtt_solver_run(mtt_solver* h, int iters);
int mtt_solver_iterations(const mtt_solver* h);
/* Average strategy in the reference layout [infoset][MTT_MAX_ACTIONS][1216]
 * (normalised, uniform where the strategy sum is 0); n_floats must be at
 * least n_infosets*MTT_MAX_ACTIONS*1216. */
int mtt_solver_get_strategy(mtt_solver* h, float* host_out, size_t n_floats);
/* FP64 best response on the GPU against the average strategy:
 * out[0] = BR0 value, out[1] = BR1 value, out[2], out[3] = the profile's
 * own values, out[4] = pair mass M, out[5] = exploitability in % of the
 * starting pot = 100*(BR0 + BR1 - pot*M) / (2*M*pot). */
int mtt_solver_exploitability_fp64(mtt_solver* h, double* out);
/* Raw regrets and strategy sums in the reference layout (either path). */
int mtt_solver_get_state(mtt_solver* h, float* R_out, float* S_out, size_t n_floats);
/* Per-kernel timing with HIP events on the handle's stream (for benchmarks):
 * enable clears the totals; kernel_times returns entry idx (-1 past the end). */
int mtt_solver_timing(mtt_solver* h, int enable);
int mtt_solver_kernel_times(mtt_solver* h, int idx, char* name, int name_cap, double* total_ms, int64_t* launches);
/* info: [0] nodes [1] infosets [2] alive combos [3] device bytes
 * [4] decision nodes [5] showdown nodes [6] fold nodes [7] matrices */
int mtt_solver_info(mtt_solver* h, int64_t* info);
/* Generated sweeps (no counterpart in the reference, whose sweeps are fixed
 * kernels): in MTT_MODE_BF16_SPLIT the turn and river sweeps are emitted as
 * straight-line HIP per street template and compiled with hipRTC at create
 * (environment MTT_JIT=0 selects the interpreted sweeps).  active = 1 when
 * the handle runs generated sweeps; compile_seconds = hipRTC compile and load
 * time (0 when the same code was already built in this process). */
int mtt_solver_jit_info(mtt_solver* h, int* active, double* compile_seconds);
/* The generated sweep source of a game's street templates, and a
 * compile-only hipRTC build of it for gfx950 (neither needs a device).
 * mtt_hu_sweep_source writes at most cap-1 chars plus a terminator and the
 * full length to *len. */
int mtt_hu_sweep_source(const mtt_hu_game* game, char* buf, size_t cap, size_t* len);
int mtt_hu_sweep_compile(const mtt_hu_game* game, size_t* code_bytes, double* seconds);
/* River showdown values of the batched mode (no counterpart in the
 * reference: it multiplies by equity matrices,
 * honest_multistreet_solver.cu:128-196; these two are for inspection and
 * tests).  The batched mode computes them by a sign x blocker GEMM, or, with
 * MTT_RIVER_SD=sweep in the environment at create, by an exact sweep in
 * strength order driven by a per-board program.  mtt_river_sweep_program
 * writes the sweep program of one river board -- n live combos in strength
 * order, cards[2 * n], non-decreasing group ranks grank[n] -- as 16-word
 * records, the first a header (words 1-3: the first three column blocks
 * visited): at most cap words to out (may be null), the full length to
 * *len.  mtt_solver_river_showdown copies, for river board `board` of a
 * batched handle, the last iteration's showdown columns (opponent reach,
 * reach[ncol][1088]) and values (values[ncol][1088], weaker minus stronger
 * non-blocked reach) and the board's combo order (cards[1088][2],
 * grank[1088]); *nlive and *ncol receive the live combo and column counts.
 * Null arrays query the counts only. */
int mtt_river_sweep_program(const uint8_t* cards, const uint16_t* grank, int n, uint32_t* out, size_t cap,
                            size_t* len);
int mtt_solver_river_showdown(mtt_solver* h, int board, float* reach, float* values, uint8_t* cards,
                              uint16_t* grank, int* nlive, int* ncol);
/* The per-board tables of the prefix-scan river showdown (MTT_RIVER_SD=scan)
 * for the 1081 combos of a river board in strength order (cards[1081][2],
 * non-decreasing grank[1081]), as byte offsets (4 x index) into the kernel's
 * LDS arrays: htab[3][1088] = per position 4 gs | 4 ge << 16, then for each
 * of its two cards 4 (list * 47 + lo) | 4 (list * 47 + hi) << 16, [lo, hi)
 * being the entries of the card's 46-combo list inside the position's tie
 * group (padding positions: the zero slots 4 * 1089 and 4 * 2209);
 * ltab[2][12][64] = 4 x the positions of each live card's
 * 46 combos (cards in ascending order, list c in round c / 32 on lanes
 * 2 (c % 32) for entries 0..22 and 2 (c % 32) + 1 for 23..45), two per word.
 * Host only. */
int mtt_river_scan_tables(const uint8_t* cards, const uint16_t* grank, int n, uint32_t* htab, uint32_t* ltab);
int mtt_solver_destroy(mtt_solver* h);

/* ----------------------------------------------- batched flop subgames -- */
/* Many independent heads-up flop subgames solved together (new: the
 * reference has no batch driver).  Every instance
 * is a game with n_board = 3, root_chance = 0, last_street = 0 (flop
 * betting, showdown over the 990 turn/river runouts) and its own board,
 * pot, stack and menu[0].  ranges: [n][2][1326] (global combo order;
 * finite weights >= 0, each range totalling less than 32768).
 * Iterations run the per-instance street sweeps over all instances at once
 * and the showdowns as a batched GEMM of exact W - L counts (2 x BF16)
 * against BF16 hi+lo reach on MFMA; after the first call, iterations are
 * replayed from a captured hipGraph.  cfg->mode is ignored. */
typedef struct mtt_batch mtt_batch;
int mtt_batch_create(const mtt_hu_game* games, int n, const float* ranges, const mtt_solver_cfg* cfg,
                     mtt_batch** out);
int mtt_batch_run(mtt_batch* h, int iters);
int mtt_batch_iterations(const mtt_batch* h);
int mtt_batch_tree_info(mtt_batch* h, int instance, int* n_nodes, int* n_infosets);
/* Raw regrets / strategy sums of one instance in that instance's tree
 * layout [infoset][MTT_MAX_ACTIONS][1216] (flop-alive combo order). */
int mtt_batch_get_state(mtt_batch* h, int instance, float* R_out, float* S_out, size_t n_floats);
/* Average strategy of one instance in the reference layout (strategy sums
 * normalised per decision and combo, uniform where they sum to 0). */
int mtt_batch_get_strategy(mtt_batch* h, int instance, float* out, size_t n_floats);
/* FP64 best-response exploitability of one instance's average strategy
 * (new: the reference has no best response), out as for
 * mtt_solver_exploitability_fp64: BR0, BR1, the profile's values, pair
 * mass, exploitability in % of the instance's starting pot. */
int mtt_batch_exploitability_fp64(mtt_batch* h, int instance, double* out);
/* 0: launch kernels one by one (also while timing is on), 1: graph replay. */
int mtt_batch_set_graph(mtt_batch* h, int on);
/* Autotuner's choice for the showdown GEMM as a form code (2 or 4: hero
 * tiles per workgroup of the LDS-staged form; 10 + HT: the register-streamed
 * form with HT = 2, 3 or 4 hero tiles per wave) and, in gemm_ms[5], the
 * measured times of the candidates 2, 4, 12, 13, 14 (0 when the choice was
 * read from the tuning file or forced by MTT_FLOP_GEMM). */
int mtt_batch_tuning(mtt_batch* h, int* hero_tiles, double* gemm_ms);
/* The iteration form in use: 0 = forward, showdown GEMM and backward as
 * three launches per iteration (graph-replayed); 1 = one persistent launch
 * per iteration whose workgroups take forward, GEMM and backward items of
 * all instances from a queue (same bodies, results identical to form 0);
 * -1 = not chosen yet.  Form 1 is switched off in this build (its GPU runs
 * did not finish; MTT_FLOP_PERSISTENT=1 makes mtt_batch_run fail with
 * MTT_ERR_ARG), so the mode is 0 after the first run; ms[0], ms[1] hold the
 * per-iteration times of a tuning run of both forms (0 when none ran). ms
 * may be null. */
int mtt_batch_iter_mode(mtt_batch* h, int* mode, double* ms);
int mtt_batch_timing(mtt_batch* h, int enable);
int mtt_batch_kernel_times(mtt_batch* h, int idx, char* name, int name_cap, double* total_ms, int64_t* launches);
int mtt_batch_destroy(mtt_batch* h);

/* -------------------------------------------------- fused DCFR step -- */
/* kernel_dcfr_step, preflop_grid_generator.cu:387-410, over n entries with
 * action stride `stride` (the reference: n = 169 classes, stride 192):
 * v = sum_a sigma_a cfv_a; R_a <- (R_a + cfv_a - v) * (> 0 ? w_pos : w_neg);
 * S_a += w_strat * reach * sigma_a; sigma <- R+ / sum R+ (uniform when the
 * sum is <= 1e-7).  w_neg = 0 is the reference's floor-then-scale form.
 * Device pointers: R, sigma, S, cfv [num_actions][stride], reach [n]. */
int mtt_dcfr_step(int num_actions, float* d_regrets, float* d_current_strat, float* d_strat_sums,
                  const float* d_action_cfvs, const float* d_reach, float w_pos, float w_strat, float w_neg, int n,
                  int stride, hipStream_t stream);

/* ------------------------------------------------------- 4-way engine -- */
/* Single-street 4-player flop game (multiway_4way_engine.cu).  Per-action
 * state uses the reference's stride of 3 actions:
 * regrets/strategy/sums [infoset][3][1216], reach/cfv [node][4][1216]. */
#define MTT_4WAY_MAX_ACTIONS 3
typedef struct mtt_4way_game {
    float pot0[4];        /* chips each seat has in the pot before the street */
    float stack[4];       /* chips behind at the start of the street */
    mtt_street_menu menu;
} mtt_4way_game;
/* Betting tree (seat order; bets capped at the largest stack that can call
 * them, sizes at or above that cap become all-in, calls all-in for less
 * leave side pots; breadth-first layout, one infoset per decision node). */
int mtt_4way_tree_size(const mtt_4way_game* game, int* n_nodes, int* n_infosets);
int mtt_4way_tree_build(const mtt_4way_game* game, mtt_flat_node4* nodes, int capacity, int* n_nodes);

/* Flop board tables for the multiway showdown: the reference's
 * combo_rank_pos (multiway_4way_engine.cu:85,137-146) plus what the card
 * removal and ties need.  Strength = the reference's score of the 5 flop
 * cards; positions ascend in strength. */
typedef struct mtt_board4 {
    uint8_t c1[1216], c2[1216];        /* combo cards (1176 alive) */
    uint16_t rank_pos[1216];           /* combo -> position (= combo_rank_pos) */
    uint16_t pos_combo[1216];          /* position -> combo */
    uint16_t grp_lo[1216], grp_hi[1216];  /* tie group of a combo: [lo, hi) positions */
    uint16_t card_list[52][48];        /* combos holding a card, ascending position */
    uint8_t n_card[52];
    uint8_t card_lt[1216][2];          /* entries of each card's list strictly weaker */
    uint8_t card_le[1216][2];          /* entries weaker or tied */
} mtt_board4;
int mtt_board4_build(const uint8_t flop[3], mtt_board4* out);

/* kernel_4way_forward_sweep, multiway_4way_engine.cu:41-83. */
int mtt_4way_forward_sweep(const int* d_depth_node_indices, int num_nodes_at_depth, const mtt_flat_node4* d_nodes,
                           const float* d_cumulative_regrets, float* d_current_strategy, float* d_reach_probs,
                           hipStream_t stream);
/* kernel_4way_evaluate_terminals, multiway_4way_engine.cu:85-217, with
 * card removal against the hero's hand, split pots for ties and side pots
 * for unequal commitments:
 * fold: pay * prod_{o != p} W_o; showdown, player p still in: the pot split
 * into layers at the distinct commitments of the players still in, p taking
 * each layer it reaches unless an opponent reaching it is stronger (shared
 * with the tied ones), net chips scaled by bf[p][p] when negative, summed
 * over the 3^k weaker/tied/stronger outcomes of its k opponents still in,
 * each weighted by prod F_o / (G_o - F_o) / (W_o - G_o), times the folded
 * seats' prod W_o; folded p: -bf[p][p] c_p prod W_o.
 * F/G/W = opponent reach strictly weaker / weaker-or-tied / total, each
 * without combos that share a card with the hero.
 * Same first seven arguments as the reference: d_combo_rank_pos (device,
 * 1216 uint16) is the flop's ascending strength order, checked against the
 * evaluator's (MTT_ERR_ARG if it is not one); the board tables are built
 * from `flop` (host, 3 cards).  d_bubble_factors 4x4 floats.  Synchronous. */
int mtt_4way_evaluate_terminals(const int* d_terminal_node_indices, int num_terminals, const mtt_flat_node4* d_nodes,
                                const float* d_reach_probs, const uint16_t* d_combo_rank_pos,
                                const float* d_bubble_factors, float* d_cfvs, const uint8_t* flop,
                                hipStream_t stream);
/* The same with the board tables already on the device (d_board: a device
 * copy of mtt_board4_build's output); asynchronous on `stream`. */
int mtt_4way_evaluate_terminals_board(const int* d_terminal_node_indices, int num_terminals,
                                      const mtt_flat_node4* d_nodes, const float* d_reach_probs,
                                      const mtt_board4* d_board, const float* d_bubble_factors, float* d_cfvs,
                                      hipStream_t stream);
/* kernel_4way_backward_cfr_plus, multiway_4way_engine.cu:219-257. */
int mtt_4way_backward_cfr_plus(const int* d_depth_node_indices, int num_nodes_at_depth, const mtt_flat_node4* d_nodes,
                               const float* d_current_strategy, const float* d_reach_probs, float* d_cfvs,
                               float* d_cumulative_regrets, float* d_cumulative_strategy_sums, float iter_weight,
                               hipStream_t stream);
/* The same with the DCFR-generalised update (w_pos, w_neg, w_strat). */
int mtt_4way_backward_dcfr(const int* d_depth_node_indices, int num_nodes_at_depth, const mtt_flat_node4* d_nodes,
                           const float* d_current_strategy, const float* d_reach_probs, float* d_cfvs,
                           float* d_cumulative_regrets, float* d_cumulative_strategy_sums, float w_pos, float w_neg,
                           float w_strat, hipStream_t stream);

/* 4-way solves of one tree over many flops (new: the reference's main is a
 * stub, multiway_4way_engine.cu:259-262).  ranges: [n_flops][4][1216] in
 * each flop's alive-combo order; bubble_factors 4x4 (row p, column q). */
typedef struct mtt_multi mtt_multi;
int mtt_multi_create(const mtt_4way_game* game, int n_flops, const uint8_t* flops, const float* ranges,
                     const float* bubble_factors, const mtt_solver_cfg* cfg, mtt_multi** out);
int mtt_multi_run(mtt_multi* h, int iters);
/* Raw regrets/strategy sums of one flop [infoset][3][1216] and its root
 * values [4][1216]. */
int mtt_multi_get_state(mtt_multi* h, int flop, float* R_out, float* S_out, float* root_cfv_out);
int mtt_multi_timing(mtt_multi* h, int enable);
int mtt_multi_kernel_times(mtt_multi* h, int idx, char* name, int name_cap, double* total_ms, int64_t* launches);
int mtt_multi_destroy(mtt_multi* h);

/* ------------------------------------------------------------- ICM -- */
/* Malmuth-Harville equities in FP64 on the GPU (absent from the reference,
 * whose bubble_factors are an input, multiway_4way_engine.cu:85,128,201,204).
 * Host arrays: stacks[n] (n <= 11), payouts[k]; out_equity[n]. */
int mtt_icm_mh_fp64(const double* stacks, int n, const double* payouts, int k, double* out_equity);
/* Batched on device: d_stacks [count][n] -> d_equity [count][n]. */
int mtt_icm_mh_fp64_batch(const double* d_stacks, int count, int n, const double* d_payouts, int k,
                          double* d_equity, hipStream_t stream);
/* Bubble factors of the m players in_pot[] (host arrays): for p != q with
 * e = min(S_p, S_q), bf[p][q] = (EQ_p - EQ_p after losing e to q) /
 * (EQ_p after winning e from q - EQ_p); bf[p][p] = mean of bf[p][q] over
 * the others in the pot.  out_bf [m][m]. */
int mtt_icm_bubble_factors_fp64(const double* stacks, int n, const double* payouts, int k, const int* in_pot, int m,
                                double* out_bf);

#ifdef __cplusplus
}
#endif
#endif /* MTT_H */
