// HU solver handle: builds the tree and device state, generates the equity
// matrices the plain path reads, and drives iterations of either path.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "solver.h"

struct mtt_solver {
    mtt::SolverImpl impl;
};

namespace mtt {

constexpr int NP = kPadFlop;
constexpr int AS = MTT_MAX_ACTIONS;

SolverImpl::~SolverImpl() {
    if (batched) batched_destroy(batched);
    batched = nullptr;
    arena.release();
    if (stream) (void)hipStreamDestroy(stream);
}

// --------------------------------------------------------- matrices -----
// scores[b][h] of root-alive combo h on 5-card board b (0 if h holds a
// board card).
__global__ void k_board_scores(const uint8_t* __restrict__ boards, const mtt_alive_combo* __restrict__ combos,
                               int n_alive, uint32_t* __restrict__ out) {
    int b = blockIdx.y;
    const uint8_t* bc = boards + 5 * b;
    uint64_t m = 0, dead = 0;
#pragma unroll
    for (int k = 0; k < 5; ++k) { m |= card_bits16(bc[k]); dead |= 1ull << bc[k]; }
    for (int h = blockIdx.x * blockDim.x + threadIdx.x; h < NP; h += gridDim.x * blockDim.x) {
        uint32_t s = 0;
        if (h < n_alive) {
            mtt_alive_combo c = combos[h];
            if (!(c.mask & dead)) s = score_packed(m | card_bits16(c.c1) | card_bits16(c.c2));
        }
        out[(size_t)b * NP + h] = s;
    }
}

// E[j][h] = mean over the sub-boards valid for both hands of
// 1 (h beats j) / 0.5 (tie) / 0; zero when blocked or no valid sub-board.
// Sub-boards of matrix m: score rows rows[off[m] .. off[m+1]).
__global__ void k_sd_matrix(const uint32_t* __restrict__ scores, const int* __restrict__ rows,
                            const int* __restrict__ off, const mtt_alive_combo* __restrict__ combos, int n_alive,
                            float* __restrict__ out) {
    int m = blockIdx.z, j = blockIdx.y;
    float* row = out + ((size_t)m * NP + j) * NP;
    int r0 = off[m], r1 = off[m + 1];
    uint64_t vm = j < n_alive ? combos[j].mask : 0;
    for (int h = blockIdx.x * blockDim.x + threadIdx.x; h < NP; h += gridDim.x * blockDim.x) {
        float e = 0.0f;
        if (j < n_alive && h < n_alive && !(combos[h].mask & vm)) {
            int w = 0, t = 0, n = 0;
            for (int k = r0; k < r1; ++k) {
                const uint32_t* s = scores + (size_t)rows[k] * NP;
                uint32_t a = s[h], v = s[j];
                if (!a || !v) continue;
                w += a > v;
                t += a == v;
                n++;
            }
            e = n ? ((float)w + 0.5f * (float)t) / (float)n : 0.0f;
        }
        row[h] = e;
    }
}

int solver_gen_matrices(SolverImpl& s, float* d_out) {
    const mtt_hu_game& g = s.game;
    int nm = s.tree.n_matrices;
    // sub-boards per matrix (full 5-card boards)
    std::vector<uint8_t> boards;
    std::vector<int> rows, off(1, 0);
    auto board_with = [&](int x, int y) {
        uint8_t b[5];
        int n = 0;
        for (int k = 0; k < g.n_board; ++k) b[n++] = g.board[k];
        if (x >= 0) b[n++] = (uint8_t)x;
        if (y >= 0) b[n++] = (uint8_t)y;
        for (int k = 0; k < 5; ++k) boards.push_back(b[k]);
        return (int)boards.size() / 5 - 1;
    };
    const int dn = (int)s.deck.size();
    const bool turn_river = g.n_board == 3 && g.root_chance;
    if (turn_river && g.last_street == 2) {
        // river boards first so turn matrices can reuse their rows
        std::vector<int> rb(49 * 49, -1);
        for (int t = 0; t < dn; ++t)
            for (int r = 0; r < dn; ++r)
                if (r != t) rb[t * 49 + r] = board_with(s.deck[t], s.deck[r]);
        for (int m = 0; m < nm; ++m) {
            if (m < 49) {
                for (int r = 0; r < dn; ++r)
                    if (r != m) rows.push_back(rb[m * 49 + r]);
            } else {
                int q = m - 49, t = q / 48, rk = q % 48, r = rk < t ? rk : rk + 1;
                rows.push_back(rb[t * 49 + r]);
            }
            off.push_back((int)rows.size());
        }
    } else if (turn_river) {  // turn only, showdown over rivers
        std::vector<int> rb(49 * 49, -1);
        for (int t = 0; t < dn; ++t)
            for (int r = 0; r < dn; ++r)
                if (r != t) rb[t * 49 + r] = board_with(s.deck[t], s.deck[r]);
        for (int m = 0; m < nm; ++m) {
            for (int r = 0; r < dn; ++r)
                if (r != m) rows.push_back(rb[m * 49 + r]);
            off.push_back((int)rows.size());
        }
    } else if (g.n_board == 4 && g.last_street == 2) {
        std::vector<int> rb(dn);
        for (int r = 0; r < dn; ++r) rb[r] = board_with(s.deck[r], -1);
        for (int m = 0; m < nm; ++m) {
            if (m < 48) rows.push_back(rb[m]);
            else for (int r = 0; r < dn; ++r) rows.push_back(rb[r]);
            off.push_back((int)rows.size());
        }
    } else if (g.n_board == 5) {
        rows.push_back(board_with(-1, -1));
        off.push_back(1);
    } else if (g.n_board == 4) {  // turn board, river to come
        for (int r = 0; r < dn; ++r) rows.push_back(board_with(s.deck[r], -1));
        off.push_back((int)rows.size());
    } else {  // flop board, turn and river to come
        for (int x = 0; x < dn; ++x)
            for (int y = x + 1; y < dn; ++y) rows.push_back(board_with(s.deck[x], s.deck[y]));
        off.push_back((int)rows.size());
    }
    int nb = (int)boards.size() / 5;
    DeviceArena tmp;
    uint8_t* d_b = nullptr;
    int *d_rows = nullptr, *d_off = nullptr;
    uint32_t* d_sc = nullptr;
    int rc = tmp.upload(boards, &d_b);
    rc = rc ? rc : tmp.upload(rows, &d_rows);
    rc = rc ? rc : tmp.upload(off, &d_off);
    rc = rc ? rc : tmp.alloc((size_t)nb * NP, &d_sc);
    if (rc) return rc;
    for (int b0 = 0; b0 < nb; b0 += 65535) {
        int cnt = std::min(65535, nb - b0);
        hipLaunchKernelGGL(k_board_scores, dim3(5, cnt), dim3(256), 0, s.stream, d_b + 5 * b0, s.d_combos, s.n_alive,
                           d_sc + (size_t)b0 * NP);
    }
    for (int m0 = 0; m0 < nm; m0 += 65535) {
        int cnt = std::min(65535, nm - m0);
        hipLaunchKernelGGL(k_sd_matrix, dim3(5, NP, cnt), dim3(256), 0, s.stream, d_sc, d_rows, d_off + m0,
                           s.d_combos, s.n_alive, d_out + (size_t)m0 * NP * NP);
    }
    MTT_HIP(hipGetLastError());
    MTT_HIP(hipStreamSynchronize(s.stream));
    return MTT_OK;
}

void iteration_weights(const mtt_solver_cfg& cfg, int t, float& wp, float& wn, float& ws) {
    double tt = (double)t;
    if (cfg.scheme == MTT_SCHEME_CFR_PLUS) {
        wp = 1.0f;
        wn = 0.0f;
        ws = (float)tt;
        return;
    }
    double ta = std::pow(tt, (double)cfg.alpha);
    wp = (float)(ta / (ta + 1.0));
    if (cfg.beta <= -1e30f) wn = 0.0f;
    else {
        double tb = std::pow(tt, (double)cfg.beta);
        wn = (float)(tb / (tb + 1.0));
    }
    ws = (float)std::pow(tt, (double)cfg.gamma);
}

// ------------------------------------------------------------- plain ----
int solver_alloc_plain(SolverImpl& s) {
    size_t nn = s.tree.flat.size();
    size_t ni = (size_t)s.tree.n_infosets;
    int rc = MTT_OK;
    rc = rc ? rc : s.arena.alloc(nn * 2 * NP, &s.d_reach);
    rc = rc ? rc : s.arena.alloc(nn * 2 * NP, &s.d_cfv);
    rc = rc ? rc : s.arena.alloc(ni * AS * NP, &s.d_R);
    rc = rc ? rc : s.arena.alloc(ni * AS * NP, &s.d_S);
    rc = rc ? rc : s.arena.alloc(ni * AS * NP, &s.d_sigma);
    rc = rc ? rc : s.arena.alloc((size_t)s.tree.n_matrices * NP * NP, &s.d_mats);
    rc = rc ? rc : s.arena.alloc(s.terminals.size() * 2 * 64, &s.d_sums);
    if (rc) return rc;
    MTT_HIP(hipMemsetAsync(s.d_reach, 0, nn * 2 * NP * sizeof(float), s.stream));
    MTT_HIP(hipMemsetAsync(s.d_cfv, 0, nn * 2 * NP * sizeof(float), s.stream));
    MTT_HIP(hipMemsetAsync(s.d_R, 0, ni * AS * NP * sizeof(float), s.stream));
    MTT_HIP(hipMemsetAsync(s.d_S, 0, ni * AS * NP * sizeof(float), s.stream));
    MTT_HIP(hipMemsetAsync(s.d_sigma, 0, ni * AS * NP * sizeof(float), s.stream));
    // root reach = ranges
    std::vector<float> root(2 * NP, 0.0f);
    for (int p = 0; p < 2; ++p)
        for (int i = 0; i < s.n_alive; ++i) root[p * NP + i] = s.range[p][i];
    MTT_HIP(hipMemcpyAsync(s.d_reach, root.data(), root.size() * sizeof(float), hipMemcpyHostToDevice, s.stream));
    rc = solver_gen_matrices(s, s.d_mats);
    if (rc) return rc;
    rc = build_terminal_plan(s.tree.flat, s.terminals, s.n_alive, s.c1, s.c2, s.plan, s.arena);
    return rc;
}

int solver_run_plain(SolverImpl& s, int iters) {
    int D = (int)s.depth_lists.size();
    for (int it = 0; it < iters; ++it) {
        int t = s.iters_done + 1;
        float wp, wn, ws;
        iteration_weights(s.cfg, t, wp, wn, ws);
        int rc = MTT_OK;
        KernelTimer::Mark m;
        if (s.timer.enabled) s.timer.begin("plain_forward", s.stream, m);
        for (int d = 0; d < D && !rc; ++d)
            rc = launch_hu_forward(s.d_depth[d], (int)s.depth_lists[d].size(), s.d_nodes, s.d_deck, s.d_combos, s.d_R,
                                   s.d_sigma, s.d_reach, s.n_alive, s.stream);
        if (s.timer.enabled) s.timer.end(m, s.stream);
        if (s.timer.enabled) s.timer.begin("plain_terminals", s.stream, m);
        if (!rc)
            rc = launch_hu_terminals_grouped(s.plan, s.d_nodes, s.d_deck, s.d_combos, s.d_reach, s.d_mats, s.d_sums,
                                             s.d_cfv, s.n_alive, s.stream);
        if (s.timer.enabled) s.timer.end(m, s.stream);
        if (s.timer.enabled) s.timer.begin("plain_backward", s.stream, m);
        for (int d = D - 1; d >= 0 && !rc; --d)
            rc = launch_hu_backward(s.d_depth[d], (int)s.depth_lists[d].size(), s.d_nodes, s.d_deck, s.d_combos,
                                    s.d_sigma, s.d_reach, s.d_cfv, s.d_R, s.d_S, wp, wn, ws, s.n_alive, s.stream);
        if (s.timer.enabled) s.timer.end(m, s.stream);
        if (rc) return rc;
        s.iters_done = t;
    }
    return MTT_OK;
}

// Average strategy (normalised) from strategy sums in the reference layout.
__global__ void k_avg_strategy(const mtt_flat_node* __restrict__ nodes, const int* __restrict__ dec_nodes, int n,
                               const float* __restrict__ S, float* __restrict__ out) {
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= NP) return;
    for (int k = blockIdx.y; k < n; k += gridDim.y) {
        mtt_flat_node x = nodes[dec_nodes[k]];
        size_t base = (size_t)x.infoset_offset * AS * NP + tid;
        float v[AS], sum = 0.0f;
#pragma unroll
        for (int a = 0; a < AS; ++a) {
            v[a] = a < x.num_actions ? S[base + (size_t)a * NP] : 0.0f;
            sum += v[a];
        }
#pragma unroll
        for (int a = 0; a < AS; ++a)
            out[base + (size_t)a * NP] = a < x.num_actions ? (sum > 0.0f ? v[a] / sum : 1.0f / x.num_actions) : 0.0f;
    }
}

static int avg_strategy_device(SolverImpl& s, float* d_out) {
    std::vector<int> dec;
    for (int i = 0; i < (int)s.tree.flat.size(); ++i)
        if (s.tree.flat[i].type == MTT_NODE_DECISION) dec.push_back(i);
    DeviceArena tmp;
    int* d_dec = nullptr;
    int rc = tmp.upload(dec, &d_dec);
    if (rc) return rc;
    const float* S = s.d_S;
    float* Sref = nullptr;
    if (s.cfg.mode == MTT_MODE_BF16_SPLIT) {
        rc = tmp.alloc((size_t)s.tree.n_infosets * AS * NP, &Sref);
        if (rc) return rc;
        rc = batched_strategy_sums(s, Sref);
        if (rc) return rc;
        S = Sref;
    }
    MTT_HIP(hipMemsetAsync(d_out, 0, (size_t)s.tree.n_infosets * AS * NP * sizeof(float), s.stream));
    int n = (int)dec.size();
    if (n)
        hipLaunchKernelGGL(k_avg_strategy, dim3(NP / 256 + 1, std::min(n, 65535)), dim3(256), 0, s.stream, s.d_nodes,
                           d_dec, n, S, d_out);
    MTT_HIP(hipGetLastError());
    MTT_HIP(hipStreamSynchronize(s.stream));
    return MTT_OK;
}

// ---------------------------------------------------------------- create --
static int solver_create(const mtt_hu_game* game, const float* r0, const float* r1, const mtt_solver_cfg* cfg,
                         SolverImpl& s) {
    MTT_ARG(game && r0 && r1 && cfg, "mtt_solver_create: null pointer");
    MTT_ARG(cfg->mode == MTT_MODE_PLAIN_FP32 || cfg->mode == MTT_MODE_BF16_SPLIT, "mtt_solver_create: mode %d",
            cfg->mode);
    MTT_ARG(cfg->scheme == MTT_SCHEME_CFR_PLUS || cfg->scheme == MTT_SCHEME_DCFR, "mtt_solver_create: scheme %d",
            cfg->scheme);
    s.game = *game;
    s.cfg = *cfg;
    int rc = build_hu_tree(*game, s.tree);
    if (rc) return rc;
    MTT_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    uint64_t bm = 0;
    for (int k = 0; k < game->n_board; ++k) bm |= 1ull << game->board[k];
    s.n_alive = alive_combos(bm, s.c1, s.c2);
    for (int i = 0; i < s.n_alive; ++i) s.gidx[i] = combo_index(s.c1[i], s.c2[i]);
    for (int c = 0; c < 52; ++c)
        if (!(bm & (1ull << c))) s.deck.push_back((uint8_t)c);
    for (int p = 0; p < 2; ++p) {
        const float* src = p == 0 ? r0 : r1;
        s.range[p].assign(s.n_alive, 0.0f);
        double tot = 0.0;
        for (int i = 0; i < s.n_alive; ++i) {
            float v = src[s.gidx[i]];
            MTT_ARG(std::isfinite(v) && v >= 0, "mtt_solver_create: range %d has a negative or non-finite weight", p);
            s.range[p][i] = v;
            tot += v;
        }
        // the batched path sums reaches per card in signed 2^48 fixed point
        MTT_ARG(tot < kMaxRangeTotal, "mtt_solver_create: range %d totals %g, at least %g (normalise the range)", p,
                tot, kMaxRangeTotal);
    }
    // pair mass M = sum_i r0(i) * (tot1 - R1[c1] - R1[c2] + r1(i))
    double tot1 = 0, card1[52] = {0};
    for (int j = 0; j < s.n_alive; ++j) {
        tot1 += s.range[1][j];
        card1[s.c1[j]] += s.range[1][j];
        card1[s.c2[j]] += s.range[1][j];
    }
    for (int i = 0; i < s.n_alive; ++i)
        s.pair_mass += s.range[0][i] * (tot1 - card1[s.c1[i]] - card1[s.c2[i]] + s.range[1][i]);
    MTT_ARG(s.pair_mass > 0, "mtt_solver_create: the ranges have no non-conflicting pair");
    // depth lists (BFS) and terminals
    const auto& F = s.tree.flat;
    std::vector<int> depth(F.size(), 0);
    std::vector<int> q(1, 0);
    for (size_t h = 0; h < q.size(); ++h) {
        int id = q[h];
        const mtt_flat_node& x = F[id];
        if (x.type == MTT_NODE_FOLD || x.type == MTT_NODE_SHOWDOWN) {
            s.terminals.push_back(id);
            continue;
        }
        if ((int)s.depth_lists.size() <= depth[id]) s.depth_lists.resize(depth[id] + 1);
        s.depth_lists[depth[id]].push_back(id);
        for (int k = 0; k < x.num_actions; ++k) {
            int c = x.children_offset + k;
            MTT_ARG(c > id && c < (int)F.size(), "tree: bad child offset");
            depth[c] = depth[id] + 1;
            q.push_back(c);
        }
    }
    // common device data
    std::vector<mtt_alive_combo> combos(NP);
    memset(combos.data(), 0, sizeof(mtt_alive_combo) * NP);
    for (int i = 0; i < s.n_alive; ++i) combos[i] = {s.c1[i], s.c2[i], (1ull << s.c1[i]) | (1ull << s.c2[i])};
    for (int i = s.n_alive; i < NP; ++i) combos[i].mask = ~0ull;
    rc = s.arena.upload(F, &s.d_nodes);
    rc = rc ? rc : s.arena.upload(s.deck, &s.d_deck);
    rc = rc ? rc : s.arena.upload(combos, &s.d_combos);
    if (rc) return rc;
    s.d_depth.resize(s.depth_lists.size());
    for (size_t d = 0; d < s.depth_lists.size() && !rc; ++d) rc = s.arena.upload(s.depth_lists[d], &s.d_depth[d]);
    if (rc) return rc;
    if (cfg->mode == MTT_MODE_PLAIN_FP32) return solver_alloc_plain(s);
    return batched_create(s);
}

}  // namespace mtt

using namespace mtt;

namespace mtt {

int solver_exploitability_of(const mtt_hu_game& game, const float* r0_1326, const float* r1_1326,
                             const float* host_strat, size_t n_floats, int n_infosets, double* out6) {
    mtt_solver_cfg cfg;
    memset(&cfg, 0, sizeof(cfg));
    cfg.mode = MTT_MODE_PLAIN_FP32;
    cfg.scheme = MTT_SCHEME_CFR_PLUS;
    mtt_solver* h = nullptr;
    int rc = mtt_solver_create(&game, r0_1326, r1_1326, &cfg, &h);
    if (rc) return rc;
    SolverImpl& si = h->impl;
    const size_t need = (size_t)si.tree.n_infosets * AS * NP;
    if (si.tree.n_infosets != n_infosets || n_floats < need) {
        mtt_solver_destroy(h);
        set_error("exploitability: the strategy has %d infosets, the game's tree %d", n_infosets, si.tree.n_infosets);
        return MTT_ERR_ARG;
    }
    DeviceArena tmp;
    float* d = nullptr;
    rc = tmp.alloc(need, &d);
    if (!rc) {
        hipError_t e = hipMemcpyAsync(d, host_strat, need * sizeof(float), hipMemcpyHostToDevice, si.stream);
        if (e != hipSuccess) rc = hip_fail(e, "exploitability: strategy upload", __FILE__, __LINE__);
    }
    rc = rc ? rc : br_exploitability(si, d, out6);
    mtt_solver_destroy(h);
    return rc;
}

}  // namespace mtt

extern "C" {

int mtt_solver_create(const mtt_hu_game* game, const float* range0, const float* range1, const mtt_solver_cfg* cfg,
                      mtt_solver** out) {
    (void)hipGetLastError();   // an error left by an earlier call is not this call's
    MTT_ARG(out, "mtt_solver_create: null out");
    *out = nullptr;
    mtt_solver* h = new (std::nothrow) mtt_solver;
    if (!h) { set_error("mtt_solver_create: out of host memory"); return MTT_ERR_OOM; }
    int rc = solver_create(game, range0, range1, cfg, h->impl);
    if (rc) { delete h; return rc; }
    *out = h;
    return MTT_OK;
}

int mtt_solver_run(mtt_solver* h, int iters) {
    MTT_ARG(h && iters >= 0, "mtt_solver_run: bad arguments");
    int rc = h->impl.cfg.mode == MTT_MODE_PLAIN_FP32 ? solver_run_plain(h->impl, iters) : batched_run(h->impl, iters);
    if (rc) return rc;
    MTT_HIP(hipStreamSynchronize(h->impl.stream));
    return MTT_OK;
}

int mtt_solver_iterations(const mtt_solver* h) { return h ? h->impl.iters_done : -1; }

int mtt_solver_get_strategy(mtt_solver* h, float* host_out, size_t n_floats) {
    MTT_ARG(h && host_out, "mtt_solver_get_strategy: null pointer");
    size_t need = (size_t)h->impl.tree.n_infosets * AS * NP;
    MTT_ARG(n_floats >= need, "mtt_solver_get_strategy: need %zu floats", need);
    DeviceArena tmp;
    float* d = nullptr;
    int rc = tmp.alloc(need, &d);
    rc = rc ? rc : avg_strategy_device(h->impl, d);
    if (rc) return rc;
    MTT_HIP(hipMemcpy(host_out, d, need * sizeof(float), hipMemcpyDeviceToHost));
    return MTT_OK;
}

int mtt_solver_exploitability_fp64(mtt_solver* h, double* out) {
    MTT_ARG(h && out, "mtt_solver_exploitability_fp64: null pointer");
    size_t need = (size_t)h->impl.tree.n_infosets * AS * NP;
    DeviceArena tmp;
    float* d = nullptr;
    int rc = tmp.alloc(need, &d);
    rc = rc ? rc : avg_strategy_device(h->impl, d);
    rc = rc ? rc : br_exploitability(h->impl, d, out);
    return rc;
}

int mtt_solver_get_state(mtt_solver* h, float* R_out, float* S_out, size_t n_floats) {
    MTT_ARG(h, "mtt_solver_get_state: null handle");
    size_t need = (size_t)h->impl.tree.n_infosets * AS * NP;
    MTT_ARG(n_floats >= need, "mtt_solver_get_state: need %zu floats", need);
    if (h->impl.cfg.mode == MTT_MODE_PLAIN_FP32) {
        if (R_out) MTT_HIP(hipMemcpy(R_out, h->impl.d_R, need * sizeof(float), hipMemcpyDeviceToHost));
        if (S_out) MTT_HIP(hipMemcpy(S_out, h->impl.d_S, need * sizeof(float), hipMemcpyDeviceToHost));
        return MTT_OK;
    }
    DeviceArena tmp;
    float* d = nullptr;
    int rc = tmp.alloc(need, &d);
    for (int w = 0; w < 2 && !rc; ++w) {
        float* dst = w ? R_out : S_out;
        if (!dst) continue;
        rc = batched_export(h->impl, d, w);
        if (!rc) MTT_HIP(hipMemcpy(dst, d, need * sizeof(float), hipMemcpyDeviceToHost));
    }
    return rc;
}

int mtt_solver_timing(mtt_solver* h, int enable) {
    MTT_ARG(h, "mtt_solver_timing: null handle");
    h->impl.timer.collect();
    h->impl.timer.enabled = enable != 0;
    if (enable) {
        h->impl.timer.names.clear();
        h->impl.timer.ms.clear();
        h->impl.timer.count.clear();
    }
    return MTT_OK;
}

int mtt_solver_kernel_times(mtt_solver* h, int idx, char* name, int name_cap, double* total_ms, int64_t* launches) {
    MTT_ARG(h && name && total_ms && launches && name_cap > 0, "mtt_solver_kernel_times: bad arguments");
    MTT_HIP(hipStreamSynchronize(h->impl.stream));
    h->impl.timer.collect();
    if (idx < 0 || idx >= (int)h->impl.timer.names.size()) return -1;
    snprintf(name, name_cap, "%s", h->impl.timer.names[idx].c_str());
    *total_ms = h->impl.timer.ms[idx];
    *launches = h->impl.timer.count[idx];
    return MTT_OK;
}

int mtt_solver_info(mtt_solver* h, int64_t* info) {
    MTT_ARG(h && info, "mtt_solver_info: null pointer");
    const auto& F = h->impl.tree.flat;
    int64_t dec = 0, sd = 0, fo = 0;
    for (auto& x : F) {
        dec += x.type == MTT_NODE_DECISION;
        sd += x.type == MTT_NODE_SHOWDOWN;
        fo += x.type == MTT_NODE_FOLD;
    }
    info[0] = (int64_t)F.size();
    info[1] = h->impl.tree.n_infosets;
    info[2] = h->impl.n_alive;
    info[3] = (int64_t)h->impl.arena.bytes();
    info[4] = dec;
    info[5] = sd;
    info[6] = fo;
    info[7] = h->impl.tree.n_matrices;
    return MTT_OK;
}

int mtt_solver_jit_info(mtt_solver* h, int* active, double* compile_seconds) {
    MTT_ARG(h, "mtt_solver_jit_info: null handle");
    return mtt::batched_jit_info(h->impl, active, compile_seconds);
}

int mtt_solver_river_showdown(mtt_solver* h, int board, float* reach, float* values, uint8_t* cards, uint16_t* grank,
                              int* nlive, int* ncol) {
    MTT_ARG(h, "mtt_solver_river_showdown: null handle");
    return mtt::batched_river_showdown(h->impl, board, reach, values, cards, grank, nlive, ncol);
}

int mtt_solver_destroy(mtt_solver* h) {
    delete h;
    return MTT_OK;
}

}  // extern "C"
