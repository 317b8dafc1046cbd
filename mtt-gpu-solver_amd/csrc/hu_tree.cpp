// Heads-up tree builder (see hu_tree.h).
//
// Betting rules per street (project-defined; the reference has no builder):
//   unopened  : check, each pot-fraction bet that stays below the stack,
//               all-in;
//   facing bet: fold, call, and - while fewer than `raise_cap` raises were
//               made this street - a raise to raise_mult x the facing bet
//               (if below the stack) and an all-in raise;
//   check-check or a call closes the street; a called all-in goes to
//   showdown with the remaining cards dealt out; after the last street the
//   showdown is on the full board.
// With these rules the 10bb/20bb river game of 50%/100%/all-in has 12
// decisions, 10 folds and 11 showdowns, and the 6bb/40bb turn+river game of
// 33%/75%/2.5x/all-in has per turn card 16 decisions, 14 folds, 6 all-in
// showdowns and 9 lines to the river.
#include "hu_tree.h"

#include <cmath>
#include <cstring>
#include <functional>

#include "mtt_common.h"

namespace mtt {

namespace {

struct BState {
    int street;
    int actor;
    double c[2];
    double bet[2];
    int raises;
    int checks;
};

struct Builder {
    const mtt_hu_game& g;
    std::vector<Template>& T;
    double eps = 1e-6;
    explicit Builder(const mtt_hu_game& g_, std::vector<Template>& t) : g(g_), T(t) {}

    struct Act { int kind; double amt; };  // 0 check 1 bet 2 allin 3 fold 4 call 5 raise

    std::vector<Act> actions(const BState& s) {
        const mtt_street_menu& m = g.menu[s.street];
        int a = s.actor, o = 1 - a;
        double S = g.stack, pot = g.pot + s.c[0] + s.c[1];
        std::vector<Act> v;
        if (s.bet[0] == s.bet[1]) {
            v.push_back({0, 0});
            for (int k = 0; k < m.n_sizes; ++k) {
                double amt = (double)m.sizes[k] * pot;
                if (amt <= eps || s.c[a] + amt >= S - eps) continue;
                bool dup = false;
                for (auto& x : v) dup |= (x.kind == 1 && std::fabs(x.amt - amt) < eps);
                if (!dup) v.push_back({1, amt});
            }
            if (m.allin && s.c[a] < S - eps) v.push_back({2, 0});
        } else {
            v.push_back({3, 0});
            v.push_back({4, 0});
            bool can_raise = s.raises < m.raise_cap && s.c[o] < S - eps && s.c[a] < S - eps;
            if (can_raise && m.raise_mult > 0) {
                double to = (double)m.raise_mult * s.bet[o];
                if (s.c[a] - s.bet[a] + to < S - eps) v.push_back({5, to});
            }
            if (can_raise && m.allin) v.push_back({2, 0});
        }
        return v;
    }

    // Template for a street starting with commitments c; returns its index.
    int make_template(int street, double c0, double c1) {
        int id = (int)T.size();
        T.emplace_back();
        T[id].street = street;
        T[id].c0 = (float)c0;
        T[id].c1 = (float)c1;
        BState root{street, 0, {c0, c1}, {0, 0}, 0, 0};
        T[id].nodes.push_back(TNode{});
        fill(id, 0, root, -1, -1, 0);
        // leaf templates are built after this street is complete
        Template& t = T[id];
        std::vector<std::pair<double, double>> leaf_c;
        for (int l = 0; l < t.n_leaf; ++l) {
            const TNode& n = t.nodes[t.leaf_node[l]];
            leaf_c.push_back({n.c[0], n.c[1]});
        }
        std::vector<int> nexts;
        for (auto& lc : leaf_c) nexts.push_back(make_template(street + 1, lc.first, lc.second));
        T[id].leaf_next = nexts;
        return id;
    }

    void leaf_or_showdown(int id, int idx, const BState& s, bool allin) {
        Template& t = T[id];
        TNode& n = t.nodes[idx];
        if (allin || s.street == g.last_street) {
            n.type = T_SHOWDOWN;
            n.term = (int16_t)t.n_showdown++;
            t.showdown_node.push_back(idx);
        } else {
            n.type = T_LEAF;
            n.term = (int16_t)t.n_leaf++;
            t.leaf_node.push_back(idx);
        }
    }

    void fill(int id, int idx, const BState& s, int parent, int act, int depth) {
        {
            Template& t = T[id];
            TNode& n = t.nodes[idx];
            n.parent = (int16_t)parent;
            n.act = (int16_t)act;
            n.depth = (uint8_t)depth;
            n.pot = (float)(g.pot + s.c[0] + s.c[1]);
            n.c[0] = (float)s.c[0];
            n.c[1] = (float)s.c[1];
            if (depth > t.max_depth) t.max_depth = depth;
        }
        std::vector<Act> acts = actions(s);
        int first;
        {
            Template& t = T[id];
            TNode& n = t.nodes[idx];
            n.type = T_DECISION;
            n.player = (uint8_t)s.actor;
            n.nact = (uint8_t)acts.size();
            n.slot = (int16_t)t.n_slots;
            t.n_slots += (int)acts.size();
            t.n_decisions++;
            first = (int)t.nodes.size();
            n.first_child = (int16_t)first;
            t.nodes.resize(t.nodes.size() + acts.size());
        }
        int a = s.actor, o = 1 - a;
        for (size_t k = 0; k < acts.size(); ++k) {
            int ci = first + (int)k;
            BState c = s;
            const Act& x = acts[k];
            auto set_common = [&](const BState& cs) {
                Template& t = T[id];
                TNode& n = t.nodes[ci];
                n.parent = (int16_t)idx;
                n.act = (int16_t)k;
                n.depth = (uint8_t)(depth + 1);
                n.pot = (float)(g.pot + cs.c[0] + cs.c[1]);
                n.c[0] = (float)cs.c[0];
                n.c[1] = (float)cs.c[1];
                n.nact = 0;
                n.first_child = -1;
                n.slot = -1;
                if (depth + 1 > t.max_depth) t.max_depth = depth + 1;
            };
            switch (x.kind) {
            case 0:  // check
                if (s.checks == 1 || a == 1) {
                    set_common(c);
                    leaf_or_showdown(id, ci, c, false);
                } else {
                    c.actor = o;
                    c.checks = s.checks + 1;
                    fill(id, ci, c, idx, (int)k, depth + 1);
                }
                break;
            case 1:  // bet
                c.c[a] += x.amt;
                c.bet[a] += x.amt;
                c.actor = o;
                fill(id, ci, c, idx, (int)k, depth + 1);
                break;
            case 5:  // sized raise
                c.c[a] += x.amt - c.bet[a];
                c.bet[a] = x.amt;
                c.raises++;
                c.actor = o;
                fill(id, ci, c, idx, (int)k, depth + 1);
                break;
            case 2: {  // all-in (a raise when facing a bet)
                double add = g.stack - c.c[a];
                c.c[a] = g.stack;
                c.bet[a] += add;
                if (s.bet[o] > s.bet[a]) c.raises++;
                c.actor = o;
                fill(id, ci, c, idx, (int)k, depth + 1);
                break;
            }
            case 3: {  // fold: the opponent wins
                set_common(c);
                Template& t = T[id];
                TNode& n = t.nodes[ci];
                n.type = T_FOLD;
                n.player = (uint8_t)o;
                n.term = (int16_t)t.n_fold++;
                t.fold_node.push_back(ci);
                break;
            }
            case 4: {  // call
                c.c[a] = c.c[o];
                c.bet[a] = c.bet[o];
                set_common(c);
                leaf_or_showdown(id, ci, c, c.c[a] >= g.stack - eps);
                break;
            }
            }
        }
    }
};

}  // namespace

int matrix_index(int n_deal_streets, int dealt, int t, int rk) {
    if (n_deal_streets == 0) return 0;
    if (dealt == 1) return t;
    return 49 + t * 48 + rk;
}

int build_hu_tree(const mtt_hu_game& g, HuTree& out, bool flat) {
    out = HuTree{};
    out.game = g;
    MTT_ARG(g.n_board >= 3 && g.n_board <= 5, "hu game: n_board %d not in 3..5", g.n_board);
    uint64_t seen = 0;
    for (int k = 0; k < g.n_board; ++k) {
        MTT_ARG(g.board[k] < 52, "hu game: board card %d out of range", (int)g.board[k]);
        MTT_ARG(!(seen & (1ull << g.board[k])), "hu game: repeated board card");
        seen |= 1ull << g.board[k];
    }
    MTT_ARG(g.root_chance == 0 || g.root_chance == 1, "hu game: root_chance must be 0 or 1");
    MTT_ARG(!(g.root_chance && g.n_board == 5), "hu game: no card left to deal at the root");
    int first = g.n_board - 3 + g.root_chance;
    MTT_ARG(g.last_street >= first && g.last_street <= 2, "hu game: last_street %d before first street %d",
            g.last_street, first);
    MTT_ARG(g.pot > 0 && g.stack > 0 && std::isfinite(g.pot) && std::isfinite(g.stack), "hu game: bad pot/stack");
    // streets beyond the known board need chance nodes; only turn and river
    // cards are dealt (from a flop board), which bounds the expansion.
    MTT_ARG(!(g.n_board == 3 && !g.root_chance && g.last_street > 0),
            "hu game: a flop betting street followed by more streets is not supported (use last_street=0)");
    for (int s = first; s <= g.last_street; ++s) {
        const mtt_street_menu& m = g.menu[s];
        MTT_ARG(m.n_sizes >= 0 && m.n_sizes <= 3, "hu game: street %d has %d sizes (max 3)", s, m.n_sizes);
        for (int k = 0; k < m.n_sizes; ++k)
            MTT_ARG(m.sizes[k] > 0 && std::isfinite(m.sizes[k]), "hu game: bad bet size");
        MTT_ARG(m.raise_cap >= 0 && m.raise_cap <= 8, "hu game: raise cap out of range");
        MTT_ARG(m.raise_mult == 0 || (m.raise_mult > 1 && std::isfinite(m.raise_mult)),
                "hu game: raise multiple must be 0 or a finite value above 1");
    }
    out.first_street = first;
    out.board_cards = g.n_board;
    Builder b(g, out.templates);
    b.make_template(first, 0.0, 0.0);
    for (auto& t : out.templates)
        for (auto& n : t.nodes)
            MTT_ARG(n.type != T_DECISION || n.nact <= MTT_MAX_ACTIONS, "hu game: %d actions exceed %d", n.nact,
                    MTT_MAX_ACTIONS);
    if (!flat) return MTT_OK;

    // ---- flat expansion (children contiguous, allocated per parent) ----
    // Deck = cards not on the root board, ascending.  d1/d2 are the deck
    // indices of the cards dealt below the root on the current path.
    const int deck_n = 52 - g.n_board;
    const bool turn_river = g.n_board == 3 && g.root_chance;  // flop board, root deals the turn
    if (turn_river) out.n_matrices = g.last_street == 2 ? 49 + 49 * 48 : 49;
    else if (g.n_board == 4 && g.last_street == 2) out.n_matrices = g.root_chance ? 48 : 49;
    else out.n_matrices = 1;
    auto matrix_of = [&](int nd, int d1, int d2) -> int {
        if (turn_river) return nd == 1 ? d1 : 49 + d1 * 48 + (d2 < d1 ? d2 : d2 - 1);
        if (g.n_board == 4 && g.last_street == 2) return nd == 0 ? 48 : d1;
        return 0;
    };
    std::vector<mtt_flat_node>& F = out.flat;
    F.reserve(1024);
    int infosets = 0;
    std::function<void(int, int, int, int, int, int)> expand = [&](int tid, int ni, int fi, int nd, int d1, int d2) {
        const Template& tp = out.templates[tid];
        const TNode& n = tp.nodes[ni];
        mtt_flat_node x;
        memset(&x, 0, sizeof(x));
        x.pot_size = n.pot;
        x.committed[0] = n.c[0];
        x.committed[1] = n.c[1];
        x.turn_idx = (int16_t)d1;
        x.river_idx = (int16_t)((d1 >= 0 ? d1 + 1 : 0) + (d2 >= 0 ? 64 * (d2 + 1) : 0));
        x.children_offset = -1;
        x.infoset_offset = -1;
        x.active_mask = 3;
        if ((int)out.flat_tmpl.size() < (int)F.size()) {
            out.flat_tmpl.resize(F.size(), -1);
            out.flat_tnode.resize(F.size(), -1);
        }
        out.flat_tmpl[fi] = tid;
        out.flat_tnode[fi] = ni;
        if (n.type == T_DECISION) {
            x.type = MTT_NODE_DECISION;
            x.player = n.player;
            x.num_actions = n.nact;
            x.infoset_offset = infosets++;
            int co = (int)F.size();
            x.children_offset = co;
            F[fi] = x;
            F.resize(F.size() + n.nact);
            for (int a = 0; a < n.nact; ++a) expand(tid, n.first_child + a, co + a, nd, d1, d2);
            return;
        }
        if (n.type == T_FOLD) {
            x.type = MTT_NODE_FOLD;
            x.player = n.player;
            x.active_mask = (uint8_t)(1u << n.player);
            F[fi] = x;
            return;
        }
        if (n.type == T_SHOWDOWN) {
            x.type = MTT_NODE_SHOWDOWN;
            // turn_idx = equity-matrix index (honest_multistreet_solver.cu:172)
            x.turn_idx = (int16_t)matrix_of(nd, d1, d2);
            F[fi] = x;
            return;
        }
        // T_LEAF: chance node dealing the next street
        int next = tp.leaf_next[n.term];
        x.type = MTT_NODE_CHANCE;
        x.player = 0;
        int nchild = deck_n - nd;
        x.num_actions = (uint8_t)nchild;
        x.turn_idx = (int16_t)(nd >= 1 ? d1 : -1);
        int co = (int)F.size();
        x.children_offset = co;
        F[fi] = x;
        F.resize(F.size() + nchild);
        for (int k = 0; k < nchild; ++k) {
            int di = (nd >= 1 && k >= d1) ? k + 1 : k;
            if (nd == 0) expand(next, 0, co + k, 1, di, -1);
            else expand(next, 0, co + k, 2, d1, di);
        }
    };
    F.resize(1);
    if (g.root_chance) {
        mtt_flat_node x;
        memset(&x, 0, sizeof(x));
        x.type = MTT_NODE_CHANCE;
        x.active_mask = 3;
        x.pot_size = g.pot;
        x.num_actions = (uint8_t)deck_n;
        x.turn_idx = -1;
        x.river_idx = 0;
        x.infoset_offset = -1;
        x.children_offset = 1;
        F[0] = x;
        F.resize(1 + deck_n);
        for (int k = 0; k < deck_n; ++k) expand(0, 0, 1 + k, 1, k, -1);
    } else {
        expand(0, 0, 0, 0, -1, -1);
    }
    out.n_infosets = infosets;
    return MTT_OK;
}

}  // namespace mtt

// ================================================================ C ABI ===
extern "C" {

int mtt_hu_tree_size(const mtt_hu_game* g, int* n_nodes, int* n_infosets, int* n_matrices) {
    MTT_ARG(g && n_nodes && n_infosets, "mtt_hu_tree_size: null pointer");
    mtt::HuTree t;
    int rc = mtt::build_hu_tree(*g, t);
    if (rc) return rc;
    *n_nodes = (int)t.flat.size();
    *n_infosets = t.n_infosets;
    if (n_matrices) *n_matrices = t.n_matrices;
    return MTT_OK;
}

int mtt_hu_tree_build(const mtt_hu_game* g, mtt_flat_node* nodes, int capacity, int* n_nodes) {
    MTT_ARG(g && nodes && n_nodes, "mtt_hu_tree_build: null pointer");
    mtt::HuTree t;
    int rc = mtt::build_hu_tree(*g, t);
    if (rc) return rc;
    MTT_ARG((int)t.flat.size() <= capacity, "mtt_hu_tree_build: capacity %d < %d nodes", capacity, (int)t.flat.size());
    memcpy(nodes, t.flat.data(), sizeof(mtt_flat_node) * t.flat.size());
    *n_nodes = (int)t.flat.size();
    return MTT_OK;
}

}  // extern "C"
