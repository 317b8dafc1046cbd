// 4-way single-street tree builder (the reference's 4-way engine takes a
// flat tree from an absent host, multiway_4way_engine.cu:20-24,41-257).
//
// Betting rules (project-defined):
//   players act in seat order 0..3, starting with seat 0;
//   the actor's cap is the smaller of its own stack and the largest stack
//   of another player still in the hand (chips nobody can call are never
//   bet);
//   unopened  : check, each pot-fraction bet (a size at or above the cap
//               becomes a bet of the cap, i.e. all-in), all-in when the menu
//               offers it;
//   facing bet: fold, call (all-in for less when the stack is short of the
//               bet: the excess of the others forms side pots), and - while
//               fewer than raise_cap raises were made - a raise to
//               raise_mult x the current bet (the cap when that is at or
//               above it) and an all-in raise when offered;
//   a player with no chips behind no longer acts, nor does one facing no
//   bet when nobody else in the hand can still act;
//   the street closes when every player still in the hand has acted since
//   the last bet or raise: one player left is a fold terminal (active_mask
//   = the winner's bit), otherwise a showdown among the players left
//   (active_mask = their bits).
// Nodes are laid out breadth first with the children of a decision node
// contiguous (children_offset + a), one infoset per decision node.
#include <cmath>
#include <cstring>
#include <deque>
#include <vector>

#include "mtt_common.h"

namespace mtt {

namespace {

struct S4 {
    double bet[4];   // street commitments
    int active;      // players still in the hand
    int need;        // players who still have to act
    int actor;
    int raises;
};

struct Builder4 {
    const mtt_4way_game& g;
    double eps = 1e-6;
    std::vector<mtt_flat_node4> nodes;
    int n_info = 0;
    explicit Builder4(const mtt_4way_game& g_) : g(g_) {}

    // Largest commitment the actor can make that someone else can match.
    double cap(const S4& s, int a) const {
        double other = 0.0;
        for (int p = 0; p < 4; ++p)
            if (p != a && (s.active & (1 << p))) other = std::fmax(other, (double)g.stack[p]);
        return std::fmin((double)g.stack[a], other);
    }
    double level(const S4& s) const {
        double l = 0.0;
        for (int p = 0; p < 4; ++p) l = std::fmax(l, s.bet[p]);
        return l;
    }
    double pot(const S4& s) const {
        double v = 0.0;
        for (int p = 0; p < 4; ++p) v += (double)g.pot0[p] + s.bet[p];
        return v;
    }
    bool can_act(const S4& s, int p) const {
        return (s.active & (1 << p)) && (double)g.stack[p] - s.bet[p] > eps;
    }
    // Someone other than p still has chips behind.
    bool others_can_act(const S4& s, int p) const {
        for (int o = 0; o < 4; ++o)
            if (o != p && can_act(s, o)) return true;
        return false;
    }
    int next_actor(const S4& s, int from) const {
        for (int k = 1; k <= 4; ++k) {
            int p = (from + k) & 3;
            if (!(s.need & (1 << p)) || !can_act(s, p)) continue;
            if (level(s) <= s.bet[p] + eps && !others_can_act(s, p)) continue;  // nothing left to contest
            return p;
        }
        return -1;
    }

    struct Act { int kind; double to; };  // 0 check, 1 bet/raise to `to`, 2 fold, 3 call

    std::vector<Act> actions(const S4& s) const {
        const mtt_street_menu& m = g.menu;
        const int a = s.actor;
        const double L = level(s), C = cap(s, a), P = pot(s);
        std::vector<Act> v;
        auto add_to = [&](double to) {
            if (to > C) to = C;                       // at or above the cap: all-in
            if (to <= L + eps) return;
            for (auto& x : v)
                if (x.kind == 1 && std::fabs(x.to - to) < eps) return;
            v.push_back({1, to});
        };
        if (L <= s.bet[a] + eps) {
            v.push_back({0, 0});
            for (int k = 0; k < m.n_sizes; ++k) add_to((double)m.sizes[k] * P);
            if (m.allin) add_to(C);
        } else {
            v.push_back({2, 0});
            v.push_back({3, 0});
            if (s.raises < m.raise_cap && C > L + eps) {
                if (m.raise_mult > 0) add_to((double)m.raise_mult * L);
                if (m.allin) add_to(C);
            }
        }
        return v;
    }

    void fill_node(mtt_flat_node4& n, const S4& s) const {
        n.pot_size = (float)pot(s);
        for (int p = 0; p < 4; ++p) n.committed[p] = (float)((double)g.pot0[p] + s.bet[p]);
    }

    int build() {
        std::deque<std::pair<S4, int>> q;
        S4 root;
        memset(&root, 0, sizeof(root));
        root.active = 0xF;
        root.need = 0;
        for (int p = 0; p < 4; ++p)
            if (can_act(root, p)) root.need |= 1 << p;
        root.actor = next_actor(root, 3);
        nodes.emplace_back();
        q.push_back({root, 0});
        while (!q.empty()) {
            S4 s = q.front().first;
            int id = q.front().second;
            q.pop_front();
            mtt_flat_node4 n;
            memset(&n, 0, sizeof(n));
            fill_node(n, s);
            int nact = __builtin_popcount((unsigned)s.active);
            if (nact == 1) {
                n.type = MTT_NODE_FOLD;
                n.active_mask = (uint8_t)s.active;
                n.children_offset = -1;
                n.infoset_offset = -1;
                nodes[id] = n;
                continue;
            }
            if (s.actor < 0) {
                n.type = MTT_NODE_SHOWDOWN;
                n.active_mask = (uint8_t)s.active;
                n.children_offset = -1;
                n.infoset_offset = -1;
                nodes[id] = n;
                continue;
            }
            std::vector<Act> acts = actions(s);
            if (acts.size() > 3) return -1;
            n.type = MTT_NODE_DECISION;
            n.player = (uint8_t)s.actor;
            n.num_actions = (uint8_t)acts.size();
            n.active_mask = (uint8_t)s.active;
            n.infoset_offset = n_info++;
            n.children_offset = (int)nodes.size();
            nodes[id] = n;
            for (size_t k = 0; k < acts.size(); ++k) nodes.emplace_back();
            for (size_t k = 0; k < acts.size(); ++k) {
                S4 c = s;
                const int a = s.actor;
                const Act& x = acts[k];
                c.need &= ~(1 << a);
                if (x.kind == 2) {
                    c.active &= ~(1 << a);
                } else if (x.kind == 3) {
                    c.bet[a] = std::fmin(level(s), (double)g.stack[a]);   // all-in for less when short
                } else if (x.kind == 1) {
                    if (level(s) > eps) c.raises++;  // a raise over a bet (the opening bet is not one)
                    c.bet[a] = x.to;
                    c.need = 0;
                    for (int p = 0; p < 4; ++p)
                        if (p != a && can_act(c, p)) c.need |= 1 << p;
                }
                for (int p = 0; p < 4; ++p)
                    if (!can_act(c, p)) c.need &= ~(1 << p);
                c.actor = __builtin_popcount((unsigned)c.active) > 1 ? next_actor(c, a) : -1;
                q.push_back({c, n.children_offset + (int)k});
            }
        }
        return (int)nodes.size();
    }
};

int check_game4(const mtt_4way_game* g) {
    MTT_ARG(g, "4-way game: null");
    for (int p = 0; p < 4; ++p) {
        MTT_ARG(g->stack[p] > 0 && std::isfinite(g->stack[p]), "4-way game: stack %d must be positive", p);
        MTT_ARG(g->pot0[p] >= 0 && std::isfinite(g->pot0[p]), "4-way game: pot share %d must be >= 0", p);
    }
    const mtt_street_menu& m = g->menu;
    MTT_ARG(m.n_sizes >= 0 && m.n_sizes <= 3, "4-way game: n_sizes out of range");
    for (int k = 0; k < m.n_sizes; ++k)
        MTT_ARG(m.sizes[k] > 0 && std::isfinite(m.sizes[k]), "4-way game: bet size %d must be finite and > 0", k);
    MTT_ARG(m.raise_cap >= 0 && m.raise_cap <= 4, "4-way game: raise cap out of range");
    MTT_ARG(m.raise_mult == 0 || (m.raise_mult > 1 && std::isfinite(m.raise_mult)),
            "4-way game: raise multiple must be 0 or a finite value above 1");
    return MTT_OK;
}

}  // namespace

int build_4way_tree(const mtt_4way_game& g, std::vector<mtt_flat_node4>& out, int& n_info) {
    int rc = check_game4(&g);
    if (rc) return rc;
    Builder4 b(g);
    int n = b.build();
    if (n < 0) {
        set_error("4-way game: a decision has more than 3 actions (the 4-way action stride)");
        return MTT_ERR_TREE;
    }
    out = std::move(b.nodes);
    n_info = b.n_info;
    return MTT_OK;
}

}  // namespace mtt

using namespace mtt;

extern "C" {

int mtt_4way_tree_size(const mtt_4way_game* game, int* n_nodes, int* n_infosets) {
    MTT_ARG(game && n_nodes && n_infosets, "mtt_4way_tree_size: null pointer");
    std::vector<mtt_flat_node4> v;
    int ni = 0;
    int rc = build_4way_tree(*game, v, ni);
    if (rc) return rc;
    *n_nodes = (int)v.size();
    *n_infosets = ni;
    return MTT_OK;
}

int mtt_4way_tree_build(const mtt_4way_game* game, mtt_flat_node4* nodes, int capacity, int* n_nodes) {
    MTT_ARG(game && nodes && n_nodes, "mtt_4way_tree_build: null pointer");
    std::vector<mtt_flat_node4> v;
    int ni = 0;
    int rc = build_4way_tree(*game, v, ni);
    if (rc) return rc;
    MTT_ARG((int)v.size() <= capacity, "mtt_4way_tree_build: capacity %d < %zu nodes", capacity, v.size());
    memcpy(nodes, v.data(), v.size() * sizeof(mtt_flat_node4));
    *n_nodes = (int)v.size();
    return MTT_OK;
}

}  // extern "C"
