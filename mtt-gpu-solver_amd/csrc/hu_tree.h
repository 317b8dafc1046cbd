// Heads-up game trees: bet-menu rules, street templates and the flat
// FlatNode expansion.  The reference ships no tree builder (its HU solver's
// main is a stub, honest_multistreet_solver.cu:251-254); these rules are the
// project's, chosen so that the node counts match the bet trees the
// benchmark configurations describe.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/mtt.h"

namespace mtt {

enum TType : uint8_t { T_DECISION = 0, T_FOLD = 1, T_SHOWDOWN = 2, T_LEAF = 3 };

// One node of a betting-round template.  Children are contiguous.
struct TNode {
    uint8_t type, player, nact, depth;  // player: actor (decision) / winner (fold)
    int16_t parent, act;                // parent node and action index from it
    int16_t first_child;
    int16_t slot;                       // decision: first action slot
    int16_t term;                       // showdown / fold / leaf ordinal
    float pot, c[2];
};

struct Template {
    int street = 0;                     // 0 flop, 1 turn, 2 river
    float c0 = 0, c1 = 0;               // commitments at the template root
    std::vector<TNode> nodes;           // preorder (parent before child)
    int n_slots = 0, n_decisions = 0, n_showdown = 0, n_fold = 0, n_leaf = 0;
    std::vector<int> leaf_next;         // leaf ordinal -> template of the next street
    std::vector<int> showdown_node, fold_node, leaf_node;  // ordinal -> node
    int max_depth = 0;
};

struct HuTree {
    mtt_hu_game game;
    std::vector<Template> templates;    // templates[0] is the first betting street
    int first_street = 0;
    // flat expansion (reference FlatNode layout)
    std::vector<mtt_flat_node> flat;
    std::vector<int32_t> flat_tmpl, flat_tnode;  // template and template node of each flat node
    int n_infosets = 0;
    int n_matrices = 0;                 // matrices referenced by showdown turn_idx
    int board_cards = 0;                // cards known at the root
};

// Validates the game and builds templates plus the flat tree.
// Returns MTT_OK or an error code with mtt_last_error() set.
int build_hu_tree(const mtt_hu_game& g, HuTree& out, bool flat = true);

// Matrix index used by a showdown on a board with `dealt` extra cards:
//   no chance cards: 0;  turn dealt (t): t for a turn showdown (when the
//   river is still to come) or 49 + t*48 + rk for the river board (rk = the
//   river's child index under the turn).
int matrix_index(int n_deal_streets, int dealt, int t, int rk);

}  // namespace mtt
