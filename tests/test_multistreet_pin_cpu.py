"""Pins of the oracle's multi-street (turn + river) recurrences by a brute
force that shares none of its code.

The oracle (oracle/or_cfr_impl.h, oracle/or_cfr.c) evaluates a chance node
by summing its children over the cards a combo does not hold and scaling by
one constant per street (honest_multistreet_solver.cu:205-219 divides by 47;
the corrected form divides by the exact number of cards left given both
hands).  Here the game is instead evaluated over explicit deals: every
(opponent hand, turn card, river card) with the probability of each card
counted from the deck given the board and both hands, an explicit max per
(hand, public history) for a best responder, and the 7-card evaluator only
for showdown outcomes.  A wrong chance divisor, a dropped card-removal term
or a swapped payoff changes the values below.

Tiny ranges (a few combos per player, with shared cards so that card removal
matters) keep the enumeration to seconds; the tree is the benchmark's
turn+river street structure with an all-in-only menu (21,610 nodes).
"""
import numpy as np
import pytest

N_PAD = 1216
A_STRIDE = 4


class Brute:
    """Explicit-deal evaluation of a heads-up FlatNode tree rooted at a turn
    chance node over a 3-card flop."""

    def __init__(self, oracle, nodes, flop, hands, weights, strat, alive_pos):
        self.O = oracle
        self.nodes = nodes
        self.flop = [int(c) for c in flop]
        self.hands = hands            # [p] -> list of (c1, c2)
        self.w = weights              # [p] -> np.array
        self.strat = strat            # reference layout [(info*4 + a)*1216 + alive]
        self.pos = alive_pos          # [p] -> alive index of each hand
        self.score_cache = {}
        n0, n1 = len(hands[0]), len(hands[1])
        self.disjoint = np.array([[not (set(hands[0][i]) & set(hands[1][j])) for j in range(n1)]
                                  for i in range(n0)])

    # ---------------------------------------------------------- primitives
    def holds(self, p, card):
        return np.array([card in h for h in self.hands[p]])

    def cards_left(self, board):
        """Per pair: number of deck cards neither on the board nor in either hand."""
        used = set(board)
        n0, n1 = len(self.hands[0]), len(self.hands[1])
        out = np.zeros((n0, n1))
        for i in range(n0):
            for j in range(n1):
                out[i, j] = len(set(range(52)) - used - set(self.hands[0][i]) - set(self.hands[1][j]))
        return out

    def scores(self, board, p):
        key = (tuple(sorted(board)), p)
        if key not in self.score_cache:
            cards = np.array([[*board, a, b] for a, b in self.hands[p]], np.uint8)
            sc = self.O.eval7_batch(cards).astype(np.int64)
            bad = np.array([a in board or b in board for a, b in self.hands[p]])
            sc[bad] = -1          # never used: such pairs carry zero weight
            self.score_cache[key] = sc
        return self.score_cache[key]

    def showdown_payoffs(self, board, pot, c0, c1):
        """Chips won by each player for every pair on a 5-card board."""
        s0 = self.scores(board, 0)[:, None]
        s1 = self.scores(board, 1)[None, :]
        win0 = (s0 > s1).astype(float)
        tie = (s0 == s1).astype(float)
        win1 = (s1 > s0).astype(float)
        return pot * (win0 + 0.5 * tie) - c0, pot * (win1 + 0.5 * tie) - c1

    # ------------------------------------------------------------ recursion
    def rec(self, node, board, W, mode, record):
        """Returns (V0 over player 0's hands, V1 over player 1's hands).
        W[i, j] = weight of the pair (i, j) reaching this node: both range
        weights, the probabilities of the cards dealt, and the strategy
        products the mode includes."""
        x = self.nodes[node]
        t = int(x["type"])
        if t == 1:                                     # fold
            winner = 0 if int(x["active_mask"]) == 1 else 1
            P = float(x["pot_size"])
            c = [float(x["committed"][0]), float(x["committed"][1])]
            pay0 = (P - c[0]) if winner == 0 else -c[0]
            pay1 = (P - c[1]) if winner == 1 else -c[1]
            return W.sum(1) * pay0, W.sum(0) * pay1
        if t == 2:                                     # showdown
            P = float(x["pot_size"])
            c0, c1 = float(x["committed"][0]), float(x["committed"][1])
            if len(board) == 5:
                p0, p1 = self.showdown_payoffs(board, P, c0, c1)
            else:                                      # all-in before the river: every river, counted
                p0 = np.zeros(W.shape)
                p1 = np.zeros(W.shape)
                left = self.cards_left(board)
                for card in range(52):
                    if card in board:
                        continue
                    ok = ~self.holds(0, card)[:, None] & ~self.holds(1, card)[None, :]
                    q0, q1 = self.showdown_payoffs(board + [card], P, c0, c1)
                    prob = np.where(ok, 1.0 / np.maximum(left, 1), 0.0)
                    p0 += prob * q0
                    p1 += prob * q1
            return (W * p0).sum(1), (W * p1).sum(0)
        kids = int(x["children_offset"])
        na = int(x["num_actions"])
        if t == 3:                                     # chance: remaining cards in ascending order
            rest = [c for c in range(52) if c not in board]
            assert len(rest) == na, (node, len(rest), na)
            left = self.cards_left(board)
            V0 = np.zeros(W.shape[0])
            V1 = np.zeros(W.shape[1])
            for k, card in enumerate(rest):
                ok = ~self.holds(0, card)[:, None] & ~self.holds(1, card)[None, :]
                Wc = W * np.where(ok, 1.0 / np.maximum(left, 1), 0.0)
                if not Wc.any():
                    continue
                a0, a1 = self.rec(kids + k, board + [card], Wc, mode, record)
                V0 += a0
                V1 += a1
            return V0, V1
        # decision
        p = int(x["player"])
        info = int(x["infoset_offset"])
        sig = np.array([[self.strat[(info * A_STRIDE + a) * N_PAD + k] for k in self.pos[p]] for a in range(na)],
                       np.float64)
        kind, who = mode
        child = []
        for a in range(na):
            if kind == "profile" or p != who:
                Wa = W * (sig[a][:, None] if p == 0 else sig[a][None, :])
            else:
                Wa = W
            child.append(self.rec(kids + a, board, Wa, mode, record))
        if record is not None and node in record:
            record[node] = [c[p].copy() for c in child]
        V0 = np.zeros(W.shape[0])
        V1 = np.zeros(W.shape[1])
        for a in range(na):
            if kind == "br" and p == who:
                own = child[a][p]
                if a == 0:
                    best = own.copy()
                else:
                    best = np.maximum(best, own)
            elif kind == "cfv" and p == who:
                (V0 if p == 0 else V1)[:] += sig[a] * child[a][p]
            else:
                (V0 if p == 0 else V1)[:] += child[a][p]
            (V1 if p == 0 else V0)[:] += child[a][1 - p]
        if kind == "br" and p == who:
            (V0 if p == 0 else V1)[:] = best
        return V0, V1

    def root_values(self, mode, record=None):
        W = self.w[0][:, None] * self.w[1][None, :] * self.disjoint
        return self.rec(0, list(self.flop), W, mode, record)


def _tiny_game(solver_api, seed, n_hands=6):
    """Benchmark street structure (turn chance, turn betting, river chance,
    river betting), all-in-only menu; n_hands combos per player with
    random weights, drawn so that the two ranges share cards."""
    g, _, _ = solver_api.config3_small(seed)
    nodes, ni, _ = solver_api.tree(g)
    flop = [int(c) for c in list(g.board)[:3]]
    gi, cards = solver_api.alive_index(flop)
    rng = np.random.default_rng(seed + 100)
    live = [c for c in range(52) if c not in flop]
    pool = rng.choice(live, 7, replace=False).tolist()   # few cards: many blocked pairs
    pairs = sorted({tuple(sorted(rng.choice(pool, 2, replace=False).tolist())) for _ in range(200)})
    rng.shuffle(pairs)
    hands = [pairs[:n_hands], pairs[n_hands:2 * n_hands] or pairs[:n_hands]]
    alive_of = {(int(a), int(b)): k for k, (a, b) in enumerate(cards)}
    pos = [[alive_of[h] for h in hands[p]] for p in range(2)]
    w = [rng.random(len(hands[p])) + 0.1 for p in range(2)]
    r = [np.zeros(len(gi)), np.zeros(len(gi))]
    for p in range(2):
        r[p][pos[p]] = w[p]
    return g, nodes, ni, flop, hands, w, pos, r


def _random_strategy(nodes, ni, n_alive, rng, chunk=1 << 14):
    """Random strategies in the reference layout [(info*4 + a)*1216 + combo]:
    each probability is k/256 with k >= 1 and the k of an infoset's actions
    adding to 256, so a combo's probabilities are exact in FP32 and sum to
    exactly 1 (the constant-sum identity then holds to FP64 rounding)."""
    dec = nodes[nodes["type"] == 0]
    na = np.zeros(ni, np.int64)
    na[dec["infoset_offset"]] = dec["num_actions"]
    st = np.zeros((ni, A_STRIDE, N_PAD), np.float32)
    for lo in range(0, ni, chunk):
        hi = min(ni, lo + chunk)
        n = na[lo:hi]
        live = np.arange(A_STRIDE)[None, :, None] < n[:, None, None]
        raw = rng.integers(1, 64, (hi - lo, A_STRIDE, n_alive)).astype(np.int64) * live
        tot = np.maximum(raw.sum(1, keepdims=True), 1)
        k = np.maximum((raw * 256) // tot, live.astype(np.int64))
        k[:, :1] += np.where(n[:, None, None] > 0, 256 - k.sum(1, keepdims=True), 0)
        st[lo:hi, :, :n_alive] = k.astype(np.float32) / np.float32(256)
    assert (st >= 0).all()
    return st.reshape(-1)


@pytest.mark.parametrize("seed", [0, 3])
def test_best_response_and_profile_values_brute_force(solver_api, oracle, seed):
    """or_hu_exploitability's BR values, profile values and pair mass equal
    the explicit-deal evaluation under a random strategy profile."""
    g, nodes, ni, flop, hands, w, pos, r = _tiny_game(solver_api, seed)
    h = oracle.HU(nodes, ni, flop, r[0].astype(np.float32), r[1].astype(np.float32), 0)
    st = _random_strategy(nodes, ni, h.N, np.random.default_rng(seed))
    w32 = [r[p][pos[p]].astype(np.float32).astype(np.float64) for p in range(2)]
    bf = Brute(oracle, nodes, flop, hands, w32, st, pos)
    out = h.exploitability(st, float(g.pot))
    br0 = bf.root_values(("br", 0))[0].sum()
    br1 = bf.root_values(("br", 1))[1].sum()
    v0, v1 = bf.root_values(("profile", -1))
    mass = (w32[0][:, None] * w32[1][None, :] * bf.disjoint).sum()
    scale = float(g.pot) * mass
    assert abs(out[4] - mass) <= 1e-12 * mass
    assert abs(out[2] - v0.sum()) <= 1e-10 * scale, (out[2], v0.sum())
    assert abs(out[3] - v1.sum()) <= 1e-10 * scale, (out[3], v1.sum())
    assert abs(out[0] - br0) <= 1e-10 * scale, (out[0], br0)
    assert abs(out[1] - br1) <= 1e-10 * scale, (out[1], br1)
    # the brute force itself is constant-sum: chips won add up to the starting pot
    assert abs(v0.sum() + v1.sum() - scale) <= 1e-10 * scale
    assert br0 >= v0.sum() - 1e-12 * scale and br1 >= v1.sum() - 1e-12 * scale


def test_first_iteration_regrets_brute_force(solver_api, oracle):
    """After one CFR+ iteration (uniform strategies) the regrets at turn
    decisions equal the positive parts of the explicit-deal counterfactual
    action values minus the node value.  Counterfactual values below a dealt
    turn card are conditional on that card: the brute force starts at the
    turn decision with the turn card on the board and the pairs that hold
    it removed, and deals the river itself.  A check leads to river lines
    and an all-in to turn-level folds and all-in showdowns, so a wrong river
    chance step changes the regrets."""
    g, nodes, ni, flop, hands, w, pos, r = _tiny_game(solver_api, 1)
    h = oracle.HU(nodes, ni, flop, r[0].astype(np.float32), r[1].astype(np.float32), 0)
    h.iterate(1)
    R, _ = h.state()
    uniform = np.zeros(ni * A_STRIDE * N_PAD, np.float32)
    for x in nodes[nodes["type"] == 0]:
        k, na = int(x["infoset_offset"]), int(x["num_actions"])
        for a in range(na):
            uniform[(k * A_STRIDE + a) * N_PAD:(k * A_STRIDE + a) * N_PAD + h.N] = 1.0 / na
    w32 = [r[p][pos[p]].astype(np.float32).astype(np.float64) for p in range(2)]
    bf = Brute(oracle, nodes, flop, hands, w32, uniform, pos)
    root = nodes[0]
    turn_cards = [c for c in range(52) if c not in flop]
    assert len(turn_cards) == int(root["num_actions"])
    checked = 0
    for k in range(0, len(turn_cards), 4):
        n, tc = int(root["children_offset"]) + k, turn_cards[k]
        W = bf.w[0][:, None] * bf.w[1][None, :] * bf.disjoint
        W = W * (~bf.holds(0, tc)[:, None] & ~bf.holds(1, tc)[None, :])
        probe = [n] + [int(nodes[n]["children_offset"]) + a for a in range(int(nodes[n]["num_actions"]))]
        probe = [q for q in probe if nodes[q]["type"] == 0]
        for who in (0, 1):
            rec = {q: None for q in probe if int(nodes[q]["player"]) == who}
            bf.rec(n, list(flop) + [tc], W, ("cfv", who), rec)
            for q, vals in rec.items():
                if vals is None:
                    continue
                info, na = int(nodes[q]["infoset_offset"]), int(nodes[q]["num_actions"])
                # the pair weights include the responder's own range weight;
                # counterfactual values do not
                va = np.array(vals) / bf.w[who][None, :]   # [action][hand]
                reg_bf = np.maximum(va - va.mean(0), 0.0)
                reg_or = np.array([[R[(info * A_STRIDE + a) * N_PAD + c] for c in pos[who]] for a in range(na)],
                                  np.float64)
                scale = max(np.abs(va).max(), 1e-300)
                assert np.allclose(reg_or, reg_bf, rtol=0, atol=1e-6 * scale), (q, who, reg_or, reg_bf)
                checked += int((reg_bf > 1e-3 * scale).sum())
    assert checked >= 20


@pytest.mark.parametrize("which", ["config3_small", pytest.param("config3", marks=pytest.mark.slow)])
def test_constant_sum_multistreet(solver_api, oracle, which):
    """Chips won by the two players under any profile add to the starting
    pot times the pair mass (v0 + v1 = pot * M), for random strategies and for
    the CFR+ average after a few iterations.  A chance step that does not
    average over exactly the cards left breaks the identity."""
    g, r0, r1 = getattr(solver_api, which)(2)
    nodes, ni, _ = solver_api.tree(g)
    board = list(g.board)[:g.n_board]
    gi, _ = solver_api.alive_index(board)
    h = oracle.HU(nodes, ni, board, r0[gi], r1[gi], 0)
    pot = float(g.pot)
    st = _random_strategy(nodes, ni, h.N, np.random.default_rng(7))
    e = h.exploitability(st, pot)
    # the FlatNode layout stores pots and commitments in FP32, so at each
    # terminal pot - c0 - c1 equals the starting pot to FP32 rounding
    assert abs(e[2] + e[3] - pot * e[4]) <= 1e-7 * pot * e[4], (e[2] + e[3], pot * e[4])
    assert e[0] >= e[2] and e[1] >= e[3]
    if which == "config3_small":
        # the CFR+ average strategy is normalised in FP32: its probabilities
        # sum to 1 within a few ulps, so the identity holds to ~1e-6
        h.iterate(3)
        e = h.exploitability(h.strategy(), pot)
        assert abs(e[2] + e[3] - pot * e[4]) <= 1e-6 * pot * e[4]
