"""CPU tests of the 4-way engine's pieces: the tree builder, the board
tables, the oracle's multiway terminal values against a brute-force
enumeration of opponent hands, Malmuth-Harville ICM against an enumeration
of finishing orders and closed forms, bubble factors, and the DCFR step
against hand-derived values and the reference's floor form."""
import itertools
import math
import os

import numpy as np
import pytest

from conftest import PKG, _load


@pytest.fixture(scope="module")
def M(mtt):
    return _load("mtt_multi", os.path.join(PKG, "multi.py"))


def _game(M, stacks=(30.0, 45.0, 12.0, 60.0)):
    return M.game4([2.5] * 4, list(stacks), M.config5_menu())


# ------------------------------------------------------------------ tree --
SEED0_BEHIND = (50.272126549109075, 22.734003532290274, 5.573014295214602, 3.739572664639682)  # config 5, seed 0


def _check_tree4(nodes, ni, stacks):
    dec = nodes[nodes["type"] == 0]
    assert dec["num_actions"].min() >= 2 and dec["num_actions"].max() <= 3
    assert np.array_equal(np.sort(dec["infoset_offset"]), np.arange(ni))
    seen = np.zeros(len(nodes), int)
    for n in np.nonzero(nodes["type"] == 0)[0]:
        x = nodes[n]
        p = int(x["player"])
        assert x["committed"][p] - 2.5 < stacks[p] - 1e-6          # the actor has chips behind
        for a in range(x["num_actions"]):
            c = x["children_offset"] + a
            seen[c] += 1
            ch = nodes[c]
            # pot accounting: the pot is the sum of the commitments
            assert abs(ch["pot_size"] - ch["committed"].sum()) < 1e-4
            # only the actor's commitment changes, and never beyond its stack
            d = ch["committed"] - x["committed"]
            assert np.all(np.abs(np.delete(d, p)) < 1e-6)
            assert d[p] >= -1e-6 and ch["committed"][p] - 2.5 <= stacks[p] + 1e-4
    assert seen[0] == 0 and np.all(seen[1:] == 1)  # a tree: every node but the root has one parent
    for x in nodes[nodes["type"] == 1]:
        assert bin(int(x["active_mask"])).count("1") == 1
    n_side = 0
    for x in nodes[nodes["type"] == 2]:
        m = int(x["active_mask"])
        assert bin(m).count("1") >= 2
        live = [p for p in range(4) if m >> p & 1]
        top = max(float(x["committed"][p]) for p in live)
        # everyone still in matched the largest commitment or is all-in for less
        for p in live:
            c = float(x["committed"][p])
            assert abs(c - top) < 1e-4 or abs(c - 2.5 - stacks[p]) < 1e-4
        n_side += len({round(float(x["committed"][p]), 4) for p in live}) > 1
        # no chips above what a player still in can win back
        assert max(float(v) for v in x["committed"]) <= top + 1e-4
    return n_side


def test_tree4_invariants(M):
    stacks = np.array([30.0, 45.0, 12.0, 60.0])
    nodes, ni = M.tree4(_game(M, stacks))
    assert len(nodes) == 677 and ni == 308
    assert _check_tree4(nodes, ni, stacks) > 0          # the 12bb stack is called all-in for less


def test_tree4_config5_seed0_side_pots(M):
    """Config 5's seed-0 stacks (5.57 and 3.74bb behind for seats 2 and 3):
    bets the short stacks cannot match are called all-in for less, leaving
    side pots; sizes at or above a short actor's stack collapse into one
    all-in, so the tree has 470 nodes (677 with deep stacks)."""
    stacks = np.array(SEED0_BEHIND)
    nodes, ni = M.tree4(_game(M, stacks))
    assert len(nodes) == 470 and ni == 211
    assert _check_tree4(nodes, ni, stacks) == 120
    deep, _ = M.tree4(_game(M, (200.0,) * 4))
    assert len(deep) == 677


def test_tree4_check_around_and_fold_out(M):
    nodes, _ = M.tree4(_game(M))
    # check, check, check, check -> showdown of all four at the starting pot
    n = 0
    for _ in range(4):
        assert nodes[n]["type"] == 0
        n = nodes[n]["children_offset"]  # action 0 = check
    assert nodes[n]["type"] == 2 and nodes[n]["active_mask"] == 0xF and abs(nodes[n]["pot_size"] - 10.0) < 1e-6


def test_board_tables_match_oracle(M, oracle):
    for flop in ([0, 1, 2], [12, 25, 51], [3, 20, 44]):
        a, b = M.board4(flop), oracle.board4(flop)
        for k in M.BOARD4_DTYPE.names:
            assert np.array_equal(a[k], b[k]), k
        t = b[0]
        # positions form a permutation consistent with scores
        assert sorted(t["pos_combo"][:1176].tolist()) == list(range(1176))
        sc = np.array([oracle.evaln([*flop, t["c1"][i], t["c2"][i]]) for i in range(1176)])
        assert np.all(np.diff(sc[t["pos_combo"][:1176]]) >= 0)
        assert np.all(t["n_card"][list(flop)] == 0) and np.all(np.delete(t["n_card"], flop) == 48)


# --------------------------------------------------------- terminal values --
def _sparse_reach(rng, k=10):
    r = np.zeros((4, 1216), np.float32)
    for p in range(4):
        idx = rng.choice(1176, k, replace=False)
        r[p, idx] = rng.uniform(0.05, 1.0, k)
    return r


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_terminal_prefix_scan_matches_brute_force(M, oracle, seed):
    rng = np.random.default_rng(seed)
    nodes, _ = M.tree4(_game(M))
    flop = np.sort(rng.choice(52, 3, replace=False))
    board = oracle.board4(flop)
    t = board[0]
    score = np.array([oracle.evaln([*flop, t["c1"][i], t["c2"][i]]) for i in range(1176)], np.uint32)
    bf = rng.uniform(1.0, 1.6, (4, 4))
    reach = _sparse_reach(rng)
    # one terminal of every kind: a fold win per seat, showdowns of 2, 3, 4 players
    picks = {}
    for n in np.nonzero(nodes["type"] != 0)[0]:
        x = nodes[n]
        key = (int(x["type"]), bin(int(x["active_mask"])).count("1") if x["type"] == 2 else int(x["active_mask"]))
        picks.setdefault(key, n)
    assert {k[1] for k in picks if k[0] == 2} >= {2, 3, 4}
    for key, n in sorted(picks.items()):
        a = oracle.terminal4(board, nodes[n:n + 1], reach, bf)
        b = oracle.terminal4_brute(board, nodes[n:n + 1], reach, bf, score)
        assert np.allclose(a, b, rtol=1e-9, atol=1e-12), key


@pytest.mark.parametrize("seed", [0, 1])
def test_terminal_side_pots_match_brute_force(M, oracle, seed):
    """Showdowns with unequal commitments (short stacks all-in for less):
    the layered payoff tables against chips resolved layer by layer for
    every tuple of opponent hands."""
    rng = np.random.default_rng(seed + 10)
    nodes, _ = M.tree4(_game(M, SEED0_BEHIND))
    flop = np.sort(rng.choice(52, 3, replace=False))
    board = oracle.board4(flop)
    t = board[0]
    score = np.array([oracle.evaln([*flop, t["c1"][i], t["c2"][i]]) for i in range(1176)], np.uint32)
    bf = rng.uniform(1.0, 1.6, (4, 4))
    reach = _sparse_reach(rng, 8)
    picks = {}
    for n in np.nonzero(nodes["type"] == 2)[0]:
        x = nodes[n]
        live = [p for p in range(4) if int(x["active_mask"]) >> p & 1]
        levels = len({round(float(x["committed"][p]), 4) for p in live})
        if levels > 1:
            picks.setdefault((len(live), levels), n)
    assert {k[0] for k in picks} >= {2, 3, 4} and max(k[1] for k in picks) >= 3
    for key, n in sorted(picks.items()):
        a = oracle.terminal4(board, nodes[n:n + 1], reach, bf)
        b = oracle.terminal4_brute(board, nodes[n:n + 1], reach, bf, score)
        assert np.allclose(a, b, rtol=1e-9, atol=1e-12), key


def test_terminal_side_pot_hand_cases(M, oracle):
    """A short stack all-in for 5 against two players at 20, a fourth seat
    folded after putting in 8: main pot 4 x 5 = 20 for all three, side pot
    15 + 15 + 3 = 33 for the two deep players."""
    flop = [card("2c"), card("7d"), card("Kh")]
    board = oracle.board4(flop)
    t = board[0]
    idx = {(int(t["c1"][i]), int(t["c2"][i])): i for i in range(1176)}
    def combo(a, b):
        a, b = sorted((card(a), card(b)))
        return idx[(a, b)]
    node = np.zeros(1, [("type", "u1"), ("player", "u1"), ("num_actions", "u1"), ("active_mask", "u1"),
                        ("children_offset", "<i4"), ("infoset_offset", "<i4"), ("pot_size", "<f4"),
                        ("committed", "<f4", (4,))])
    node["type"] = 2
    node["active_mask"] = 0b0111
    node["pot_size"] = 53.0
    node["committed"] = [5.0, 20.0, 20.0, 8.0]
    bf = np.full((4, 4), 1.5)
    reach = np.zeros((4, 1216), np.float32)
    short, mid, weak = combo("Kc", "Kd"), combo("Ac", "Ad"), combo("3s", "4s")
    reach[0, short] = reach[1, mid] = reach[2, weak] = 1.0
    reach[3, combo("5h", "6h")] = 1.0
    v = oracle.terminal4(board, node, reach, bf)
    # the short stack's set of kings wins the main pot: 20 - 5
    assert abs(v[0, short] - 15.0) < 1e-9
    # aces lose the main pot but win the side pot: 33 - 20
    assert abs(v[1, mid] - 13.0) < 1e-9
    # the weakest hand loses both: -20, scaled by its bubble factor
    assert abs(v[2, weak] - (-1.5 * 20.0)) < 1e-9
    # aces against a weaker short stack win both pots: 53 - 20
    reach[0] = 0
    reach[0, combo("3c", "4c")] = 1.0
    v = oracle.terminal4(board, node, reach, bf)
    assert abs(v[1, mid] - 33.0) < 1e-9
    # a tie for the side pot between the deep players, short stack best
    reach[0] = 0
    reach[0, short] = 1.0
    reach[2] = 0
    reach[2, combo("Ah", "As")] = 1.0
    v = oracle.terminal4(board, node, reach, bf)
    assert abs(v[1, mid] - (16.5 - 20.0) * 1.5) < 1e-9


def test_terminal_ties_split_and_card_removal(M, oracle):
    """Hand-made cases: a pot split three ways and an opponent hand that
    shares a card with the hero (it must not count)."""
    flop = [card("2c"), card("7d"), card("Kh")]
    board = oracle.board4(flop)
    t = board[0]
    idx = {(int(t["c1"][i]), int(t["c2"][i])): i for i in range(1176)}
    def combo(a, b):
        a, b = sorted((card(a), card(b)))
        return idx[(a, b)]
    node = np.zeros(1, [("type", "u1"), ("player", "u1"), ("num_actions", "u1"), ("active_mask", "u1"),
                        ("children_offset", "<i4"), ("infoset_offset", "<i4"), ("pot_size", "<f4"),
                        ("committed", "<f4", (4,))])
    node["type"] = 2
    node["active_mask"] = 0b0111  # seats 0, 1, 2 at showdown; seat 3 folded
    node["pot_size"] = 30.0
    node["committed"] = [9.0, 9.0, 9.0, 3.0]
    bf = np.full((4, 4), 1.25)
    reach = np.zeros((4, 1216), np.float32)
    hero = combo("As", "3s")           # ace high
    reach[1, combo("Ac", "3d")] = 1.0  # same ranks: tie
    reach[2, combo("Ad", "3h")] = 1.0  # tie
    reach[3, combo("4c", "5d")] = 1.0  # folded seat, only its mass matters
    v = oracle.terminal4(board, node, reach, bf)
    assert abs(v[0, hero] - (30.0 / 3 - 9.0)) < 1e-9        # three-way split: +1
    # seat 3 holds a hand that shares the hero's card: card removal zeroes it
    reach[3] = 0
    reach[3, combo("As", "4d")] = 1.0
    v = oracle.terminal4(board, node, reach, bf)
    assert v[0, hero] == 0.0
    # a loss is scaled by the bubble factor; folded seats lose their chips
    reach[3] = 0
    reach[3, combo("4c", "5d")] = 1.0
    reach[1] = 0
    reach[1, combo("Kc", "Kd")] = 1.0  # set of kings beats ace high
    v = oracle.terminal4(board, node, reach, bf)
    assert abs(v[0, hero] - (-1.25 * 9.0)) < 1e-9
    reach[0, hero] = 1.0  # the folded seat's value is weighted by the others' (card-removed) mass
    v = oracle.terminal4(board, node, reach, bf)
    assert abs(v[3, combo("4c", "5d")] - (-1.25 * 3.0)) < 1e-9


def card(s):
    return "23456789TJQKA".index(s[0]) * 4 + "cdhs".index(s[1])


# -------------------------------------------------------------------- ICM --
@pytest.mark.parametrize("seed", range(6))
def test_icm_dp_matches_permutation_enumeration(oracle, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 9))
    st = rng.uniform(1.0, 80.0, n)
    k = int(rng.integers(1, min(n, 4) + 1))
    pay = np.sort(rng.uniform(0.1, 1.0, k))[::-1]
    a, b = oracle.icm(st, pay), oracle.icm(st, pay, brute=True)
    assert np.allclose(a, b, rtol=1e-12, atol=1e-14)
    assert abs(a.sum() - pay.sum()) < 1e-12


def test_icm_closed_forms(oracle):
    # two players: P(first) = S_i / T
    e = oracle.icm([30.0, 10.0], [0.7, 0.3])
    assert np.allclose(e, [0.7 * 0.75 + 0.3 * 0.25, 0.7 * 0.25 + 0.3 * 0.75], rtol=1e-14)
    # equal stacks: equal equity
    e = oracle.icm([20.0] * 6, [0.5, 0.3, 0.2])
    assert np.allclose(e, 1.0 / 6, rtol=1e-14)
    # winner-take-all: chip EV
    st = np.array([5.0, 17.0, 40.0, 8.0])
    assert np.allclose(oracle.icm(st, [1.0]), st / st.sum(), rtol=1e-14)
    # three players, two paid places: hand-expanded Malmuth-Harville
    s = np.array([50.0, 30.0, 20.0])
    T = s.sum()
    p1 = s / T
    p2 = np.array([sum(s[j] / T * s[i] / (T - s[j]) for j in range(3) if j != i) for i in range(3)])
    assert np.allclose(oracle.icm(s, [0.65, 0.35]), 0.65 * p1 + 0.35 * p2, rtol=1e-14)
    # a busted player gets nothing while enough positive stacks remain
    e = oracle.icm([0.0, 10.0, 10.0, 10.0], [0.5, 0.3, 0.2])
    assert e[0] == 0.0 and abs(e[1:].sum() - 1.0) < 1e-14


def test_bubble_factors(oracle):
    st = [20.0, 20.0, 20.0, 20.0, 5.0, 40.0, 60.0, 35.0]
    bf = oracle.bubble_factors(st, [0.5, 0.3, 0.2], [0, 1, 2, 3])
    # equal stacks in the pot: symmetric, and ICM makes losing cost more than winning gains
    assert np.allclose(bf, bf.T, rtol=1e-12)
    assert np.all(bf > 1.0)
    # diagonal = mean of the row's off-diagonal entries
    for p in range(4):
        assert abs(bf[p, p] - np.mean(np.delete(bf[p], p))) < 1e-12
    # winner-take-all is chip EV: every bubble factor is 1
    bf1 = oracle.bubble_factors(st, [1.0], [0, 1, 2, 3])
    assert np.allclose(bf1, 1.0, rtol=1e-12)


# -------------------------------------------------------------- DCFR step --
def test_dcfr_step_hand_values(oracle):
    """Two actions, constant cfv = (1, 0), reach 1, from uniform sigma.
    Step 1: v = 0.5, r = (+0.5, -0.5): R = (0.5 wp, -0.5 wn), sigma = (1, 0),
    S = ws * (0.5, 0.5).  Step 2: v = 1, r = (0, -1): R = (0.5 wp^2, (-0.5 wn - 1) wn)."""
    wp, wn, ws = 0.8, 0.4, 2.0
    R = np.zeros((2, 4), np.float32)
    sg = np.full((2, 4), 0.5, np.float32)
    S = np.zeros((2, 4), np.float32)
    cfv = np.zeros((2, 4), np.float32)
    cfv[0] = 1.0
    reach = np.ones(4, np.float32)
    oracle.dcfr_step(2, R, sg, S, cfv, reach, wp, wn, ws)
    assert np.allclose(R[:, 0], [0.5 * wp, -0.5 * wn]) and np.allclose(sg[:, 0], [1.0, 0.0])
    assert np.allclose(S[:, 0], [0.5 * ws, 0.5 * ws])
    oracle.dcfr_step(2, R, sg, S, cfv, reach, wp, wn, ws)
    assert np.allclose(R[:, 0], [0.5 * wp * wp, (-0.5 * wn - 1.0) * wn])
    assert np.allclose(S[:, 0], [1.5 * ws, 0.5 * ws])


def test_dcfr_step_floor_form_is_reference(oracle):
    """w_neg = 0 reproduces preflop_grid_generator.cu:402's
    R <- w_pos * max(0, R + r) on random data."""
    rng = np.random.default_rng(3)
    na, n, stride = 3, 169, 192
    R = rng.normal(0, 1, (na, stride)).astype(np.float32)
    sg = rng.dirichlet([1.0] * na, stride).T.astype(np.float32).copy()
    S = rng.uniform(0, 1, (na, stride)).astype(np.float32)
    cfv = rng.normal(0, 1, (na, stride)).astype(np.float32)
    reach = rng.uniform(0, 1, n).astype(np.float32)
    R0, sg0 = R.copy(), sg.copy()
    oracle.dcfr_step(na, R, sg, S, cfv, reach, 0.75, 0.0, 4.0)
    v = (sg0[:, :n] * cfv[:, :n]).sum(0)
    expect = 0.75 * np.maximum(0.0, R0[:, :n] + (cfv[:, :n] - v))
    assert np.allclose(R[:, :n], expect, rtol=1e-6, atol=1e-6)
    pos = np.maximum(expect, 0)
    ss = pos.sum(0)
    expect_sig = np.where(ss > 1e-7, pos / np.where(ss > 0, ss, 1), 1.0 / na)
    assert np.allclose(sg[:, :n], expect_sig, rtol=1e-5, atol=1e-6)
    assert np.array_equal(R[:, n:], R0[:, n:])  # padding untouched


# --------------------------------------------------------- 4-way CFR loop --
def test_fourway_cfr_loop_invariants(M, oracle):
    """CFR+ keeps regrets >= 0; the strategy sums grow by the iteration
    weight times reach, so after t iterations each infoset's sum over
    actions equals sum_k k * reach_k (bounded by t(t+1)/2 for reach <= 1);
    root values stay finite."""
    nodes, ni = M.tree4(_game(M))
    rng = np.random.default_rng(0)
    flop = np.sort(rng.choice(52, 3, replace=False))
    ranges = np.zeros((4, 1216), np.float32)
    for p in range(4):
        w = rng.uniform(0, 1, 1176)
        ranges[p, :1176] = w / w.sum()
    h = oracle.FourWay(nodes, ni, flop, ranges, np.ones((4, 4)))
    h.iterate(5, scheme=0)
    R, S, root = h.state()
    assert R.min() >= 0.0 and np.isfinite(root).all()
    S = S.reshape(ni, 3, 1216)
    # the root infoset is reached with the range itself every iteration: sum_a S = (1+..+5) * r
    root_info = int(nodes[0]["infoset_offset"])
    assert np.allclose(S[root_info].sum(0)[:1176], 15.0 * ranges[int(nodes[0]["player"])][:1176], rtol=1e-5)
    assert S.sum(1).max() <= 15.0 * ranges.max() + 1e-6
