"""CPU tests of the HU tree builder, the seeded inputs and the CFR oracle
(terminal values by brute force, a closed-form toy game, convergence and
accounting invariants)."""
import collections
import itertools

import numpy as np
import pytest


def card(s):
    return "23456789TJQKA".index(s[0]) * 4 + "cdhs".index(s[1])


def counts(nodes):
    c = collections.Counter(nodes["type"].tolist())
    return c[0], c[1], c[2], c[3]


def test_tree_config1_counts(solver_api):
    g, _, _ = solver_api.config1(0)
    nodes, ni, nm = solver_api.tree(g)
    assert len(nodes) == 33 and ni == 12 and nm == 1
    assert counts(nodes) == (12, 10, 11, 0)


def test_tree_config3_counts(solver_api):
    g, _, _ = solver_api.config3(0)
    nodes, ni, nm = solver_api.tree(g)
    # per turn card 16 decisions / 14 folds / 6 all-in showdowns / 9 river lines;
    # 329,280 river decisions and 308,112 river showdowns over 49x48 boards
    assert len(nodes) == 926542
    assert counts(nodes) == (330064, 287630, 308406, 442)
    assert nm == 49 + 49 * 48
    dec = nodes[nodes["type"] == 0]
    assert dec["num_actions"].max() == 4


def test_tree_structure_invariants(solver_api):
    g, _, _ = solver_api.config3_small(0)
    nodes, ni, _ = solver_api.tree(g)
    seen = np.zeros(len(nodes), bool)
    seen[0] = True
    infos = set()
    for i, x in enumerate(nodes):
        if x["type"] in (0, 3):
            co, na = int(x["children_offset"]), int(x["num_actions"])
            assert co > i and co + na <= len(nodes)
            assert not seen[co:co + na].any()
            seen[co:co + na] = True
        if x["type"] == 0:
            infos.add(int(x["infoset_offset"]))
            # pot = start pot + commitments; commitments never exceed the stack
            assert abs(x["pot_size"] - (6.0 + x["committed"].sum())) < 1e-4
        if x["type"] == 2:
            assert abs(x["committed"][0] - x["committed"][1]) < 1e-5
    assert seen.all() and infos == set(range(ni))


def splitmix(seed):
    s = seed & (2**64 - 1)
    while True:
        s = (s + 0x9E3779B97F4A7C15) & (2**64 - 1)
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2**64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2**64 - 1)
        yield z ^ (z >> 31)


def test_seeded_inputs_restated(solver_api):
    """Board draw and uniform ranges against an independent restatement."""
    rng = splitmix(0)
    u = lambda: (next(rng) >> 11) * 2.0**-53
    deck = list(range(52))
    board = []
    for i in range(5):
        j = i + int(u() * (52 - i))
        deck[i], deck[j] = deck[j], deck[i]
        board.append(deck[i])
    ranges = []
    for p in range(2):
        w = []
        for a, b in itertools.combinations(range(52), 2):
            w.append(0.0 if (a in board or b in board) else u())
        w = np.array(w)
        ranges.append(w / w.sum())
    b2, r0, r1 = solver_api.seeded_inputs(0, 5, 0)
    assert list(b2) == board
    assert np.allclose(r0, ranges[0], rtol=1e-6, atol=1e-12) and np.allclose(r1, ranges[1], rtol=1e-6, atol=1e-12)
    # Dirichlet: positive on unblocked combos, zero on blocked, normalised
    f, q0, q1 = solver_api.seeded_inputs(7, 3, 1, 0.5)
    blocked = np.array([a in f or b in f for a, b in itertools.combinations(range(52), 2)])
    assert (q0[blocked] == 0).all() and (q0[~blocked] > 0).all() and abs(q0.sum() - 1) < 1e-5


def _oracle_for(solver_api, oracle, g, r0, r1, prec=0):
    nodes, ni, _ = solver_api.tree(g)
    board = list(g.board)[:g.n_board]
    gi, cards = solver_api.alive_index(board)
    return oracle.HU(nodes, ni, board, r0[gi], r1[gi], prec), nodes, gi, cards


def _brute_values(oracle, board, cards, ro, pot, committed, nd_cards=()):
    """P*(E r)(i) - c*(nb r)(i) by enumerating pairs (and rivers)."""
    full = list(board) + list(nd_cards)
    n = len(cards)
    live = np.array([not (a in full or b in full) for a, b in cards])
    boards = [full] if len(full) == 5 else [full + [c] for c in range(52) if c not in full]
    m = np.array([(1 << int(a)) | (1 << int(b)) for a, b in cards], dtype=object)
    nb = np.array([[not (m[i] & m[j]) for j in range(n)] for i in range(n)])
    nb &= live[:, None] & live[None, :]
    eq = np.zeros(n)
    cnt = np.zeros((n, n))
    for bd in boards:
        ok = np.array([not (a in bd or b in bd) for a, b in cards])
        sc = np.zeros(n, np.int64)
        idx = np.where(ok)[0]
        sc[idx] = oracle.eval7_batch(np.array([[*bd, cards[i][0], cards[i][1]] for i in idx], np.uint8))
        e = (sc[:, None] > sc[None, :]) + 0.5 * (sc[:, None] == sc[None, :])
        both = ok[:, None] & ok[None, :] & nb
        eq += (e * both) @ np.where(True, ro, 0)
        cnt += both
    if len(boards) > 1:
        eq = np.zeros(n)
        acc = np.zeros((n, n))
        for bd in boards:
            ok = np.array([not (a in bd or b in bd) for a, b in cards])
            sc = np.zeros(n, np.int64)
            idx = np.where(ok)[0]
            sc[idx] = oracle.eval7_batch(np.array([[*bd, cards[i][0], cards[i][1]] for i in idx], np.uint8))
            e = (sc[:, None] > sc[None, :]) + 0.5 * (sc[:, None] == sc[None, :])
            acc += e * (ok[:, None] & ok[None, :] & nb)
        eq = (acc / 44.0) @ ro
    mass = nb.astype(float) @ ro
    v = pot * eq - committed * mass
    return np.where(live, v, 0.0), np.where(live, mass, 0.0)


def test_oracle_river_terminals_brute_force(solver_api, oracle):
    g, r0, r1 = solver_api.config1(3)
    h, nodes, gi, cards = _oracle_for(solver_api, oracle, g, r0, r1)
    rng = np.random.default_rng(0)
    a = rng.random(h.N)
    b = rng.random(h.N)
    board = list(g.board)[:5]
    for nid in np.where(nodes["type"] == 2)[0][:3]:
        x = nodes[nid]
        v0, v1 = h.terminal(int(nid), [], a, b)
        w0, _ = _brute_values(oracle, board, cards, b, x["pot_size"], x["committed"][0])
        w1, _ = _brute_values(oracle, board, cards, a, x["pot_size"], x["committed"][1])
        assert np.allclose(v0[:h.N], w0, atol=1e-9) and np.allclose(v1[:h.N], w1, atol=1e-9)
    for nid in np.where(nodes["type"] == 1)[0][:3]:
        x = nodes[nid]
        w = 0 if x["active_mask"] == 1 else 1
        v0, v1 = h.terminal(int(nid), [], a, b)
        _, m0 = _brute_values(oracle, board, cards, b, 0, 0)
        _, m1 = _brute_values(oracle, board, cards, a, 0, 0)
        gain = x["pot_size"] - x["committed"][w]
        loss = -x["committed"][1 - w]
        want0 = (gain if w == 0 else loss) * m0
        want1 = (gain if w == 1 else loss) * m1
        assert np.allclose(v0[:h.N], want0, atol=1e-9) and np.allclose(v1[:h.N], want1, atol=1e-9)


@pytest.mark.slow
def test_oracle_turn_allin_brute_force(solver_api, oracle):
    """Turn all-in showdown = mean over the 44 rivers (restricted ranges)."""
    g, r0, r1 = solver_api.config3_small(2)
    h, nodes, gi, cards = _oracle_for(solver_api, oracle, g, r0, r1)
    rng = np.random.default_rng(1)
    keep = rng.random(h.N) < 0.25
    a = rng.random(h.N) * keep
    b = rng.random(h.N) * keep
    flop = list(g.board)[:3]
    deck = [c for c in range(52) if c not in flop]
    sd = [i for i in range(len(nodes)) if nodes[i]["type"] == 2 and nodes[i]["turn_idx"] < 49]
    nid = sd[0]
    d1 = int(nodes[nid]["river_idx"]) - 1
    tc = deck[d1]
    live = np.array([tc not in (int(c[0]), int(c[1])) for c in cards])
    v0, _ = h.terminal(nid, [d1], a * live, b * live)
    w0, _ = _brute_values(oracle, flop, cards, b * live, nodes[nid]["pot_size"], nodes[nid]["committed"][0], [tc])
    assert np.allclose(v0[:h.N], w0, atol=1e-9)


def test_clairvoyance_closed_form(solver_api, oracle):
    """Polarised bettor (nuts / air) vs a bluff-catcher on a river, pot-size
    bet, disjoint card sets: the bettor bluffs air with probability 1/2 and
    the caller calls 1/2 (bluff:value = s/(P+s), call = P/(P+s))."""
    board = [card(c) for c in ("2c", "7d", "9h", "Jc", "Ks")]
    g = solver_api.make_game(board, 10.0, 20.0, {2: solver_api.menu((1.0,), 0.0, 0, 0)}, 0, 2)
    pairs = list(itertools.combinations(range(52), 2))
    r0 = np.zeros(1326, np.float32)
    r1 = np.zeros(1326, np.float32)
    nuts = pairs.index(tuple(sorted((card("Kh"), card("Kd")))))
    air = pairs.index(tuple(sorted((card("3c"), card("4d")))))
    bc = pairs.index(tuple(sorted((card("9c"), card("9d")))))
    r0[nuts] = r0[air] = 0.5
    r1[bc] = 1.0
    h, nodes, gi, cards = _oracle_for(solver_api, oracle, g, r0, r1)
    h.iterate(3000)
    st = h.strategy().reshape(-1, 4, 1216)
    ia = {int(k): i for i, k in enumerate(gi)}
    root = nodes[0]
    assert root["num_actions"] == 2
    bet_node = nodes[int(root["children_offset"]) + 1]
    assert bet_node["type"] == 0 and bet_node["player"] == 1 and bet_node["num_actions"] == 2
    p_bluff = st[0, 1, ia[air]]
    p_value = st[0, 1, ia[nuts]]
    p_call = st[int(bet_node["infoset_offset"]), 1, ia[bc]]
    assert abs(p_value - 1.0) < 0.01
    assert abs(p_bluff - 0.5) < 0.02
    assert abs(p_call - 0.5) < 0.02
    e = h.exploitability(h.strategy(), 10.0)
    assert e[5] < 1.0


def test_oracle_config1_convergence_and_accounting(solver_api, oracle):
    g, r0, r1 = solver_api.config1(0)
    h, nodes, gi, cards = _oracle_for(solver_api, oracle, g, r0, r1)
    h.iterate(100)
    e100 = h.exploitability(h.strategy(), 10.0)
    h.iterate(900)
    e = h.exploitability(h.strategy(), 10.0)
    assert e[5] < e100[5] and e[5] < 0.5            # <= 0.5% pot at 1000 iterations
    assert e[0] >= e[2] - 1e-9 and e[1] >= e[3] - 1e-9  # best response >= profile value
    # constant-sum: values add to pot * mass (strategies are float32: 1e-6)
    assert abs(e[2] + e[3] - 10.0 * e[4]) < 1e-6 * e[4]
    R, S = h.state()
    assert (R >= 0).all()                            # CFR+ floors regrets
    st = h.strategy().reshape(-1, 4, 1216)
    for x in nodes[nodes["type"] == 0]:
        s = st[int(x["infoset_offset"]), :int(x["num_actions"]), :h.N].sum(0)
        assert np.allclose(s, 1.0, atol=1e-5)


def test_oracle_fp32_tracks_fp64(solver_api, oracle):
    g, r0, r1 = solver_api.config1(0)
    h64, _, _, _ = _oracle_for(solver_api, oracle, g, r0, r1, 0)
    h32, _, _, _ = _oracle_for(solver_api, oracle, g, r0, r1, 1)
    h64.iterate(200)
    h32.iterate(200)
    e64 = h64.exploitability(h64.strategy(), 10.0)[5]
    e32 = h64.exploitability(h32.strategy(), 10.0)[5]
    assert abs(e64 - e32) < 0.02


def test_oracle_dcfr_converges(solver_api, oracle):
    g, r0, r1 = solver_api.config1(1)
    h, _, _, _ = _oracle_for(solver_api, oracle, g, r0, r1)
    h.iterate(100, scheme=1, alpha=1.5, beta=0.0, gamma=2.0)
    e1 = h.exploitability(h.strategy(), 10.0)[5]
    h.iterate(500, scheme=1, alpha=1.5, beta=0.0, gamma=2.0)
    e2 = h.exploitability(h.strategy(), 10.0)[5]
    assert e2 < e1 and e2 < 0.5


def test_dcfr_weights_telescope():
    """gamma = 2 averaging with weights t^2 equals the per-iteration discount
    (t/(t+1))^2 of the running average (preflop_grid_generator.cu:588-591)."""
    T = 50
    rng = np.random.default_rng(0)
    x = rng.random(T)
    direct = sum((t + 1) ** 2 * x[t] for t in range(T))
    run = 0.0
    for t in range(T):
        run = run * (t / (t + 1)) ** 2 + x[t] if t else x[t]
    assert abs(direct / T**2 - run) < 1e-9 * direct
