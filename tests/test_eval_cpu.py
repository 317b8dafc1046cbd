"""CPU tests: the oracle evaluator against the reference's own data and
closed forms, and the library's host evaluator against the oracle."""
import itertools
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

# Hand counts per category over all C(52,7) hands (standard combinatorics;
# the reference's categories are the standard ones, mtt_ev_dumper.cu:53-83).
CLOSED_FORM = [23294460, 58627800, 31433400, 6461620, 6180020, 4047644, 3473184, 224848, 41584]


def card(s):
    r = "23456789TJQKA".index(s[0])
    return r * 4 + "cdhs".index(s[1])


def hand(*cs):
    return np.array([[card(c) for c in cs]], np.uint8)


def test_category_counts_closed_form(oracle):
    assert oracle.category_counts().tolist() == CLOSED_FORM


def test_quirks_known_answers(oracle):
    ev = lambda *cs: int(oracle.eval7_batch(hand(*cs))[0])
    # quads carry no kicker: AAAA+K ties AAAA+Q
    assert ev("Ac", "Ad", "Ah", "As", "Kc", "2d", "3h") == ev("Ac", "Ad", "Ah", "As", "Qc", "2d", "3h")
    # pair carries no kickers
    assert ev("Ac", "Ad", "Kh", "Qs", "9c", "5d", "3h") == ev("Ac", "Ad", "Jh", "Ts", "8c", "4d", "2h")
    # high card compares all seven ranks
    assert ev("Ac", "Kd", "Qh", "9s", "7c", "4d", "3h") > ev("Ac", "Kd", "Qh", "9s", "7c", "4d", "2h")
    # wheel = 3, six-high straight = 4
    assert ev("Ac", "2d", "3h", "4s", "5c", "9d", "Jh") == (4 << 24) | 3
    assert ev("6c", "2d", "3h", "4s", "5c", "9d", "Jh") == (4 << 24) | 4
    # two trips make a full house of (top trips, second trips)
    assert ev("Kc", "Kd", "Kh", "7s", "7c", "7d", "2h") == (6 << 24) | (11 << 4) | 5
    # steel wheel straight flush
    assert ev("Ac", "2c", "3c", "4c", "5c", "Kd", "Kh") == (8 << 24) | 3
    # flush nibbles: top five of six suited cards
    assert ev("Ac", "Jc", "9c", "7c", "5c", "3c", "Kd") == (5 << 24) | (12 << 16) | (9 << 12) | (7 << 8) | (5 << 4) | 3


def test_distinct_values_per_category(oracle):
    rng = np.random.default_rng(7)
    cards = np.argsort(rng.random((400000, 52)), axis=1)[:, :7].astype(np.uint8)
    s = oracle.eval7_batch(cards)
    cat = s >> 24
    # upper bounds implied by the encoding (mtt_ev_dumper.cu:63-82)
    bound = {8: 10, 7: 13, 6: 156, 4: 10, 3: 13, 2: 78, 1: 13}
    for c, b in bound.items():
        assert len(np.unique(s[cat == c])) <= b


def test_canonical_flops(oracle, mtt):
    f = oracle.canonical_flops()
    assert f[:, 3].astype(int).sum() == 22100  # C(52,3)
    g = np.zeros((1755, 4), np.uint8)
    assert mtt.lib().mtt_canonical_flops(g.ctypes.data) == 1755
    assert (f == g).all()
    # every flop is three distinct cards, and all flops are distinct
    assert len({tuple(sorted(x[:3])) for x in f}) == 1755


def test_class_map(oracle, mtt):
    counts = np.zeros(169, int)
    for a, b in itertools.combinations(range(52), 2):
        c = oracle.lib().or_canonical_class(a, b)
        assert c == mtt.lib().mtt_canonical_class(a, b)
        counts[c] += 1
    assert (counts[:13] == 6).all() and (counts[13:91] == 4).all() and (counts[91:] == 12).all()


def test_host_evaluator_matches_oracle(oracle, mtt):
    rng = np.random.default_rng(3)
    cards = np.argsort(rng.random((300000, 52)), axis=1)[:, :7].astype(np.uint8)
    want = oracle.eval7_batch(cards)
    got = np.zeros_like(want)
    mtt.call("mtt_eval7_host_batch", cards.ctypes.data, cards.shape[0], got.ctypes.data)
    assert (want == got).all()


@pytest.mark.slow
def test_host_evaluator_exhaustive(oracle, mtt):
    """All C(52,7) hands, in chunks by the first two cards."""
    rest = {}
    for a in range(52):
        for b in range(a + 1, 52):
            others = np.arange(b + 1, 52, dtype=np.uint8)
            if len(others) < 5:
                continue
            combos = np.array(list(itertools.combinations(others, 5)), np.uint8)
            cards = np.concatenate([np.full((len(combos), 1), a, np.uint8), np.full((len(combos), 1), b, np.uint8), combos], 1)
            want = oracle.eval7_batch(cards)
            got = np.zeros_like(want)
            mtt.call("mtt_eval7_host_batch", cards.ctypes.data, cards.shape[0], got.ctypes.data)
            assert (want == got).all(), (a, b)


def test_oracle_ev_dump_matches_reference_slice(oracle):
    z = np.load(os.path.join(GOLDEN, "ev_dump_slice.npz"))
    idx, vals = z["flop_index"], z["values"]
    for k in idx[:12]:
        got = oracle.ev_dump(int(k), 1)[0]
        assert np.array_equal(got, vals[list(idx).index(k)]), k


def test_ev_dump_fixture_meta():
    meta = json.load(open(os.path.join(GOLDEN, "ev_dump_meta.json")))
    assert meta["ev_dump_bytes"] == 1755 * 4 * 2 * 169 * 4
    z = np.load(os.path.join(GOLDEN, "ev_dump_slice.npz"))
    v = z["values"]
    # measured properties of the reference dump: p1 = -p0, scenario ratios
    assert np.abs(v[:, :, 1, :] + v[:, :, 0, :]).max() < 1e-6
    # per-combo products are summed per class, so the ratio holds to rounding
    for s in (1, 2, 3):
        assert np.abs(v[:, s, 0, :] - (1 + 0.1 * s) * v[:, 0, 0, :]).max() < 1e-6


def test_flop_counts_oracle_invariants(oracle, mtt):
    flop = [card("Kc"), card("8d"), card("3h")]
    W, T = oracle.flop_counts(flop)
    alive = [i for i, (a, b) in enumerate(itertools.combinations(range(52), 2)) if a not in flop and b not in flop]
    A = np.ix_(alive, alive)
    Wa, Ta = W[A], T[A]
    pairs = list(itertools.combinations(range(52), 2))
    masks = np.array([(1 << pairs[i][0]) | (1 << pairs[i][1]) for i in alive], dtype=object)
    blocked = np.array([[bool(masks[i] & masks[j]) for j in range(len(alive))] for i in range(len(alive))])
    tot = Wa + Wa.T + Ta
    assert (tot[~blocked] == 990).all()
    assert (Wa[blocked] == 0).all() and (Ta[blocked] == 0).all()
    assert (Ta == Ta.T).all()
    # rows/cols of flop-blocked combos are zero
    dead = [i for i in range(1326) if i not in set(alive)]
    assert (W[dead].sum() == 0) and (W[:, dead].sum() == 0)
    # brute force on sampled pairs
    rng = np.random.default_rng(5)
    deck = [c for c in range(52) if c not in flop]
    for _ in range(6):
        i, j = rng.choice(len(alive), 2, replace=False)
        if blocked[i, j]:
            continue
        hi, hj = pairs[alive[i]], pairs[alive[j]]
        rest = [c for c in deck if c not in hi and c not in hj]
        w = t = 0
        for x, y in itertools.combinations(rest, 2):
            a = oracle.eval7_batch(np.array([[*flop, x, y, *hi]], np.uint8))[0]
            b = oracle.eval7_batch(np.array([[*flop, x, y, *hj]], np.uint8))[0]
            w += a > b
            t += a == b
        assert Wa[i, j] == w and Ta[i, j] == t


def test_library_exports_declared_symbols(mtt):
    import re
    hdr = open(os.path.join(os.path.dirname(GOLDEN), "..", "include", "mtt.h")).read()
    names = set(re.findall(r"\b(mtt_[a-z0-9_]+)\s*\(", hdr))
    L = mtt.lib()
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_flop_matrix_entry_quarter_tie(oracle):
    """preflop_grid_generator.cu:176-180 counts wins += 2, ties += 1 and
    divides (wins + 0.5*ties) by total (+= 2 per runout): two hands with the
    same ranks that can never make different hands tie on every runout, so
    the entry is exactly 1/4; a blocked pair is 0."""
    c = lambda s: "23456789TJQKA".index(s[0]) * 4 + "cdhs".index(s[1])
    flop = sorted([c("2h"), c("7s"), c("9h")])  # no clubs or diamonds: neither hand can flush
    assert oracle.flop_matrix_entry(flop, (c("Ac"), c("Kc")), (c("Ad"), c("Kd"))) == 0.25
    assert oracle.flop_matrix_entry(flop, (c("Ac"), c("Kc")), (c("Ac"), c("Qd"))) == 0.0
    # aces against a hand drawing dead to nothing but runner-runner: between 0 and 1
    e = oracle.flop_matrix_entry(flop, (c("Ac"), c("Ad")), (c("3c"), c("4d")))
    assert 0.5 < e < 1.0


def test_flop_checksum_fixture_matches_oracle(oracle):
    """The committed per-flop checksums (tests/golden/flop_checksums.npz)
    are the oracle's: recomputed here for two flops."""
    import importlib.util
    from conftest import GOLDEN
    spec = importlib.util.spec_from_file_location("make_flop_checksums", os.path.join(GOLDEN, "make_flop_checksums.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    fx = np.load(os.path.join(GOLDEN, "flop_checksums.npz"))
    assert fx["checksums"].shape == (1755, 4)
    assert np.array_equal(fx["flops"], oracle.canonical_flops()[:, :3])
    w1, w2 = mk.splitmix_weights(1), mk.splitmix_weights(2)
    for f in (0, 1333):
        W, T = oracle.flop_counts(fx["flops"][f])
        assert np.array_equal(mk.checksums(W, T, w1, w2), fx["checksums"][f]), f
