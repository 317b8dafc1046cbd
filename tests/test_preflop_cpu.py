"""CPU tests of the oracle's preflop pieces (preflop_grid_generator.cu):
representative hands of the 169 classes, and the Monte-Carlo all-in matrix
against the exact expectation of its own board-sampling process."""
import math

import numpy as np
import pytest


def test_class_to_cards_maps_back(oracle):
    """Each class's representative hand (class_to_cards,
    preflop_grid_generator.cu:79-92, with its rank-block walk corrected: the
    reference's maps the last class of every block to a hand of two equal
    ranks and never ends for 32s / 32o) belongs to that class under the
    canonical class map (pinned separately by closed forms)."""
    seen = set()
    for cls in range(169):
        a, b = oracle.class_to_cards(cls)
        assert 0 <= a < 52 and 0 <= b < 52 and a != b
        assert oracle.lib().or_canonical_class(a, b) == cls
        seen.add((a, b))
    assert len(seen) == 169


@pytest.mark.parametrize("h,v", [(0, 168), (13, 91), (57, 120), (12, 1)])
def test_preflop_mc_matches_exact_distribution(oracle, h, v):
    """Many samples of the Monte-Carlo cell (kernel_gen_preflop_matrix,
    preflop_grid_generator.cu:95-127) estimate the exact expectation of the
    same sampling process (every sequence of the five picks enumerated,
    repeated cards included) within 4.5 standard errors.  A change to the
    deck, the collision rule, the pick range or the evaluator moves the
    expectation by more than that."""
    n = 400000
    w, t = oracle.preflop_cell(h, v, n)
    est = (w + 0.5 * t) / n
    mean, var = oracle.preflop_exact(h, v)
    se = math.sqrt(var / n)
    assert abs(est - mean) <= 4.5 * se, (h, v, est, mean, se)


def test_preflop_matrix_structure(oracle):
    """Diagonal 0.5; cells equal (wins + ties/2) / samples; a class against
    the same classes gives the same counts whichever order the matrix is
    walked in (per-cell streams)."""
    w, t, e = oracle.preflop_matrix(3000)
    assert np.all(np.diag(e) == 0.5)
    off = ~np.eye(169, dtype=bool)
    want = ((w.astype(np.float32) + np.float32(0.5) * t.astype(np.float32)) / np.float32(3000)).astype(np.float32)
    assert np.array_equal(e[off], want[off])
    assert np.all(w + t <= 3000) and w.min() >= 0
    for h, v in ((3, 77), (160, 2)):
        assert oracle.preflop_cell(h, v, 3000) == (w[h, v], t[h, v])
    # the two orders of a matchup estimate complementary equities
    assert np.abs(e + e.T - 1.0)[off].max() < 0.1


# ------------------------------------------------- preflop co-solver ------
def _members():
    import itertools
    return itertools.combinations(range(52), 2)


def test_flop_class_matrix_brute_force(oracle):
    """Class equities (kernel_reduce_combo_to_class_matrices,
    preflop_grid_generator.cu:189-223, over all 1,326 combos): entries
    rebuilt from the per-pair flop entries (or_flop_matrix_entry, the
    reference's :130-183 entry) by an independent loop over the members of
    each class, hero members outer, villain inner, FP32 sum; M[v][h]."""
    L = oracle.lib()
    flop = [0, 5, 10]
    M = oracle.flop_class_matrix(flop)
    mem = {c: [] for c in range(169)}
    for a, b in _members():
        mem[L.or_canonical_class(a, b)].append((a, b))
    fm = set(flop)
    for h, v in [(0, 168), (13, 91), (57, 120), (12, 1), (5, 5), (100, 30), (168, 0)]:
        s, n = np.float32(0), 0
        for hc in mem[h]:
            if set(hc) & fm:
                continue
            for vc in mem[v]:
                if set(vc) & (fm | set(hc)):
                    continue
                s = np.float32(s + np.float32(oracle.flop_matrix_entry(flop, hc, vc)))
                n += 1
        want = np.float32(s / np.float32(n)) if n else np.float32(0.5)
        assert M[v, h] == want, (h, v, M[v, h], want)


def test_flop_class_matrix_no_valid_pair(oracle):
    """On a flop of three aces no AA combo avoids the flop: the AA row and
    column are 0.5 (preflop_grid_generator.cu:221); every other entry is a
    mean of equities in [0, 1] and the quarter-tie rule makes a class against
    its own rank-twin classes tie at most to 1/4."""
    A = 12 * 4
    M = oracle.flop_class_matrix([A, A + 1, A + 2])
    assert np.all(M[0, :] == 0.5) and np.all(M[:, 0] == 0.5)
    rest = M[1:, 1:]
    assert rest.min() >= 0.0 and rest.max() <= 1.0
    # KK against KK ties on every runout that neither improves: at most 1/4
    assert M[1, 1] <= 0.25 + 1e-6


def _root():
    return np.array([6] * 13 + [4] * 78 + [12] * 78, np.float32) / np.float32(1326)


@pytest.mark.parametrize("iters", [1, 2, 50])
def test_preflop_strategy_sums_closed_form(oracle, iters):
    """Strategy sums of the DCFR loop (preflop_grid_generator.cu:387-410,
    588-614): each iteration adds t^2 x root reach x sigma, and sigma sums to
    one, so the sums over actions equal root reach x sum t^2 at every node;
    2-action nodes leave their third row zero.  On coin-flip data (all
    equities 1/2) the big blind facing an open jam has fold -1.35bb against
    call 0.5 x 22.5 - 0.5 x 1.35 x 20 = -2.25bb, so from the second
    iteration on it folds: fold sums root x (1/2 + sum_{t>=2} t^2), call
    root x 1/2."""
    M = np.full((2, 169, 169), 0.5, np.float32)
    S = oracle.preflop_solve(M, np.array([1.0, 3.0], np.float32), np.full((169, 169), 0.5, np.float32),
                             iters=iters)
    root = _root()
    tot = sum(t * t for t in range(1, iters + 1))
    for node, na in enumerate((3, 3, 2, 2)):
        np.testing.assert_allclose(S[node, :na].sum(0), root * tot, rtol=2e-6)
        if na < 3:
            assert np.all(S[node, 2] == 0)
    np.testing.assert_allclose(S[3, 1], root * 0.5, rtol=1e-6)
    np.testing.assert_allclose(S[3, 0], root * (tot - 0.5), rtol=1e-6)


def test_preflop_monotone_game(oracle):
    """A game where class index orders strength (equity 1 against every
    weaker class, 0 against every stronger, 1/2 against itself, both on the
    flop and all-in): after 500 iterations stronger classes call jams at
    least as often (facing the open jam and facing the 3-bet jam), and the
    weakest class folds most of the time at both first decisions while the
    strongest never folds.  An all-in matrix read with hero and villain
    swapped makes the weakest hands call; a flop matrix read that way makes
    the weakest hands raise and defend."""
    idx = np.arange(169)
    A = np.where(idx[:, None] < idx[None, :], 1.0, 0.0).astype(np.float32)
    np.fill_diagonal(A, 0.5)
    M = np.stack([np.ascontiguousarray(A.T)] * 2)       # M[v][h] = equity of h against v
    S = oracle.preflop_solve(M, np.array([1.0, 1.0], np.float32), A, iters=500)
    F = [S[n] / np.maximum(S[n].sum(0, keepdims=True), 1e-30) for n in range(4)]
    assert np.all(np.diff(F[3][1]) <= 1e-4)
    assert np.all(np.diff(F[2][1]) <= 1e-4)
    for n in (0, 1):
        assert F[n][0][168] > 0.5 and F[n][0][0] < 1e-3, (n, F[n][0][[0, 168]])
