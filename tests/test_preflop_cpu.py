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
