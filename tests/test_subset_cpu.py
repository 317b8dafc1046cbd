"""CPU tests of the flop-subset selection (scripts/optimize_subset.py:28-65
restated in mtt-gpu-solver_amd/subset.py): the reference's committed subset
parses in its format, and the non-negative matching pursuit recovers a
planted exact combination of flops."""
import json
import os

import numpy as np

from conftest import GOLDEN, PKG, _load


def _S():
    return _load("mtt_subset", os.path.join(PKG, "subset.py"))


def test_reference_subset_format():
    S = _S()
    meta = json.load(open(os.path.join(GOLDEN, "ev_dump_meta.json")))
    data = bytes.fromhex(meta["subset_bin_hex"])
    cards, w = S.read_subset(data)
    assert len(data) == 4 + 8 * 100 and cards.shape == (100, 3)
    canon = {tuple(int(x) for x in c[:3]) for c in S.canonical_flops()}
    assert all(tuple(int(x) for x in c) in canon for c in cards)
    assert abs(float(w.sum()) - 1755.0) < 1e-2 and (w >= 0).all()
    # the text file lists the same flops and weights to 4 decimals
    lines = meta["subset_weights_txt"].splitlines()
    assert len(lines) == 100
    for (c, wt), line in zip(zip(cards, w), lines):
        name, val = line.split(":")
        assert name == "".join("23456789TJQKA"[x // 4] + "cdhs"[x % 4] for x in c)
        assert abs(float(val) - float(wt)) < 6e-5


def test_matching_pursuit_recovers_planted_combination():
    """Only five flops have nonzero EV columns (random, so linearly
    independent): the orbit-weighted mean of all flops is then exactly
    sum_k w_k col_k / sum(w), a non-negative combination of those five, so the
    pursuit picks exactly them and the rescaled NNLS weights are
    1755 * w_k / sum_k w_k."""
    S = _S()
    rng = np.random.default_rng(3)
    planted = [17, 400, 901, 1300, 1754]
    dump = np.zeros((1755, 4, 2, 169), np.float32)
    dump[planted] = rng.normal(size=(5, 4, 2, 169)).astype(np.float32)
    idx, wt = S.select_subset(dump, k=5)
    assert sorted(idx.tolist()) == planted
    w = S.canonical_flops()[:, 3].astype(np.float64)
    want = 1755.0 * w[idx] / w[idx].sum()
    assert np.allclose(wt, want, rtol=1e-6)
    # the bytes written follow SubsetFlop (preflop_grid_generator.cu:30)
    cards, w32 = S.read_subset(S.subset_bytes(idx, wt))
    assert np.array_equal(cards, S.canonical_flops()[idx, :3]) and np.allclose(w32, wt, rtol=1e-6)
