"""Error paths of the C ABI that need no GPU: games, trees and host inputs
that the library must refuse with MTT_ERR_ARG and a message, before any
device work.  (The device-side counterparts -- bad ranges to the solve
handles, an allocation that fails -- are in test_errors_gpu.py.)"""
import ctypes
import math
import os

import numpy as np
import pytest

from conftest import PKG, _load


@pytest.fixture(scope="module")
def S(mtt):
    return _load("mtt_solver", os.path.join(PKG, "solver.py"))


@pytest.fixture(scope="module")
def M(mtt):
    return _load("mtt_multi", os.path.join(PKG, "multi.py"))


def _refused(mtt, fn, *args):
    with pytest.raises(mtt.MTTError) as e:
        mtt.call(fn, *args)
    assert e.value.status == mtt.MTT_ERR_ARG, str(e.value)
    assert mtt.last_error(), "no message"
    return str(e.value)


def _hu_variants(S):
    """(label, game) pairs, each broken in one field."""
    def base():
        g, _, _ = S.config1(0)
        return g
    out = []
    g = base(); g.pot = float("nan"); out.append(("NaN pot", g))
    g = base(); g.stack = -1.0; out.append(("negative stack", g))
    g = base(); g.stack = float("inf"); out.append(("infinite stack", g))
    g = base(); g.board[1] = g.board[0]; out.append(("repeated board card", g))
    g = base(); g.board[2] = 52; out.append(("card 52", g))
    g = base(); g.n_board = 6; out.append(("six board cards", g))
    g = base(); g.root_chance = 1; out.append(("root chance with five cards", g))
    g = base(); g.last_street = 5; out.append(("last street 5", g))
    g = base(); g.menu[2].n_sizes = 4; out.append(("four sizes", g))
    g = base(); g.menu[2].sizes[0] = float("nan"); out.append(("NaN size", g))
    g = base(); g.menu[2].sizes[1] = -0.5; out.append(("negative size", g))
    g = base(); g.menu[2].raise_mult = 0.5; out.append(("raise multiple 0.5", g))
    g = base(); g.menu[2].raise_mult = float("inf"); out.append(("infinite raise multiple", g))
    g = base(); g.menu[2].raise_cap = 9; out.append(("raise cap 9", g))
    return out


def test_bad_hu_games_refused(mtt, S):
    n, ni, nm = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    g, _, _ = S.config1(0)
    mtt.call("mtt_hu_tree_size", ctypes.byref(g), ctypes.byref(n), ctypes.byref(ni), ctypes.byref(nm))
    assert n.value == 33
    for label, bad in _hu_variants(S):
        msg = _refused(mtt, "mtt_hu_tree_size", ctypes.byref(bad), ctypes.byref(n), ctypes.byref(ni),
                       ctypes.byref(nm))
        assert "hu game" in msg, (label, msg)
    # a buffer too small for the tree
    arr = np.zeros(10, S.FLAT_DTYPE)
    nn = ctypes.c_int()
    _refused(mtt, "mtt_hu_tree_build", ctypes.byref(g), arr.ctypes.data, 10, ctypes.byref(nn))
    _refused(mtt, "mtt_hu_tree_size", None, ctypes.byref(n), ctypes.byref(ni), ctypes.byref(nm))


def test_bad_4way_games_refused(mtt, M):
    n, ni = ctypes.c_int(), ctypes.c_int()
    ok = M.game4([2.5] * 4, [30.0, 45.0, 12.0, 60.0], M.config5_menu())
    mtt.call("mtt_4way_tree_size", ctypes.byref(ok), ctypes.byref(n), ctypes.byref(ni))
    assert n.value > 0
    bad = []
    g = M.game4([2.5] * 4, [30.0, 0.0, 12.0, 60.0], M.config5_menu()); bad.append(g)
    g = M.game4([2.5] * 4, [30.0, float("nan"), 12.0, 60.0], M.config5_menu()); bad.append(g)
    g = M.game4([2.5, -1.0, 2.5, 2.5], [30.0, 45.0, 12.0, 60.0], M.config5_menu()); bad.append(g)
    m = M.config5_menu(); m.n_sizes = 4
    bad.append(M.game4([2.5] * 4, [30.0, 45.0, 12.0, 60.0], m))
    m = M.config5_menu(); m.sizes[0] = float("nan")
    bad.append(M.game4([2.5] * 4, [30.0, 45.0, 12.0, 60.0], m))
    m = M.config5_menu(); m.sizes[0] = 0.0
    bad.append(M.game4([2.5] * 4, [30.0, 45.0, 12.0, 60.0], m))
    m = M.config5_menu(); m.raise_mult = 0.9
    bad.append(M.game4([2.5] * 4, [30.0, 45.0, 12.0, 60.0], m))
    m = M.config5_menu(); m.raise_cap = 5
    bad.append(M.game4([2.5] * 4, [30.0, 45.0, 12.0, 60.0], m))
    for g in bad:
        msg = _refused(mtt, "mtt_4way_tree_size", ctypes.byref(g), ctypes.byref(n), ctypes.byref(ni))
        assert "4-way game" in msg, msg


def test_bad_cards_and_flops_refused(mtt):
    L = mtt.lib()
    assert L.mtt_canonical_class(0, 1) == 12
    for a, b in ((0, 0), (-1, 3), (3, 52), (60, 2)):
        assert L.mtt_canonical_class(a, b) == -1
    W = np.zeros(4, np.int32)
    for flop in ([3, 3, 9], [0, 1, 52], [200, 1, 2]):
        f = np.array(flop, np.uint8)
        _refused(mtt, "mtt_flop_counts", f.ctypes.data, 1, W.ctypes.data, W.ctypes.data, 0, None)
        _refused(mtt, "mtt_flop_class_matrices", f.ctypes.data, 1, W.ctypes.data, None)
    ok = np.array([0, 1, 2], np.uint8)
    _refused(mtt, "mtt_flop_counts", ok.ctypes.data, -1, W.ctypes.data, W.ctypes.data, 0, None)
    _refused(mtt, "mtt_flop_counts", ok.ctypes.data, 1, None, W.ctypes.data, 0, None)
    _refused(mtt, "mtt_flop_counts", ok.ctypes.data, 1, W.ctypes.data, W.ctypes.data, 7, None)
    _refused(mtt, "mtt_flop_class_matrices", ok.ctypes.data, 5000, W.ctypes.data, None)
    _refused(mtt, "mtt_gen_preflop_matrix", W.ctypes.data, None, None, 0, None)
    _refused(mtt, "mtt_class_cards", 169, W.ctypes.data, W.ctypes.data)


def test_bad_preflop_solve_arguments_refused(mtt):
    """Checked on the host before anything reaches the device."""
    buf = np.zeros(16, np.float32)
    p = buf.ctypes.data
    for w, n, iters, stack, bf, dead in (
            ([float("nan")], 1, 10, 20.0, 1.35, 2.5),   # NaN subset weight
            ([-1.0], 1, 10, 20.0, 1.35, 2.5),           # negative weight
            ([1.0], 0, 10, 20.0, 1.35, 2.5),            # no flops
            ([1.0], 1, -1, 20.0, 1.35, 2.5),            # negative iterations
            ([1.0], 1, 10, 0.0, 1.35, 2.5),             # no stack
            ([1.0], 1, 10, 20.0, float("inf"), 2.5),    # infinite bubble factor
            ([1.0], 1, 10, 20.0, 1.35, -1.0)):          # negative dead pot
        wt = np.array(w, np.float32)
        _refused(mtt, "mtt_preflop_solve", p, wt.ctypes.data, n, p, iters, stack, bf, dead, p, None)
    wt = np.array([1.0], np.float32)
    _refused(mtt, "mtt_preflop_solve", None, wt.ctypes.data, 1, p, 10, 20.0, 1.35, 2.5, p, None)


def test_scan_and_sweep_tables_refuse_bad_boards(mtt, S):
    cards = np.array([[0, 1], [2, 3]], np.uint8)
    rank = np.array([0, 1], np.uint16)
    h = np.zeros(3 * 1088, np.uint32)
    lt = np.zeros(2 * 12 * 64, np.uint32)
    _refused(mtt, "mtt_river_scan_tables", cards.ctypes.data, rank.ctypes.data, 2, h.ctypes.data, lt.ctypes.data)
    ln = ctypes.c_size_t()
    bad_rank = np.array([1, 0], np.uint16)
    _refused(mtt, "mtt_river_sweep_program", cards.ctypes.data, bad_rank.ctypes.data, 2, None, 0, ctypes.byref(ln))
