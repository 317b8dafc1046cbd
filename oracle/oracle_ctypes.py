"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker; never by the product.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None
_P = ctypes.c_void_p


def lib():
    """The oracle library; MTT_ORACLE_LIB names another build of the same
    sources (the sanitizer test's instrumented one)."""
    global _lib
    if _lib is None:
        path = os.environ.get("MTT_ORACLE_LIB", LIB_PATH)
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run make -C oracle")
        L = ctypes.CDLL(path)
        L.or_eval7.restype = ctypes.c_uint32
        L.or_eval7.argtypes = [_P]
        L.or_eval7_batch.argtypes = [_P, ctypes.c_int64, _P]
        L.or_category_counts.argtypes = [_P]
        L.or_canonical_class.argtypes = [ctypes.c_int, ctypes.c_int]
        L.or_canonical_flops.argtypes = [_P]
        L.or_ev_dump.argtypes = [ctypes.c_int, ctypes.c_int, _P]
        L.or_combo_index.argtypes = [ctypes.c_int, ctypes.c_int]
        L.or_flop_counts.argtypes = [_P, _P, _P]
        L.or_turn_matrix.argtypes = [_P, ctypes.c_int, _P]
        L.or_set_threads.argtypes = [ctypes.c_int]
        L.or_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_P)


def eval7_batch(cards):
    cards = np.ascontiguousarray(cards, dtype=np.uint8)
    out = np.zeros(cards.shape[0], np.uint32)
    lib().or_eval7_batch(_p(cards), cards.shape[0], _p(out))
    return out


def set_threads(n):
    """OpenMP threads for the oracle's parallel loops."""
    lib().or_set_threads(int(n))


def max_threads():
    return int(lib().or_max_threads())


def class_to_cards(cls):
    a, b = ctypes.c_uint8(), ctypes.c_uint8()
    lib().or_class_to_cards(int(cls), ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def preflop_cell(h, v, samples):
    w, t = ctypes.c_int32(), ctypes.c_int32()
    lib().or_preflop_cell(int(h), int(v), int(samples), ctypes.byref(w), ctypes.byref(t))
    return w.value, t.value


def preflop_matrix(samples=3000):
    """(wins, ties, equity) [169][169] of the Monte-Carlo all-in matrix."""
    w = np.zeros((169, 169), np.int32)
    t = np.zeros((169, 169), np.int32)
    e = np.zeros((169, 169), np.float32)
    lib().or_preflop_matrix(int(samples), _p(w), _p(t), _p(e))
    return w, t, e


def preflop_exact(h, v):
    """Exact mean and variance of one sample's score (1 win, 1/2 tie) under
    the Monte-Carlo kernel's board-sampling process."""
    m, s = ctypes.c_double(), ctypes.c_double()
    lib().or_preflop_exact(int(h), int(v), ctypes.byref(m), ctypes.byref(s))
    return m.value, s.value


def category_counts():
    out = np.zeros(9, np.uint64)
    lib().or_category_counts(_p(out))
    return out


def canonical_flops():
    out = np.zeros((1755, 4), np.uint8)
    n = lib().or_canonical_flops(_p(out))
    assert n == 1755
    return out


def ev_dump(first, count):
    out = np.zeros((count, 4, 2, 169), np.float32)
    lib().or_ev_dump(first, count, _p(out))
    return out


def flop_counts(flop):
    f = np.asarray(flop, np.uint8)
    W = np.zeros((1326, 1326), np.int32)
    T = np.zeros((1326, 1326), np.int32)
    lib().or_flop_counts(_p(f), _p(W), _p(T))
    return W, T


def turn_matrix(flop, turn_card):
    f = np.asarray(flop, np.uint8)
    E = np.zeros((1216, 1216), np.float32)
    lib().or_turn_matrix(_p(f), int(turn_card), _p(E))
    return E


class HU:
    """Oracle heads-up CFR over a FlatNode tree (numpy structured array with
    the 32-byte node layout).  precision 0 = FP64, 1 = FP32."""

    def __init__(self, nodes, n_infosets, board, r0_alive, r1_alive, precision=0):
        L = lib()
        L.or_hu_create.restype = _P
        L.or_hu_create.argtypes = [_P, ctypes.c_int, ctypes.c_int, _P, ctypes.c_int, _P, _P, ctypes.c_int]
        L.or_hu_destroy.argtypes = [_P]
        L.or_hu_n_combos.argtypes = [_P]
        L.or_hu_iterate.argtypes = [_P, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.or_hu_avg_strategy.argtypes = [_P, _P]
        L.or_hu_state.argtypes = [_P, _P, _P]
        L.or_hu_exploitability.argtypes = [_P, _P, ctypes.c_double, _P]
        L.or_hu_terminal.argtypes = [_P, ctypes.c_int, _P, ctypes.c_int, _P, _P, _P, _P]
        self.nodes = np.ascontiguousarray(nodes)
        self.n_infosets = n_infosets
        self.board = np.ascontiguousarray(board, np.uint8)
        self.r0 = np.ascontiguousarray(r0_alive, np.float32)
        self.r1 = np.ascontiguousarray(r1_alive, np.float32)
        self.h = L.or_hu_create(self.nodes.ctypes.data, len(self.nodes), n_infosets, self.board.ctypes.data,
                                len(self.board), self.r0.ctypes.data, self.r1.ctypes.data, precision)
        self.N = L.or_hu_n_combos(self.h)

    def iterate(self, n, scheme=0, alpha=1.5, beta=0.0, gamma=2.0):
        lib().or_hu_iterate(self.h, n, scheme, alpha, beta, gamma)

    def strategy(self):
        out = np.zeros(self.n_infosets * 4 * 1216, np.float32)
        lib().or_hu_avg_strategy(self.h, _p(out))
        return out

    def state(self):
        R = np.zeros(self.n_infosets * 4 * 1216, np.float32)
        S = np.zeros_like(R)
        lib().or_hu_state(self.h, _p(R), _p(S))
        return R, S

    def exploitability(self, strat, pot):
        out = np.zeros(6, np.float64)
        s = np.ascontiguousarray(strat, np.float32)
        lib().or_hu_exploitability(self.h, _p(s), pot, _p(out))
        return out

    def terminal(self, node, dealt, r0, r1):
        d = np.ascontiguousarray(dealt, np.uint8)
        a = np.ascontiguousarray(r0, np.float64)
        b = np.ascontiguousarray(r1, np.float64)
        v0 = np.zeros(1216)
        v1 = np.zeros(1216)
        lib().or_hu_terminal(self.h, node, _p(d) if len(d) else None, len(d), _p(a), _p(b), _p(v0), _p(v1))
        return v0, v1

    def close(self):
        if self.h:
            lib().or_hu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------ 4-way (or_multi.c)
BOARD4_DTYPE = np.dtype([("c1", "u1", (1216,)), ("c2", "u1", (1216,)), ("rank_pos", "<u2", (1216,)),
                         ("pos_combo", "<u2", (1216,)), ("grp_lo", "<u2", (1216,)), ("grp_hi", "<u2", (1216,)),
                         ("card_list", "<u2", (52, 48)), ("n_card", "u1", (52,)), ("card_lt", "u1", (1216, 2)),
                         ("card_le", "u1", (1216, 2))])


def _multi_sigs():
    L = lib()
    if getattr(L, "_multi_ready", False):
        return L
    L.or_evaln.argtypes = [_P, ctypes.c_int]
    L.or_evaln.restype = ctypes.c_uint32
    L.or_4way_board.argtypes = [_P, _P]
    L.or_4way_terminal.argtypes = [_P, _P, _P, _P, _P]
    L.or_4way_terminal_brute.argtypes = [_P, _P, _P, _P, _P, _P]
    L.or_dcfr_step.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P, _P, _P, _P, ctypes.c_float,
                               ctypes.c_float, ctypes.c_float]
    L.or_icm.argtypes = [_P, ctypes.c_int, _P, ctypes.c_int, _P]
    L.or_icm_brute.argtypes = [_P, ctypes.c_int, _P, ctypes.c_int, _P]
    L.or_bubble_factors.argtypes = [_P, ctypes.c_int, _P, ctypes.c_int, _P, ctypes.c_int, _P]
    L.or_4way_create.argtypes = [_P, ctypes.c_int, ctypes.c_int, _P, _P, _P]
    L.or_4way_create.restype = _P
    L.or_4way_destroy.argtypes = [_P]
    L.or_4way_iterate.argtypes = [_P, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                  ctypes.c_int]
    L.or_4way_state.argtypes = [_P, _P, _P, _P]
    L._multi_ready = True
    return L


def evaln(cards):
    c = np.ascontiguousarray(cards, np.uint8)
    return int(_multi_sigs().or_evaln(c.ctypes.data, len(c)))


def board4(flop):
    b = np.zeros(1, BOARD4_DTYPE)
    f = np.ascontiguousarray(flop, np.uint8)
    _multi_sigs().or_4way_board(f.ctypes.data, b.ctypes.data)
    return b


def terminal4(board, node, reach, bf):
    """node: one record of the 4-way node dtype; reach [4][1216] float32;
    bf [4][4] float64.  Returns [4][1216] float64."""
    out = np.zeros(4 * 1216, np.float64)
    nd = np.ascontiguousarray(node)
    r = np.ascontiguousarray(reach, np.float32)
    b = np.ascontiguousarray(bf, np.float64)
    _multi_sigs().or_4way_terminal(board.ctypes.data, nd.ctypes.data, r.ctypes.data, b.ctypes.data, out.ctypes.data)
    return out.reshape(4, 1216)


def terminal4_brute(board, node, reach, bf, score):
    out = np.zeros(4 * 1216, np.float64)
    nd = np.ascontiguousarray(node)
    r = np.ascontiguousarray(reach, np.float32)
    b = np.ascontiguousarray(bf, np.float64)
    sc = np.ascontiguousarray(score, np.uint32)
    _multi_sigs().or_4way_terminal_brute(board.ctypes.data, nd.ctypes.data, r.ctypes.data, b.ctypes.data,
                                         sc.ctypes.data, out.ctypes.data)
    return out.reshape(4, 1216)


def dcfr_step(num_actions, R, sigma, S, cfv, reach, w_pos, w_neg, w_strat):
    """In place on float32 arrays [num_actions][stride]; reach [n]."""
    n = len(reach)
    stride = R.shape[1]
    _multi_sigs().or_dcfr_step(num_actions, n, stride, R.ctypes.data, sigma.ctypes.data, S.ctypes.data,
                               cfv.ctypes.data, reach.ctypes.data, w_pos, w_neg, w_strat)


def icm(stacks, payouts, brute=False):
    st = np.ascontiguousarray(stacks, np.float64)
    pay = np.ascontiguousarray(payouts, np.float64)
    out = np.zeros(len(st), np.float64)
    fn = _multi_sigs().or_icm_brute if brute else _multi_sigs().or_icm
    fn(st.ctypes.data, len(st), pay.ctypes.data, len(pay), out.ctypes.data)
    return out


def bubble_factors(stacks, payouts, in_pot):
    st = np.ascontiguousarray(stacks, np.float64)
    pay = np.ascontiguousarray(payouts, np.float64)
    ip = np.ascontiguousarray(in_pot, np.int32)
    out = np.zeros((len(ip), len(ip)), np.float64)
    _multi_sigs().or_bubble_factors(st.ctypes.data, len(st), pay.ctypes.data, len(pay), ip.ctypes.data, len(ip),
                                    out.ctypes.data)
    return out


class FourWay:
    """FP32 4-way CFR+/DCFR loop of one flop (or_multi.c)."""

    def __init__(self, nodes, n_infosets, flop, ranges, bf):
        L = _multi_sigs()
        self._nodes = np.ascontiguousarray(nodes)
        self.n_info = n_infosets
        f = np.ascontiguousarray(flop, np.uint8)
        r = np.ascontiguousarray(ranges, np.float32)
        b = np.ascontiguousarray(bf, np.float64)
        self.h = L.or_4way_create(self._nodes.ctypes.data, len(self._nodes), n_infosets, f.ctypes.data, r.ctypes.data,
                                  b.ctypes.data)
        self.t = 0

    def iterate(self, n, scheme=1, alpha=1.5, beta=0.0, gamma=2.0):
        _multi_sigs().or_4way_iterate(self.h, n, self.t + 1, alpha, beta, gamma, scheme)
        self.t += n

    def state(self):
        k = self.n_info * 3 * 1216
        R = np.zeros(k, np.float32)
        S = np.zeros(k, np.float32)
        root = np.zeros(4 * 1216, np.float32)
        _multi_sigs().or_4way_state(self.h, R.ctypes.data, S.ctypes.data, root.ctypes.data)
        return R, S, root

    def close(self):
        if self.h:
            _multi_sigs().or_4way_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def flop_matrix_entry(flop, hero, villain):
    L = lib()
    L.or_flop_matrix_entry.argtypes = [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.or_flop_matrix_entry.restype = ctypes.c_float
    f = np.ascontiguousarray(flop, np.uint8)
    return float(L.or_flop_matrix_entry(f.ctypes.data, int(hero[0]), int(hero[1]), int(villain[0]), int(villain[1])))


def flop_class_matrix(flop):
    """[169 (villain class)][169 (hero class)] float32 class equities of one
    flop (or_flop_class_matrix)."""
    M = np.zeros((169, 169), np.float32)
    f = np.asarray(flop, np.uint8)
    lib().or_flop_class_matrix(_p(f), _p(M))
    return M


def preflop_solve(M, weights, allin, iters=500, stack_bb=20.0, bf=1.35, dead_pot=2.5):
    """Strategy sums [4 nodes][3 actions][169] of the preflop DCFR loop
    (or_preflop_solve) over class matrices M [n][169][169]."""
    M = np.ascontiguousarray(M, np.float32)
    w = np.ascontiguousarray(weights, np.float32)
    a = np.ascontiguousarray(allin, np.float32)
    out = np.zeros((4, 3, 169), np.float32)
    lib().or_preflop_solve(int(M.shape[0]), _p(w), _p(M), _p(a), int(iters), ctypes.c_float(stack_bb),
                           ctypes.c_float(bf), ctypes.c_float(dead_pot), _p(out))
    return out
