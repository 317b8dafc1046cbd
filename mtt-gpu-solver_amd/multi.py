"""Host-side helpers for the 4-way engine, ICM and the DCFR step over the
C ABI: the 4-way game description, trees, flop board tables, the inputs of
the 4-way benchmark configuration, and a small wrapper of the batched
4-way solver.

Loaded by the tests and the benchmark next to mtt.py (same directory).
"""
import ctypes
import importlib.util
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
if "mtt_solver" in sys.modules and hasattr(sys.modules["mtt_solver"], "menu"):
    S = sys.modules["mtt_solver"]
else:
    _spec = importlib.util.spec_from_file_location("mtt_solver", os.path.join(_HERE, "solver.py"))
    S = importlib.util.module_from_spec(_spec)
    sys.modules["mtt_solver"] = S
    _spec.loader.exec_module(S)
mtt = S.mtt

_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
N_PAD, N_ALIVE, A4, P4 = 1216, 1176, 3, 4


class Game4(ctypes.Structure):
    _fields_ = [("pot0", ctypes.c_float * 4), ("stack", ctypes.c_float * 4), ("menu", S.Menu)]


for name, args in {
    "mtt_4way_tree_size": [_P, _P, _P],
    "mtt_4way_tree_build": [_P, _P, _I, _P],
    "mtt_board4_build": [_P, _P],
    "mtt_4way_forward_sweep": [_P, _I, _P, _P, _P, _P, _P],
    "mtt_4way_evaluate_terminals": [_P, _I, _P, _P, _P, _P, _P, _P, _P],
    "mtt_4way_evaluate_terminals_board": [_P, _I, _P, _P, _P, _P, _P, _P],
    "mtt_4way_backward_cfr_plus": [_P, _I, _P, _P, _P, _P, _P, _P, _F, _P],
    "mtt_4way_backward_dcfr": [_P, _I, _P, _P, _P, _P, _P, _P, _F, _F, _F, _P],
    "mtt_multi_create": [_P, _I, _P, _P, _P, _P, _P],
    "mtt_multi_run": [_P, _I],
    "mtt_multi_get_state": [_P, _I, _P, _P, _P],
    "mtt_multi_timing": [_P, _I],
    "mtt_multi_kernel_times": [_P, _I, ctypes.c_char_p, _I, _P, _P],
    "mtt_multi_destroy": [_P],
    "mtt_dcfr_step": [_I, _P, _P, _P, _P, _P, _F, _F, _F, _I, _I, _P],
    "mtt_icm_mh_fp64": [_P, _I, _P, _I, _P],
    "mtt_icm_mh_fp64_batch": [_P, _I, _I, _P, _I, _P, _P],
    "mtt_icm_bubble_factors_fp64": [_P, _I, _P, _I, _P, _I, _P],
}.items():
    mtt.register(name, args)

NODE4_DTYPE = np.dtype([("type", "u1"), ("player", "u1"), ("num_actions", "u1"), ("active_mask", "u1"),
                        ("children_offset", "<i4"), ("infoset_offset", "<i4"), ("pot_size", "<f4"),
                        ("committed", "<f4", (4,))])
assert NODE4_DTYPE.itemsize == 32

BOARD4_DTYPE = np.dtype([("c1", "u1", (1216,)), ("c2", "u1", (1216,)), ("rank_pos", "<u2", (1216,)),
                         ("pos_combo", "<u2", (1216,)), ("grp_lo", "<u2", (1216,)), ("grp_hi", "<u2", (1216,)),
                         ("card_list", "<u2", (52, 48)), ("n_card", "u1", (52,)), ("card_lt", "u1", (1216, 2)),
                         ("card_le", "u1", (1216, 2))])


def game4(pot0, stack, m):
    g = Game4()
    for p in range(4):
        g.pot0[p] = pot0[p]
        g.stack[p] = stack[p]
    g.menu = m
    return g


def tree4(game):
    n, ni = ctypes.c_int(), ctypes.c_int()
    mtt.call("mtt_4way_tree_size", ctypes.byref(game), ctypes.byref(n), ctypes.byref(ni))
    arr = np.zeros(n.value, NODE4_DTYPE)
    nn = ctypes.c_int()
    mtt.call("mtt_4way_tree_build", ctypes.byref(game), arr.ctypes.data, n.value, ctypes.byref(nn))
    return arr, ni.value


def board4(flop):
    b = np.zeros(1, BOARD4_DTYPE)
    f = np.asarray(flop, np.uint8)
    mtt.call("mtt_board4_build", f.ctypes.data, b.ctypes.data)
    return b


def icm(stacks, payouts):
    st = np.ascontiguousarray(stacks, np.float64)
    pay = np.ascontiguousarray(payouts, np.float64)
    out = np.zeros(len(st), np.float64)
    mtt.call("mtt_icm_mh_fp64", st.ctypes.data, len(st), pay.ctypes.data, len(pay), out.ctypes.data)
    return out


def bubble_factors(stacks, payouts, in_pot):
    st = np.ascontiguousarray(stacks, np.float64)
    pay = np.ascontiguousarray(payouts, np.float64)
    ip = np.ascontiguousarray(in_pot, np.int32)
    out = np.zeros((len(ip), len(ip)), np.float64)
    mtt.call("mtt_icm_bubble_factors_fp64", st.ctypes.data, len(st), pay.ctypes.data, len(pay), ip.ctypes.data,
             len(ip), out.ctypes.data)
    return out


def config5_menu():
    """33%/75% pot bets and one 2.5x raise, no all-in option."""
    return S.menu((0.33, 0.75), 2.5, 1, 0)


def config5(seed=0, n_flops=100):
    """4-way flop engine with ICM payouts: eight stacks Uniform[5,80]bb, the
    first four seats in the pot with 2.5bb each; payouts 50/30/20% of the
    prize pool; flops drawn without replacement per instance; ranges
    Uniform[0,1] over each flop's 1176 combos, normalised.  Inputs come from
    numpy's PCG64 with the given seed (host-generated, shared by every path)."""
    rng = np.random.default_rng(seed)
    stacks8 = rng.uniform(5.0, 80.0, 8)
    payouts = np.array([0.5, 0.3, 0.2])
    pot0 = [2.5] * 4
    behind = [float(stacks8[p]) - 2.5 for p in range(4)]
    game = game4(pot0, behind, config5_menu())
    flops = np.zeros((n_flops, 3), np.uint8)
    ranges = np.zeros((n_flops, 4, N_PAD), np.float32)
    for f in range(n_flops):
        flops[f] = np.sort(rng.choice(52, 3, replace=False))
        for p in range(4):
            w = rng.uniform(0.0, 1.0, N_ALIVE).astype(np.float32)
            ranges[f, p, :N_ALIVE] = w / w.sum()
    bf = bubble_factors(stacks8, payouts, [0, 1, 2, 3]).astype(np.float32)
    return game, flops, ranges, bf, stacks8, payouts


class Multi:
    """Batched 4-way solve of one tree over many flops."""

    def __init__(self, game, flops, ranges, bf, scheme=S.SCHEME_DCFR, alpha=1.5, beta=0.0, gamma=2.0):
        self.game = game
        self.flops = np.ascontiguousarray(flops, np.uint8)
        self.ranges = np.ascontiguousarray(ranges, np.float32)
        self.bf = np.ascontiguousarray(bf, np.float32)
        cfg = S.SolverCfg(0, scheme, alpha, beta, gamma)
        h = ctypes.c_void_p()
        mtt.call("mtt_multi_create", ctypes.byref(game), len(self.flops), self.flops.ctypes.data,
                 self.ranges.ctypes.data, self.bf.ctypes.data, ctypes.byref(cfg), ctypes.byref(h))
        self.h = h
        self.nodes, self.n_info = tree4(game)

    def run(self, iters):
        mtt.call("mtt_multi_run", self.h, iters)

    def state(self, flop):
        n = self.n_info * A4 * N_PAD
        R = np.zeros(n, np.float32)
        Sm = np.zeros(n, np.float32)
        root = np.zeros(P4 * N_PAD, np.float32)
        mtt.call("mtt_multi_get_state", self.h, flop, R.ctypes.data, Sm.ctypes.data, root.ctypes.data)
        return R, Sm, root

    def timing(self, on):
        mtt.call("mtt_multi_timing", self.h, 1 if on else 0)

    def kernel_times(self):
        out = {}
        buf = ctypes.create_string_buffer(64)
        k = 0
        while True:
            ms, n = ctypes.c_double(), ctypes.c_int64()
            rc = mtt.lib().mtt_multi_kernel_times(self.h, k, buf, 64, ctypes.byref(ms), ctypes.byref(n))
            if rc != 0:
                break
            out[buf.value.decode()] = (ms.value, n.value)
            k += 1
        return out

    def close(self):
        if self.h:
            mtt.lib().mtt_multi_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
