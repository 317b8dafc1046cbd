"""Host-side helpers over the solver C ABI: game descriptions of the
benchmark configurations, seeded inputs, and a small Solver wrapper.

Loaded by the tests and the benchmark next to mtt.py (same directory).
"""
import ctypes
import importlib.util
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
if "mtt" in sys.modules and hasattr(sys.modules["mtt"], "lib"):
    mtt = sys.modules["mtt"]
else:
    _spec = importlib.util.spec_from_file_location("mtt", os.path.join(_HERE, "mtt.py"))
    mtt = importlib.util.module_from_spec(_spec)
    sys.modules["mtt"] = mtt
    _spec.loader.exec_module(mtt)

_P = ctypes.c_void_p
_I = ctypes.c_int

MODE_PLAIN_FP32, MODE_BF16_SPLIT = 0, 1
SCHEME_CFR_PLUS, SCHEME_DCFR = 0, 1


class Menu(ctypes.Structure):
    _fields_ = [("n_sizes", ctypes.c_int), ("sizes", ctypes.c_float * 3), ("raise_mult", ctypes.c_float),
                ("raise_cap", ctypes.c_int), ("allin", ctypes.c_int)]


class Game(ctypes.Structure):
    _fields_ = [("board", ctypes.c_uint8 * 5), ("n_board", ctypes.c_int), ("root_chance", ctypes.c_int),
                ("last_street", ctypes.c_int), ("pot", ctypes.c_float), ("stack", ctypes.c_float),
                ("menu", Menu * 3)]


class SolverCfg(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("scheme", ctypes.c_int), ("alpha", ctypes.c_float),
                ("beta", ctypes.c_float), ("gamma", ctypes.c_float)]


for name, args in {
    "mtt_seeded_cards": [ctypes.c_uint64, _I, _P],
    "mtt_seeded_inputs": [ctypes.c_uint64, _I, _I, ctypes.c_float, _P, _P, _P],
    "mtt_hu_tree_size": [_P, _P, _P, _P],
    "mtt_hu_tree_build": [_P, _P, _I, _P],
    "mtt_solver_create": [_P, _P, _P, _P, _P],
    "mtt_solver_run": [_P, _I],
    "mtt_solver_iterations": [_P],
    "mtt_solver_get_strategy": [_P, _P, ctypes.c_size_t],
    "mtt_solver_exploitability_fp64": [_P, _P],
    "mtt_solver_info": [_P, _P],
    "mtt_solver_jit_info": [_P, _P, _P],
    "mtt_hu_sweep_source": [_P, _P, ctypes.c_size_t, _P],
    "mtt_hu_sweep_compile": [_P, _P, _P],
    "mtt_solver_river_showdown": [_P, ctypes.c_int, _P, _P, _P, _P, _P, _P],
    "mtt_river_sweep_program": [_P, _P, ctypes.c_int, _P, ctypes.c_size_t, _P],
    "mtt_river_scan_tables": [_P, _P, ctypes.c_int, _P, _P],
    "mtt_solver_get_state": [_P, _P, _P, ctypes.c_size_t],
    "mtt_solver_timing": [_P, _I],
    "mtt_solver_kernel_times": [_P, _I, ctypes.c_char_p, _I, _P, _P],
    "mtt_solver_destroy": [_P],
    "mtt_hu_forward_sweep": [_P, _I, _P, _P, _P, _P, _P, _P, _I, _I, _P],
    "mtt_hu_evaluate_terminals": [_P, _I, _P, _P, _P, _P, _P, _P, _I, _P],
    "mtt_hu_backward_sweep": [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, ctypes.c_float, _I, _I, _P],
    "mtt_hu_backward_sweep_dcfr": [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, ctypes.c_float, ctypes.c_float,
                                   ctypes.c_float, _I, _I, _P],
}.items():
    mtt.register(name, args)

FLAT_DTYPE = np.dtype([("type", "u1"), ("player", "u1"), ("num_actions", "u1"), ("active_mask", "u1"),
                       ("children_offset", "<i4"), ("infoset_offset", "<i4"), ("turn_idx", "<i2"),
                       ("river_idx", "<i2"), ("pot_size", "<f4"), ("committed", "<f4", (2,)), ("_pad", "V4")])
assert FLAT_DTYPE.itemsize == 32


def menu(sizes=(), raise_mult=0.0, raise_cap=0, allin=1):
    m = Menu()
    m.n_sizes = len(sizes)
    for k, s in enumerate(sizes):
        m.sizes[k] = s
    m.raise_mult = raise_mult
    m.raise_cap = raise_cap
    m.allin = allin
    return m


def seeded_inputs(seed, n_board, dist, alpha=1.0):
    board = np.zeros(5, np.uint8)
    r0 = np.zeros(1326, np.float32)
    r1 = np.zeros(1326, np.float32)
    mtt.call("mtt_seeded_inputs", seed, n_board, dist, alpha, board.ctypes.data, r0.ctypes.data, r1.ctypes.data)
    return board[:n_board].copy(), r0, r1


def make_game(board, pot, stack, menus, root_chance=0, last_street=2):
    g = Game()
    for k, c in enumerate(board):
        g.board[k] = int(c)
    g.n_board = len(board)
    g.root_chance = root_chance
    g.last_street = last_street
    g.pot = pot
    g.stack = stack
    for s, m in menus.items():
        g.menu[s] = m
    return g


def config1(seed=0):
    """HU river: 5-card seeded board, uniform[0,1] ranges over unblocked
    combos, pot 10, stack 20, bet 50%/100%/all-in, one raise (all-in)."""
    board, r0, r1 = seeded_inputs(seed, 5, 0)
    g = make_game(board, 10.0, 20.0, {2: menu((0.5, 1.0), 0.0, 1, 1)}, 0, 2)
    return g, r0, r1


def config3(seed=0, menu_turn=None, menu_river=None):
    """HU turn+river: seeded flop, root deals the turn; pot 6, stack 40; per
    street 33%/75%, one 2.5x raise, all-in; Dirichlet(0.5) ranges."""
    board, r0, r1 = seeded_inputs(seed, 3, 1, 0.5)
    m = menu((0.33, 0.75), 2.5, 1, 1)
    g = make_game(board, 6.0, 40.0, {1: menu_turn or m, 2: menu_river or m}, 1, 2)
    return g, r0, r1


def config3_small(seed=0):
    """Same street structure as config 3 with an all-in-only menu (for
    oracle-sized multi-street parity)."""
    m = menu((), 0.0, 1, 1)
    return config3(seed, m, m)


def tree(game):
    n, ni, nm = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    mtt.call("mtt_hu_tree_size", ctypes.byref(game), ctypes.byref(n), ctypes.byref(ni), ctypes.byref(nm))
    arr = np.zeros(n.value, FLAT_DTYPE)
    nn = ctypes.c_int()
    mtt.call("mtt_hu_tree_build", ctypes.byref(game), arr.ctypes.data, n.value, ctypes.byref(nn))
    return arr, ni.value, nm.value


def sweep_source(game):
    """Generated sweep source of a game's street templates (batched path)."""
    n = ctypes.c_size_t()
    mtt.call("mtt_hu_sweep_source", ctypes.byref(game), None, 0, ctypes.byref(n))
    buf = ctypes.create_string_buffer(n.value + 1)
    mtt.call("mtt_hu_sweep_source", ctypes.byref(game), buf, n.value + 1, ctypes.byref(n))
    return buf.value.decode()


def sweep_compile(game):
    """hipRTC compile-only build of the generated sweeps: (code bytes, seconds)."""
    n, t = ctypes.c_size_t(), ctypes.c_double()
    mtt.call("mtt_hu_sweep_compile", ctypes.byref(game), ctypes.byref(n), ctypes.byref(t))
    return n.value, t.value


def alive_index(board):
    """Root-alive combo list (lexicographic) and its 1326 global indices."""
    bs = set(int(c) for c in board)
    pairs = [(a, b) for a in range(52) for b in range(a + 1, 52)]
    alive = [(k, a, b) for k, (a, b) in enumerate(pairs) if a not in bs and b not in bs]
    return np.array([k for k, _, _ in alive]), np.array([[a, b] for _, a, b in alive], np.uint8)


class Solver:
    def __init__(self, game, r0, r1, mode=MODE_PLAIN_FP32, scheme=SCHEME_CFR_PLUS, alpha=1.5, beta=0.0, gamma=2.0):
        self.game = game
        cfg = SolverCfg(mode, scheme, alpha, beta, gamma)
        h = ctypes.c_void_p()
        r0 = np.ascontiguousarray(r0, np.float32)
        r1 = np.ascontiguousarray(r1, np.float32)
        mtt.call("mtt_solver_create", ctypes.byref(game), r0.ctypes.data, r1.ctypes.data, ctypes.byref(cfg),
                 ctypes.byref(h))
        self.h = h
        info = np.zeros(8, np.int64)
        mtt.call("mtt_solver_info", self.h, info.ctypes.data)
        self.info = info

    def run(self, iters):
        mtt.call("mtt_solver_run", self.h, iters)

    def jit_info(self):
        """(generated sweeps active, hipRTC compile + load seconds)."""
        a, t = ctypes.c_int(), ctypes.c_double()
        mtt.call("mtt_solver_jit_info", self.h, ctypes.byref(a), ctypes.byref(t))
        return bool(a.value), t.value

    def river_showdown(self, board):
        """Last iteration's river showdown of one board (batched mode):
        (reach [ncol][1088], values [ncol][1088], cards [1088][2],
        group ranks [1088], live combo count)."""
        nl, nc = ctypes.c_int(), ctypes.c_int()
        mtt.call("mtt_solver_river_showdown", self.h, board, None, None, None, None, ctypes.byref(nl),
                 ctypes.byref(nc))
        reach = np.zeros((nc.value, mtt.NPR), np.float32)
        vals = np.zeros((nc.value, mtt.NPR), np.float32)
        cards = np.zeros((mtt.NPR, 2), np.uint8)
        grank = np.zeros(mtt.NPR, np.uint16)
        mtt.call("mtt_solver_river_showdown", self.h, board, reach.ctypes.data, vals.ctypes.data, cards.ctypes.data,
                 grank.ctypes.data, ctypes.byref(nl), ctypes.byref(nc))
        return reach, vals, cards, grank, nl.value

    @property
    def iterations(self):
        return mtt.lib().mtt_solver_iterations(self.h)

    def strategy(self):
        n = int(self.info[1]) * mtt.MAX_ACTIONS * mtt.N_PAD
        out = np.zeros(n, np.float32)
        mtt.call("mtt_solver_get_strategy", self.h, out.ctypes.data, n)
        return out

    def state(self):
        n = int(self.info[1]) * mtt.MAX_ACTIONS * mtt.N_PAD
        R = np.zeros(n, np.float32)
        S = np.zeros(n, np.float32)
        mtt.call("mtt_solver_get_state", self.h, R.ctypes.data, S.ctypes.data, n)
        return R, S

    def timing(self, enable):
        mtt.call("mtt_solver_timing", self.h, 1 if enable else 0)

    def kernel_times(self):
        """{kernel name: (total ms, launches)} recorded since timing(True)."""
        out = {}
        k = 0
        while True:
            name = ctypes.create_string_buffer(64)
            ms = ctypes.c_double()
            n = ctypes.c_int64()
            rc = mtt.lib().mtt_solver_kernel_times(self.h, k, name, 64, ctypes.byref(ms), ctypes.byref(n))
            if rc != 0:
                break
            out[name.value.decode()] = (ms.value, n.value)
            k += 1
        return out

    def exploitability(self):
        out = np.zeros(6, np.float64)
        mtt.call("mtt_solver_exploitability_fp64", self.h, out.ctypes.data)
        return out

    def close(self):
        if self.h:
            mtt.lib().mtt_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------- batched flop subgames
for name, args in {
    "mtt_batch_create": [_P, _I, _P, _P, _P],
    "mtt_batch_run": [_P, _I],
    "mtt_batch_iterations": [_P],
    "mtt_batch_tree_info": [_P, _I, _P, _P],
    "mtt_batch_get_state": [_P, _I, _P, _P, ctypes.c_size_t],
    "mtt_batch_get_strategy": [_P, _I, _P, ctypes.c_size_t],
    "mtt_batch_exploitability_fp64": [_P, _I, _P],
    "mtt_batch_set_graph": [_P, _I],
    "mtt_batch_tuning": [_P, _P, _P],
    "mtt_batch_iter_mode": [_P, _P, _P],
    "mtt_batch_timing": [_P, _I],
    "mtt_batch_kernel_times": [_P, _I, ctypes.c_char_p, _I, _P, _P],
    "mtt_batch_destroy": [_P],
}.items():
    mtt.register(name, args)

CONFIG4_SIZES = (0.25, 0.33, 0.5, 0.75, 1.0, 1.5)


def config4(seed=0, n=512):
    """Batched uneven heads-up flop subgames: per instance a seeded flop and
    Uniform[0,1] ranges over its unblocked combos (the project generator,
    seed * 100003 + k), stacks Uniform[5,60]bb, pot Uniform[2,12]bb, a menu
    of 1-3 sizes from {25,33,50,75,100,150}% pot with one 2.5x raise size
    and a raise cap of 1-3 (no all-in option); flop betting, then showdown
    over the turn and river.  Menu/stack/pot draws from numpy's PCG64."""
    rng = np.random.default_rng(seed)
    games, ranges = [], np.zeros((n, 2, 1326), np.float32)
    for k in range(n):
        board, r0, r1 = seeded_inputs(seed * 100003 + k, 3, 0)
        stack = float(rng.uniform(5.0, 60.0))
        pot = float(rng.uniform(2.0, 12.0))
        ns = int(rng.integers(1, 4))
        sizes = tuple(sorted(rng.choice(CONFIG4_SIZES, ns, replace=False).tolist()))
        cap = int(rng.integers(1, 4))
        games.append(make_game(board, pot, stack, {0: menu(sizes, 2.5, cap, 0)}, 0, 0))
        ranges[k, 0], ranges[k, 1] = r0, r1
    return games, ranges


class Batch:
    """Many heads-up flop subgames solved together (mtt_batch_*)."""

    def __init__(self, games, ranges, scheme=SCHEME_DCFR, alpha=1.5, beta=0.0, gamma=2.0):
        self.games = (Game * len(games))(*games)
        self.ranges = np.ascontiguousarray(ranges, np.float32)
        cfg = SolverCfg(MODE_BF16_SPLIT, scheme, alpha, beta, gamma)
        h = ctypes.c_void_p()
        mtt.call("mtt_batch_create", self.games, len(games), self.ranges.ctypes.data, ctypes.byref(cfg),
                 ctypes.byref(h))
        self.h = h
        self.n = len(games)

    def run(self, iters):
        mtt.call("mtt_batch_run", self.h, iters)

    def set_graph(self, on):
        mtt.call("mtt_batch_set_graph", self.h, 1 if on else 0)

    def tree_info(self, k):
        nn, ni = ctypes.c_int(), ctypes.c_int()
        mtt.call("mtt_batch_tree_info", self.h, k, ctypes.byref(nn), ctypes.byref(ni))
        return nn.value, ni.value

    def state(self, k):
        _, ni = self.tree_info(k)
        n = ni * mtt.MAX_ACTIONS * 1216
        R = np.zeros(n, np.float32)
        S = np.zeros(n, np.float32)
        mtt.call("mtt_batch_get_state", self.h, k, R.ctypes.data, S.ctypes.data, n)
        return R, S

    @property
    def iterations(self):
        return mtt.lib().mtt_batch_iterations(self.h)

    def strategy(self, k):
        """Average strategy of instance k, reference layout [infoset][4][1216]."""
        _, ni = self.tree_info(k)
        n = ni * mtt.MAX_ACTIONS * 1216
        out = np.zeros(n, np.float32)
        mtt.call("mtt_batch_get_strategy", self.h, k, out.ctypes.data, n)
        return out

    def exploitability(self, k):
        """FP64 best response of instance k: BR0, BR1, profile values, pair
        mass, exploitability in % of the instance's starting pot."""
        out = np.zeros(6, np.float64)
        mtt.call("mtt_batch_exploitability_fp64", self.h, k, out.ctypes.data)
        return out

    def tuning(self):
        ht = ctypes.c_int()
        ms = (ctypes.c_double * 5)()
        mtt.call("mtt_batch_tuning", self.h, ctypes.byref(ht), ms)
        return ht.value, tuple(ms)

    def iter_mode(self):
        """(form, (ms per iteration of form 0, form 1)): 0 = three launches
        per iteration, 1 = the persistent launch, -1 = not chosen yet."""
        m = ctypes.c_int()
        ms = (ctypes.c_double * 2)()
        mtt.call("mtt_batch_iter_mode", self.h, ctypes.byref(m), ms)
        return m.value, (ms[0], ms[1])

    def timing(self, on):
        mtt.call("mtt_batch_timing", self.h, 1 if on else 0)

    def kernel_times(self):
        out = {}
        buf = ctypes.create_string_buffer(64)
        k = 0
        while True:
            ms, n = ctypes.c_double(), ctypes.c_int64()
            rc = mtt.lib().mtt_batch_kernel_times(self.h, k, buf, 64, ctypes.byref(ms), ctypes.byref(n))
            if rc != 0:
                break
            out[buf.value.decode()] = (ms.value, n.value)
            k += 1
        return out

    def close(self):
        if self.h:
            mtt.lib().mtt_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
