"""River showdown sweep programs (host side, no GPU).

The batched path computes river showdown values C = S r -- S the
sign x blocker matrix of the board's strength order, as in the reference's
terminal evaluation (honest_multistreet_solver.cu:128-196, equity matrix x
reach) -- by a sweep in strength order driven by a per-board program of
16-word records.  These tests replay the library's programs with the
kernel's semantics (SINGLE and READ batches issue all their reads, then
their updates in member order; HALF batches add atomically, a tie group in
two passes around its READs; rows come from the current 32-row block of the
columns, one block per visit of the program; values go to a 16-row output
tile flushed by the FLUSH flag) on integer reach in 2^48
fixed point, and require the exact integers of the dense product: every
hero's (weaker - stronger) non-blocked reach, on evaluator-ranked river
boards and on synthetic extremes (no combos, one tie group, no ties, random
group sizes)."""
import ctypes
import itertools

import numpy as np
import pytest

REC = 16
OP_NOP, OP_SINGLE, OP_HALF, OP_READ, OP_END, OP_VISIT = range(6)
F_FLUSH = 1 << 9
RB, OB = 32, 16           # rows per column block, per output block
NO_BLOCK = 0xFF


@pytest.fixture(scope="module")
def lib(mtt):
    mtt.lib()  # raises if the library has not been built
    L = ctypes.CDLL(mtt.LIB_PATH)  # own handle: argtypes below stay local
    L.mtt_eval7.restype = ctypes.c_uint32
    L.mtt_eval7.argtypes = [ctypes.c_char_p]
    L.mtt_river_sweep_program.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                          ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
    return L


def program(lib, cards, grank):
    n = len(grank)
    c = np.ascontiguousarray(cards, np.uint8).reshape(-1)
    g = np.ascontiguousarray(grank, np.uint16)
    ln = ctypes.c_size_t()
    assert lib.mtt_river_sweep_program(c.ctypes.data, g.ctypes.data, n, None, 0, ctypes.byref(ln)) == 0
    out = np.zeros(ln.value, np.uint32)
    assert lib.mtt_river_sweep_program(c.ctypes.data, g.ctypes.data, n, out.ctypes.data, ln.value,
                                       ctypes.byref(ln)) == 0
    assert ln.value % REC == 0
    return out.reshape(-1, REC)


def replay(prog, cards, n, R):
    """The kernel's arithmetic on int64 columns R [ncol][n] (exact)."""
    ncol = R.shape[0]
    csT = np.zeros((52, ncol), np.int64)
    for j in range(n):
        csT[cards[j, 0]] += R[:, j]
        csT[cards[j, 1]] += R[:, j]
    H = -csT
    HP = -(csT.sum(0) // 2)
    out = np.zeros((ncol, n), np.int64)
    tile, tile_rows = {}, set()
    ring, queue = int(prog[0, 1]), [int(prog[0, 2]), int(prog[0, 3])]  # record 0: first three blocks
    oblk, pos, visit, done = 0, 1, 0, False
    while not done:
        vh = prog[pos]
        assert int(vh[0]) == OP_VISIT and 2 <= int(vh[2]) <= 16
        if visit:  # the crossing: next block into the ring, the header's block requested
            ring = queue.pop(0)
            queue.append(int(vh[1]))
        assert ring != NO_BLOCK
        for rec in prog[pos + 1:pos + int(vh[2])]:
            h = int(rec[0])
            op, k = h & 15, (h >> 4) & 15
            if op == OP_END:
                done = True
                break
            mem = [(int(w) & 0x7FF, (int(w) >> 11) & 63, (int(w) >> 17) & 63, int(w) >> 31) for w in rec[1:1 + k]]
            assert op in (OP_SINGLE, OP_HALF, OP_READ)
            assert k == 5 or (k == 15 and op != OP_SINGLE)
            valid = [v for *_, v in mem]
            assert sum(valid) >= (1 if k == 5 else 6) and all(valid[valid.index(1):]), "padding: invalid copies first"
            for row, a, b, v in mem:
                assert row < n and {a, b} == {int(cards[row, 0]), int(cards[row, 1])}
                assert v or (row, a, b) == mem[valid.index(1)][:3], "padding copies the first member"
                if op != OP_READ:
                    assert row // RB == ring, "row outside the ring block"
            if op == OP_SINGLE:  # reads, then writes: cards must not repeat (HALF adds atomically)
                used = [c for _, a, b, v in mem if v for c in (a, b)]
                assert len(used) == len(set(used)) and len(used) <= 8, "cards repeat within a SINGLE batch"
            va = [H[a].copy() for _, a, _, _ in mem]
            vb = [H[b].copy() for _, _, b, _ in mem]
            for m, (row, a, b, v) in enumerate(mem):
                r = R[:, row] if v else 0 * R[:, row]
                if op == OP_READ:
                    val = HP - va[m] - vb[m]
                elif op == OP_SINGLE:
                    val = HP - va[m] - vb[m] - r
                    HP = HP + 2 * r
                    H[a] = va[m] + 2 * r
                    H[b] = vb[m] + 2 * r
                else:  # atomic adds: no value read before the batch
                    HP = HP + r
                    H[a] = H[a] + r
                    H[b] = H[b] + r
                if op != OP_HALF:
                    assert row // OB == oblk, "value outside the output block"
                    tile[row % OB] = val
                    tile_rows.add(row)
            if h & F_FLUSH:
                nrows = ((h >> 10) & 15) + 1
                assert {oblk * OB + x for x in range(nrows)} == tile_rows
                for x in range(nrows):
                    out[:, oblk * OB + x] = tile[x]
                tile, tile_rows = {}, set()
                oblk += 1
        pos += int(vh[2])
        visit += 1
    assert oblk * OB >= n and (n == 0 or oblk == (n + OB - 1) // OB), "every row flushed once"
    return out


def dense(cards, grank, R):
    n = len(grank)
    g = np.asarray(grank, np.int64)
    sign = np.sign(g[:, None] - g[None, :])
    cm = np.zeros((n, 52), bool)
    cm[np.arange(n), cards[:, 0]] = True
    cm[np.arange(n), cards[:, 1]] = True
    blocked = (cm.astype(np.int32) @ cm.T.astype(np.int32)) > 0
    S = np.where(blocked, 0, sign).astype(np.int64)
    return R @ S.T  # [ncol][hero]


def fixed_point(r):
    return np.rint(r.astype(np.float64) * 2.0 ** 48).astype(np.int64)


def river_board(lib, board):
    live = [c for c in range(52) if c not in board]
    sc = sorted((lib.mtt_eval7(bytes(list(board) + [a, b])), a, b) for a, b in itertools.combinations(live, 2))
    cards = np.array([[a, b] for _, a, b in sc], np.int64)
    grank = np.zeros(len(sc), np.int64)
    for i in range(1, len(sc)):
        grank[i] = grank[i - 1] + (sc[i][0] != sc[i - 1][0])
    return cards, grank


def columns(rng, n, ncol):
    r = (rng.random((ncol, n)) * (rng.random((ncol, n)) < 0.8)).astype(np.float32)
    r[0] = 0.0
    r[1] = 1.0
    r[2] = np.float32(1e-7) * rng.random(n).astype(np.float32)  # near the fixed-point resolution
    return fixed_point(r)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_sweep_matches_dense_on_river_boards(lib, seed):
    rng = np.random.default_rng(seed)
    board = [int(x) for x in rng.choice(52, 5, replace=False)]
    cards, grank = river_board(lib, board)
    n = len(grank)
    assert n == 1081
    R = columns(rng, n, 6)
    prog = program(lib, cards, grank)
    assert np.array_equal(replay(prog, cards, n, R), dense(cards, grank, R))


def synthetic(rng, n, kind):
    combos = list(itertools.combinations(range(52), 2))
    pick = rng.choice(len(combos), n, replace=False)
    cards = np.array([combos[i] for i in pick], np.int64).reshape(-1, 2)
    if kind == "one_group":
        grank = np.zeros(n, np.int64)
    elif kind == "no_ties":
        grank = np.arange(n, dtype=np.int64)
    else:
        grank = np.sort(rng.integers(0, max(1, n // 7), n))
    return cards, grank


@pytest.mark.parametrize("kind,n", [("one_group", 1081), ("no_ties", 1081), ("random", 1081), ("random", 37),
                                    ("one_group", 1), ("no_ties", 17), ("random", 0)])
def test_sweep_matches_dense_synthetic(lib, kind, n):
    rng = np.random.default_rng(n * 7 + len(kind))
    cards, grank = synthetic(rng, n, kind)
    R = columns(rng, n, 4) if n else np.zeros((4, 0), np.int64)
    prog = program(lib, cards, grank)
    assert np.array_equal(replay(prog, cards, n, R), dense(cards, grank, R))


def test_sweep_program_rejects_bad_input(lib):
    cards = np.array([[0, 1], [2, 3]], np.uint8)
    ln = ctypes.c_size_t()
    bad_rank = np.array([1, 0], np.uint16)
    assert lib.mtt_river_sweep_program(cards.ctypes.data, bad_rank.ctypes.data, 2, None, 0, ctypes.byref(ln)) != 0
    same = np.array([[4, 4], [2, 3]], np.uint8)
    ok_rank = np.array([0, 1], np.uint16)
    assert lib.mtt_river_sweep_program(same.ctypes.data, ok_rank.ctypes.data, 2, None, 0, ctypes.byref(ln)) != 0
    assert lib.mtt_river_sweep_program(cards.ctypes.data, ok_rank.ctypes.data, 1089, None, 0,
                                       ctypes.byref(ln)) != 0


# ------------------------------------------------------- prefix scan -----
# The scan form (MTT_RIVER_SD=scan): per column, doubled strength prefixes
# DP[x] = 2 P(x) - Ptot and, per live card, doubled prefixes over its 46
# combos D_c[k] = 2 ex_c(k) - tot_c; a hero's doubled value is six lookups.
# Card lists: list c in round c // 32 on lanes 2 (c % 32) (entries 0..22) and
# 2 (c % 32) + 1 (entries 23..45).
# The replay below follows the kernel's arithmetic on integer columns.

def scan_tables(lib, cards, grank):
    lib.mtt_river_scan_tables.argtypes = [ctypes.c_void_p] * 2 + [ctypes.c_int] + [ctypes.c_void_p] * 2
    c = np.ascontiguousarray(cards, np.uint8).reshape(-1)
    g = np.ascontiguousarray(grank, np.uint16)
    h = np.zeros((3, 1088), np.uint32)
    lt = np.zeros((2, 12, 64), np.uint32)
    rc = lib.mtt_river_scan_tables(c.ctypes.data, g.ctypes.data, len(grank), h.ctypes.data, lt.ctypes.data)
    return rc, h, lt


def replay_scan(h, lt, n, R):
    ncol = R.shape[0]
    Rp = np.zeros((ncol, 1088), np.int64)
    Rp[:, :n] = R
    P = np.zeros((ncol, 1089), np.int64)
    P[:, 1:] = np.cumsum(Rp, axis=1)
    tot = P[:, 1088:1089]
    DP = 2 * P - tot
    pos = np.zeros((47, 46), np.int64)
    for c in range(47):
        r, l2 = divmod(c, 32)
        for k in range(46):
            h_, m = divmod(k, 23)
            pos[c, k] = ((int(lt[r, m >> 1, 2 * l2 + h_]) >> (16 * (m & 1))) & 0xFFFF) // 4
    g = Rp[:, pos]                                     # [ncol][47][46]
    ex = np.zeros((ncol, 47, 47), np.int64)
    ex[:, :, 1:] = np.cumsum(g, axis=2)
    Dl = np.zeros((ncol, 47 * 47 + 1), np.int64)       # + the zero slot
    Dl[:, :47 * 47] = (2 * ex - ex[:, :, 46:47]).reshape(ncol, 47 * 47)
    DPz = np.zeros((ncol, 1090), np.int64)             # DP[0..1088] + the zero slot 1089
    DPz[:, :1089] = DP
    assert np.all(h % 4 == 0) and np.all((h >> 16) % 4 == 0)
    lo16 = lambda w: (w & 0xFFFF).astype(np.int64) // 4
    hi16 = lambda w: (w >> 16).astype(np.int64) // 4
    two_c = (DPz[:, lo16(h[0])] + DPz[:, hi16(h[0])] - Dl[:, lo16(h[1])] - Dl[:, hi16(h[1])]
             - Dl[:, lo16(h[2])] - Dl[:, hi16(h[2])])
    assert np.all(two_c % 2 == 0)
    assert np.all(two_c[:, n:] == 0)                    # padding positions read the zero slots
    assert np.all(lo16(h[0])[n:] == 1089) and np.all(hi16(h[0])[:n] <= 1088)
    return two_c[:, :n] // 2


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_scan_matches_dense_on_river_boards(lib, seed):
    rng = np.random.default_rng(seed + 10)
    board = [int(x) for x in rng.choice(52, 5, replace=False)]
    cards, grank = river_board(lib, board)
    rc, h, lt = scan_tables(lib, cards, grank)
    assert rc == 0
    R = columns(rng, 1081, 6)
    assert np.array_equal(replay_scan(h, lt, 1081, R), dense(cards, grank, R))


@pytest.mark.parametrize("kind", ["one_group", "no_ties", "random", "pairs"])
def test_scan_matches_dense_synthetic_groups(lib, kind):
    """The 1081 combos of a river board in a shuffled order with synthetic
    tie structures: every combo tied, no ties, random group sizes, groups of
    two (which put two combos of one card in a group)."""
    rng = np.random.default_rng(len(kind))
    board = [3, 17, 28, 40, 51]
    live = [c for c in range(52) if c not in board]
    cards = np.array(list(itertools.combinations(live, 2)), np.int64)[rng.permutation(1081)]
    if kind == "one_group":
        grank = np.zeros(1081, np.int64)
    elif kind == "no_ties":
        grank = np.arange(1081, dtype=np.int64)
    elif kind == "pairs":
        grank = np.arange(1081, dtype=np.int64) // 2
    else:
        grank = np.sort(rng.integers(0, 150, 1081))
    rc, h, lt = scan_tables(lib, cards, grank)
    assert rc == 0
    R = columns(rng, 1081, 4)
    assert np.array_equal(replay_scan(h, lt, 1081, R), dense(cards, grank, R))


def test_scan_tables_reject_bad_input(lib):
    cards, grank = river_board(lib, [0, 9, 18, 27, 36])
    assert scan_tables(lib, cards, grank)[0] == 0
    assert scan_tables(lib, cards[:1080], grank[:1080])[0] != 0          # not a river board's 1081
    bad = grank.copy()
    bad[5], bad[900] = bad[900], bad[5]
    assert scan_tables(lib, cards, bad)[0] != 0                           # ranks not sorted
    dup = cards.copy()
    dup[7] = dup[8]                                                       # a card list of 47 and one of 45
    assert scan_tables(lib, dup, grank)[0] != 0
