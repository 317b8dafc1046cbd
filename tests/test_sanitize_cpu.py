"""The CPU oracle built with AddressSanitizer and UndefinedBehaviorSanitizer
and exercised on every kind of input the tests give it (evaluator, flop
counts, turn and flop-matrix entries, HU CFR on a river and a two-street
tree with its best response, the 4-way engine with side pots, ICM, the DCFR
step, the preflop pieces).  Trees, tables and ranges come from the library in
this process; the instrumented oracle runs in a child process with the
sanitizer runtime preloaded and must finish without a report."""
import os
import subprocess
import sys
import textwrap

import numpy as np
import pytest

from conftest import ORACLE, PKG, _load


def _asan_runtime():
    r = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True)
    path = r.stdout.strip()
    return path if r.returncode == 0 and os.path.isabs(path) and os.path.exists(path) else None


CHILD = r"""
import sys
import numpy as np
sys.path.insert(0, ORACLE)
import oracle_ctypes as o

d = np.load(NPZ)
o.set_threads(2)
rng = np.random.default_rng(0)
# evaluator: random 7-card hands of distinct cards
hands = np.array([rng.choice(52, 7, replace=False) for _ in range(20000)], np.uint8)
o.eval7_batch(hands)
o.canonical_flops()
flop = d["flop"]
W, T = o.flop_counts(flop)
assert W.shape == (1326, 1326)
o.turn_matrix(flop, int(d["turn"]))
o.flop_matrix_entry(flop, (0, 5), (9, 13))
# HU CFR + best response: a river tree and a two-street tree
for tag in ("river", "two"):
    h = o.HU(d[tag + "_nodes"], int(d[tag + "_ni"]), d[tag + "_board"], d[tag + "_r0"], d[tag + "_r1"], 1)
    h.iterate(2)
    st = h.strategy()
    h.exploitability(st, float(d[tag + "_pot"]))
    h.state()
    h.close()
# 4-way: terminals with side pots, two iterations
b4 = o.board4(flop)
n4 = d["nodes4"]
bf = d["bf"]
reach = d["reach4"]
for k in np.nonzero(n4["type"] != 0)[0][:40]:
    o.terminal4(b4, n4[k:k + 1], reach, bf)
fw = o.FourWay(n4, int(d["ni4"]), flop, d["ranges4"], bf)
fw.iterate(2)
fw.state()
fw.close()
o.icm([30.0, 20.0, 10.0, 5.0, 2.0], [0.5, 0.3, 0.2])
o.icm([30.0, 20.0, 10.0, 5.0], [0.6, 0.4], brute=True)
o.bubble_factors([30.0, 20.0, 10.0, 5.0, 2.0], [0.5, 0.3, 0.2], [0, 1, 3])
R = np.zeros((3, 192), np.float32); sg = np.full((3, 192), 1 / 3, np.float32); S = np.zeros_like(R)
o.dcfr_step(3, R, sg, S, rng.random((3, 192)).astype(np.float32), np.ones(169, np.float32), 0.5, 0.0, 1.0)
# preflop
for c in range(169):
    o.class_to_cards(c)
o.preflop_cell(0, 168, 500)
M = o.flop_class_matrix(flop)
A = rng.random((169, 169)).astype(np.float32)
o.preflop_solve(np.stack([M, M]), np.array([1.0, 2.0], np.float32), A, iters=20)
print("sanitized oracle run: OK", flush=True)
"""


def test_oracle_under_sanitizers(tmp_path, mtt):
    rt = _asan_runtime()
    if rt is None:
        pytest.skip("gcc has no AddressSanitizer runtime here")
    S = _load("mtt_solver", os.path.join(PKG, "solver.py"))
    M4 = _load("mtt_multi", os.path.join(PKG, "multi.py"))
    # instrumented build of the oracle's sources
    so = tmp_path / "liboracle_san.so"
    srcs = [os.path.join(ORACLE, f) for f in ("or_eval.c", "or_cfr.c", "or_multi.c", "or_preflop.c")]
    cmd = ["gcc", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-std=c11",
           "-ffp-contract=off", "-shared", "-o", str(so), *srcs, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    # inputs from the library
    g1, r0, r1 = S.config1(0)
    g2, q0, q1 = S.config3_small(1)
    data = {}
    for tag, (g, a, b) in {"river": (g1, r0, r1), "two": (g2, q0, q1)}.items():
        nodes, ni, _ = S.tree(g)
        board = list(g.board)[:g.n_board]
        gi, _ = S.alive_index(board)
        data.update({tag + "_nodes": nodes, tag + "_ni": ni, tag + "_board": np.array(board, np.uint8),
                     tag + "_r0": a[gi], tag + "_r1": b[gi], tag + "_pot": g.pot})
    flop = np.array([4, 21, 46], np.uint8)
    nodes4, ni4 = M4.tree4(M4.game4([2.5] * 4, [50.27, 22.73, 5.57, 3.74], M4.config5_menu()))
    rng = np.random.default_rng(1)
    reach4 = np.zeros((4, 1216), np.float32)
    for p in range(4):
        idx = rng.choice(1176, 8, replace=False)
        reach4[p, idx] = rng.uniform(0.1, 1.0, 8)
    ranges4 = np.zeros((4, 1216), np.float32)
    ranges4[:, :1176] = rng.random((4, 1176))
    data.update(flop=flop, turn=np.int64(30), nodes4=nodes4, ni4=ni4, reach4=reach4, ranges4=ranges4,
                bf=rng.uniform(1.0, 1.5, (4, 4)))
    npz = tmp_path / "inputs.npz"
    np.savez(npz, **data)
    script = tmp_path / "child.py"
    script.write_text(f"ORACLE = {ORACLE!r}\nNPZ = {str(npz)!r}\n" + textwrap.dedent(CHILD))
    env = dict(os.environ)
    env.update(LD_PRELOAD=rt, MTT_ORACLE_LIB=str(so), OMP_NUM_THREADS="2",
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([sys.executable, str(script)], capture_output=True, text=True, env=env, timeout=900)
    report = r.stderr[-4000:]
    assert r.returncode == 0, report
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, report
    assert "sanitized oracle run: OK" in r.stdout
