"""GPU parity of the evaluator and equity builders (through the C ABI)
against the CPU oracle and the reference's committed EV dump."""
import itertools
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _alive(flop):
    fl = set(int(c) for c in flop)
    return [(a, b) for a, b in itertools.combinations(range(52), 2) if a not in fl and b not in fl]


def _combo_struct(mtt, flop):
    arr = (mtt.AliveCombo * mtt.N_PAD)()
    for i, (a, b) in enumerate(_alive(flop)):
        arr[i].c1, arr[i].c2, arr[i].mask = a, b, (1 << a) | (1 << b)
    return np.frombuffer(bytes(arr), np.uint8)


def _deck(flop):
    return np.array([c for c in range(52) if c not in set(int(x) for x in flop)], np.uint8)


def test_gpu_eval_batch(mtt, oracle, torch_dev):
    import torch
    rng = np.random.default_rng(11)
    cards = np.argsort(rng.random((1 << 20, 52)), axis=1)[:, :7].astype(np.uint8)
    want = oracle.eval7_batch(cards)
    d_cards = torch.from_numpy(cards).to(torch_dev)
    d_out = torch.zeros(cards.shape[0], dtype=torch.int32, device=torch_dev)
    mtt.call("mtt_eval7_batch", d_cards.data_ptr(), cards.shape[0], d_out.data_ptr(), None)
    torch.cuda.synchronize()
    got = d_out.cpu().numpy().view(np.uint32)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("flop_idx", [0, 40, 1000])
def test_gpu_flop_counts_bitexact(mtt, oracle, torch_dev, flop_idx):
    import torch
    flops = oracle.canonical_flops()
    flop = flops[flop_idx, :3].copy()
    print("flop", flop_idx, flop.tolist(), flush=True)
    W0, T0 = oracle.flop_counts(flop)
    dW = torch.full((1326 * 1326,), -7, dtype=torch.int32, device=torch_dev)
    dT = torch.full((1326 * 1326,), -7, dtype=torch.int32, device=torch_dev)
    mtt.call("mtt_flop_counts", flop.ctypes.data, 1, dW.data_ptr(), dT.data_ptr(), mtt.LAYOUT_GLOBAL1326, None)
    torch.cuda.synchronize()
    W = dW.cpu().numpy().reshape(1326, 1326)
    T = dT.cpu().numpy().reshape(1326, 1326)
    assert np.array_equal(W, W0)
    assert np.array_equal(T, T0)


def test_gpu_flop_counts_all_1755(mtt, torch_dev):
    """Config 2 at full size: exhaustive W/T counts of all 1,755 canonical
    flops (global 1326 layout) against the oracle's per-flop checksums
    (sum W, sum T and splitmix-weighted sums, exact int64 arithmetic:
    tests/golden/flop_checksums.npz), plus the invariants on every entry:
    W + W^T + T = 990 on unblocked pairs of alive combos, T symmetric,
    zero elsewhere."""
    import importlib.util
    import time
    import torch
    spec = importlib.util.spec_from_file_location("make_flop_checksums", os.path.join(GOLDEN, "make_flop_checksums.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    fx = np.load(os.path.join(GOLDEN, "flop_checksums.npz"))
    flops, want = np.ascontiguousarray(fx["flops"]), fx["checksums"]
    assert flops.shape == (1755, 3)
    canon = np.zeros((1755, 4), np.uint8)
    assert mtt.lib().mtt_canonical_flops(canon.ctypes.data) == 1755
    assert np.array_equal(canon[:, :3], flops)
    w1 = torch.from_numpy(mk.splitmix_weights(1)).to(torch_dev)
    w2 = torch.from_numpy(mk.splitmix_weights(2)).to(torch_dev)
    pairs = np.array(list(itertools.combinations(range(52), 2)), np.int64)
    cm = torch.from_numpy((np.int64(1) << pairs[:, 0]) | (np.int64(1) << pairs[:, 1])).to(torch_dev)
    disjoint = (cm[:, None] & cm[None, :]) == 0
    B = 64
    W = torch.empty((B, 1326, 1326), dtype=torch.int32, device=torch_dev)
    T = torch.empty_like(W)
    t_kernels = 0.0
    for f0 in range(0, 1755, B):
        n = min(B, 1755 - f0)
        print(f"config2 flops {f0}..{f0 + n - 1}", flush=True)
        W.fill_(-7)
        T.fill_(-7)
        torch.cuda.synchronize()
        t = time.time()
        mtt.call("mtt_flop_counts", flops[f0:f0 + n].ctypes.data, n, W.data_ptr(), T.data_ptr(),
                 mtt.LAYOUT_GLOBAL1326, None)
        torch.cuda.synchronize()
        t_kernels += time.time() - t
        Wn, Tn = W[:n].long(), T[:n].long()
        got = torch.stack([Wn.sum((1, 2)), Tn.sum((1, 2)), (Wn * w1).sum((1, 2)), (Tn * w2).sum((1, 2))], 1)
        bad = np.nonzero((got.cpu().numpy() != want[f0:f0 + n]).any(1))[0]
        assert len(bad) == 0, f"flops {(f0 + bad).tolist()[:10]} differ from the oracle checksums"
        fm = torch.from_numpy((np.int64(1) << flops[f0:f0 + n].astype(np.int64)).sum(1)).to(torch_dev)
        alive = (cm[None, :] & fm[:, None]) == 0
        live = alive[:, :, None] & alive[:, None, :] & disjoint[None]
        assert torch.equal(Tn, Tn.transpose(1, 2))
        assert bool(((Wn + Wn.transpose(1, 2) + Tn) == 990)[live].all())
        assert not bool(Wn[~live].any()) and not bool(Tn[~live].any())
        del Wn, Tn, got, live, alive
    print(f"config2: 1755 flops, {t_kernels:.3f} s in mtt_flop_counts (batches of {B})", flush=True)


def test_gpu_flop_counts_unaligned_outputs(mtt, torch_dev):
    """Output buffers that are not 16-byte aligned (offset by one int32) give
    the same counts as aligned ones."""
    import torch
    canon = np.zeros((1755, 4), np.uint8)
    mtt.lib().mtt_canonical_flops(canon.ctypes.data)
    fl = np.ascontiguousarray(canon[[7, 901], :3])
    S = 1326
    n = 2 * S * S
    Wa = torch.full((n,), -7, dtype=torch.int32, device=torch_dev)
    Ta = torch.full((n,), -7, dtype=torch.int32, device=torch_dev)
    Wu = torch.full((n + 1,), -7, dtype=torch.int32, device=torch_dev)
    Tu = torch.full((n + 1,), -7, dtype=torch.int32, device=torch_dev)
    mtt.call("mtt_flop_counts", fl.ctypes.data, 2, Wa.data_ptr(), Ta.data_ptr(), mtt.LAYOUT_GLOBAL1326, None)
    mtt.call("mtt_flop_counts", fl.ctypes.data, 2, Wu.data_ptr() + 4, Tu.data_ptr() + 4, mtt.LAYOUT_GLOBAL1326, None)
    torch.cuda.synchronize()
    assert torch.equal(Wa, Wu[1:]) and torch.equal(Ta, Tu[1:])
    assert int(Wu[0]) == -7 and int(Tu[0]) == -7


def test_gpu_flop_counts_batch_and_alive_layout(mtt, oracle, torch_dev):
    import torch
    flops = oracle.canonical_flops()
    sel = [3, 500, 1754]
    fl = np.ascontiguousarray(flops[sel, :3])
    S = mtt.N_PAD
    dW = torch.zeros(len(sel) * S * S, dtype=torch.int32, device=torch_dev)
    dT = torch.zeros(len(sel) * S * S, dtype=torch.int32, device=torch_dev)
    mtt.call("mtt_flop_counts", fl.ctypes.data, len(sel), dW.data_ptr(), dT.data_ptr(), mtt.LAYOUT_ALIVE1216, None)
    torch.cuda.synchronize()
    W = dW.cpu().numpy().reshape(len(sel), S, S)
    T = dT.cpu().numpy().reshape(len(sel), S, S)
    for k, fi in enumerate(sel):
        W0, T0 = oracle.flop_counts(fl[k])
        g = [oracle.lib().or_combo_index(a, b) for a, b in _alive(fl[k])]
        A = np.ix_(g, g)
        assert np.array_equal(W[k, :1176, :1176], W0[A])
        assert np.array_equal(T[k, :1176, :1176], T0[A])
        assert not W[k, 1176:].any() and not W[k, :, 1176:].any()


def test_gpu_ev_dump_matches_reference(mtt, torch_dev):
    z = np.load(os.path.join(GOLDEN, "ev_dump_slice.npz"))
    idx, vals = z["flop_index"], z["values"]
    for k, f in enumerate(idx[:16]):
        out = np.zeros((1, 4, 2, 169), np.float32)
        mtt.call("mtt_ev_dump_tensor", int(f), 1, out.ctypes.data)
        assert np.array_equal(out[0], vals[k]), f


def test_gpu_ev_dump_file_all_1755(mtt, torch_dev, tmp_path):
    """bin/ev_dumper end to end (mtt_ev_dumper.cu:118-176): mtt_ev_dump
    writes the whole float32[1755][4][2][169] tensor to a file of the
    reference's size; every player-0 scenario-0 value equals the reference's
    data/mtt_ev_dump_1755.bin bit for bit (tests/golden/ev_dump_p0s0.npz),
    player 1 is -player 0 to FP32 rounding as in the reference, and the
    scenarios are the 1.1/1.2/1.3 rescalings of the same per-combo values."""
    import time
    ref = np.load(os.path.join(GOLDEN, "ev_dump_p0s0.npz"))["values"]
    path = str(tmp_path / "mtt_ev_dump_1755.bin")
    t = time.time()
    mtt.call("mtt_ev_dump", 1755, path.encode())
    dt = time.time() - t
    raw = open(path, "rb").read()
    assert len(raw) == 1755 * 4 * 2 * 169 * 4
    dump = np.frombuffer(raw, "<f4").reshape(1755, 4, 2, 169)
    d = np.abs(dump[:, 0, 0, :].astype(np.float64) - ref)
    exact = float(np.mean(dump[:, 0, 0, :] == ref))
    print(f"EV dump of 1755 flops in {dt:.2f} s; player 0 scenario 0 vs reference: max |diff| {d.max():.3g}, "
          f"{exact:.4f} of the values bit-equal", flush=True)
    assert np.array_equal(dump[:, 0, 0, :], ref)
    assert np.abs(dump[:, :, 1, :].astype(np.float64) + dump[:, :, 0, :]).max() <= 1e-6
    for sc in (1, 2, 3):
        want = dump[:, 0, 0, :].astype(np.float64) * (1 + 0.1 * sc)
        assert np.abs(dump[:, sc, 0, :] - want).max() <= 1e-5


def test_gpu_turn_matrices(mtt, oracle, torch_dev):
    import torch
    flop = np.array([50, 33, 2], np.uint8)
    combos = torch.from_numpy(_combo_struct(mtt, flop)).to(torch_dev)
    deck = torch.from_numpy(_deck(flop)).to(torch_dev)
    dfl = torch.from_numpy(flop).to(torch_dev)
    out = torch.zeros(49 * 1216 * 1216, dtype=torch.float32, device=torch_dev)
    mtt.call("mtt_gen_turn_matrices", combos.data_ptr(), deck.data_ptr(), dfl.data_ptr(), out.data_ptr(), None)
    torch.cuda.synchronize()
    E = out.view(49, 1216, 1216)
    dk = _deck(flop)
    for t in (0, 17, 48):
        want = oracle.turn_matrix(flop, int(dk[t]))
        assert np.array_equal(E[t].cpu().numpy(), want), t


def test_gpu_river_matrices(mtt, oracle, torch_dev):
    import torch
    flop = np.array([0, 21, 42], np.uint8)
    combos = torch.from_numpy(_combo_struct(mtt, flop)).to(torch_dev)
    dk = _deck(flop)
    deck = torch.from_numpy(dk).to(torch_dev)
    dfl = torch.from_numpy(flop).to(torch_dev)
    boards = np.array([[0, 1], [10, 40], [48, 3]], np.int16)
    out = torch.zeros(len(boards) * 1216 * 1216, dtype=torch.float32, device=torch_dev)
    mtt.call("mtt_gen_river_matrices", combos.data_ptr(), deck.data_ptr(), dfl.data_ptr(),
             boards.ctypes.data, len(boards), out.data_ptr(), None)
    torch.cuda.synchronize()
    E = out.view(len(boards), 1216, 1216).cpu().numpy()
    al = _alive(flop)
    for k, (t, r) in enumerate(boards):
        bc = {int(dk[t]), int(dk[r])}
        cards = np.array([[*flop, dk[t], dk[r], a, b] for a, b in al], np.uint8)
        s = oracle.eval7_batch(cards).astype(np.int64)
        ok = np.array([a not in bc and b not in bc for a, b in al])
        m = np.array([(1 << a) | (1 << b) for a, b in al], dtype=object)
        blk = np.array([[bool(m[i] & m[j]) for j in range(len(al))] for i in range(len(al))])
        # E[j][h]: 1 if h beats j, 1/2 tie
        want = np.where(s[None, :] > s[:, None], 1.0, np.where(s[None, :] == s[:, None], 0.5, 0.0)).astype(np.float32)
        want[blk] = 0
        want[~ok, :] = 0
        want[:, ~ok] = 0
        assert np.array_equal(E[k, :1176, :1176], want)
        assert not E[k, 1176:].any() and not E[k, :, 1176:].any()


def test_gpu_flop_matrices(mtt, oracle, torch_dev):
    """kernel_gen_100_flop_matrices restated: GPU entries bit-equal to the
    oracle on sampled pairs; rows of heroes that hold a flop card untouched."""
    import torch
    combos_np = np.zeros(1216, dtype=[("c1", "u1"), ("c2", "u1"), ("pad", "V6"), ("mask", "<u8")])
    k = 0
    for a in range(52):
        for b in range(a + 1, 52):
            if k < 1176:  # the reference's preflop list: the first 1176 of the 1326 lexicographic combos
                combos_np[k] = (a, b, b"\0" * 6, (1 << a) | (1 << b))
                k += 1
    flops = np.zeros(2, dtype=[("c1", "u1"), ("c2", "u1"), ("c3", "u1"), ("pad", "u1"), ("w", "<f4")])
    flops[0] = (3, 22, 45, 0, 1.0)
    flops[1] = (0, 1, 2, 0, 1.0)
    d_c = torch.from_numpy(combos_np.view(np.uint8).copy()).to(torch_dev)
    d_f = torch.from_numpy(flops.view(np.uint8).copy()).to(torch_dev)
    out = torch.full((2 * 1216 * 1216,), -7.0, device=torch_dev)
    mtt.call("mtt_gen_flop_matrices", d_f.data_ptr(), 2, d_c.data_ptr(), out.data_ptr(), None)
    torch.cuda.synchronize()
    E = out.view(2, 1216, 1216).cpu().numpy()
    rng = np.random.default_rng(0)
    for f in range(2):
        fl = [int(flops[f]["c1"]), int(flops[f]["c2"]), int(flops[f]["c3"])]
        for _ in range(300):
            h, j = rng.integers(0, 1176, 2)
            hero = (int(combos_np[h]["c1"]), int(combos_np[h]["c2"]))
            vil = (int(combos_np[j]["c1"]), int(combos_np[j]["c2"]))
            if set(hero) & set(fl):
                assert E[f, j, h] == -7.0  # not written
                continue
            assert E[f, j, h] == np.float32(oracle.flop_matrix_entry(fl, hero, vil)), (f, h, j)
    assert (E[:, 1176:, :] == -7.0).all() and (E[:, :, 1176:] == -7.0).all()
