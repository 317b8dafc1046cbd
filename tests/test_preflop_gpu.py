"""GPU parity of the preflop pieces (through the C ABI) against the oracle:
the Monte-Carlo all-in matrix bit for bit (win and tie counts of every
class matchup, and the equities)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_class_cards_match_oracle(mtt, oracle):
    for cls in range(169):
        a, b = ctypes.c_uint8(), ctypes.c_uint8()
        mtt.call("mtt_class_cards", cls, ctypes.byref(a), ctypes.byref(b))
        assert (a.value, b.value) == oracle.class_to_cards(cls), cls


def test_preflop_matrix_bitexact(mtt, oracle, torch_dev):
    """kernel_gen_preflop_matrix (preflop_grid_generator.cu:95-127) with the
    reference's 3000 samples: the same wins, ties and equities as the oracle
    in all 28,561 cells."""
    import time
    import torch
    eq = torch.zeros(169 * 169, dtype=torch.float32, device=torch_dev)
    w = torch.zeros(169 * 169, dtype=torch.int32, device=torch_dev)
    t = torch.zeros_like(w)
    torch.cuda.synchronize()
    t0 = time.time()
    mtt.call("mtt_gen_preflop_matrix", eq.data_ptr(), w.data_ptr(), t.data_ptr(), 3000, None)
    torch.cuda.synchronize()
    dt = time.time() - t0
    W0, T0, E0 = oracle.preflop_matrix(3000)
    print(f"preflop Monte-Carlo matrix: {dt * 1e3:.2f} ms on the GPU (28,561 cells x 3000 samples)", flush=True)
    assert np.array_equal(w.cpu().numpy().reshape(169, 169), W0)
    assert np.array_equal(t.cpu().numpy().reshape(169, 169), T0)
    assert np.array_equal(eq.cpu().numpy().reshape(169, 169), E0)


# ------------------------------------------------- preflop co-solver ------
def _flops_three():
    # a rainbow low flop, three aces (AA has no valid combo), a monotone flop
    # given out of order (the entry sorts the cards)
    return np.array([[0, 5, 10], [48, 49, 50], [40, 4, 16]], np.uint8)


def test_flop_class_matrices_bitexact(mtt, oracle, torch_dev):
    """kernel_gen_100_flop_matrices + kernel_reduce_combo_to_class_matrices
    (preflop_grid_generator.cu:130-223) over all 1,326 combos: the GPU's
    class matrices equal the oracle's bit for bit on three flops."""
    import torch
    fl = _flops_three()
    out = torch.full((3, 169, 169), -1.0, dtype=torch.float32, device=torch_dev)
    mtt.call("mtt_flop_class_matrices", fl.ctypes.data, 3, out.data_ptr(), None)
    got = out.cpu().numpy()
    for f in range(3):
        want = oracle.flop_class_matrix(fl[f])
        bad = np.argwhere(got[f] != want)
        assert bad.size == 0, (f, bad[:5], got[f][tuple(bad[0])], want[tuple(bad[0])])


def test_flop_class_matrices_batch_positions(mtt, torch_dev):
    """Eleven flops in one call (the XCD-ordered grid covers flop slots
    0..15 and leaves 11..15 idle) give the same matrices as one flop per
    call; zero flops is a no-op and a flop with a repeated card is refused."""
    import torch
    flops = np.array([[0, 5, 10], [48, 49, 50], [4, 16, 40], [1, 2, 3], [7, 20, 33], [11, 12, 51], [8, 9, 44],
                      [14, 27, 39], [21, 22, 23], [30, 31, 47], [6, 18, 50]], np.uint8)
    n = len(flops)
    all_ = torch.empty((n, 169, 169), dtype=torch.float32, device=torch_dev)
    mtt.call("mtt_flop_class_matrices", flops.ctypes.data, n, all_.data_ptr(), None)
    one = torch.empty((169, 169), dtype=torch.float32, device=torch_dev)
    for f in (0, 7, 10):
        mtt.call("mtt_flop_class_matrices", flops[f].ctypes.data, 1, one.data_ptr(), None)
        assert torch.equal(all_[f], one), f
    mtt.call("mtt_flop_class_matrices", flops.ctypes.data, 0, all_.data_ptr(), None)
    bad = np.array([[3, 3, 9]], np.uint8)
    with pytest.raises(mtt.MTTError):
        mtt.call("mtt_flop_class_matrices", bad.ctypes.data, 1, all_.data_ptr(), None)


def _allin(mtt, torch_dev):
    import torch
    A = torch.empty(169 * 169, dtype=torch.float32, device=torch_dev)
    mtt.call("mtt_gen_preflop_matrix", A.data_ptr(), None, None, 3000, None)
    return A


def test_preflop_solve_matches_oracle(mtt, oracle, torch_dev):
    """The DCFR loop (preflop_grid_generator.cu:228-421, 588-614), 500
    iterations on the class matrices of three flops and the 3000-sample
    all-in matrix: the GPU's strategy sums equal the oracle's bit for bit
    (both compute in FP32 with the same operation order and explicit fused
    multiply-adds, and take the discount weights from the host's libm)."""
    import torch
    fl = _flops_three()
    M = torch.empty((3, 169, 169), dtype=torch.float32, device=torch_dev)
    mtt.call("mtt_flop_class_matrices", fl.ctypes.data, 3, M.data_ptr(), None)
    A = _allin(mtt, torch_dev)
    w = np.array([4.0, 12.0, 24.0], np.float32)
    out = torch.empty((4, 3, 169), dtype=torch.float32, device=torch_dev)
    mtt.call("mtt_preflop_solve", M.data_ptr(), w.ctypes.data, 3, A.data_ptr(), 500, 20.0, 1.35, 2.5,
             out.data_ptr(), None)
    got = out.cpu().numpy()
    want = oracle.preflop_solve(M.cpu().numpy(), w, A.cpu().numpy().reshape(169, 169), iters=500)
    diff = np.abs(got.astype(np.float64) - want)
    print(f"preflop solve, 3 flops: max |GPU - oracle| = {diff.max():.3e}", flush=True)
    assert np.array_equal(got, want), np.argwhere(got != want)[:5]


def _reference_subset():
    import json
    import os
    from conftest import GOLDEN
    meta = json.load(open(os.path.join(GOLDEN, "ev_dump_meta.json")))
    return bytes.fromhex(meta["subset_bin_hex"])


def test_preflop_cosolve_reference_subset(mtt, oracle, torch_dev):
    """The reference program's configuration (preflop_grid_generator.cu:
    426-614): its 100-flop subset (data/mtt_subset.bin, committed as a
    fixture), 3000-sample all-in matrix, 20bb, bubble factor 1.35, dead pot
    2.5bb, 500 iterations.  mtt_preflop_cosolve equals the oracle's loop run
    on the GPU's class matrices bit for bit; the sums over actions are root
    reach x sum t^2; premium pairs never fold at the two first decisions."""
    import time
    import torch
    import importlib.util
    from conftest import PKG
    spec = importlib.util.spec_from_file_location("mtt_preflop", f"{PKG}/preflop.py")
    pf = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(pf)
    data = _reference_subset()
    cards, w = pf.read_subset(data)
    assert len(cards) == 100
    A = _allin(mtt, torch_dev)
    sub = np.frombuffer(data, np.uint8, offset=4).copy()
    out = torch.empty((4, 3, 169), dtype=torch.float32, device=torch_dev)
    mtt.call("mtt_preflop_timing", 1)
    torch.cuda.synchronize()
    t0 = time.time()
    mtt.call("mtt_preflop_cosolve", sub.ctypes.data, 100, A.data_ptr(), 500, 20.0, 1.35, 2.5, out.data_ptr(), None)
    dt = time.time() - t0
    times = {}
    for k in range(8):
        name = ctypes.create_string_buffer(32)
        ms, n = ctypes.c_double(), ctypes.c_int64()
        if mtt.lib().mtt_preflop_kernel_times(k, name, 32, ctypes.byref(ms), ctypes.byref(n)) != 0:
            break
        times[name.value.decode()] = (ms.value, n.value)
    mtt.call("mtt_preflop_timing", 0)
    print(f"preflop co-solve, 100 flops x 500 iterations: {dt * 1e3:.1f} ms wall; kernels {times}", flush=True)
    got = out.cpu().numpy()
    M = pf.class_matrices(cards)
    want = oracle.preflop_solve(M.cpu().numpy(), w, A.cpu().numpy().reshape(169, 169), iters=500)
    assert np.array_equal(got, want), np.argwhere(got != want)[:5]
    root = np.array([6] * 13 + [4] * 78 + [12] * 78, np.float32) / np.float32(1326)
    tot = sum(t * t for t in range(1, 501))
    for node, na in enumerate((3, 3, 2, 2)):
        np.testing.assert_allclose(got[node, :na].sum(0), root * tot, rtol=1e-5)
    avg = pf.average_strategy(got)
    print(pf.snapshot(got), flush=True)
    assert avg[0, 0, :3].max() < 0.01 and avg[1, 0, :3].max() < 0.01
