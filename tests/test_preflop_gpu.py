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
