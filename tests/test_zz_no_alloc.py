"""mtt_flop_counts_noalloc (no device allocation, static per-GPU workspace)
against mtt_flop_counts through the C ABI: identical outputs on two calls in
a row, the second with fewer flops (the workspace keeps the first call's
contents), in both output layouts; more flops than the workspace holds are
refused before any GPU work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _counts(mtt, name, flops, layout, dev):
    import torch
    S = 1326 if layout == mtt.LAYOUT_GLOBAL1326 else 1216
    n = len(flops)
    W = torch.full((n, S, S), -7, dtype=torch.int32, device=dev)
    T = torch.full((n, S, S), -7, dtype=torch.int32, device=dev)
    f = np.ascontiguousarray(flops, np.uint8)
    mtt.call(name, f.ctypes.data, n, W.data_ptr(), T.data_ptr(), layout, None)
    torch.cuda.synchronize()
    return W, T


@pytest.mark.parametrize("layout", ["global", "alive"])
def test_noalloc_matches_allocating_entry(mtt, torch_dev, layout):
    import torch
    lay = mtt.LAYOUT_GLOBAL1326 if layout == "global" else mtt.LAYOUT_ALIVE1216
    canon = np.zeros((1755, 4), np.uint8)
    assert mtt.lib().mtt_canonical_flops(canon.ctypes.data) == 1755
    rng = np.random.default_rng(11)
    first = canon[rng.choice(1755, 16, replace=False), :3]
    second = canon[rng.choice(1755, 5, replace=False), :3]
    for k, flops in enumerate((first, second)):
        print(f"no-alloc flop counts, layout {layout}, call {k + 1}: {len(flops)} flops", flush=True)
        Wa, Ta = _counts(mtt, "mtt_flop_counts_noalloc", flops, lay, torch_dev)
        Wb, Tb = _counts(mtt, "mtt_flop_counts", flops, lay, torch_dev)
        assert torch.equal(Wa, Wb) and torch.equal(Ta, Tb), (layout, k)
        assert int(Wa.min()) >= 0 and int(Ta.min()) >= 0      # every entry written
        del Wa, Ta, Wb, Tb


def test_noalloc_refuses_oversized_call(mtt, torch_dev):
    flops = np.tile(np.array([[0, 1, 2]], np.uint8), (17, 1))
    W = ctypes.c_void_p(1)    # never dereferenced: the size check comes first
    rc = mtt.lib().mtt_flop_counts_noalloc(flops.ctypes.data, 17, W, W, mtt.LAYOUT_GLOBAL1326, None)
    assert rc == mtt.MTT_ERR_ARG
    assert "16-flop workspace" in mtt.last_error()
