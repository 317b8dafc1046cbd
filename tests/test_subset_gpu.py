"""End to end on the GPU: the EV dump of all 1,755 flops (mtt_ev_dump_tensor,
bit-equal to the reference's data/mtt_ev_dump_1755.bin) through the flop
subset selection (scripts/optimize_subset.py:28-65) gives the reference's
committed data/mtt_subset.bin and data/mtt_subset_weights.txt."""
import json
import os
import time

import numpy as np
import pytest

from conftest import GOLDEN, PKG, _load

pytestmark = pytest.mark.gpu


def test_subset_from_gpu_dump_matches_reference(mtt, torch_dev):
    S = _load("mtt_subset", os.path.join(PKG, "subset.py"))
    meta = json.load(open(os.path.join(GOLDEN, "ev_dump_meta.json")))
    t = time.time()
    dump = np.zeros((1755, 4, 2, 169), np.float32)
    mtt.call("mtt_ev_dump_tensor", 0, 1755, dump.ctypes.data)
    t_dump = time.time() - t
    t = time.time()
    idx, wt = S.select_subset(dump)
    t_sel = time.time() - t
    print(f"EV dump {t_dump:.2f} s (GPU), subset selection {t_sel:.2f} s (host)", flush=True)
    assert S.subset_bytes(idx, wt) == bytes.fromhex(meta["subset_bin_hex"])
    assert S.subset_text(idx, wt) == meta["subset_weights_txt"]
