"""Writes the golden fixtures under tests/golden/ from the reference's data.

Run in the build container (the reference tree is not on the GPU box):
    python tests/golden/make_fixtures.py
Fixtures are data only: slices of the reference's committed
data/mtt_ev_dump_1755.bin (written by src/tools/mtt_ev_dumper.cu:172-173),
its size and sha256, the contents of data/mtt_subset.bin and
data/mtt_subset_weights.txt, and the
player-0 scenario-0 plane of the dump for all 1,755 flops
(ev_dump_p0s0.npz; in the reference file player 1 is -player 0 to FP32
rounding, at most 4.8e-7 apart, checked here, and scenarios 1-3 are host
rescalings of the same per-combo values, compared in full by the slice).
"""
import hashlib
import json
import os

import numpy as np

REF = "/root/reference/data"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    raw = open(os.path.join(REF, "mtt_ev_dump_1755.bin"), "rb").read()
    dump = np.frombuffer(raw, dtype="<f4").reshape(1755, 4, 2, 169)
    idx = sorted(set([0, 1, 2, 12, 13, 14, 15, 324, 325, 326, 327, 328, 329, 1754]
                     + np.linspace(16, 1750, 26).astype(int).tolist()))
    np.savez_compressed(os.path.join(HERE, "ev_dump_slice.npz"),
                        flop_index=np.array(idx, np.int32), values=dump[idx])
    p1_dev = float(np.abs(dump[:, :, 1, :] + dump[:, :, 0, :]).max())
    assert p1_dev <= 1e-6
    np.savez_compressed(os.path.join(HERE, "ev_dump_p0s0.npz"), values=dump[:, 0, 0, :].copy())
    sub = open(os.path.join(REF, "mtt_subset.bin"), "rb").read()
    sub_txt = open(os.path.join(REF, "mtt_subset_weights.txt")).read()
    meta = {
        "ev_dump_bytes": len(raw),
        "ev_dump_sha256": hashlib.sha256(raw).hexdigest(),
        "ev_dump_flop_index": idx,
        "subset_bin_hex": sub.hex(),
        "subset_weights_txt": sub_txt,
        "ev_dump_max_abs_p0_plus_p1": p1_dev,
    }
    with open(os.path.join(HERE, "ev_dump_meta.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print("wrote", len(idx), "flops")


if __name__ == "__main__":
    main()
