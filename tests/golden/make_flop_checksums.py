"""Writes tests/golden/flop_checksums.npz: per canonical flop (the 1,755 of
generate_all_1755_flops, mtt_ev_dumper.cu:34-51, in that order) four int64
checksums of the exhaustive win/tie counts W, T in the global 1326 x 1326
layout, computed by the CPU oracle (oracle/liboracle.so, or_flop_counts):

    sum(W), sum(T), sum(W * w1), sum(T * w2)      (int64, wrapping)

with w1, w2 = splitmix64 of the flattened index (seeds 1 and 2) viewed as
int64.  One wrong count changes a weighted sum by a nonzero weight, and the
sums are exact integer arithmetic, so the GPU side can recompute them with
any summation order.

Run in the build container (about 25 minutes on 8 cores):
    python tests/golden/make_flop_checksums.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "flop_checksums.npz")


def splitmix_weights(seed, n=1326 * 1326):
    """splitmix64(seed * 2^32 + k) for k = 0..n-1, as int64."""
    with np.errstate(over="ignore"):
        z = (np.arange(n, dtype=np.uint64) + np.uint64(seed << 32)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z.view(np.int64).reshape(1326, 1326)


def checksums(W, T, w1, w2):
    with np.errstate(over="ignore"):
        return np.array([W.sum(dtype=np.int64), T.sum(dtype=np.int64),
                         (W.astype(np.int64) * w1).sum(dtype=np.int64),
                         (T.astype(np.int64) * w2).sum(dtype=np.int64)], np.int64)


def main():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_ctypes as O
    flops = O.canonical_flops()
    w1, w2 = splitmix_weights(1), splitmix_weights(2)
    out = np.zeros((len(flops), 4), np.int64)
    t0 = time.time()
    for f in range(len(flops)):
        W, T = O.flop_counts(flops[f, :3])
        out[f] = checksums(W, T, w1, w2)
        if f % 100 == 0:
            print(f"flop {f}: {time.time() - t0:.0f}s", flush=True)
    np.savez_compressed(OUT, flops=flops[:, :3].copy(), checksums=out)
    print(f"wrote {OUT} ({len(flops)} flops, {time.time() - t0:.0f}s)")


if __name__ == "__main__":
    main()
