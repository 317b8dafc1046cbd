"""Flop-subset selection over the EV dump (scripts/optimize_subset.py:28-65):
a non-negative orthogonal matching pursuit picks the flops whose EV
columns best rebuild the orbit-weighted mean EV of all 1,755 canonical
flops, with non-negative least squares refits; the subset is written in the
reference's formats (mtt_subset.bin: uint32 count, then per flop
{u8 c1, c2, c3, pad; float32 weight}; mtt_subset_weights.txt:
"KcKdKh:35.5963" lines).  The weights are rescaled to sum to the number of
canonical flops.

Host-side tool: the EV dump it reads comes from the GPU (mtt_ev_dump /
mtt_ev_dump_tensor); the selection itself is small dense linear algebra
(1,352 features x 1,755 flops) done with numpy and scipy's NNLS, the routine
the reference uses.
"""
import ctypes
import importlib.util
import os
import struct
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

N_FLOPS = 1755
N_FEATURES = 4 * 2 * 169
RANKS = "23456789TJQKA"
SUITS = "cdhs"


def canonical_flops():
    """[1755][4] uint8: three cards and the orbit weight, in the reference's
    order (generate_all_1755_flops, mtt_ev_dumper.cu:34-51)."""
    out = np.zeros((N_FLOPS, 4), np.uint8)
    n = mtt.lib().mtt_canonical_flops(out.ctypes.data)
    if n != N_FLOPS:
        raise mtt.MTTError(f"mtt_canonical_flops returned {n}")
    return out


def select_subset(dump, k=100):
    """dump: float32 [1755][4][2][169] (the EV-dump tensor).  Returns the
    indices of the k chosen flops in selection order and their weights
    (non-negative, summing to 1755)."""
    from scipy.optimize import nnls
    A = np.ascontiguousarray(np.asarray(dump, np.float32).reshape(N_FLOPS, N_FEATURES).T)
    w = canonical_flops()[:, 3].astype(np.float64)
    target = (A @ w) / w.sum()
    norms = np.linalg.norm(A, axis=0)
    norms[norms == 0] = 1.0
    An = A / norms
    chosen = []
    residual = target.copy()
    for _ in range(k):
        score = An.T @ residual
        score[chosen] = -np.inf
        chosen.append(int(np.argmax(score)))
        coef, _ = nnls(A[:, chosen], target)
        residual = target - A[:, chosen] @ coef
    coef, _ = nnls(A[:, chosen], target)
    return np.array(chosen, np.int64), coef * (N_FLOPS / coef.sum())


def subset_bytes(indices, weights):
    """The mtt_subset.bin layout (SubsetFlop, preflop_grid_generator.cu:30)."""
    fl = canonical_flops()
    out = [struct.pack("<I", len(indices))]
    for i, wt in zip(indices, weights):
        c = fl[i]
        out.append(struct.pack("<BBBBf", int(c[0]), int(c[1]), int(c[2]), 0, float(wt)))
    return b"".join(out)


def subset_text(indices, weights):
    fl = canonical_flops()
    name = lambda c: RANKS[c // 4] + SUITS[c % 4]
    return "".join(f"{name(int(fl[i][0]))}{name(int(fl[i][1]))}{name(int(fl[i][2]))}:{wt:.4f}\n"
                   for i, wt in zip(indices, weights))


def write_subset(indices, weights, bin_path="mtt_subset.bin", txt_path="mtt_subset_weights.txt"):
    with open(bin_path, "wb") as f:
        f.write(subset_bytes(indices, weights))
    with open(txt_path, "w") as f:
        f.write(subset_text(indices, weights))


def read_subset(data):
    """Parse mtt_subset.bin bytes: ([n][3] cards, [n] float32 weights)."""
    n = struct.unpack_from("<I", data, 0)[0]
    cards = np.zeros((n, 3), np.uint8)
    w = np.zeros(n, np.float32)
    for k in range(n):
        c1, c2, c3, _, wt = struct.unpack_from("<BBBBf", data, 4 + 8 * k)
        cards[k] = (c1, c2, c3)
        w[k] = wt
    return cards, w
