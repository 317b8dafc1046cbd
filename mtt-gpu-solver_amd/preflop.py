"""Preflop co-solver driver (src/preflop/preflop_grid_generator.cu main,
:426-656) over the GPU entry points: the Monte-Carlo all-in matrix
(mtt_gen_preflop_matrix), class equity matrices of the subset flops
(mtt_flop_class_matrices) and the DCFR loop (mtt_preflop_solve), with the
reference's defaults (20bb, bubble factor 1.35, dead pot 2.5bb, 500
iterations, 3000 all-in samples) and its snapshot table of the average
button / big blind strategies.

PyTorch holds the device buffers; every computation runs in libmtt.so.
"""
import ctypes
import importlib.util
import os
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

N_CLASSES = 169
NODES = ("button root", "big blind vs min-raise", "button vs 3-bet jam", "big blind vs open jam")
ACTIONS = (("fold", "min-raise", "jam"), ("fold", "call", "3-bet jam"), ("fold", "call"), ("fold", "call"))

# Snapshot rows of the reference's table (preflop_grid_generator.cu:626-628),
# labelled by what the class indices are (the reference calls 17 "A5s"; it
# is A9s).
SNAPSHOT = ((0, "AA"), (1, "KK"), (2, "QQ"), (3, "JJ"), (4, "TT"), (5, "99"), (6, "88"), (12, "22"),
            (13, "AKs"), (17, "A9s"), (72, "T9s"), (85, "76s"), (91, "AKo"), (104, "QJo"), (168, "32o"))


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise mtt.MTTError("the preflop co-solver runs on the GPU: no device visible to torch")
    return torch


def allin_matrix(samples=3000, device="cuda:0"):
    """[169][169] float32 device tensor: hero-class row equities."""
    torch = _torch()
    eq = torch.empty(N_CLASSES * N_CLASSES, dtype=torch.float32, device=device)
    mtt.call("mtt_gen_preflop_matrix", eq.data_ptr(), None, None, int(samples), None)
    return eq.view(N_CLASSES, N_CLASSES)


def class_matrices(flops, device="cuda:0"):
    """flops: [n][3] cards.  Returns [n][169 villain][169 hero] float32 on
    the device."""
    torch = _torch()
    f = np.ascontiguousarray(np.asarray(flops, np.uint8).reshape(-1, 3))
    out = torch.empty((f.shape[0], N_CLASSES, N_CLASSES), dtype=torch.float32, device=device)
    mtt.call("mtt_flop_class_matrices", f.ctypes.data, int(f.shape[0]), out.data_ptr(), None)
    return out


def solve(class_mats, weights, allin, iters=500, stack_bb=20.0, bubble_factor=1.35, dead_pot=2.5):
    """The DCFR loop.  Returns the strategy sums [4][3][169] (numpy)."""
    torch = _torch()
    w = np.ascontiguousarray(np.asarray(weights, np.float32))
    M = class_mats.contiguous()
    A = allin.contiguous()
    if M.dim() != 3 or M.shape[1:] != (N_CLASSES, N_CLASSES) or M.shape[0] != w.shape[0]:
        raise mtt.MTTError(f"class matrices {tuple(M.shape)} do not match {w.shape[0]} weights")
    out = torch.empty((4, 3, N_CLASSES), dtype=torch.float32, device=M.device)
    mtt.call("mtt_preflop_solve", M.data_ptr(), w.ctypes.data, int(w.shape[0]), A.data_ptr(), int(iters),
             float(stack_bb), float(bubble_factor), float(dead_pot), out.data_ptr(), None)
    return out.cpu().numpy()


def read_subset(data):
    """mtt_subset.bin bytes -> ([n][3] cards, [n] weights)."""
    import struct
    n = struct.unpack_from("<I", data, 0)[0]
    cards = np.zeros((n, 3), np.uint8)
    w = np.zeros(n, np.float32)
    for k in range(n):
        c1, c2, c3, _, wt = struct.unpack_from("<BBBBf", data, 4 + 8 * k)
        cards[k] = (c1, c2, c3)
        w[k] = wt
    return cards, w


def average_strategy(sums):
    """Strategy sums -> average strategies, normalised per class as the
    reference's table (total + 1e-7)."""
    s = np.asarray(sums, np.float32)
    return s / (s.sum(axis=1, keepdims=True) + np.float32(1e-7))


def snapshot(sums):
    """The reference's table: button fold / min-raise / jam and big blind
    fold / defend / 3-bet jam for 15 classes."""
    avg = average_strategy(sums)
    lines = ["Hand   BTN Fold    BTN MinRaise  BTN Jam     BB Fold     BB Defend     BB 3B-Jam"]
    for c, name in SNAPSHOT:
        b, d = avg[0, :, c], avg[1, :, c]
        lines.append(f"{name:<7}{b[0]:<12.3f}{b[1]:<14.3f}{b[2]:<12.3f}{d[0]:<12.3f}{d[1]:<14.3f}{d[2]:<12.3f}")
    return "\n".join(lines)


def run(subset, iters=500, stack_bb=20.0, bubble_factor=1.35, dead_pot=2.5, samples=3000):
    """The reference program end to end from mtt_subset.bin bytes (or a
    path): all-in matrix, class matrices of the subset, DCFR.  Returns the
    strategy sums."""
    if isinstance(subset, (str, os.PathLike)):
        with open(subset, "rb") as f:
            subset = f.read()
    cards, w = read_subset(subset)
    A = allin_matrix(samples)
    M = class_matrices(cards)
    return solve(M, w, A, iters, stack_bb, bubble_factor, dead_pot)
