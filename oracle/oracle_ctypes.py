"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker; never by the product.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None
_P = ctypes.c_void_p


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run make -C oracle")
        L = ctypes.CDLL(LIB_PATH)
        L.or_eval7.restype = ctypes.c_uint32
        L.or_eval7.argtypes = [_P]
        L.or_eval7_batch.argtypes = [_P, ctypes.c_int64, _P]
        L.or_category_counts.argtypes = [_P]
        L.or_canonical_class.argtypes = [ctypes.c_int, ctypes.c_int]
        L.or_canonical_flops.argtypes = [_P]
        L.or_ev_dump.argtypes = [ctypes.c_int, ctypes.c_int, _P]
        L.or_combo_index.argtypes = [ctypes.c_int, ctypes.c_int]
        L.or_flop_counts.argtypes = [_P, _P, _P]
        L.or_turn_matrix.argtypes = [_P, ctypes.c_int, _P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_P)


def eval7_batch(cards):
    cards = np.ascontiguousarray(cards, dtype=np.uint8)
    out = np.zeros(cards.shape[0], np.uint32)
    lib().or_eval7_batch(_p(cards), cards.shape[0], _p(out))
    return out


def category_counts():
    out = np.zeros(9, np.uint64)
    lib().or_category_counts(_p(out))
    return out


def canonical_flops():
    out = np.zeros((1755, 4), np.uint8)
    n = lib().or_canonical_flops(_p(out))
    assert n == 1755
    return out


def ev_dump(first, count):
    out = np.zeros((count, 4, 2, 169), np.float32)
    lib().or_ev_dump(first, count, _p(out))
    return out


def flop_counts(flop):
    f = np.asarray(flop, np.uint8)
    W = np.zeros((1326, 1326), np.int32)
    T = np.zeros((1326, 1326), np.int32)
    lib().or_flop_counts(_p(f), _p(W), _p(T))
    return W, T


def turn_matrix(flop, turn_card):
    f = np.asarray(flop, np.uint8)
    E = np.zeros((1216, 1216), np.float32)
    lib().or_turn_matrix(_p(f), int(turn_card), _p(E))
    return E
