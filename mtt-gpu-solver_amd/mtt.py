"""ctypes binding of libmtt.so (the C ABI declared in include/mtt.h).

This is the thin host-side layer the tests and the benchmark use.  Device
buffers are passed as raw addresses (e.g. ``torch.Tensor.data_ptr()``); the
library never allocates through, or depends on, PyTorch.  Every call checks
the returned status and raises :class:`MTTError` with ``mtt_last_error()``.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmtt.so")

MTT_OK = 0
MTT_ERR_ARG, MTT_ERR_OOM, MTT_ERR_HIP, MTT_ERR_TREE, MTT_ERR_STATE = 1, 2, 3, 4, 5
LAYOUT_GLOBAL1326 = 0
LAYOUT_ALIVE1216 = 1
N_ALIVE = 1176
N_PAD = 1216
NPR = 1088          # river combos per board in the batched layout (1081 live, padded)
MAX_ACTIONS = 4

NODE_DECISION, NODE_FOLD, NODE_SHOWDOWN, NODE_CHANCE = 0, 1, 2, 3


class MTTError(RuntimeError):
    """A status other than MTT_OK; .status holds it (MTT_ERR_ARG, ...)."""
    status = None


class FlatNode(ctypes.Structure):
    """mtt_flat_node: honest_multistreet_solver.cu:16-21 layout (32 bytes)."""
    _fields_ = [("type", ctypes.c_uint8), ("player", ctypes.c_uint8),
                ("num_actions", ctypes.c_uint8), ("active_mask", ctypes.c_uint8),
                ("children_offset", ctypes.c_int32), ("infoset_offset", ctypes.c_int32),
                ("turn_idx", ctypes.c_int16), ("river_idx", ctypes.c_int16),
                ("pot_size", ctypes.c_float), ("committed", ctypes.c_float * 2),
                ("_pad", ctypes.c_uint8 * 4)]


class FlatNode4(ctypes.Structure):
    """mtt_flat_node4: multiway_4way_engine.cu:20-24 layout (32 bytes)."""
    _fields_ = [("type", ctypes.c_uint8), ("player", ctypes.c_uint8),
                ("num_actions", ctypes.c_uint8), ("active_mask", ctypes.c_uint8),
                ("children_offset", ctypes.c_int32), ("infoset_offset", ctypes.c_int32),
                ("pot_size", ctypes.c_float), ("committed", ctypes.c_float * 4)]


class AliveCombo(ctypes.Structure):
    """mtt_alive_combo: mtt_ev_dumper.cu:24 layout (16 bytes)."""
    _fields_ = [("c1", ctypes.c_uint8), ("c2", ctypes.c_uint8), ("_pad", ctypes.c_uint8 * 6),
                ("mask", ctypes.c_uint64)]


assert ctypes.sizeof(FlatNode) == 32 and ctypes.sizeof(FlatNode4) == 32
assert ctypes.sizeof(AliveCombo) == 16

_lib = None
_P = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64
_F = ctypes.c_float
_D = ctypes.c_double

# name -> argtypes (restype int unless listed in _RESTYPES)
_SIGS = {
    "mtt_last_error": [],
    "mtt_version": [],
    "mtt_eval7": [_P],
    "mtt_eval7_host_batch": [_P, _I64, _P],
    "mtt_eval7_batch": [_P, _I64, _P, _P],
    "mtt_canonical_class": [_I, _I],
    "mtt_canonical_flops": [_P],
    "mtt_combo_index": [_I, _I],
    "mtt_flop_counts": [_P, _I, _P, _P, _I, _P],
    "mtt_flop_counts_noalloc": [_P, _I, _P, _P, _I, _P],
    "mtt_flop_counts_timing": [_I],
    "mtt_flop_counts_kernel_times": [_I, ctypes.c_char_p, _I, _P, _P],
    "mtt_evaluate_flop_ev": [_P, _P, _P, _P, _P, _P, _P, _P],
    "mtt_ev_dump_tensor": [_I, _I, _P],
    "mtt_ev_dump": [_I, ctypes.c_char_p],
    "mtt_gen_turn_matrices": [_P, _P, _P, _P, _P],
    "mtt_gen_river_matrices": [_P, _P, _P, _P, _I, _P, _P],
    "mtt_gen_flop_matrices": [_P, _I, _P, _P, _P],
    "mtt_gen_preflop_matrix": [_P, _P, _P, _I, _P],
    "mtt_class_cards": [_I, _P, _P],
    "mtt_flop_class_matrices": [_P, _I, _P, _P],
    "mtt_preflop_solve": [_P, _P, _I, _P, _I, _F, _F, _F, _P, _P],
    "mtt_preflop_cosolve": [_P, _I, _P, _I, _F, _F, _F, _P, _P],
    "mtt_preflop_timing": [_I],
    "mtt_preflop_kernel_times": [_I, ctypes.c_char_p, _I, _P, _P],
}
_RESTYPES = {"mtt_last_error": ctypes.c_char_p, "mtt_version": ctypes.c_char_p,
             "mtt_eval7": ctypes.c_uint32}


def lib():
    """Load libmtt.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MTTError(f"{LIB_PATH} is missing: run __graft_entry__.build() (make -C mtt-gpu-solver_amd)")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, ctypes.c_int)
        _lib = L
    return _lib


def declared_symbols():
    return list(_SIGS)


def register(name, argtypes, restype=ctypes.c_int):
    """Declare one more entry point (used by modules that extend the ABI)."""
    _SIGS[name] = argtypes
    if restype is not ctypes.c_int:
        _RESTYPES[name] = restype
    if _lib is not None:
        fn = getattr(_lib, name)
        fn.argtypes = argtypes
        fn.restype = restype


def last_error():
    return lib().mtt_last_error().decode(errors="replace")


def check(status, what=""):
    if status != MTT_OK:
        e = MTTError(f"{what} failed with status {status}: {last_error()}")
        e.status = status
        raise e
    return status


def call(name, *args):
    """Call an int-returning entry point and raise on a non-zero status."""
    return check(getattr(lib(), name)(*args), name)


def ptr(x):
    """Address of a torch tensor, numpy array or int."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    if hasattr(x, "ctypes"):
        return x.ctypes.data
    raise TypeError(type(x))
