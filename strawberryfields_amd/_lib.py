"""ctypes binding of libsfamd.so (the C ABI declared in include/sfamd.h).

The library is the compiled HIP code in ``strawberryfields_amd/csrc``; it is
built in-tree by ``strawberryfields_amd/build.py``.  There is no fallback:
if the library is missing or fails to load, every entry point raises.
"""
import ctypes
import os

import torch  # noqa: F401  (loads the HIP runtime the library binds to)

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libsfamd.so")

C64, C128 = 0, 1
(
    OP_ROTATION,
    OP_KERR,
    OP_DISPLACEMENT,
    OP_SQUEEZE,
    OP_BEAMSPLITTER,
    OP_LOSS,
    OP_MIX,
    OP_TWO_MODE_SQUEEZE,
    OP_MZGATE,
    OP_CROSS_KERR,
    OP_CUBIC_PHASE,
) = range(11)
OP_KINDS = {
    "R": OP_ROTATION,
    "K": OP_KERR,
    "D": OP_DISPLACEMENT,
    "S": OP_SQUEEZE,
    "BS": OP_BEAMSPLITTER,
    "LOSS": OP_LOSS,
    "MIX": OP_MIX,
    "S2": OP_TWO_MODE_SQUEEZE,
    "MZ": OP_MZGATE,
    "CK": OP_CROSS_KERR,
    "V": OP_CUBIC_PHASE,
}
TWO_MODE = ("BS", "S2", "MZ", "CK")
MAX_CUTOFF = 64
MAX_CUTOFF_BS = 40  # beamsplitter, S2, MZ, CK
MAX_CUTOFF_V = 32  # cubic phase


class SfamdError(RuntimeError):
    """A libsfamd call returned a non-zero status."""


class OpDesc(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("mode0", ctypes.c_int32),
        ("mode1", ctypes.c_int32),
        ("pad", ctypes.c_int32),
    ]


class ProgInfo(ctypes.Structure):
    _fields_ = [
        ("table_elems", ctypes.c_int64),
        ("partial_doubles", ctypes.c_int64),
        ("ws_elems", ctypes.c_int64),
        ("final_elems", ctypes.c_int64),
        ("final_mixed", ctypes.c_int32),
        ("nops", ctypes.c_int32),
        ("nsteps", ctypes.c_int32),
        ("pad", ctypes.c_int32),
    ]


_P = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64
_D = ctypes.c_double

# name -> (restype, argtypes); every exported symbol of include/sfamd.h
SIGNATURES = {
    "sfamd_version": (_I, []),
    "sfamd_status_string": (ctypes.c_char_p, [_I]),
    "sfamd_gate_table_size": (_I64, [_I, _I]),
    "sfamd_build_gate": (_I, [_I, _D, _D, _I, _I, _P, _P]),
    "sfamd_apply_1mode": (_I, [_I, _P, _P, _P, _I, _I, _I, _I, _P, _I, _P]),
    "sfamd_apply_bs": (_I, [_P, _P, _P, _I, _I, _I, _I, _I, _P, _I, _P]),
    "sfamd_apply_2mode": (_I, [_I, _P, _P, _P, _I, _I, _I, _I, _I, _P, _I, _P]),
    "sfamd_apply_matrix": (_I, [_P, _P, _P, _I, _I, _I, _I, _P, _I, _P]),
    "sfamd_apply_loss": (_I, [_P, _P, _I, _I, _I, _P, _I, _P]),
    "sfamd_mix": (_I, [_P, _P, _I, _I, _I, _P]),
    "sfamd_replace_mode": (_I, [_P, _P, _I, _I, _I, _P, _I, _P]),
    "sfamd_reduce_partials": (_I64, [_I, _I, _I]),
    "sfamd_norm": (_I, [_P, _I, _I, _I, _P, _P, _I, _P]),
    "sfamd_all_fock_probs": (_I, [_P, _P, _I, _I, _I, _I, _P]),
    "sfamd_reduced_dm_partials": (_I64, [_I, _I, _I, _I]),
    "sfamd_reduced_dm": (_I, [_P, _P, _I, _I, _I, ctypes.POINTER(_I), _I, _P, _I, _P]),
    "sfamd_diagonal_expectation": (
        _I,
        [_P, _I, _I, _I, ctypes.POINTER(_I), _I, _P, _P, _P, _I, _P],
    ),
    "sfamd_all_fock_probs_vjp": (_I, [_P, _P, _P, _I, _I, _I, _I, _P]),
    "sfamd_overlap_partials": (_I64, [_I64]),
    "sfamd_overlap": (_I, [_P, _P, _I64, _P, _P, _I, _P]),
    "sfamd_fidelity_loss": (_I, [_P, _P, _I64, _P, _P, _P, _I, _P]),
    "sfamd_fidelity_loss_noalloc": (_I, [_P, _P, _I64, _P, _P, _I, _P]),
    "sfamd_number_loss": (_I, [_P, _P, _I, _I, _P, _P, _I, _P]),
    "sfamd_marginal_partials": (_I64, [_I, _I, _I]),
    "sfamd_marginal_probs": (_I, [_P, _I, _I, _I, ctypes.POINTER(_I), _I, _P, _P, _I, _P]),
    "sfamd_project_reset": (
        _I,
        [_P, _P, _I, _I, _I, ctypes.POINTER(_I), ctypes.POINTER(_I), _I, _D, _I, _P],
    ),
    "sfamd_homodyne_pdf": (_I, [_P, _I, _D, _D, _I, _D, _P, _I, _P]),
    "sfamd_embed_modes": (_I, [_P, _P, _P, _I, _I, ctypes.POINTER(_I), _I, _I, _I, _P]),
    "sfamd_prog_create": (_I, [ctypes.POINTER(OpDesc), _I, _I, _I, _I, _I, _I, ctypes.POINTER(_P)]),
    "sfamd_prog_info": (_I, [_P, ctypes.POINTER(ProgInfo)]),
    "sfamd_prog_state_elems": (_I64, [_P, _I]),
    "sfamd_prog_step_ops": (_I, [_P, _I, ctypes.POINTER(ctypes.c_int32), _I]),
    "sfamd_prog_destroy": (_I, [_P]),
    "sfamd_prog_build_table": (_I, [_P, _P, _P, _P]),
    "sfamd_prog_forward": (_I, [_P, _I, _I, _P, _P, _P, ctypes.POINTER(_P), _P, _P]),
    "sfamd_prog_adjoint": (
        _I,
        [_P, _I, _I, _P, ctypes.POINTER(_P), _P, _P, _P, _P, _P, _P],
    ),
    "sfamd_graph_begin": (_I, [_P]),
    "sfamd_graph_end": (_I, [_P, ctypes.POINTER(_P)]),
    "sfamd_graph_launch": (_I, [_P, _P]),
    "sfamd_graph_destroy": (_I, [_P]),
}

_lib = None


def load():
    """Load libsfamd.so (once) and return the ctypes handle; raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "libsfamd.so not found at %s: build it with `python -m strawberryfields_amd.build` "
            "(or __graft_entry__.build()); there is no non-HIP fallback" % LIB_PATH
        )
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status, what=""):
    """Raise SfamdError for a non-zero status."""
    if status != 0:
        msg = load().sfamd_status_string(status).decode()
        if status == -1:
            raise ValueError("%s: %s" % (what, msg))
        raise SfamdError("%s failed with status %d (%s)" % (what, status, msg))


def call(name, *args):
    """Call an int-returning entry point and raise on a non-zero status."""
    st = getattr(load(), name)(*args)
    check(st, name)


def dtype_code(dtype):
    if dtype == torch.complex64:
        return C64
    if dtype == torch.complex128:
        return C128
    raise ValueError("state dtype must be torch.complex64 or torch.complex128, got %s" % dtype)


def real_dtype(dtype):
    return torch.float32 if dtype == torch.complex64 else torch.float64


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def stream_ptr(device=None):
    """The current torch HIP stream of `device` as a void*."""
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
