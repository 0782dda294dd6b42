"""ctypes binding of libqcf.so (the C ABI declared in include/qcf.h).

The library is the only compute path: if it is missing or fails to load, every
entry point raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# QCF_LIBRARY may name another build of the same library (the sanitized one
# the CPU tests can run against); there is no non-HIP path either way.
LIB_PATH = os.environ.get("QCF_LIBRARY") or os.path.join(_HERE, "libqcf.so")


class QCFError(RuntimeError):
    """Base class of errors reported by libqcf."""


class DimensionMismatch(QCFError, ValueError):
    """Mirror of Julia's DimensionMismatch (src/mul_leftright.jl:138)."""


class BoundsError(QCFError, IndexError):
    """Mirror of Julia's BoundsError."""


class ArgumentError(QCFError, ValueError):
    """Mirror of Julia's ArgumentError / NoZeroQubit (src/throws.jl:1)."""


class MethodError(QCFError, TypeError):
    """Mirror of Julia's MethodError: no PauliFrame method for an operation."""


class HIPError(QCFError):
    """HIP runtime failure."""


_ERRORS = {1: DimensionMismatch, 2: BoundsError, 3: ArgumentError, 4: MethodError,
           5: HIPError, 6: MemoryError}

# qcf_op (include/qcf.h)
OP_DTYPE = np.dtype([("kind", "<u4"), ("q1", "<u4"), ("q2", "<u4"), ("bit", "<u4"),
                     ("arg_off", "<u4"), ("arg_len", "<u4"), ("p", "<f8", (3,))],
                    align=True)
assert OP_DTYPE.itemsize == 48


class ProgramInfo(C.Structure):
    _fields_ = [("nqubits", C.c_uint32), ("nbits", C.c_uint32), ("nplanes", C.c_uint32),
                ("ninstr", C.c_uint32), ("nsites", C.c_uint32), ("nclasses", C.c_uint32),
                ("nreplay", C.c_uint32), ("lds_bytes", C.c_uint32),
                ("bitops_per_shot", C.c_double)]


class Tune(C.Structure):
    _fields_ = [("waves_per_cu", C.c_uint32), ("tile_blocks", C.c_uint32),
                ("lds_cap", C.c_uint32), ("jit_max_planes", C.c_uint32),
                ("streams", C.c_uint32), ("concurrent", C.c_uint32),
                ("steal_planes", C.c_uint32)]

    def as_list(self):
        return [getattr(self, f) for f, _ in self._fields_]


_P = C.c_void_p
_SZ = C.c_size_t
_U32 = C.c_uint32
_U64 = C.c_uint64
_I = C.c_int

# name -> (restype, argtypes)
_SIGNATURES = {
    "qcf_last_error": (C.c_char_p, []),
    "qcf_version": (C.c_char_p, []),
    "qcf_mul_left": (_I, [_P, _P, _P, _P, _SZ, _I, _P]),
    "qcf_mul_right": (_I, [_P, _P, _P, _P, _SZ, _I, _P]),
    "qcf_mul_vec": (_I, [_P, _P, _SZ, _I, _I, _P, _P]),
    "qcf_mul_rows": (_I, [_P, _P, _SZ, _SZ, _SZ, _P, _P, _SZ, _P, _I, _I, _P]),
    "qcf_canonicalize": (_I, [_P, _P, _SZ, _SZ, _SZ, _SZ, _I, _P, _P]),
    "qcf_canonicalize_noalloc": (_I, [_P, _P, _SZ, _SZ, _SZ, _SZ, _I, _P, _P]),
    "qcf_program_build": (_I, [_P, _SZ, _P, _SZ, _U32, _U32, C.POINTER(_P)]),
    "qcf_program_free": (_I, [_P]),
    "qcf_program_get_info": (_I, [_P, C.POINTER(ProgramInfo)]),
    "qcf_program_gauge_planes": (_I, [_P, _P, _SZ]),
    "qcf_pf_sample": (_I, [_P, _U64, _U64, _U64, _P, _P, _P, _SZ, _U32, _P]),
    "qcf_plane_words": (_SZ, [_U64]),
    "qcf_batch_create": (_I, [_P, _SZ, _P, _P, _P, _P, C.POINTER(Tune), C.POINTER(_P)]),
    "qcf_batch_launch": (_I, [_P, _P]),
    "qcf_batch_launch_graph": (_I, [_P, _P]),
    "qcf_batch_free": (_I, [_P]),
    "qcf_pf_sample_batched": (_I, [_P, _SZ, _P, _P, _P, _P, C.POINTER(Tune), _P]),
    "qcf_autotune": (_I, [_P, _SZ, _P, C.POINTER(Tune), _P]),
    "qcf_debug_philox": (_I, [_P, _P, _SZ, _P]),
    "qcf_debug_wave_scan": (_I, [_P, _P, _SZ, _P]),
    "qcf_debug_xoshiro128ss": (_I, [_P, _P, _SZ, _U32, _P]),
    "qcf_debug_gauge": (_I, [_P, _P, _SZ, _U32, _P]),
    "qcf_debug_gauge_stream": (_I, [_U64, _P, _P, _SZ, _U32, _P]),
    "qcf_debug_check_launch": (_I, [C.c_uint]),
    "qcf_debug_jit_source": (_I, [_P, _I, _P, _SZ, C.POINTER(_SZ)]),
    "qcf_debug_jit_compile": (_I, [_P, _I, C.POINTER(_SZ)]),
    "qcf_record_to_bool": (_I, [_P, _SZ, _U64, _U32, _P, _SZ, _P]),
    "qcf_record_from_bool": (_I, [_P, _SZ, _U64, _U32, _P, _SZ, _P]),
    "qcf_frame_to_reference": (_I, [_P, _SZ, _U64, _U32, _P, _SZ, _P]),
    "qcf_frame_from_reference": (_I, [_P, _SZ, _U64, _U32, _P, _SZ, _P]),
    "qcf_program_bitops": (C.c_double, [_P]),
    "qcf_dem_build": (_I, [_P, _P, _SZ, _P, _SZ, C.POINTER(_P)]),
    "qcf_dem_free": (_I, [_P]),
    "qcf_dem_size": (_I, [_P, C.POINTER(_SZ), C.POINTER(_U32), C.POINTER(_U32)]),
    "qcf_dem_mechanism": (_I, [_P, _SZ, _P, _SZ, C.POINTER(_SZ), C.POINTER(_U64),
                               C.POINTER(C.c_double)]),
    "qcf_matcher_create": (_I, [_P, C.POINTER(_P)]),
    "qcf_matcher_free": (_I, [_P]),
    "qcf_matcher_info": (_I, [_P, C.POINTER(_SZ), C.POINTER(_U32)]),
    "qcf_matcher_decode": (_I, [_P, _P, _U64, _P, _P, _I]),
    "qcf_matcher_decode_packed": (_I, [_P, _P, _SZ, _U64, _P, _I]),
    "qcf_debug_max_weight_matching": (_I, [_I, _P, _P]),
    "qcf_mlp_create": (_I, [_P, _U32, _U32, _U32, _U32, _U32, _U32, _U32, C.POINTER(_P)]),
    "qcf_mlp_free": (_I, [_P]),
    "qcf_mlp_nparams": (_I, [_P, C.POINTER(_SZ)]),
    "qcf_mlp_set_params": (_I, [_P, _P, _P]),
    "qcf_mlp_get_params": (_I, [_P, _P, _P]),
    "qcf_mlp_train_step": (_I, [_P, _P, _SZ, _U64, C.c_float, C.c_float, C.c_float,
                                C.c_float, _P, _P]),
    "qcf_mlp_gradients": (_I, [_P, _P, _SZ, _U64, _P, _P, _P]),
    "qcf_mlp_infer": (_I, [_P, _P, _SZ, _U64, _P, _P]),
    "qcf_debug_gemm_bf16": (_I, [_I, _I, _I, _I, _P, _P, _P, _P]),
}

_lib = None


def lib():
    """Load libqcf.so once; raise if it is absent (there is no other path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with quantumclifford.jl_amd/build.py "
                "(there is no non-HIP fallback)")
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def exported_symbols():
    return list(_SIGNATURES)


def check(status: int) -> None:
    if status == 0:
        return
    msg = lib().qcf_last_error().decode(errors="replace")
    raise _ERRORS.get(status, QCFError)(msg or f"qcf status {status}")


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))
