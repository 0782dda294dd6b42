"""CPU oracle for the frame sampler and the Pauli product.

TEST INFRASTRUCTURE ONLY: a ctypes wrapper of liboracle.so (qcf_oracle.c,
a C restatement of QuantumClifford.jl v0.11.7) plus numpy helpers that move
data between the reference layouts the oracle uses and the shot-packed
layouts of libqcf.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this package; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")
_lib = None

_P = C.c_void_p
_SZ = C.c_size_t


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            subprocess.run(["make", "-C", _HERE, "liboracle.so"], check=True,
                           capture_output=True)
        h = C.CDLL(_SO)
        h.oracle_mul_ordered.argtypes = [_P, _P, _SZ, C.c_int, _P, _P]
        h.oracle_mul_left_vec.argtypes = [_P, _P, _SZ, C.c_int]
        h.oracle_mul_left_vec.restype = C.c_uint8
        h.oracle_mul_right_vec.argtypes = [_P, _P, _SZ, C.c_int]
        h.oracle_mul_right_vec.restype = C.c_uint8
        h.oracle_mul_left_pauli.argtypes = [_P, _P, _P, _P, _SZ, C.c_int]
        h.oracle_mul_right_pauli.argtypes = [_P, _P, _P, _P, _SZ, C.c_int]
        h.oracle_mul_rows.argtypes = [_P, _P, _SZ, _SZ, _SZ, _P, _P, _SZ, _P, C.c_int, C.c_int]
        h.oracle_canonicalize.argtypes = [_P, _P, _SZ, _SZ, _SZ, _SZ, C.c_int, _P]
        h.oracle_prodphase.argtypes = [_P, _P, _SZ]
        h.oracle_prodphase.restype = C.c_uint8
        h.oracle_comm.argtypes = [_P, _P, _SZ]
        h.oracle_comm.restype = C.c_uint8
        h.oracle_philox4x32_10.argtypes = [_P, _P, _P]
        h.oracle_pf_run.argtypes = [_P, _SZ, _P, _SZ, C.c_uint32, C.c_uint32, C.c_uint64, _P, _P,
                                    _P, _SZ, C.c_int, C.c_uint64, C.c_int]
        h.oracle_pf_run.restype = C.c_int
        _lib = h
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ---- Pauli product ------------------------------------------------------
def mul_left_vec(r: np.ndarray, l: np.ndarray, phases=True) -> int:
    """In place on r (uint64 [x;z]); returns the UInt8 product phase."""
    return int(lib().oracle_mul_left_vec(_p(r), _p(l), r.size // 2, int(phases)))


def mul_right_vec(l: np.ndarray, r: np.ndarray, phases=True) -> int:
    return int(lib().oracle_mul_right_vec(_p(l), _p(r), l.size // 2, int(phases)))


def mul_left_pauli(r, rphase: np.ndarray, l, lphase: np.ndarray, phases=True):
    lib().oracle_mul_left_pauli(_p(r), _p(rphase), _p(l), _p(lphase), r.size // 2, int(phases))


def mul_right_pauli(l, lphase, r, rphase, phases=True):
    lib().oracle_mul_right_pauli(_p(l), _p(lphase), _p(r), _p(rphase), l.size // 2, int(phases))


def mul_rows(xzs, phases, src, src_phase, skip=None, row_mask=None, phases_on=True, right=False):
    """xzs: (nrows, ld) uint64 row-contiguous; src: (2W,) uint64."""
    nrows, ld = xzs.shape
    lib().oracle_mul_rows(_p(xzs), _p(phases), ld, src.size // 2, nrows, _p(src), _p(src_phase),
                          2**64 - 1 if skip is None else skip, _p(row_mask), int(phases_on),
                          int(right))


def canonicalize(xzs, phases, nqubits, phases_on=True):
    """canonicalize! in place on xzs (nrows, ld) uint64 row-contiguous and
    phases (nrows,) uint8; returns (X-stage pivots, all pivots)."""
    nrows, ld = xzs.shape
    ranks = np.zeros(2, np.uint64)
    lib().oracle_canonicalize(_p(xzs), _p(phases), ld, ld // 2, nrows, nqubits,
                              int(phases_on), _p(ranks))
    return int(ranks[0]), int(ranks[1])


def prodphase(l, r) -> int:
    return int(lib().oracle_prodphase(_p(l), _p(r), l.size // 2))


def comm(l, r) -> int:
    return int(lib().oracle_comm(_p(l), _p(r), l.size // 2))


def philox4x32_10(ctr, key) -> np.ndarray:
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    o = np.zeros(4, np.uint32)
    lib().oracle_philox4x32_10(_p(c), _p(k), _p(o))
    return o


# ---- frames ---------------------------------------------------------------
def pf_run(ops: np.ndarray, args: np.ndarray, nq: int, nbits: int, shots: int,
           masks: np.ndarray | None = None, fresh=True, xzs=None, meas=None, seed=0,
           nthreads=1):
    """Run the reference frame semantics.  Returns (xzs (2W, shots) uint64,
    meas (nbits, shots) uint8).  `masks` is (nplanes, ld) shot-packed uint64."""
    W = (nq + 63) // 64
    if xzs is None:
        xzs = np.zeros((2 * W, shots), np.uint64)
    if meas is None:
        meas = np.zeros((max(nbits, 1), shots), np.uint8)
    xzs = np.ascontiguousarray(xzs, np.uint64)
    meas = np.ascontiguousarray(meas, np.uint8)
    ld = 0
    if masks is not None:
        masks = np.ascontiguousarray(masks, np.uint64)
        ld = masks.shape[1]
    args = np.ascontiguousarray(args, np.uint32)
    st = lib().oracle_pf_run(_p(ops), len(ops), _p(args) if len(args) else None, len(args),
                             nq, nbits, shots, _p(xzs), _p(meas), _p(masks), ld, int(fresh),
                             seed, nthreads)
    if st != 0:
        raise RuntimeError(f"oracle status {st}")
    return xzs, meas


# ---- layout helpers (numpy) ----------------------------------------------
def unpack_shots(planes: np.ndarray, shots: int) -> np.ndarray:
    """(nplanes, ld) shot-packed uint64 -> (nplanes, shots) uint8."""
    p = np.ascontiguousarray(planes, np.uint64)
    bits = np.unpackbits(p.view(np.uint8).reshape(p.shape[0], -1), axis=1, bitorder="little")
    return bits[:, :shots]


def pack_shots(bits: np.ndarray) -> np.ndarray:
    """(nplanes, shots) 0/1 -> (nplanes, ceil(shots/64)) uint64."""
    n, s = bits.shape
    ld = (s + 63) // 64
    pad = np.zeros((n, ld * 64), np.uint8)
    pad[:, :s] = bits
    return np.packbits(pad, axis=1, bitorder="little").view(np.uint64).reshape(n, ld)


def frame_to_reference(frame_planes: np.ndarray, nq: int, shots: int) -> np.ndarray:
    """Shot-packed frame (2nq, ld) -> reference xzs (2W, shots) (pad bits zero)."""
    W = (nq + 63) // 64
    bits = unpack_shots(frame_planes, shots).astype(np.uint64)  # (2nq, shots)
    out = np.zeros((2 * W, shots), np.uint64)
    for half in range(2):
        for q in range(nq):
            out[half * W + q // 64] |= bits[half * nq + q] << np.uint64(q % 64)
    return out


def reference_to_frame(xzs: np.ndarray, nq: int, shots: int) -> np.ndarray:
    W = (nq + 63) // 64
    bits = np.zeros((2 * nq, shots), np.uint8)
    for half in range(2):
        for q in range(nq):
            bits[half * nq + q] = (xzs[half * W + q // 64] >> np.uint64(q % 64)) & np.uint64(1)
    return pack_shots(bits)
