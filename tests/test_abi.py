"""C-ABI checks that need no GPU: the library loads, exports every symbol the
header declares, and the host compiler reports the reference's errors."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import qcf
from qcf import _lib, circuits

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    src = open(os.path.join(ROOT, "include", "qcf.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(qcf_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    h = C.CDLL(_lib.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(h, s), s
    # the Python binding declares the same set
    assert set(syms) == set(_lib.exported_symbols())


def test_plane_words_aligns_tiles():
    """The recommended plane stride holds every shot and is a whole number of
    256-byte tiles (32 words per 2048 shots); Python and C agree."""
    f = qcf.lib().qcf_plane_words
    for shots in (0, 1, 63, 64, 2047, 2048, 2049, 10**7, 2**33 + 5):
        w = f(shots)
        assert w == qcf.frames.plane_words(shots)
        assert w % 32 == 0 and w * 64 >= shots and w >= 32
        assert w - 32 < max(32, qcf.frames.words(shots))


def test_version():
    assert b"gfx950" in qcf.lib().qcf_version()


@pytest.mark.parametrize("bad,exc", [
    (lambda: qcf.Program([qcf.sCNOT(1, 2), qcf.sMZ(3, 5)], nqubits=2, nbits=2), qcf.BoundsError),
    (lambda: qcf.Program([qcf.sMZ(4, 1)], nqubits=2, nbits=2), qcf.BoundsError),
    (lambda: qcf.Program([qcf.ClassicalXOR((), 1)]), qcf.BoundsError),
    (lambda: qcf.Program([qcf.ClassicalXOR((1, 9), 2)], nbits=4), qcf.BoundsError),
    (lambda: qcf.Program([qcf.sMY(1, 1)]), qcf.MethodError),
    (lambda: qcf.sCNOT(1, 1), qcf.ArgumentError),
    (lambda: qcf.sHadamard(0), qcf.ArgumentError),
])
def test_errors_mirror_reference(bad, exc):
    with pytest.raises(exc):
        bad()


def test_hidden_ancilla_qubit_is_addressable():
    # the frame has nqubits+1 qubits (src/pauli_frames.jl:30-37)
    p = qcf.Program([qcf.sCNOT(1, 3)], nqubits=2, nbits=1)
    assert p.nq == 3


def test_bitops_accounting():
    """Frame bit-op weights: CNOT 2, H 2, DEPOLARIZE1 2, DEPOLARIZE2 4, X-flip 1,
    sMZ 1, sMRZ 3 (1 record + 2 reset), XOR of k bits k-1."""
    assert qcf.Program([qcf.sCNOT(1, 2)]).bitops_per_shot == 2
    assert qcf.Program([qcf.sHadamard(1)]).bitops_per_shot == 2
    assert qcf.Program([qcf.sPhase(1)]).bitops_per_shot == 1
    assert qcf.Program([qcf.sSWAP(1, 2)]).bitops_per_shot == 4
    assert qcf.Program([qcf.PauliError(1, 0.1)]).bitops_per_shot == 2
    assert qcf.Program([qcf.PauliError(1, 0.1, 0, 0)]).bitops_per_shot == 1
    assert qcf.Program([qcf.Depolarize2(1, 2, 0.1)]).bitops_per_shot == 4
    assert qcf.Program([qcf.sMZ(1, 1)]).bitops_per_shot == 1
    assert qcf.Program([qcf.sMRZ(1, 1)]).bitops_per_shot == 3
    assert qcf.Program([qcf.sMZ(1, 1), qcf.sMZ(2, 2), qcf.ClassicalXOR((1, 2), 3)]
                       ).bitops_per_shot == 3


def test_surface_code_program_fits_on_chip():
    c = circuits.rotated_surface_code(5, 25, 1e-3)
    p = qcf.Program(c.ops)
    assert p.info.lds_bytes <= 64 * 1024
    assert p.info.nclasses == 1
    # replay planes: initZ! (50) + sMRZ z-flips + 2 per single-qubit site + 4 per DEP2
    assert p.nreplay == 12447
    assert abs(p.bitops_per_shot - 18217) < 1e-9


def test_sites_of_all_rates_share_one_thinned_class():
    circ = [qcf.PauliError(1, p) for p in np.geomspace(1e-5, 1e-1, 9)]
    p = qcf.Program(circ)
    assert p.info.nclasses == 1 and p.info.nsites == 9
    # a zero-rate site needs no class
    assert qcf.Program([qcf.PauliError(1, 0.0)]).info.nclasses == 0


@pytest.mark.parametrize("nrows,nwords,ld,nqubits,exc", [
    (16385, 1, 2, 64, qcf.QCFError),            # more rows than the elimination kernel holds
    (10, 1, 2, 65, qcf.DimensionMismatch),      # 65 qubits do not fit one word per half
    (10, 2, 3, 100, qcf.DimensionMismatch),     # row stride shorter than the row
])
def test_canonicalize_argument_errors(nrows, nwords, ld, nqubits, exc):
    """qcf_canonicalize rejects bad shapes before it touches the device (so
    host pointers are never dereferenced here)."""
    xzs = np.zeros(max(1, nrows * ld), np.uint64)
    ph = np.zeros(max(1, nrows), np.uint8)
    with pytest.raises(exc):
        _lib.call("qcf_canonicalize", xzs.ctypes.data_as(C.c_void_p),
                  ph.ctypes.data_as(C.c_void_p), ld, nwords, nrows, nqubits, 1, None, None)
