"""Pins the oracle's Pauli-frame semantics (no GPU): gate bit maps against
unitary conjugation, measurement/reset/XOR semantics of src/pauli_frames.jl,
and circuit-level invariants."""
import itertools

import numpy as np
import pytest
import scipy.linalg as sla

import oracle
import qcf
from qcf import circuits

I2 = np.eye(2, dtype=complex)
X = np.array([[0, 1], [1, 0]], dtype=complex)
Z = np.array([[1, 0], [0, -1]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
SIG = {(0, 0): I2, (1, 0): X, (0, 1): Z, (1, 1): Y}
H = np.array([[1, 1], [1, -1]], dtype=complex) / np.sqrt(2)
S = np.diag([1, 1j])
SQX = 0.5 * np.array([[1 + 1j, 1 - 1j], [1 - 1j, 1 + 1j]])
SQY = 0.5 * np.array([[1 + 1j, -1 - 1j], [1 + 1j, 1 + 1j]])
CXYZ = sla.expm(-1j * np.pi / 3 * (X + Y + Z) / np.sqrt(3))
SWAP = np.eye(4)[[0, 2, 1, 3]].astype(complex)
CNOT = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=complex)
CZ = np.diag([1, 1, 1, -1]).astype(complex)
ISWAP = np.array([[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]])


def ctrl(Pc, Pt):
    """Controlled-Pt with the control read in the Pc basis."""
    return np.kron((I2 + Pc) / 2, I2) + np.kron((I2 - Pc) / 2, Pt)


UNITARIES_1Q = {
    qcf.sHadamard: H, qcf.sPhase: S, qcf.sInvPhase: S.conj().T, qcf.sX: X, qcf.sY: Y,
    qcf.sZ: Z, qcf.sId1: I2, qcf.sSQRTX: SQX, qcf.sInvSQRTX: SQX.conj().T, qcf.sSQRTY: SQY,
    qcf.sInvSQRTY: SQY.conj().T, qcf.sHadamardXY: (X + Y) / np.sqrt(2),
    qcf.sHadamardYZ: (Y + Z) / np.sqrt(2), qcf.sCXYZ: CXYZ, qcf.sCZYX: CXYZ.conj().T,
}
UNITARIES_2Q = {
    qcf.sCNOT: CNOT, qcf.sZCX: CNOT, qcf.sCPHASE: CZ, qcf.sZCZ: CZ, qcf.sSWAP: SWAP,
    qcf.sISWAP: ISWAP, qcf.sInvISWAP: ISWAP.conj().T, qcf.sZCY: ctrl(Z, Y),
    qcf.sXCX: ctrl(X, X), qcf.sXCY: ctrl(X, Y), qcf.sXCZ: ctrl(X, Z), qcf.sYCX: ctrl(Y, X),
    qcf.sYCY: ctrl(Y, Y), qcf.sYCZ: ctrl(Y, Z),
    qcf.sSQRTZZ: sla.expm(1j * np.pi / 4 * np.kron(Z, Z)),
    qcf.sInvSQRTZZ: sla.expm(-1j * np.pi / 4 * np.kron(Z, Z)),
    qcf.sSQRTXX: sla.expm(1j * np.pi / 4 * np.kron(X, X)),
    qcf.sInvSQRTXX: sla.expm(-1j * np.pi / 4 * np.kron(X, X)),
    qcf.sSQRTYY: sla.expm(1j * np.pi / 4 * np.kron(Y, Y)),
    qcf.sInvSQRTYY: sla.expm(-1j * np.pi / 4 * np.kron(Y, Y)),
    qcf.sSWAPCX: CNOT @ SWAP, qcf.sInvSWAPCX: (CNOT @ SWAP).conj().T,
    qcf.sCXSWAP: SWAP @ CNOT, qcf.sCZSWAP: SWAP @ CZ,
}


def decode(m, n):
    for bits in itertools.product((0, 1), repeat=2 * n):
        xs, zs = bits[:n], bits[n:]
        b = np.array([[1]], dtype=complex)
        for x, z in zip(xs, zs):
            b = np.kron(b, SIG[(x, z)])
        c = np.trace(b.conj().T @ m) / 2**n
        if abs(abs(c) - 1) < 1e-9 and np.allclose(m, c * b):
            return xs, zs
    raise ValueError("not a Pauli")


def run_gate_on_all_paulis(op, n):
    """Apply `op` with the oracle to 4^n frames holding every n-qubit Pauli."""
    combos = list(itertools.product((0, 1), repeat=2 * n))
    shots = len(combos)
    nq = n + 1
    bits = np.zeros((2 * nq, shots), np.uint8)
    for s, c in enumerate(combos):
        for q in range(n):
            bits[q, s] = c[q]            # x of qubit q+1
            bits[nq + q, s] = c[n + q]   # z of qubit q+1
    xzs = oracle.frame_to_reference(oracle.pack_shots(bits), nq, shots)
    ops, args = qcf.lower_circuit([op], n)
    xzs, _ = oracle.pf_run(ops, args, nq, 1, shots, fresh=False, xzs=xzs)
    out = oracle.unpack_shots(oracle.reference_to_frame(xzs, nq, shots), shots)
    return combos, out


@pytest.mark.parametrize("gate", list(UNITARIES_1Q), ids=lambda g: g.__name__)
def test_single_qubit_bitmaps_match_unitaries(gate):
    U = UNITARIES_1Q[gate]
    combos, out = run_gate_on_all_paulis(gate(1), 1)
    for s, (x, z) in enumerate(combos):
        xs, zs = decode(U @ SIG[(x, z)] @ U.conj().T, 1)
        assert (out[0, s], out[2, s]) == (xs[0], zs[0]), (gate.__name__, x, z)


@pytest.mark.parametrize("gate", list(UNITARIES_2Q), ids=lambda g: g.__name__)
def test_two_qubit_bitmaps_match_unitaries(gate):
    U = UNITARIES_2Q[gate]
    combos, out = run_gate_on_all_paulis(gate(1, 2), 2)
    for s, c in enumerate(combos):
        P = np.kron(SIG[(c[0], c[2])], SIG[(c[1], c[3])])
        xs, zs = decode(U @ P @ U.conj().T, 2)
        got = (out[0, s], out[1, s], out[3, s], out[4, s])
        assert got == (xs[0], xs[1], zs[0], zs[1]), (gate.__name__, c)


@pytest.mark.parametrize("gate", qcf.ops.TWO_QUBIT_GATES, ids=lambda g: g.__name__)
def test_all_two_qubit_maps_are_symplectic(gate):
    """Covers sZCrY / sInvZCrY too, whose unitaries are not written out here:
    the image of every pair of Paulis must keep their commutation."""
    combos, out = run_gate_on_all_paulis(gate(1, 2), 2)
    img = [(out[0, s], out[1, s], out[3, s], out[4, s]) for s in range(len(combos))]

    def comm(a, b):  # a = (x1, x2, z1, z2) ordering of combos: (x1,x2,z1,z2)
        return (a[0] * b[2] + a[2] * b[0] + a[1] * b[3] + a[3] * b[1]) % 2

    for i in range(len(combos)):
        for j in range(len(combos)):
            ci = (combos[i][0], combos[i][1], combos[i][2], combos[i][3])
            cj = (combos[j][0], combos[j][1], combos[j][2], combos[j][3])
            assert comm(ci, cj) == comm(img[i], img[j])


def _frame_run(circ, nqubits, nbits, shots, masks=None, fresh=True, xzs=None, meas=None):
    ops, args = qcf.lower_circuit(circ, nqubits)
    return oracle.pf_run(ops, args, nqubits + 1, nbits, shots, masks=masks, fresh=fresh,
                         xzs=xzs, meas=meas)


def test_measure_semantics():
    """sMZ(q,0) records nothing; sMRZ(q,0) still resets (src/pauli_frames.jl:85-116);
    measurements are assignments and ClassicalXOR overwrites its store (:63-72)."""
    shots = 64
    nq = 3
    rng = np.random.default_rng(0)
    bits = rng.integers(0, 2, (2 * nq, shots)).astype(np.uint8)
    xzs = oracle.frame_to_reference(oracle.pack_shots(bits), nq, shots)
    meas = rng.integers(0, 2, (4, shots)).astype(np.uint8)
    circ = [qcf.sMZ(1, 0), qcf.sMZ(1, 1), qcf.sMRZ(2, 0), qcf.sMZ(2, 2),
            qcf.ClassicalXOR((1, 2, 3), 3)]
    nmasks = nq + 1
    masks = oracle.pack_shots(np.zeros((nmasks, shots), np.uint8))
    x2, m2 = _frame_run(circ, 2, 4, shots, masks=masks, fresh=False, xzs=xzs,
                        meas=meas.copy())
    assert np.array_equal(m2[0], bits[0])            # x of qubit 1
    assert not m2[1].any()                            # qubit 2 was reset
    assert np.array_equal(m2[2], bits[0] ^ 0 ^ meas[2])  # XOR of bits 1,2,3 (old bit 3)
    assert np.array_equal(m2[3], meas[3])             # untouched bit


def test_ghz_invariants():
    """test/test_pauliframe.jl:24-43: GHZ measurements have row sums divisible by
    the qubit count and about half ones."""
    n, shots = 3, 20000
    circ = [qcf.sHadamard(1), qcf.sCNOT(1, 2), qcf.sCNOT(1, 3)] + [qcf.sMZ(i, i) for i in (1, 2, 3)]
    rng = np.random.default_rng(1)
    ops, args = qcf.lower_circuit(circ, n)
    prog_masks = rng.integers(0, 2, (n + 1, shots)).astype(np.uint8)
    _, m = oracle.pf_run(ops, args, n + 1, 3, shots, masks=oracle.pack_shots(prog_masks))
    s = m.sum(axis=0)
    assert np.all(s % 3 == 0)
    assert 0.48 < m.mean() < 0.52


@pytest.mark.parametrize("d", [3, 5])
def test_noiseless_surface_code_detectors_are_zero(d):
    c = circuits.rotated_surface_code(d, d, p=0.0, noise=False)
    shots = 256
    ops, args = qcf.lower_circuit(c.ops, c.nqubits)
    prog = qcf.Program(c.ops)
    rng = np.random.default_rng(d)
    masks = oracle.pack_shots(rng.integers(0, 2, (prog.nreplay, shots)).astype(np.uint8))
    _, m = oracle.pf_run(ops, args, c.nqubits + 1, c.nbits, shots, masks=masks)
    det = m[np.array(c.det_bits) - 1]
    assert not det.any()
    assert not m[np.array(c.obs_bits) - 1].any()
    # X-type first-round measurements are random (gauge), Z-type are not
    assert m[: c.nmeas].any()


def test_surface_code_counts():
    c = circuits.rotated_surface_code(5, 25, 1e-3)
    assert c.nqubits == 49 and c.nmeas == 625 and len(c.detectors) == 600
    assert len(c.observables) == 1
    p = qcf.Program(c.ops)
    assert p.nq == 50 and p.nbits == 1226


def test_repetition_code_counts():
    c = circuits.repetition_code(3, 3, 1e-3)
    assert c.nmeas == 9 and len(c.detectors) == 8 and c.nqubits == 5


def test_single_fault_flips_expected_detectors():
    """An X error on a data qubit before the final measurement flips exactly the
    Z-checks it touches and the observable iff it is on the logical line."""
    d = 3
    base = circuits.rotated_surface_code(d, 2, p=0.0, noise=False)
    data, xm, zm, idx = circuits.rotated_surface_code_layout(d)
    shots = 1
    ops, args = qcf.lower_circuit(base.ops, base.nqubits)
    prog = qcf.Program(base.ops)
    masks = np.zeros((prog.nreplay, 64), np.uint64)
    _, m0 = oracle.pf_run(ops, args, base.nqubits + 1, base.nbits, shots,
                          masks=masks.copy())
    assert not m0[np.array(base.det_bits) - 1].any()
    # inject X on data qubit (3,3) through the in-place frame right before the final M
    k = max(i for i, op in enumerate(base.ops) if isinstance(op, qcf.sMRZ))
    pre, post = base.ops[: k + 1], base.ops[k + 1:]
    xzs, m1 = oracle.pf_run(*qcf.lower_circuit(pre, base.nqubits), base.nqubits + 1,
                            base.nbits, shots, masks=masks.copy())
    q = idx[(3, 3)]
    xzs[(q - 1) // 64, 0] ^= np.uint64(1 << ((q - 1) % 64))
    ops2, args2 = qcf.lower_circuit(post, base.nqubits)
    _, m2 = oracle.pf_run(ops2, args2, base.nqubits + 1, base.nbits, shots,
                          masks=np.zeros((1, 64), np.uint64), fresh=False, xzs=xzs, meas=m1)
    flipped = m2[np.array(base.det_bits) - 1, 0]
    touching = [m for m in zm if (3, 3) in [(m[0] + a, m[1] + b) for a, b in
                                           ((1, 1), (-1, 1), (1, -1), (-1, -1))]]
    assert flipped.sum() == len(touching) == 2
