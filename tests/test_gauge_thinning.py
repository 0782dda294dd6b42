"""The gauge draws a fresh, record-only launch leaves out
(qcf_program_gauge_planes), checked against the oracle.

Replaying one gauge draw alone through the oracle (every other draw zero)
gives the record bits that draw reaches: one column of G, where the record of
a shot is R = G g + N over GF(2), g the uniform gauge bits and N the noise
part.  The draws the library keeps must be linearly independent and span
every column of G.  Then G g is uniform on the same space with or without the
others, so the distribution of every record is unchanged; the GPU tests check
the records of such launches bit for bit against the oracle replaying the
masks the interpreter exports, and their marginals against the exact ones."""
import numpy as np
import pytest

import qcf
from faulthelp import inject, program_faults
from qcf import circuits
from test_gpu_frames import random_circuit


def gf2_rank(cols):
    """Rank over GF(2) of the columns of a (rows, m) 0/1 matrix."""
    basis = {}
    rank = 0
    for j in range(cols.shape[1]):
        v = int.from_bytes(np.packbits(cols[:, j]).tobytes(), "big")
        while v:
            top = v.bit_length() - 1
            if top not in basis:
                basis[top] = v
                rank += 1
                break
            v ^= basis[top]
    return rank


def check(prog):
    _, gauge = program_faults(prog)
    gp = prog.gauge_planes()
    assert sorted(np.flatnonzero(gp).tolist()) == sorted(gauge)  # the gauge planes, only
    cols = inject(prog, [[g] for g in gauge]).astype(np.uint8)  # (nbits, draws)
    kept = [i for i, g in enumerate(gauge) if gp[g] == 1]
    r_all = gf2_rank(cols)
    assert gf2_rank(cols[:, kept]) == len(kept) == r_all
    return len(gauge), len(kept)


def test_rotated_d5_25_rounds_needs_24_draws():
    """d = 5, 25 rounds: 699 gauge draws per shot (50 initZ!, 49 initial
    resets, 600 ancilla resets), of which 24 span what they do to the record
    -- 12 X-check patterns of the data's initZ! z bits, which every round's
    X-checks see, and 12 from the X-check qubits' own z bits, which reach the
    final data measurement -- all in the prologue, so the rounds draw none."""
    prog = qcf.Program(circuits.rotated_surface_code(5, 25, 1e-3).ops)
    assert check(prog) == (699, 24)
    gp = prog.gauge_planes()
    assert (gp[prog.nq:] != 1).all()  # no draw after the initZ! words


@pytest.mark.parametrize("d,rounds", [(3, 3), (3, 1), (7, 3)])
def test_rotated_codes(d, rounds):
    ndraws, nkept = check(qcf.Program(circuits.rotated_surface_code(d, rounds, 2e-3).ops))
    assert 0 < nkept < ndraws


def test_repetition_and_benchmark_circuits():
    check(qcf.Program(circuits.repetition_code(5, 4, 1e-2).ops))
    check(qcf.Program(circuits.unrotated_surface_code_benchmark(3, 2)))


@pytest.mark.parametrize("seed", range(6))
def test_random_circuits(seed):
    """Every gate, sMZ/sMRZ/sMX/sMRX, PauliMeasurement, classical XOR and
    noise kind, bits overwritten and read back."""
    rng = np.random.default_rng(9000 + seed)
    n, nbits = int(rng.integers(2, 8)), int(rng.integers(1, 10))
    prog = qcf.Program(random_circuit(rng, n, nbits, 120), nqubits=n, nbits=nbits)
    check(prog)


def test_no_record_needs_no_draw():
    prog = qcf.Program([qcf.sMRZ(1, 0), qcf.sHadamard(1), qcf.sCNOT(1, 2), qcf.sMRZ(2, 0)],
                       nqubits=2, nbits=1)
    gp = prog.gauge_planes()
    assert (gp[gp > 0] == 2).all() and (gp > 0).sum() == 5  # 3 initZ! + 2 resets
