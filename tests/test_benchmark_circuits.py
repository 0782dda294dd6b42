"""The reference's own benchmark circuits: text ingestion and generators.

benchmark/surface-code-circuit.jl (the circuit of the reference's GPU frame
benchmark, benchmark/benchmark-surface-code-gpu.jl) is read with the Julia
text reader and compared op for op with the generator, when the reference
checkout is present; the generator's circuit is then checked against what
stabilizer algebra fixes for a noiseless surface-code memory (deterministic
checks flip never, round-1 X checks are fair coins) on the oracle.
"""
import os

import numpy as np
import pytest

import oracle
import qcf
from qcf import circuits, jltext
from surfacehelp import check_noiseless_invariants

REF_CIRCUIT = "/root/reference/benchmark/surface-code-circuit.jl"


def test_generator_equals_reference_file():
    if not os.path.exists(REF_CIRCUIT):
        pytest.skip("reference checkout not present")
    with open(REF_CIRCUIT) as fh:
        ref = jltext.parse_circuit(fh.read())
    gen = circuits.unrotated_surface_code_benchmark(11, 12)
    assert len(ref) == len(gen) == 17802
    assert [repr(o) for o in ref] == [repr(o) for o in gen]


def test_generator_counts():
    ops = circuits.unrotated_surface_code_benchmark(11, 12)
    kinds = {}
    for o in ops:
        kinds[type(o).__name__] = kinds.get(type(o).__name__, 0) + 1
    assert kinds == {"sCNOT": 9660, "sHadamard": 2640, "sMRZ": 2751, "sMZ": 2751}
    assert max(q for o in ops for q in o.affectedqubits()) == 441
    assert sorted(b for o in ops for b in o.affectedbits()) == list(range(1, 2752))


def test_parse_ops_values_and_forms():
    text = """
    # a comment
    Any[sCNOT(1, 2), sHadamard(3), PauliError((1,4), 1e-3), PauliError(5, 0.001, 0, 0),
        ClassicalXOR(1:3, 4), ClassicalXOR([2, 5], 6), sMRZ(4, 0), sMZ(2),
        PauliMeasurement(P"-XZ", 7), NoiseOpAll(UnbiasedUncorrelatedNoise(0.5)),
        NoisyGate(sSWAP(1, 2), PauliNoise(0.1, 0.2, 0.3)),
        SingleQubitOperator(2, true, false, false, true, false, false)]
    """
    ops = jltext.parse_ops(text)
    assert [type(o).__name__ for o in ops] == [
        "sCNOT", "sHadamard", "NoiseOp", "NoiseOp", "ClassicalXOR", "ClassicalXOR", "sMRZ",
        "sMZ", "PauliMeasurement", "NoiseOpAll", "NoisyGate", "SingleQubitOperator"]
    assert (ops[0].q1, ops[0].q2) == (1, 2)
    assert ops[2].indices == (1, 4) and ops[2].noise == qcf.ops.UnbiasedUncorrelatedNoise(1e-3)
    assert ops[3].noise == qcf.ops.PauliNoise(0.001, 0.0, 0.0)
    assert ops[4].bits == (1, 2, 3) and ops[4].store == 4
    assert ops[5].bits == (2, 5) and ops[5].store == 6
    assert ops[6].bit == 0 and ops[7].bit == 0
    assert repr(ops[8].pauli) == "- XZ" and ops[8].bit == 7
    assert ops[10].noise == qcf.ops.PauliNoise(0.1, 0.2, 0.3)
    assert ops[11].m == (True, False, False, True)


@pytest.mark.parametrize("bad", ["Any[run(`ls`)]", "Any[sCNOT(1, 2)", "Any[sCNOT(1; 2)]",
                                 "Any[foo(1)]", "Any[sHadamard(x)]", "Any[sMZ(1, 2) + 1]"])
def test_parse_rejects_anything_else(bad):
    with pytest.raises(qcf.ArgumentError):
        jltext.parse_ops(bad)


def test_parse_circuit_skips_preamble():
    text = "using QuantumClifford\n\n# QuantumClifford Program\ncirc = Any[\n    sMRZ(1, 0),\n" \
           "    sHadamard(2),\n]\n"
    ops = jltext.parse_circuit(text)
    assert [repr(o) for o in ops] == ["sMRZ(1,0)", "sHadamard(2)"]
    assert [repr(o) for o in jltext.parse_circuit(text, name="circ")] == [repr(o) for o in ops]
    with pytest.raises(qcf.ArgumentError):
        jltext.parse_circuit(text, name="other")


def test_x_diag_circuits_structure():
    """benchmark/benchmarks.jl:103-110 and :161-171."""
    c = circuits.x_diag_circuit(100)
    assert len(c) == 201 and repr(c[-1]) == "sMZ(101,1)"
    assert repr(c[0]) == "sHadamard(1)" and repr(c[1]) == "sCNOT(1,101)"
    n = circuits.x_diag_circuit_noisy_measurement(1000)
    assert len(n) == 5000
    assert n[4].bits == (1, 2, 3) and n[4].store == 1          # i=1: 1:(1%6+2)
    assert n[29].bits == tuple(range(1, 3)) and n[29].store == 6  # i=6: 1:2


def test_noiseless_benchmark_invariants_on_oracle():
    d, shots = 5, 3000
    ops = circuits.unrotated_surface_code_benchmark(d, 4)
    lops, args = qcf.lower_circuit(ops)
    nq = (2 * d - 1) ** 2 + 1  # pftrajectories sizes frames one qubit past the largest
    nbits = max(b for o in ops for b in o.affectedbits())
    _, rec = oracle.pf_run(lops, args, nq, nbits, shots, None, True, seed=5)
    first = check_noiseless_invariants(rec, ops, d)
    assert abs(first.mean() - 0.5) < 0.03
    # a planted fault (an X error on one data qubit before the last round's
    # checks, applied through the oracle) breaks the invariants
    bad = list(ops)
    at = max(i for i, o in enumerate(bad) if type(o).__name__ == "sHadamard") - 1
    bad.insert(at, qcf.PauliError(1, 1.0, 0.0, 0.0))
    lops, args = qcf.lower_circuit(bad)
    _, rec = oracle.pf_run(lops, args, nq, nbits, shots, None, True, seed=5)
    with pytest.raises(AssertionError):
        check_noiseless_invariants(rec, ops, d)
