"""GPU parity on the reference's own benchmark circuits: the 441-qubit
unrotated surface code of benchmark/surface-code-circuit.jl (rebuilt by
circuits.unrotated_surface_code_benchmark) and the x_diag circuits of
benchmark/benchmarks.jl:103-110, 161-171."""
import numpy as np
import pytest

import oracle
import qcf
from qcf import circuits
from surfacehelp import check_noiseless_invariants

pytestmark = pytest.mark.gpu


def export_replay(ops, shots, seed):
    """Native draws on the GPU, exported and replayed through the oracle from a
    fresh frame: record and final frame must agree bit for bit."""
    prog = qcf.Program(ops)
    masks = qcf.alloc_masks(prog, shots)
    f = qcf.pftrajectories(prog, shots, seed=seed, export_masks=masks)
    mk = masks.cpu().numpy().view(np.uint64).reshape(-1, f.ld)[: prog.nreplay]
    xzs, om = oracle.pf_run(prog.ops, prog.args, prog.nq, prog.nbits, shots, masks=mk)
    gm = oracle.unpack_shots(f.record_host(), shots)
    gf = oracle.unpack_shots(f.frame_host(), shots)
    of = oracle.unpack_shots(oracle.reference_to_frame(xzs, prog.nq, shots), shots)
    return prog, gm, om, gf, of


@pytest.mark.parametrize("shots", [2048 + 37])
def test_reference_surface_code_circuit_export_replay(gpu, shots):
    ops = circuits.unrotated_surface_code_benchmark(11, 12)
    print(f"unrotated d=11 x 12 rounds, {shots} shots", flush=True)
    prog, gm, om, gf, of = export_replay(ops, shots, seed=19)
    assert prog.nq == 442 and prog.nbits == 2751
    assert np.array_equal(gm, om[: prog.nbits])
    assert np.array_equal(gf, of)


def test_reference_surface_code_circuit_invariants_2pow19(gpu):
    """The largest trajectory count of benchmark-surface-code-gpu.jl (2^19):
    stabilizer invariants on the packed record (no Z check ever flips, X checks
    repeat, final data parity) and fair round-1 X checks."""
    shots = 2**19
    ops = circuits.unrotated_surface_code_benchmark(11, 12)
    print(f"unrotated d=11 x 12 rounds, {shots} shots, invariants", flush=True)
    f = qcf.pftrajectories(ops, shots, seed=3, keep_frame=False)
    rec = f.record_host()  # (nbits, ld) packed words; padding words are zero
    first = check_noiseless_invariants(rec, ops, 11)
    ones = np.unpackbits(np.ascontiguousarray(first).view(np.uint8)).sum()
    frac = ones / (first.shape[0] * shots)
    assert abs(frac - 0.5) < 0.002, frac


@pytest.mark.parametrize("csize,shots", [(100, 1), (1000, 100), (1000, 1000)])
def test_x_diag_circuit(gpu, csize, shots):
    """benchmarks.jl q101_r1 / q1001_r100 / q1001_r10000 shapes."""
    print(f"x_diag_circuit({csize}), {shots} shots", flush=True)
    prog, gm, om, gf, of = export_replay(circuits.x_diag_circuit(csize), shots, seed=4)
    assert np.array_equal(gm, om[: prog.nbits])
    assert np.array_equal(gf, of)


def test_x_diag_circuit_noisy_measurement(gpu):
    """benchmarks.jl compactification case: PauliFrame(1000, 1001, 1001)."""
    print("x_diag_circuit_noisy_measurement(1000), 1000 shots", flush=True)
    ops = circuits.x_diag_circuit_noisy_measurement(1000)
    prog, gm, om, gf, of = export_replay(ops, 1000, seed=5)
    assert np.array_equal(gm, om[: prog.nbits])
    assert np.array_equal(gf, of)
    assert gm.any()
