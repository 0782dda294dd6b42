"""Per-circuit specialised kernels (compiled at run time) against the
interpreter: for native launches both draw the same events, so records and
frames must agree bit for bit -- including the partial last tile, where lanes
past the last shot store nothing and bits past it are stored as zeros.
The interpreter itself is pinned to the oracle by the export/replay tests."""
import numpy as np
import pytest
import torch

import oracle
import qcf
from qcf import circuits
from test_gpu_frames import random_circuit

pytestmark = pytest.mark.gpu


def _both(monkeypatch, run):
    monkeypatch.delenv("QCF_NO_JIT", raising=False)
    monkeypatch.setenv("QCF_REQUIRE_JIT", "1")
    a = run()
    torch.cuda.synchronize()
    monkeypatch.delenv("QCF_REQUIRE_JIT")
    monkeypatch.setenv("QCF_NO_JIT", "1")
    b = run()
    torch.cuda.synchronize()
    monkeypatch.delenv("QCF_NO_JIT")
    return a, b


@pytest.mark.parametrize("d,rounds,p,shots", [(3, 3, 1e-2, 3 * 2048 + 77),
                                               (5, 25, 1e-3, 40 * 2048 + 5),
                                               (5, 5, 3e-2, 4096)])
def test_surface_code_specialised_equals_interpreter(gpu, monkeypatch, d, rounds, p, shots):
    prog = qcf.Program(circuits.rotated_surface_code(d, rounds, p).ops)

    def run():
        f = qcf.pftrajectories(prog, shots, seed=1000 + d)
        return f.record_host(), f.frame_host()

    (ra, fa), (rb, fb) = _both(monkeypatch, run)
    assert np.array_equal(ra, rb)
    assert np.array_equal(fa, fb)
    assert oracle.unpack_shots(ra, shots).any()


@pytest.mark.parametrize("shots", [1, 100, 2047, 5 * 2048 + 3])
def test_record_only_launch_specialised_equals_interpreter(gpu, monkeypatch, shots):
    prog = qcf.Program(circuits.repetition_code(5, 5, 2e-2).ops)

    def run():
        return qcf.pftrajectories(prog, shots, seed=4, keep_frame=False).record_host()

    a, b = _both(monkeypatch, run)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("shots", [2048, 2 * 2048 + 1])
def test_random_circuits_specialised_equals_interpreter(gpu, monkeypatch, shots):
    """Every gate, measurement, reset, classical XOR and noise kind."""
    rng = np.random.default_rng(500 + shots)
    for trial in range(4):
        n, nbits = int(rng.integers(2, 9)), int(rng.integers(1, 12))
        circ = random_circuit(rng, n, nbits, 150)
        prog = qcf.Program(circ, nqubits=n, nbits=nbits)

        def run():
            f = qcf.pftrajectories(prog, shots, seed=trial)
            return f.record_host(), f.frame_host()

        (ra, fa), (rb, fb) = _both(monkeypatch, run)
        assert np.array_equal(ra, rb), trial
        assert np.array_equal(fa, fb), trial


def test_in_place_specialised_equals_interpreter(gpu, monkeypatch):
    """pftrajectories(state, circuit) on existing frames (the in-place variant)."""
    rng = np.random.default_rng(3)
    n, nbits, shots = 6, 10, 3 * 2048 + 9
    circ = random_circuit(rng, n, nbits, 120)
    prog = qcf.Program(circ, nqubits=n, nbits=nbits)
    fb = oracle.pack_shots(rng.integers(0, 2, (2 * (n + 1), shots)).astype(np.uint8))
    mb = oracle.pack_shots(rng.integers(0, 2, (nbits, shots)).astype(np.uint8))

    def run():
        f = qcf.PauliFrame(shots, n, nbits, init=False)
        f.frame.view(2 * (n + 1), f.ld)[:, : fb.shape[1]].copy_(torch.from_numpy(fb.view(np.int64)))
        f.record.view(nbits, f.ld)[:, : mb.shape[1]].copy_(torch.from_numpy(mb.view(np.int64)))
        qcf.pftrajectories_state(f, prog, seed=77)
        return f.record_host(), f.frame_host()

    (ra, fa), (rb, fb2) = _both(monkeypatch, run)
    assert np.array_equal(ra, rb)
    assert np.array_equal(fa, fb2)


@pytest.mark.parametrize("d,rounds,force", [(7, 3, True), (11, 2, False)])
def test_global_frames_specialised_equals_interpreter(gpu, monkeypatch, d, rounds, force):
    """Frames too large for LDS (or forced out of it) keep their planes in
    global memory in the specialised kernel too (QCF_JIT_GLOBAL=1 opts in)."""
    monkeypatch.setenv("QCF_JIT_GLOBAL", "1")
    if force:
        monkeypatch.setenv("QCF_FORCE_GLOBAL_FRAMES", "1")
    prog = qcf.Program(circuits.rotated_surface_code(d, rounds, 5e-3).ops)
    if not force:
        assert prog.info.lds_bytes > 160 * 1024

    def run():
        f = qcf.pftrajectories(prog, 3 * 2048 + 5, seed=21)
        return f.record_host(), f.frame_host()

    (ra, fa), (rb, fb) = _both(monkeypatch, run)
    assert np.array_equal(ra, rb)
    assert np.array_equal(fa, fb)


@pytest.mark.parametrize("case", ["unrotated_11x3", "rotated_13x2", "rotated_13x13", "rotated_11x3",
                                  "rotated_9x3", "rotated_15x2", "x_diag_1000", "x_diag_noisy_1000"])
def test_split_frames_specialised_equals_interpreter(gpu, monkeypatch, case):
    """Frames too large for LDS run the split kernel (most-used planes in
    registers, the rest in LDS, noise on register planes through slots), with
    half of a CU's LDS per wave where the registers allow (rotated d=11, and
    d=9, whose frame fits but would run one wave per CU): bit for bit the
    interpreter's result, partial tile and kept frame included.  Rotated d=15
    (1237 planes) is beyond split mode and runs half-tile passes;
    x_diag_circuit(1000) of benchmark/benchmarks.jl (1,001 qubits) and its
    noisy-measurement variant (1000 noise sites in one run) quarter tiles.
    Launches of fewer tiles than CUs spread the parts of each tile over
    workgroups."""
    if case == "x_diag_1000":
        ops, shots = circuits.x_diag_circuit(1000), 1000
    elif case == "x_diag_noisy_1000":
        ops, shots = circuits.x_diag_circuit_noisy_measurement(1000), 1000
    elif case == "rotated_15x2":
        ops, shots = circuits.rotated_surface_code(15, 2, 5e-3).ops, 3 * 2048 + 1000
    elif case == "unrotated_11x3":
        ops, shots = circuits.unrotated_surface_code_benchmark(11, 3), 3 * 2048 + 5
    elif case == "rotated_13x2":
        ops, shots = circuits.rotated_surface_code(13, 2, 5e-3).ops, 2 * 2048 + 77
    elif case == "rotated_13x13":  # 929 planes: the split kernel with 24 noise slots
        ops, shots = circuits.rotated_surface_code(13, 13, 2e-3).ops, 2048 + 77
    elif case == "rotated_11x3":
        ops, shots = circuits.rotated_surface_code(11, 3, 5e-3).ops, 3 * 2048 + 100
    else:
        ops, shots = circuits.rotated_surface_code(9, 3, 1e-2).ops, 4 * 2048 + 1
    prog = qcf.Program(ops)
    if case != "rotated_9x3":
        assert prog.info.lds_bytes > 160 * 1024
    print(f"split kernel {case}, {shots} shots", flush=True)

    def run():
        f = qcf.pftrajectories(prog, shots, seed=31)
        return f.record_host(), f.frame_host()

    (ra, fa), (rb, fb) = _both(monkeypatch, run)
    assert np.array_equal(ra, rb)
    assert np.array_equal(fa, fb)
    assert oracle.unpack_shots(ra, shots).any()


@pytest.mark.parametrize("case", ["unrotated_11x2", "rotated_15x1"])
def test_split_frames_in_place_equals_interpreter(gpu, monkeypatch, case):
    """pftrajectories(state, circuit) on an existing 442-qubit frame (split
    kernel) and 450-qubit frame (half-tile passes load and store the frame
    one half of each tile at a time)."""
    if case == "unrotated_11x2":
        ops = circuits.unrotated_surface_code_benchmark(11, 2)
    else:
        ops = circuits.rotated_surface_code(15, 1, 5e-3).ops
    prog = qcf.Program(ops)
    n, nbits, shots = prog.nq - 1, prog.nbits, 2048 + 9
    rng = np.random.default_rng(11)
    fb = oracle.pack_shots(rng.integers(0, 2, (2 * (n + 1), shots)).astype(np.uint8))
    mb = oracle.pack_shots(rng.integers(0, 2, (nbits, shots)).astype(np.uint8))
    print(f"specialised kernel in place, {case}", flush=True)

    def run():
        f = qcf.PauliFrame(shots, n, nbits, init=False)
        f.frame.view(2 * (n + 1), f.ld)[:, : fb.shape[1]].copy_(torch.from_numpy(fb.view(np.int64)))
        f.record.view(nbits, f.ld)[:, : mb.shape[1]].copy_(torch.from_numpy(mb.view(np.int64)))
        qcf.pftrajectories_state(f, prog, seed=78)
        return f.record_host(), f.frame_host()

    (ra, fa), (rb, fb2) = _both(monkeypatch, run)
    assert np.array_equal(ra, rb)
    assert np.array_equal(fa, fb2)


@pytest.mark.parametrize("shots", [333, 3 * 2048 + 9])
def test_many_bit_xor_and_pauli_measurement_specialised_equals_interpreter(gpu, monkeypatch,
                                                                          shots):
    """ClassicalXOR of 17 and 33 bits, PauliMeasurement (its ancilla the hidden
    qubit) and noise after every gate."""
    circ = [qcf.sHadamard(q) for q in range(1, 21)]
    circ += [qcf.ops.NoisyGate(qcf.sCNOT(q, q + 1), qcf.ops.UnbiasedUncorrelatedNoise(0.05))
             for q in range(1, 20)]
    circ += [qcf.PauliMeasurement(qcf.P("X" * 10 + "Z" * 5 + "Y" * 5, device="cpu"), 25)]
    circ += [qcf.sMZ(q, q) for q in range(1, 21)]
    circ += [qcf.ClassicalXOR(list(range(1, 18)), 21),
             qcf.ClassicalXOR(list(range(1, 21)) + list(range(1, 14)), 22)]
    prog = qcf.Program(circ)

    def run():
        f = qcf.pftrajectories(prog, shots, seed=88)
        return f.record_host(), f.frame_host()

    (ra, fa), (rb, fb) = _both(monkeypatch, run)
    assert np.array_equal(ra, rb) and np.array_equal(fa, fb)
    assert oracle.unpack_shots(ra, shots)[20].any()


def _three(monkeypatch, run):
    """Linear kernel (the default for noiseless record-only launches), frame
    kernel (QCF_JIT_LINEAR=0) and interpreter, in that order."""
    monkeypatch.setenv("QCF_REQUIRE_JIT", "1")
    monkeypatch.delenv("QCF_JIT_LINEAR", raising=False)
    lin = run()
    monkeypatch.setenv("QCF_JIT_LINEAR", "0")
    frame = run()
    monkeypatch.delenv("QCF_JIT_LINEAR")
    monkeypatch.delenv("QCF_REQUIRE_JIT")
    monkeypatch.setenv("QCF_NO_JIT", "1")
    interp = run()
    monkeypatch.delenv("QCF_NO_JIT")
    torch.cuda.synchronize()
    return lin, frame, interp


@pytest.mark.parametrize("shots", [1, 2047, 3 * 2048 + 77])
def test_linear_kernel_on_the_reference_benchmark_circuit(gpu, monkeypatch, shots):
    """The noiseless d=11 benchmark circuit (benchmark/surface-code-circuit.jl),
    record only: the linear kernel, the split frame kernel and the interpreter
    give the same bits, and the stabilizer checks of every round after the
    first are deterministic (the record's gauge part repeats)."""
    prog = qcf.Program(circuits.unrotated_surface_code_benchmark(11, 12))

    def run():
        return qcf.pftrajectories(prog, shots, seed=123, keep_frame=False).record_host()

    lin, frame, interp = _three(monkeypatch, run)
    assert np.array_equal(lin, frame)
    assert np.array_equal(lin, interp)
    assert oracle.unpack_shots(lin, shots).any()  # gauge-random outcomes exist


@pytest.mark.parametrize("shots", [100, 2049])
def test_zero_rate_noise_runs_as_noiseless(gpu, monkeypatch, shots):
    """Rotated d=5 x 5 with every noise op at rate 0: the linear kernel, the
    frame kernel (no event passes compiled in) and the interpreter give the
    same records; with the frame returned, the frame kernel equals the
    interpreter on both record and frame."""
    prog = qcf.Program(circuits.rotated_surface_code(5, 5, 0.0).ops)

    def run():
        return qcf.pftrajectories(prog, shots, seed=77, keep_frame=False).record_host()

    lin, frame, interp = _three(monkeypatch, run)
    assert np.array_equal(lin, frame)
    assert np.array_equal(lin, interp)

    def run_frame():
        r = qcf.pftrajectories(prog, shots, seed=78)
        return r.record_host(), r.frame_host()

    monkeypatch.setenv("QCF_REQUIRE_JIT", "1")
    jr, jf = run_frame()
    monkeypatch.delenv("QCF_REQUIRE_JIT")
    monkeypatch.setenv("QCF_NO_JIT", "1")
    ir, iff = run_frame()
    monkeypatch.delenv("QCF_NO_JIT")
    assert np.array_equal(jr, ir)
    assert np.array_equal(jf, iff)


@pytest.mark.parametrize("shots", [100, 2049])
def test_linear_kernel_on_random_noiseless_circuits(gpu, monkeypatch, shots):
    """Every gate, sMZ/sMRZ/sMX/sMRX, PauliMeasurement and classical XOR
    (noise ops removed), record only."""
    rng = np.random.default_rng(900 + shots)
    noise = (qcf.ops.NoiseOp, qcf.ops.NoiseOpAll, qcf.ops.NoisyGate)
    for trial in range(4):
        n, nbits = int(rng.integers(2, 9)), int(rng.integers(1, 12))
        circ = [op for op in random_circuit(rng, n, nbits, 150) if not isinstance(op, noise)]
        prog = qcf.Program(circ, nqubits=n, nbits=nbits)

        def run():
            return qcf.pftrajectories(prog, shots, seed=trial, keep_frame=False).record_host()

        lin, frame, interp = _three(monkeypatch, run)
        assert np.array_equal(lin, frame), trial
        assert np.array_equal(lin, interp), trial
