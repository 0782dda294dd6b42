"""Frame sampler on the GPU against the oracle.

Bit-exact: noiseless circuits from the same initial frames (the pattern of
test/test_gpu.jl:70-90), and noisy circuits under replayed masks -- both
random masks and the masks the native RNG itself drew (exported).
Statistical: closed-form rates (test/test_gpu.jl:116-134)."""
import numpy as np
import pytest
import torch

import oracle
import qcf
from qcf import circuits

pytestmark = pytest.mark.gpu

SHOT_COUNTS = [1, 31, 32, 33, 100, 2047, 2048, 2049, 4096 + 5]


def padded(planes, ld):
    """(nplanes, words) uint64 -> flat int64 tensor of nplanes planes of `ld` words."""
    out = np.zeros((planes.shape[0], ld), np.uint64)
    out[:, : planes.shape[1]] = planes
    return torch.from_numpy(out.view(np.int64).reshape(-1))


def gpu_inplace(circ, nqubits, nbits, shots, frame_bits, meas_bits, masks_bits=None):
    """Run pftrajectories(state, circuit) on the GPU from given frame bits."""
    f = qcf.PauliFrame(shots, nqubits, nbits, init=False)
    f.frame.copy_(padded(oracle.pack_shots(frame_bits), f.ld))
    f.record.copy_(padded(oracle.pack_shots(meas_bits), f.ld))
    masks = None
    if masks_bits is not None:
        masks = padded(oracle.pack_shots(masks_bits), f.ld).cuda()
    qcf.pftrajectories_state(f, circ, masks=masks)
    return (oracle.unpack_shots(f.frame_host(), shots),
            oracle.unpack_shots(f.record_host(), shots))


def oracle_inplace(circ, nqubits, nbits, shots, frame_bits, meas_bits, masks_bits=None):
    nq = nqubits + 1
    ops, args = qcf.lower_circuit(circ, nqubits)
    xzs = oracle.frame_to_reference(oracle.pack_shots(frame_bits), nq, shots)
    masks = None if masks_bits is None else oracle.pack_shots(masks_bits)
    xzs, meas = oracle.pf_run(ops, args, nq, nbits, shots, masks=masks, fresh=False, xzs=xzs,
                              meas=meas_bits.copy())
    return oracle.unpack_shots(oracle.reference_to_frame(xzs, nq, shots), shots), meas


def test_device_philox_matches_known_answers(gpu):
    recs = np.array([[0, 0, 0, 0, 0, 0], [0xFFFFFFFF] * 6,
                     [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344, 0xA4093822, 0x299F31D0]],
                    np.uint32)
    rng = np.random.default_rng(0)
    recs = np.concatenate([recs, rng.integers(0, 2**32, (100, 6), dtype=np.uint32)])
    inp = torch.from_numpy(recs.view(np.int32).reshape(-1)).cuda()
    out = torch.zeros(4 * len(recs), dtype=torch.int32, device="cuda")
    qcf._lib.call("qcf_debug_philox", qcf.pauli._ptr(inp), qcf.pauli._ptr(out), len(recs),
                  qcf.pauli._stream())
    got = out.cpu().numpy().view(np.uint32).reshape(-1, 4)
    assert got[0].tolist() == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert got[1].tolist() == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert got[2].tolist() == [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]
    for r, g in zip(recs, got):
        assert g.tolist() == oracle.philox4x32_10(r[:4], r[4:]).tolist()


GATES_1Q = qcf.ops.SINGLE_QUBIT_GATES
GATES_2Q = qcf.ops.TWO_QUBIT_GATES


@pytest.mark.parametrize("shots", [100, 2049])
def test_every_gate_from_random_frames(gpu, shots):
    n = 4
    rng = np.random.default_rng(shots)
    circ = []
    for g in GATES_1Q:
        circ.append(g(int(rng.integers(1, n + 1))))
    for g in GATES_2Q:
        a, b = rng.choice(np.arange(1, n + 1), 2, replace=False)
        circ.append(g(int(a), int(b)))
    circ.append(qcf.SingleQubitOperator(2, True, True, False, True))
    circ.append(qcf.SingleQubitOperator(3, False, True, True, False))
    circ.append(qcf.SingleQubitOperator(1, True, True, True, True))  # singular map
    for q in range(1, n + 1):
        circ.append(qcf.sMZ(q, q))
    fb = rng.integers(0, 2, (2 * (n + 1), shots)).astype(np.uint8)
    mb = rng.integers(0, 2, (n, shots)).astype(np.uint8)
    gf, gm = gpu_inplace(circ, n, n, shots, fb, mb)
    of, om = oracle_inplace(circ, n, n, shots, fb, mb)
    assert np.array_equal(gf, of)
    assert np.array_equal(gm, om)


def random_pauli_measurement(rng, n, nbits):
    """PauliMeasurement of a random Pauli on qubits 1..m; the reference uses
    qubit m+1 as its ancilla (src/pauli_frames.jl:118-134), the frame's hidden
    qubit when m = n and a circuit qubit otherwise."""
    m = n if rng.random() < 0.7 else int(rng.integers(1, n + 1))
    letters = "".join("_XZY"[int(c)] for c in rng.integers(0, 4, m))
    sign = ("", "-", "i", "-i")[int(rng.integers(4))]
    return qcf.PauliMeasurement(qcf.P(sign + letters, device="cpu"), int(rng.integers(0, nbits + 1)))


def random_circuit(rng, n, nbits, length):
    circ = []
    written = []
    for _ in range(length):
        k = rng.integers(0, 11)
        q = int(rng.integers(1, n + 1))
        if k < 3:
            g = GATES_1Q[rng.integers(len(GATES_1Q))]
            circ.append(g(q))
        elif k < 6:
            a, b = rng.choice(np.arange(1, n + 1), 2, replace=False)
            g = GATES_2Q[rng.integers(len(GATES_2Q))]
            circ.append(g(int(a), int(b)))
        elif k == 6:
            b = int(rng.integers(0, nbits + 1))
            circ.append((qcf.sMZ, qcf.sMRZ, qcf.sMX, qcf.sMRX)[rng.integers(4)](q, b))
            if b:
                written.append(b)
        elif k == 7:
            circ.append(qcf.PauliError(q, float(rng.uniform(0, 0.3))))
        elif k == 8:
            if rng.random() < 0.5:
                circ.append(qcf.PauliError(q, *rng.uniform(0, 0.1, 3)))
            else:
                a, b = rng.choice(np.arange(1, n + 1), 2, replace=False)
                circ.append(qcf.Depolarize2(int(a), int(b), float(rng.uniform(0, 0.3))))
        elif k == 9:
            ar = int(rng.integers(1, 5))
            bits = [int(x) for x in rng.integers(1, nbits + 1, ar)]
            circ.append(qcf.ClassicalXOR(bits, int(rng.integers(1, nbits + 1))))
        else:
            circ.append(random_pauli_measurement(rng, n, nbits))
    return circ


@pytest.mark.parametrize("shots", SHOT_COUNTS)
def test_random_circuits_replayed_masks(gpu, shots):
    """Noisy random circuits: the same pre-drawn masks replayed on both sides."""
    rng = np.random.default_rng(77 + shots)
    for trial in range(3):
        n, nbits = int(rng.integers(2, 9)), int(rng.integers(1, 12))
        circ = random_circuit(rng, n, nbits, 120)
        prog = qcf.Program(circ, nqubits=n, nbits=nbits)
        fb = rng.integers(0, 2, (2 * (n + 1), shots)).astype(np.uint8)
        mb = rng.integers(0, 2, (nbits, shots)).astype(np.uint8)
        masks = rng.integers(0, 2, (prog.nreplay, shots)).astype(np.uint8)
        gf, gm = gpu_inplace(circ, n, nbits, shots, fb, mb, masks)
        of, om = oracle_inplace(circ, n, nbits, shots, fb, mb, masks)
        assert np.array_equal(gf, of), trial
        assert np.array_equal(gm, om), trial


def fresh_export_and_replay(circ, shots, seed):
    """Native RNG on the GPU with every draw exported, then the oracle replays
    the exported masks from a fresh PauliFrame: records must agree bit for bit."""
    prog = qcf.Program(circ)
    masks = qcf.alloc_masks(prog, shots)
    f = qcf.pftrajectories(prog, shots, seed=seed, export_masks=masks)
    assert f.ld == qcf.frames.plane_words(shots) and f.ld % 32 == 0
    mk = masks.cpu().numpy().view(np.uint64).reshape(-1, f.ld)[: prog.nreplay]
    ops, args = prog.ops, prog.args
    xzs, om = oracle.pf_run(ops, args, prog.nq, prog.nbits, shots, masks=mk)
    gm = oracle.unpack_shots(f.record_host(), shots)
    gf = oracle.unpack_shots(f.frame_host(), shots)
    of = oracle.unpack_shots(oracle.reference_to_frame(xzs, prog.nq, shots), shots)
    return prog, f, gm, om, gf, of, mk


def test_repetition_code_native_seed0_bit_exact(gpu):
    """Distance-3 repetition memory, 10^4 shots at seed 0: detector, observable and
    raw record bits equal the oracle replaying the masks the GPU drew."""
    c = circuits.repetition_code(3, 3, 1e-3)
    prog, f, gm, om, gf, of, mk = fresh_export_and_replay(c.ops, 10**4, seed=0)
    assert np.array_equal(gm, om[: prog.nbits])
    assert np.array_equal(gf, of)
    assert mk.any()  # noise and gauge draws happened


@pytest.mark.parametrize("d,rounds,shots", [(3, 3, 5000), (5, 25, 4096 + 77), (7, 7, 3000),
                                           (9, 3, 2100)])
def test_surface_code_native_export_replay(gpu, d, rounds, shots):
    c = circuits.rotated_surface_code(d, rounds, 1e-2)
    prog, f, gm, om, gf, of, mk = fresh_export_and_replay(c.ops, shots, seed=1234 + d)
    assert np.array_equal(gm, om[: prog.nbits])
    assert np.array_equal(gf, of)
    # some detectors fire at p = 1e-2
    assert gm[np.array(c.det_bits) - 1].any()


def test_noisy_gate_and_noise_op_all_export_replay(gpu):
    """NoisyGate (the gate, then its noise on the gate's qubits) and NoiseOpAll
    (the noise on every frame qubit), src/noise.jl:174-199, with every noise
    kind: native draws exported and replayed through the oracle."""
    ops = qcf.ops
    circ = [qcf.sHadamard(1),
            ops.NoisyGate(qcf.sCNOT(1, 2), ops.UnbiasedUncorrelatedNoise(0.05)),
            ops.NoisyGate(qcf.sCNOT(2, 3), ops.PauliNoise(0.01, 0.02, 0.03)),
            ops.NoiseOpAll(ops.UnbiasedUncorrelatedNoise(0.02)),
            ops.NoisyGate(qcf.sSWAP(1, 3), ops.DepolarizationNoise(0.1)),
            ops.NoisyGate(qcf.sHadamard(2), ops.DepolarizationNoise(0.08)),
            qcf.sMZ(1, 1), qcf.sMZ(2, 2), qcf.sMZ(3, 3),
            ops.NoiseOpAll(ops.PauliNoise(0.1, 0.0, 0.0)),
            qcf.sMRZ(1, 4), qcf.sMRZ(2, 5), qcf.sMZ(3, 6), qcf.ClassicalXOR([1, 4], 7)]
    prog, f, gm, om, gf, of, mk = fresh_export_and_replay(circ, 6000, seed=31)
    assert np.array_equal(gm, om[: prog.nbits])
    assert np.array_equal(gf, of)
    assert gm[:6].any() and mk[prog.nq:].any()


@pytest.mark.parametrize("shots", [333, 4096 + 9])
def test_classical_xor_of_many_bits_export_replay(gpu, shots):
    """ClassicalXOR of 17 bits (the arity the reference's compactification no
    longer covers, test/test_sumtypecompactification.jl:2), of one bit, and of
    33 bits with repeats, over random gauge outcomes and noise."""
    circ = [qcf.sHadamard(q) for q in range(1, 21)]
    circ += [qcf.PauliError(q, 0.1) for q in range(1, 21)]
    circ += [qcf.sMZ(q, q) for q in range(1, 21)]
    circ += [qcf.ClassicalXOR(list(range(1, 18)), 21), qcf.ClassicalXOR([3], 22),
             qcf.ClassicalXOR(list(range(1, 21)) + list(range(1, 14)), 23),
             qcf.ClassicalXOR([21, 22, 23], 24)]
    prog, f, gm, om, gf, of, mk = fresh_export_and_replay(circ, shots, seed=17)
    assert np.array_equal(gm, om[: prog.nbits])
    want17 = np.bitwise_xor.reduce(gm[:17], axis=0)
    assert np.array_equal(gm[20], want17) and 0.3 < gm[20].mean() < 0.7


def test_mixed_rates_thinning_export_replay(gpu):
    """Sites of different rates share one event stream thinned per site."""
    circ = []
    for r in range(6):
        for q in range(1, 6):
            circ.append(qcf.PauliError(q, 0.02 * (q / 5)))
            circ.append(qcf.PauliError(q, 0.001, 0.003, 0.01))
        circ.append(qcf.Depolarize2(1, 2, 0.05))
        circ.append(qcf.sCNOT(1, 3))
        for q in range(1, 6):
            circ.append(qcf.sMZ(q, 5 * r + q))
    prog, f, gm, om, gf, of, mk = fresh_export_and_replay(circ, 9000, seed=42)
    assert np.array_equal(gm, om[: prog.nbits])
    assert np.array_equal(gf, of)


@pytest.mark.parametrize("shots", [1, 100, 2047, 2049, 5000])
def test_padding_inside_stride_is_zero(gpu, shots):
    """A record buffer prefilled with ones: every word past the shots, up to
    the padded stride, comes back zero; the shots themselves match a
    zero-initialised run."""
    import ctypes as C
    c = circuits.rotated_surface_code(3, 3, 5e-3)
    prog = qcf.Program(c.ops)
    ld = qcf.frames.plane_words(shots)
    rec = torch.full((prog.nbits * ld,), -1, dtype=torch.int64, device="cuda")
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    qcf._lib.call("qcf_pf_sample", prog.handle, shots, 0, 77, None, C.c_void_p(rec.data_ptr()),
                  None, ld, 1, st)
    host = rec.cpu().numpy().view(np.uint64).reshape(prog.nbits, ld)
    bits = oracle.unpack_shots(host, ld * 64)
    assert not bits[:, shots:].any()
    want = qcf.pftrajectories(prog, shots, seed=77, keep_frame=False).record_host()
    assert np.array_equal(host, want)


def test_noiseless_surface_code_native(gpu):
    c = circuits.rotated_surface_code(5, 25, 0.0, noise=False)
    f = qcf.pftrajectories(c.ops, 200_000, seed=3)
    rec = f.record_host()
    assert not rec[np.array(c.det_bits) - 1].any()
    assert not rec[np.array(c.obs_bits) - 1].any()
    # first-round X-check outcomes are uniformly random (gauge bits)
    first = oracle.unpack_shots(rec[:12], 200_000)
    assert abs(first.mean() - 0.5) < 0.01


def test_same_seed_same_bits_different_seed_differs(gpu):
    c = circuits.rotated_surface_code(3, 3, 5e-3)
    prog = qcf.Program(c.ops)
    a = qcf.pftrajectories(prog, 50_000, seed=11).record_host()
    b = qcf.pftrajectories(prog, 50_000, seed=11).record_host()
    d = qcf.pftrajectories(prog, 50_000, seed=12).record_host()
    assert np.array_equal(a, b)
    assert not np.array_equal(a, d)


def test_noise_rates_closed_form(gpu):
    """test/test_gpu.jl:116-134: rates [0, .5, 2p/3, .5]."""
    p = 0.1
    shots = 10**5
    circ = [qcf.sMZ(1, 1), qcf.sHadamard(2), qcf.sMZ(2, 2),
            qcf.NoiseOp(qcf.UnbiasedUncorrelatedNoise(p), [3, 4]), qcf.sMZ(3, 3),
            qcf.sHadamard(4), qcf.sMZ(4, 4)]
    f = qcf.PauliFrame(shots, 4, 4)
    qcf.pftrajectories_state(f, circ)
    m = qcf.measurements(f).float().mean(dim=0).cpu().numpy()
    assert np.allclose(m, [0, 0.5, 2 * p / 3, 0.5], atol=0.01)


def test_measurements_matrix_layout(gpu):
    c = circuits.repetition_code(3, 3, 0.05)
    f = qcf.pftrajectories(c.ops, 3000, seed=5)
    m = qcf.measurements(f)
    assert m.shape == (3000, c.nbits) and m.dtype == torch.bool
    bits = oracle.unpack_shots(f.record_host(), 3000)
    assert np.array_equal(m.cpu().numpy().T.astype(np.uint8), bits)


def test_absolute_measurements_from_register_bits(gpu):
    """measurements(register, frame) = bitview(register) .xor measurements(frame)
    (src/pauli_frames.jl:289): every column flips where the reference bit is 1."""
    c = circuits.repetition_code(3, 3, 0.05)
    f = qcf.pftrajectories(c.ops, 2049, seed=8)
    ref = np.random.default_rng(3).integers(0, 2, c.nbits).astype(bool)
    rel = qcf.measurements(f).cpu().numpy()
    ab = qcf.measurements(ref, f).cpu().numpy()
    assert ab.shape == rel.shape and ab.dtype == bool
    assert np.array_equal(ab, rel ^ ref[None, :])
    with pytest.raises(qcf.DimensionMismatch):
        qcf.measurements(ref[:-1], f)


def test_apply_and_applynoise_one_op_at_a_time(gpu):
    """apply!/applynoise! on a PauliFrame op by op give the same frame and record
    as the whole circuit through pftrajectories(state, circuit) (src/pauli_frames.jl:
    58-174, 234-239), replaying the same masks: the noise ops draw nothing else."""
    ops = [qcf.sHadamard(1), qcf.sCNOT(1, 2), qcf.sMZ(2, 1), qcf.sPhase(2), qcf.sMRZ(1, 2),
           qcf.sSWAP(1, 3), qcf.sMZ(3, 3)]
    shots = 3000
    a = qcf.PauliFrame(shots, 3, 3, init=False)
    b = a.copy()
    a.frame.random_(-2**62, 2**62)
    b.frame.copy_(a.frame)
    # whole circuit, with the one sMRZ z draw supplied as a replay mask
    masks = torch.randint(-2**62, 2**62, (qcf.frames.plane_words(shots),), dtype=torch.int64,
                          device=a.frame.device)
    qcf.pftrajectories(a, ops, masks=masks)
    for op in ops:
        if isinstance(op, qcf.sMRZ):
            qcf.pftrajectories_state(b, [op], masks=masks)
        else:
            qcf.apply_(b, op)
    assert np.array_equal(a.frame_host(), b.frame_host())
    assert np.array_equal(a.record_host(), b.record_host())
    # applynoise!: X flips with p = 1 flip every x bit of the qubit
    c = a.copy()
    qcf.applynoise_(c, qcf.ops.PauliNoise(1.0, 0.0, 0.0), 2)
    fx, cx = a.frame_host(), c.frame_host()
    bits = oracle.unpack_shots((fx ^ cx), shots)
    assert bits[1].all() and not np.delete(bits, 1, axis=0).any()


def test_repetition_code_syndrome_implications(gpu):
    """test/test_pauliframe.jl:2-22: encode |111> with sX and CNOTs, X error on
    qubit 2 with p = 0.75, syndrome CNOTs and sMZ.  Wherever the frame has X on
    qubit 2 (and no Z), qubits 4 and 5 carry X and both syndrome bits are set;
    the state form pftrajectories(frame, circuit) and the fresh form agree on
    that rule."""
    circuit = [qcf.sX(1), qcf.sCNOT(1, 2), qcf.sCNOT(1, 3), qcf.PauliError(2, 0.75),
               qcf.sCNOT(1, 4), qcf.sCNOT(2, 4), qcf.sCNOT(2, 5), qcf.sCNOT(3, 5),
               qcf.sMZ(4, 1), qcf.sMZ(5, 2)]
    shots = 4096 + 11
    frames = [qcf.pftrajectories(qcf.PauliFrame(shots, 5, 2), circuit, seed=1),
              qcf.pftrajectories(circuit, trajectories=shots, seed=2)]
    for f in frames:
        fx = oracle.unpack_shots(f.frame_host(), shots)  # planes: x of 1..6, z of 1..6
        m = qcf.measurements(f).cpu().numpy()
        nq = f.nq
        sel = (fx[1] == 1) & (fx[nq + 1] == 0)
        # X or Y on qubit 2 with probability 1/2 (PauliError(2, 0.75)), and its z
        # bit is a uniformly random gauge bit (initZ! propagated by the CNOTs):
        # X-without-Z in a quarter of the frames (4 sigma at 4107 shots: 0.027)
        assert abs(sel.mean() - 0.25) < 0.027
        assert fx[3][sel].all() and fx[4][sel].all()
        assert m[sel, 0].all() and m[sel, 1].all()


def test_pauli_measurement_correlations(gpu):
    """test/test_pauliframe.jl:92-115: H on three qubits, PauliMeasurement ZZ_
    and _ZZ, then sMZ of each qubit; after decoding with the two parity
    outcomes the three single-qubit outcomes agree in every frame."""
    circuit = [qcf.sHadamard(1), qcf.sHadamard(2), qcf.sHadamard(3),
               qcf.PauliMeasurement(qcf.P("ZZ_", device="cpu"), 1),
               qcf.PauliMeasurement(qcf.P("_ZZ", device="cpu"), 2),
               qcf.sMZ(1, 3), qcf.sMZ(2, 4), qcf.sMZ(3, 5)]
    n = 3000
    for m in (qcf.measurements(qcf.pftrajectories(qcf.PauliFrame(n, 3, 5), circuit, seed=3)),
              qcf.measurements(qcf.pftrajectories(circuit, trajectories=n, seed=4))):
        m = m.cpu().numpy().copy()
        m[:, 3] ^= m[:, 0]
        m[:, 4] ^= m[:, 0] ^ m[:, 1]
        assert (m[:, 2] == m[:, 3]).all() and (m[:, 3] == m[:, 4]).all()
        assert 0.4 < m[:, 2].mean() < 0.6  # the outcomes themselves are random


def test_reference_frame_layout_roundtrip(gpu):
    c = circuits.rotated_surface_code(3, 2, 0.01)
    f = qcf.pftrajectories(c.ops, 777, seed=9)
    xzs = qcf.frame_to_reference(f)
    want = oracle.frame_to_reference(f.frame_host(), f.nq, 777)
    assert np.array_equal(xzs.cpu().numpy().view(np.uint64), want)
    g = qcf.PauliFrame(777, f.nq - 1, f.nbits, init=False)
    qcf.frame_from_reference(xzs, g)
    assert np.array_equal(g.frame_host(), f.frame_host())


@pytest.mark.parametrize("batch_jit", ["1", "0"])
def test_batched_equals_single_launches(gpu, monkeypatch, batch_jit):
    """Batched records equal single launches, with the circuits' specialised
    kernels on helper streams (the default) and with every circuit on the
    persistent interpreter kernel (QCF_BATCH_JIT=0)."""
    monkeypatch.setenv("QCF_BATCH_JIT", batch_jit)
    rng = np.random.default_rng(4)
    progs, shots, seeds = [], [], []
    for k in range(12):
        d = int(rng.choice([3, 5, 7]))
        c = circuits.rotated_surface_code(d, d, float(10 ** rng.uniform(-4, -2)))
        progs.append(qcf.Program(c.ops))
        shots.append(int(10 ** rng.uniform(3, 4.5)))
        seeds.append(int(rng.integers(0, 2**62)))
    from qcf import batch
    outs = batch.sample_batched(progs, shots, seeds)
    for k, (p, s, sd, o) in enumerate(zip(progs, shots, seeds, outs)):
        single = qcf.pftrajectories(p, s, seed=sd, keep_frame=False).record_host()
        diff = np.argwhere(o != single)
        assert len(diff) == 0, (k, p.nq, p.info.lds_bytes, s, len(diff), diff[:5].tolist())


def test_batched_mixes_zero_rate_and_noisy_circuits(gpu):
    """A batch holding the same code at p = 0 (noise ops kept at rate 0: its
    kernel drops their event passes, and record-only launches of it take the
    linear kernel) and at p > 0, noiseless circuits without noise ops, and a
    d=13 circuit (split kernel, 24 noise slots): every record equals its
    single launch."""
    cs = [circuits.rotated_surface_code(5, 5, 0.0), circuits.rotated_surface_code(5, 5, 2e-3),
          circuits.rotated_surface_code(3, 3, 0.0, noise=False),
          circuits.rotated_surface_code(5, 5, 0.0), circuits.rotated_surface_code(13, 13, 1e-3),
          circuits.rotated_surface_code(7, 7, 0.0)]
    progs = [qcf.Program(c.ops) for c in cs]
    shots = [3000, 5000, 2049, 777, 4200, 2048]
    seeds = [11, 12, 13, 14, 15, 16]
    from qcf import batch
    outs = batch.sample_batched(progs, shots, seeds, graph=True)
    for k, (p, s, sd, o) in enumerate(zip(progs, shots, seeds, outs)):
        single = qcf.pftrajectories(p, s, seed=sd, keep_frame=False).record_host()
        assert np.array_equal(o, single), k


def test_global_frame_storage_matches_lds(gpu, monkeypatch):
    """Frames too large for LDS run from global memory; forcing that path on
    circuits that fit must give the same bits."""
    c = circuits.rotated_surface_code(5, 5, 3e-3)
    prog = qcf.Program(c.ops)
    a = qcf.pftrajectories(prog, 20000, seed=99)
    monkeypatch.setenv("QCF_FORCE_GLOBAL_FRAMES", "1")
    b = qcf.pftrajectories(prog, 20000, seed=99)
    torch.cuda.synchronize()
    assert np.array_equal(a.record_host(), b.record_host())
    assert np.array_equal(a.frame_host(), b.frame_host())


def test_large_circuit_global_frames_export_replay(gpu):
    """d=13 needs more than the 160 KB of LDS per tile: global-memory planes."""
    c = circuits.rotated_surface_code(13, 2, 5e-3)
    prog = qcf.Program(c.ops)
    assert prog.info.lds_bytes > 160 * 1024
    prog, f, gm, om, gf, of, mk = fresh_export_and_replay(c.ops, 2100, seed=5)
    assert np.array_equal(gm, om[: prog.nbits])
    assert np.array_equal(gf, of)


@pytest.mark.parametrize("batch_jit", ["1", "0"])
def test_batched_with_large_circuits(gpu, monkeypatch, batch_jit):
    monkeypatch.setenv("QCF_BATCH_JIT", batch_jit)
    rng = np.random.default_rng(8)
    progs, shots, seeds = [], [], []
    for d in (3, 15, 5, 13, 11, 3):
        c = circuits.rotated_surface_code(d, d, float(10 ** rng.uniform(-4, -2)))
        progs.append(qcf.Program(c.ops))
        shots.append(int(10 ** rng.uniform(3, 4)))
        seeds.append(int(rng.integers(0, 2**62)))
    from qcf import batch
    outs = batch.sample_batched(progs, shots, seeds, graph=True)
    for k, (p, s, sd, o) in enumerate(zip(progs, shots, seeds, outs)):
        single = qcf.pftrajectories(p, s, seed=sd, keep_frame=False).record_host()
        assert np.array_equal(o, single), k


@pytest.mark.parametrize("steal", [0, 1, 0xFFFFFFFF])
def test_shared_tiles_over_repeated_launches(gpu, steal):
    """Tiles of the specialised launches whose frames exceed LDS taken by both
    those launches and the persistent kernel from one counter per launch (d =
    15 with steal_planes 0, d = 13 and 15 with 1), or by the launches alone
    (0xFFFFFFFF), equal single launches; the counters reset themselves, so
    every further launch and graph replay (records refilled with ones in
    between) writes them all again."""
    rng = np.random.default_rng(31)
    progs, shots, seeds = [], [], []
    for d in (15, 13, 5, 13, 15, 3):
        c = circuits.rotated_surface_code(d, d, float(10 ** rng.uniform(-4, -2)))
        progs.append(qcf.Program(c.ops))
        shots.append(int(10 ** rng.uniform(4, 4.7)))
        seeds.append(int(rng.integers(0, 2**62)))
    from qcf import batch
    singles = [qcf.pftrajectories(p, s, seed=sd, keep_frame=False).record
               for p, s, sd in zip(progs, shots, seeds)]
    t = qcf._lib.Tune()
    t.steal_planes = steal
    b = batch.Batch(progs, shots, seeds, t)
    for run, go in enumerate((b.launch, b.launch, b.launch_graph, b.launch_graph, b.launch)):
        for r in b.records:
            r.fill_(-1)
        go()
        torch.cuda.synchronize()
        for k, (p, s) in enumerate(zip(progs, shots)):
            n = p.nbits * qcf.frames.plane_words(s)
            assert torch.equal(b.records[k][:n], singles[k][:n]), (run, k, p.nq, s)


def test_reference_converters_on_a_view_of_a_larger_frame(gpu):
    """The converters address the reference matrices with a column stride, so a
    view of rows b:b+n of a P-shot frame (src/pauli_frames.jl:17) converts in
    place: the view's columns hold the frame, every other element is left
    alone, and converting back gives the frame again."""
    import ctypes as C
    c = circuits.rotated_surface_code(3, 3, 2e-2)
    prog = qcf.Program(c.ops)
    n, P, b = 3000, 5000, 700
    f = qcf.pftrajectories(prog, n, seed=8)
    W = (f.nq + 63) // 64
    st = qcf.frames._stream()
    parent = torch.full((2 * W, P), -1, dtype=torch.int64, device="cuda")
    qcf._lib.call("qcf_frame_to_reference", qcf.frames._ptr(f.frame), f.ld, n, f.nq,
                  C.c_void_p(parent.data_ptr() + 8 * b), P, st)
    want = qcf.frame_to_reference(f)
    assert torch.equal(parent[:, b:b + n], want)
    assert (parent[:, :b] == -1).all() and (parent[:, b + n:] == -1).all()
    g = qcf.PauliFrame(n, f.nq - 1, f.nbits, init=False)
    qcf._lib.call("qcf_frame_from_reference", C.c_void_p(parent.data_ptr() + 8 * b), P, n,
                  f.nq, qcf.frames._ptr(g.frame), g.ld, st)
    assert np.array_equal(g.frame_host(), f.frame_host())
    bools = torch.full((f.nbits, P), 7, dtype=torch.uint8, device="cuda")
    qcf._lib.call("qcf_record_to_bool", qcf.frames._ptr(f.record), f.ld, n, f.nbits,
                  C.c_void_p(bools.data_ptr() + b), P, st)
    assert torch.equal(bools[:, b:b + n].t().bool(), qcf.measurements(f))
    assert (bools[:, :b] == 7).all() and (bools[:, b + n:] == 7).all()
    rec = torch.zeros_like(f.record)
    qcf._lib.call("qcf_record_from_bool", C.c_void_p(bools.data_ptr() + b), P, n, f.nbits,
                  qcf.frames._ptr(rec), f.ld, st)
    nw = f.nbits * f.ld
    assert torch.equal(rec[:nw].view(f.nbits, f.ld)[:, : (n + 63) // 64],
                       f.record[:nw].view(f.nbits, f.ld)[:, : (n + 63) // 64])
    with pytest.raises(qcf.DimensionMismatch):
        qcf._lib.call("qcf_frame_to_reference", qcf.frames._ptr(f.frame), f.ld, n, f.nq,
                      C.c_void_p(parent.data_ptr()), n - 1, st)


def test_launch_bound_is_checked_on_the_host(gpu):
    """Every launch first checks its block size against the kernel's
    __launch_bounds__ on the host; the check itself, run against the
    persistent batch kernel (bound 64) without launching it."""
    qcf._lib.call("qcf_debug_check_launch", 64)
    with pytest.raises(qcf.ArgumentError, match="pf_batch_kernel"):
        qcf._lib.call("qcf_debug_check_launch", 128)


def test_wave_scan_is_exact(gpu):
    """The DPP prefix sum that places noise events equals numpy's cumsum."""
    rng = np.random.default_rng(1)
    x = rng.integers(0, 1 << 24, (97, 64), dtype=np.uint32)
    x[0] = 1
    x[1] = np.arange(64)
    inp = torch.from_numpy(x.view(np.int32).reshape(-1)).cuda()
    out = torch.zeros_like(inp)
    qcf._lib.call("qcf_debug_wave_scan", qcf.pauli._ptr(inp), qcf.pauli._ptr(out), 97,
                  qcf.pauli._stream())
    got = out.cpu().numpy().view(np.uint32).reshape(97, 64)
    assert np.array_equal(got, np.cumsum(x.astype(np.uint64), axis=1).astype(np.uint32))


@pytest.mark.parametrize("shots", [1, 2049, 5000])
def test_fresh_launch_defines_every_word(gpu, shots):
    """pftrajectories allocates its frame and record without filling them:
    the launch itself writes every word (record columns no operation writes
    are zeroed, bits past the last shot are zero), so memory left full of ones
    by an earlier tensor gives the same result as zeroed buffers."""
    rng = np.random.default_rng(shots)
    n, nbits = 6, 12
    circ = random_circuit(rng, n, 8, 100)  # bits 9..12 are never written
    prog = qcf.Program(circ, nqubits=n, nbits=nbits)
    ld = qcf.frames.plane_words(shots)
    junk = torch.full((4 * (2 * (n + 1) + nbits) * ld,), -1, dtype=torch.int64, device="cuda")
    del junk  # the caching allocator hands this block out again
    f = qcf.pftrajectories(prog, shots, seed=12)
    zf = torch.zeros(2 * (n + 1) * ld, dtype=torch.int64, device="cuda")
    zr = torch.zeros(nbits * ld, dtype=torch.int64, device="cuda")
    qcf._lib.call("qcf_pf_sample", prog.handle, shots, 0, 12, qcf.frames._ptr(zf),
                  qcf.frames._ptr(zr), None, ld, 1, qcf.frames._stream())
    torch.cuda.synchronize()
    assert torch.equal(f.record, zr) and torch.equal(f.frame, zf)
    rec = f.record_host()
    assert not rec[8:].any()
    bits = np.unpackbits(rec.view(np.uint8), axis=1, bitorder="little")
    assert not bits[:, shots:].any()


def test_global_frames_on_two_streams_equal_sequential_runs(gpu, monkeypatch):
    """Frames in global memory use a workspace per stream: two replay launches
    queued on different streams at once give the records of the same launches
    run one after the other."""
    monkeypatch.setenv("QCF_FORCE_GLOBAL_FRAMES", "1")
    rng = np.random.default_rng(21)
    circ = random_circuit(rng, 6, 8, 150)
    prog = qcf.Program(circ, nqubits=6, nbits=8)
    shots = 3 * 2048 + 5
    runs = []
    for k in range(2):
        f = qcf.PauliFrame(shots, 6, 8, init=False)
        f.frame.random_(-2**62, 2**62)
        m = torch.randint(-2**62, 2**62, (max(1, prog.nreplay) * f.ld,), dtype=torch.int64,
                          device="cuda")
        runs.append((f, m))
    want = []
    for f, m in runs:
        g = f.copy()
        qcf.pftrajectories_state(g, prog, masks=m)
        want.append((g.frame_host(), g.record_host()))
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    for s in streams:
        s.wait_stream(torch.cuda.current_stream())
    for (f, m), s in zip(runs, streams):
        with torch.cuda.stream(s):
            qcf.pftrajectories_state(f, prog, masks=m)
    torch.cuda.synchronize()
    for (f, _), (wf, wr) in zip(runs, want):
        assert np.array_equal(f.frame_host(), wf) and np.array_equal(f.record_host(), wr)
