"""The oracle's own noise samplers (per-shot draws, as the reference's
applynoise! makes them) against the closed forms of their channels.

The frame starts at zero (pftrajectories(state, circuit) on a frame that was
never initZ!-ed), so after one noise operation the frame of every shot is
exactly the Pauli the channel drew:
* PauliError(q, p) = UnbiasedUncorrelatedNoise(p): X, Y, Z with p/3 each
  (src/noise.jl:8-13,70-72, src/pauli_frames.jl:136-155);
* PauliError(q, px, py, pz) = PauliNoise: X px, Y py, Z pz
  (src/noise.jl:56-72,161-163, src/pauli_frames.jl:157-174);
* the two-qubit depolarizing channel (no reference frame method): the 15
  non-identity two-qubit Paulis with e/15 each, the distribution of
  DepolarizationNoise(two_qubit_error_probability=e) (src/noise.jl:49-53,
  99-129).
Every category count is within 5 sigma of the binomial expectation."""
import numpy as np
import pytest

import oracle
import qcf

SHOTS = 10**7


def frame_paulis(circ, nqubits, shots, seed, nthreads=4):
    """Per-shot (x, z) bits of qubits 1..nqubits after `circ` from a zero frame."""
    ops, args = qcf.lower_circuit(circ, nqubits)
    nq = nqubits + 1
    W = (nq + 63) // 64
    xzs, _ = oracle.pf_run(ops, args, nq, 0, shots, None, fresh=False,
                           xzs=np.zeros((2 * W, shots), np.uint64), seed=seed,
                           nthreads=nthreads)
    x = [(xzs[(q - 1) // 64] >> np.uint64((q - 1) % 64)) & np.uint64(1) for q in range(1, nqubits + 1)]
    z = [(xzs[W + (q - 1) // 64] >> np.uint64((q - 1) % 64)) & np.uint64(1)
         for q in range(1, nqubits + 1)]
    return [a.astype(np.int64) for a in x], [b.astype(np.int64) for b in z]


def assert_counts(counts, probs, n):
    counts = np.asarray(counts, float)
    probs = np.asarray(probs, float)
    assert counts.sum() == n
    sd = np.sqrt(n * probs * (1 - probs))
    z = np.abs(counts - n * probs) / np.maximum(sd, 1e-12)
    assert z.max() < 5.0, (z.round(2).tolist(), (counts / n).tolist(), probs.tolist())


def pauli_code(x, z):
    """1 = X, 2 = Z, 3 = Y (the base-4 digit of src/noise.jl:112-118)."""
    return x + 2 * z


def test_unbiased_noise_closed_form():
    p = 0.3
    x, z = frame_paulis([qcf.PauliError(1, p)], 1, SHOTS, seed=11)
    c = np.bincount(pauli_code(x[0], z[0]), minlength=4)
    assert_counts(c, [1 - p, p / 3, p / 3, p / 3], SHOTS)


def test_pauli_noise_closed_form():
    px, py, pz = 0.05, 0.1, 0.15
    x, z = frame_paulis([qcf.PauliError(1, px, py, pz)], 1, SHOTS, seed=12)
    c = np.bincount(pauli_code(x[0], z[0]), minlength=4)
    assert_counts(c, [1 - px - py - pz, px, pz, py], SHOTS)


def test_two_qubit_depolarizing_closed_form():
    e = 0.3
    x, z = frame_paulis([qcf.Depolarize2(1, 2, e)], 2, SHOTS, seed=13)
    d1, d2 = pauli_code(x[0], z[0]), pauli_code(x[1], z[1])
    c = np.bincount(d1 + 4 * d2, minlength=16)
    assert_counts(c, [1 - e] + [e / 15] * 15, SHOTS)
    # what measurements of the x bits see: qubit 1 flips with 8e/15 (first
    # digit X or Y), both qubits with 4e/15
    n1 = int(x[0].sum())
    n12 = int((x[0] & x[1]).sum())
    assert_counts([n1, SHOTS - n1], [8 * e / 15, 1 - 8 * e / 15], SHOTS)
    assert_counts([n12, SHOTS - n12], [4 * e / 15, 1 - 4 * e / 15], SHOTS)


@pytest.mark.parametrize("p", [1e-3, 0.02])
def test_depolarizing_then_measure_rates(p):
    """test/test_gpu.jl:116-134 style: depolarize then sMZ flips with 2p/3."""
    ops, args = qcf.lower_circuit([qcf.PauliError(1, p), qcf.sMZ(1, 1)], 1)
    _, meas = oracle.pf_run(ops, args, 2, 1, SHOTS, None, fresh=True, seed=14, nthreads=4)
    n = int(meas[0].sum())
    assert_counts([n, SHOTS - n], [2 * p / 3, 1 - 2 * p / 3], SHOTS)


def test_exact_marginals_equal_brute_force_enumeration():
    """The marginal model of faulthelp (independent sites XORed, 1/2 wherever a
    gauge draw reaches) against the exact distribution of a tiny circuit:
    every combination of every site's choice (or none) and every gauge
    assignment, injected together into the oracle."""
    import itertools
    from faulthelp import exact_record_marginals, inject, program_faults
    circ = [qcf.sHadamard(1), qcf.Depolarize2(1, 2, 0.2), qcf.sCNOT(1, 2),
            qcf.PauliError(2, 0.02, 0.05, 0.1), qcf.sMRZ(2, 1), qcf.PauliError(1, 0.15),
            qcf.sMZ(1, 2), qcf.ClassicalXOR([1, 2], 3), qcf.sMZ(2, 4)]
    prog = qcf.Program(circ)
    faults, gauge = program_faults(prog)
    sites = sorted({s for s, _, _ in faults})
    choices = [[(None, 1 - sum(p for s2, _, p in faults if s2 == s))] +
               [(pl, p) for s2, pl, p in faults if s2 == s] for s in sites]
    combos, probs = [], []
    for pick in itertools.product(*choices):
        for g in itertools.product((0, 1), repeat=len(gauge)):
            planes = [x for pl, _ in pick if pl for x in pl] + [gp for gp, b in zip(gauge, g) if b]
            combos.append(planes)
            probs.append(np.prod([p for _, p in pick]) / 2 ** len(gauge))
    rec = inject(prog, combos)
    want = (rec * np.asarray(probs)[None, :]).sum(axis=1)
    assert np.allclose(exact_record_marginals(prog), want, rtol=1e-12, atol=1e-15)


def test_oracle_native_sampling_matches_exact_marginals():
    """The oracle's per-shot sampler on the rotated d=5 memory circuit
    (25 rounds, p = 2e-3): every record bit within 5 sigma of its exact
    marginal at 10^6 shots."""
    from faulthelp import exact_record_marginals
    from qcf import circuits
    c = circuits.rotated_surface_code(5, 25, 2e-3)
    prog = qcf.Program(c.ops)
    want = exact_record_marginals(prog)
    n = 10**6
    _, om = oracle.pf_run(prog.ops, prog.args, prog.nq, prog.nbits, n, None, True, seed=5,
                          nthreads=8)
    got = om[: prog.nbits].sum(axis=1) / n
    sd = np.sqrt(np.maximum(want * (1 - want), 1e-12) / n)
    z = np.abs(got - want) / sd
    dets = np.asarray(c.det_bits) - 1
    assert want[dets].min() > 1e-3  # every detector can fire
    assert z.max() < 5.0, (int(z.argmax()), float(z.max()))
