"""The native RNG of the GPU sampler against closed forms and exact marginals,
and the default (specialised-kernel) path against the oracle directly.

* Closed forms at 10^7 shots on a zero initial frame (the frame after one
  noise operation is the Pauli drawn): UnbiasedUncorrelatedNoise, PauliNoise
  and the two-qubit depolarizing channel, every Pauli category within 5 sigma
  (src/noise.jl:8-13,56-72,99-129; src/pauli_frames.jl:136-174).  The device
  quantizes the single-qubit Pauli choice to 2^-15 and thinning to 2^-31 of
  the site rate; both are far below the 5 sigma bound at these counts
  (2^-15 of p = 0.3 is 9e-6, sigma of a 0.05 category is 6.9e-5).
* Exact marginals of every record bit (faulthelp.exact_record_marginals,
  pinned on the CPU by brute-force enumeration) against the default launch
  of the rotated d=5 x 25-round memory circuit at p = 1e-3, 10^7 shots, and
  of d=9 x 9 (registers plus LDS) at p = 3e-3.
* A record of the default launch equals the oracle replaying the draws the
  interpreter exports for the same seed (the specialised kernels cannot
  export, so this ties them to the oracle without the interpreter's output).
* Gauge draws (the z-flip of a reset, src/pauli_frames.jl:110-113) are
  uniform and pairwise independent across consecutive draws, adjacent shots
  of a word, adjacent words and adjacent tiles.
"""
import numpy as np
import pytest
import torch

import oracle
import qcf
from faulthelp import exact_record_marginals
from qcf import circuits

pytestmark = pytest.mark.gpu

SHOTS = 10**7
M1, M2 = 0x5555555555555555, 0x3333333333333333
M4, H01 = 0x0F0F0F0F0F0F0F0F, 0x0101010101010101


def popcount_planes(t: torch.Tensor, nplanes: int, ld: int) -> np.ndarray:
    """Ones per plane of a shot-packed int64 tensor (SWAR popcount on the GPU)."""
    t = t[: nplanes * ld].view(nplanes, ld)
    out = []
    for p0 in range(0, nplanes, 64):
        x = t[p0:p0 + 64]
        x = x - ((x >> 1) & M1)
        x = (x & M2) + ((x >> 2) & M2)
        x = (x + (x >> 4)) & M4
        x = (x * H01) >> 56
        out.append(x.sum(dim=1).cpu())
    return torch.cat(out).numpy().astype(np.float64)


def assert_within_5_sigma(counts, probs, n, what):
    probs = np.asarray(probs, float)
    sd = np.sqrt(np.maximum(n * probs * (1 - probs), 1e-12))
    z = np.abs(np.asarray(counts, float) - n * probs) / sd
    k = int(z.argmax())
    assert z.max() < 5.0, (what, k, float(z[k]), float(counts[k]) / n, float(probs[k]))


def zero_frame_paulis(op, nqubits, seed):
    f = qcf.PauliFrame(SHOTS, nqubits, 1, init=False)
    qcf.pftrajectories_state(f, [op], seed=seed)
    planes = f.frame_host()  # x of 1..nq, then z of 1..nq (nq includes the ancilla)
    x = [oracle.unpack_shots(planes[q:q + 1], SHOTS)[0].astype(np.int64) for q in range(nqubits)]
    z = [oracle.unpack_shots(planes[f.nq + q:f.nq + q + 1], SHOTS)[0].astype(np.int64)
         for q in range(nqubits)]
    return x, z


def test_unbiased_noise_closed_form(gpu):
    p = 0.3
    x, z = zero_frame_paulis(qcf.PauliError(1, p), 1, seed=21)
    c = np.bincount(x[0] + 2 * z[0], minlength=4)
    assert_within_5_sigma(c, [1 - p, p / 3, p / 3, p / 3], SHOTS, "I X Z Y")


def test_pauli_noise_closed_form(gpu):
    px, py, pz = 0.05, 0.1, 0.15
    x, z = zero_frame_paulis(qcf.PauliError(1, px, py, pz), 1, seed=22)
    c = np.bincount(x[0] + 2 * z[0], minlength=4)
    assert_within_5_sigma(c, [1 - px - py - pz, px, pz, py], SHOTS, "I X Z Y")


def test_two_qubit_depolarizing_closed_form(gpu):
    e = 0.3
    x, z = zero_frame_paulis(qcf.Depolarize2(1, 2, e), 2, seed=23)
    d1, d2 = x[0] + 2 * z[0], x[1] + 2 * z[1]
    c = np.bincount(d1 + 4 * d2, minlength=16)
    assert_within_5_sigma(c, [1 - e] + [e / 15] * 15, SHOTS, "16 two-qubit Paulis")
    n1, n12 = int(x[0].sum()), int((x[0] & x[1]).sum())
    assert_within_5_sigma([n1, n12], [8 * e / 15, 4 * e / 15], SHOTS, "x1, x1 and x2")


@pytest.mark.parametrize("d,rounds,p,shots", [(5, 25, 1e-3, SHOTS), (9, 9, 3e-3, 2 * 10**6)])
def test_default_launch_matches_exact_marginals(gpu, d, rounds, p, shots):
    """Every measurement, detector and observable bit of the default launch
    (the specialised kernel) within 5 sigma of its exact marginal."""
    c = circuits.rotated_surface_code(d, rounds, p)
    prog = qcf.Program(c.ops)
    want = exact_record_marginals(prog)
    f = qcf.pftrajectories(prog, shots, seed=2026 + d, keep_frame=False)
    counts = popcount_planes(f.record, prog.nbits, f.ld)
    dets = np.asarray(c.det_bits) - 1
    assert want[dets].min() > 1e-4
    assert_within_5_sigma(counts, want, shots, f"record bits of d={d}")


def test_default_launch_equals_oracle_replay(gpu):
    """d=5 x 25, p = 1e-3: the record of the default launch equals the oracle
    replaying the draws that the interpreter exports for the same seed."""
    c = circuits.rotated_surface_code(5, 25, 1e-3)
    prog = qcf.Program(c.ops)
    shots, seed = 65536 + 77, 404
    f = qcf.pftrajectories(prog, shots, seed=seed, keep_frame=False)
    masks = qcf.alloc_masks(prog, shots)
    qcf.pftrajectories(prog, shots, seed=seed, export_masks=masks, keep_frame=False)
    mk = masks.cpu().numpy().view(np.uint64).reshape(-1, f.ld)[: prog.nreplay]
    _, om = oracle.pf_run(prog.ops, prog.args, prog.nq, prog.nbits, shots, masks=mk, nthreads=8)
    got = oracle.unpack_shots(f.record_host(), shots)
    assert np.array_equal(got, om[: prog.nbits])
    assert got[np.asarray(c.det_bits) - 1].any()


def test_gauge_draws_uniform_and_pairwise_independent(gpu):
    """R resets of one qubit, each followed by an X measurement: after sMRZ the
    frame's z bit is the gauge draw, which sMX records, so the record holds R
    consecutive gauge draws per shot.  At 10^7 shots every draw is 1 with
    probability 1/2 and the correlation of any of the pairs below is 0, each
    within 5 sigma (sigma = 1/sqrt(pairs))."""
    R = 8
    circ = []
    for k in range(R):
        circ += [qcf.sMRZ(1, 0), qcf.sMX(1, k + 1)]
    prog = qcf.Program(circ)
    f = qcf.pftrajectories(prog, SHOTS, seed=77, keep_frame=False)
    counts = popcount_planes(f.record, R, f.ld)
    assert_within_5_sigma(counts, [0.5] * R, SHOTS, "gauge draws")
    bits = oracle.unpack_shots(f.record_host()[:R], SHOTS).astype(np.float64) * 2 - 1
    pairs = {"consecutive draws": (bits[:-1], bits[1:]),
             "adjacent shots": (bits[:, :-1], bits[:, 1:]),
             "adjacent words (shot s, s + 32)": (bits[:, :-32], bits[:, 32:]),
             "adjacent tiles (shot s, s + 2048)": (bits[:, :-2048], bits[:, 2048:])}
    for what, (a, b) in pairs.items():
        c = (a * b).mean(axis=1)
        n = a.shape[1]
        assert np.abs(c).max() * np.sqrt(n) < 5.0, (what, float(np.abs(c).max() * np.sqrt(n)))
