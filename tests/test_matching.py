"""Matching decoder (host C++ in libqcf) and the detector error model it
decodes over.

* blossom solver: maximum-weight matchings equal networkx's on committed
  random-graph fixtures (tests/golden/make_matching.py), and minimum-weight
  perfect matchings equal a dynamic program over subsets on small complete
  graphs;
* error model: every single fault pushed through the oracle's replay of the
  reference operations (one fault per shot, injected through replay masks)
  gives the same detector sets, observables and merged rates;
* decoding: equals brute-force minimum-weight decoding over all subsets of
  the repetition-code error mechanisms; corrects every unambiguous
  graph-like single fault of the surface code; logical error rate grows with
  p (oracle-sampled shots);
* the AbstractSyndromeDecoder contract (src/ecc/decoder_pipeline.jl:12-24,
  :172-189): guesses times the faults matrix give the decoded observables.
"""
import ctypes as C
import itertools
import json
import os
from collections import defaultdict

import numpy as np
import pytest

import oracle
import qcf
from qcf import _lib, circuits, decoders
from qcf.ops import K

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def mwm(n, edges):
    w = np.zeros((n, n), dtype=np.int64)
    for u, v, x in edges:
        w[u, v] = w[v, u] = x
    mate = np.full(max(n, 1), -1, dtype=np.int32)
    _lib.call("qcf_debug_max_weight_matching", n, w.ctypes.data, mate.ctypes.data)
    return w, mate[:n]


def test_blossom_equals_networkx_fixtures():
    data = json.load(open(os.path.join(GOLDEN, "matching_networkx.json")))
    assert len(data["cases"]) >= 100
    for case in data["cases"]:
        n = case["n"]
        w, mate = mwm(n, case["edges"])
        pairs = set()
        for u in range(n):
            if mate[u] >= 0:
                assert mate[mate[u]] == u
                assert w[u, mate[u]] > 0
                pairs.add((min(u, mate[u]), max(u, mate[u])))
        assert sum(int(w[u, v]) for u, v in pairs) == case["weight"], case


def min_perfect_dp(W):
    n = W.shape[0]
    best = {0: 0}

    def f(mask):
        if mask in best:
            return best[mask]
        i = (mask & -mask).bit_length() - 1
        r = mask & ~(1 << i)
        v = min(W[i, j] + f(r & ~(1 << j)) for j in range(n) if r >> j & 1)
        best[mask] = v
        return v

    return f((1 << n) - 1)


def test_min_weight_perfect_matching_equals_subset_dp():
    rng = np.random.default_rng(5)
    for trial in range(60):
        n = int(rng.choice([2, 4, 6, 8, 10]))
        W = rng.integers(0, 500, (n, n))
        W = np.triu(W, 1)
        W = W + W.T
        big = int(W.max()) * n + 1
        edges = [(u, v, big - int(W[u, v])) for u in range(n) for v in range(u + 1, n)]
        _, mate = mwm(n, edges)
        assert (mate >= 0).all()
        got = sum(int(W[u, mate[u]]) for u in range(n) if mate[u] > u)
        assert got == min_perfect_dp(W), trial


# ---------------------------------------------------------------- error model
FAULTS_1Q = {1: (1, 0), 2: (0, 1), 3: (1, 1)}  # X, Z, Y -> (x flip, z flip)


def oracle_dem(circ, det_bits, obs_bits):
    """Inject every single fault through replay masks into the oracle and read
    the detector / observable flips; merge like signatures."""
    prog = qcf.Program(circ)
    ops, args = prog.ops, prog.args
    faults = []  # (list of (plane, 1)), rate
    plane = prog.nq  # initZ! planes first
    for op in ops:
        k = int(op["kind"])
        if k in (K["MRZ"], K["MRX"]):
            plane += 1
        elif k == K["PAULI_ERROR"]:
            p = float(op["p"][0])
            for pa in (1, 2, 3):
                faults.append(([plane + j for j in range(2) if FAULTS_1Q[pa][j]], p / 3))
            plane += 2
        elif k == K["PAULI_NOISE"]:
            px, py, pz = (float(x) for x in op["p"])
            for pa, q in ((1, px), (2, pz), (3, py)):
                if q > 0:
                    faults.append(([plane + j for j in range(2) if FAULTS_1Q[pa][j]], q))
            plane += 2
        elif k == K["DEPOLARIZE2"]:
            e = float(op["p"][0])
            for pa in range(1, 16):
                d1, d2 = pa & 3, pa >> 2
                pl = [plane + j for j in range(2) if d1 and FAULTS_1Q[d1][j]]
                pl += [plane + 2 + j for j in range(2) if d2 and FAULTS_1Q[d2][j]]
                faults.append((pl, e / 15))
            plane += 4
        else:
            assert k < K["MZ"] or k in (K["MZ"], K["MX"], K["CLASSICAL_XOR"]), k
    assert plane == prog.nreplay
    shots = len(faults)
    ld = qcf.frames.words(shots)
    bits = np.zeros((prog.nreplay, shots), np.uint8)
    for s, (pl, _) in enumerate(faults):
        bits[pl, s] = 1
    masks = oracle.pack_shots(bits)
    _, meas = oracle.pf_run(ops, args, prog.nq, prog.nbits, shots, masks=masks)
    dem = {}
    det_idx = np.asarray(det_bits) - 1
    obs_idx = np.asarray(obs_bits) - 1
    for s, (_, p) in enumerate(faults):
        dets = tuple(int(i) for i in np.flatnonzero(meas[det_idx, s]))
        obs = int(sum(int(meas[o, s]) << j for j, o in enumerate(obs_idx)))
        if not dets and not obs:
            continue
        q = dem.get((dets, obs), 0.0)
        dem[(dets, obs)] = q * (1 - p) + p * (1 - q)
    return dem


@pytest.mark.parametrize("make", [lambda: circuits.repetition_code(3, 3, 1e-3),
                                  lambda: circuits.repetition_code(5, 2, 3e-3),
                                  lambda: circuits.rotated_surface_code(3, 2, 1e-3),
                                  lambda: circuits.rotated_surface_code(3, 3, 2e-3)])
def test_error_model_equals_oracle_single_faults(make):
    c = make()
    ours = decoders.DetectorErrorModel(c.ops, c.det_bits, c.obs_bits)
    got = {(tuple(int(x) for x in d), o): p for d, o, p in ours.mechanisms()}
    want = oracle_dem(c.ops, c.det_bits, c.obs_bits)
    assert set(got) == set(want)
    for k in want:
        assert got[k] == pytest.approx(want[k], rel=1e-9, abs=1e-15), k


def _interleaved_noise_repetition_code():
    """The d=5, 2-round repetition code with an extra PauliError after every
    CNOT on a data qubit that CNOT leaves alone: the program compiler moves
    those sites back across the operations on other qubits into the noise
    run before them."""
    c = circuits.repetition_code(5, 2, 3e-3)
    ops, extra = [], 0
    k = 0
    for op in c.ops:
        ops.append(op)
        if isinstance(op, qcf.sCNOT):
            q = 1 + (k % 5)
            k += 1
            if q not in (op.q1, op.q2):
                ops.append(qcf.ops.PauliError(q, 2e-3))
                extra += 1
    c.ops = ops
    return c, extra


def test_error_model_with_moved_noise_equals_oracle_single_faults():
    """Noise moved back across operations on other qubits leaves every
    single fault's detectors and observable as the oracle's replay of the
    original operation order gives them."""
    c, extra = _interleaved_noise_repetition_code()

    def passes(ops):
        prog = qcf.Program(ops)
        n = C.c_size_t()
        _lib.call("qcf_debug_jit_source", prog.handle, 1 | 8, None, 0, C.byref(n))
        buf = C.create_string_buffer(n.value + 1)
        _lib.call("qcf_debug_jit_source", prog.handle, 1 | 8, buf, n.value + 1, C.byref(n))
        return buf.value.decode().split("qcf_jit_kernel")[-1].count("qcfd::group_events")

    # every inserted site joined the noise run before it: no event pass more
    # than the unmodified circuit's (without the moves, up to one per site)
    assert extra == 12
    assert passes(c.ops) == passes(circuits.repetition_code(5, 2, 3e-3).ops)
    ours = decoders.DetectorErrorModel(c.ops, c.det_bits, c.obs_bits)
    got = {(tuple(int(x) for x in d), o): p for d, o, p in ours.mechanisms()}
    want = oracle_dem(c.ops, c.det_bits, c.obs_bits)
    assert set(got) == set(want)
    for k in want:
        assert got[k] == pytest.approx(want[k], rel=1e-9, abs=1e-15), k


# ---------------------------------------------------------------- decoding
def test_repetition_code_decoding_equals_brute_force_min_weight():
    c = circuits.repetition_code(3, 3, 2e-3)
    dec = decoders.MatchingDecoder(c.ops, c.det_bits, c.obs_bits)
    mechs = dec.dem.mechanisms()
    assert all(len(d) <= 2 for d, _, _ in mechs) and len(mechs) <= 20
    m, nd = len(mechs), dec.ndet
    H = np.zeros((m, nd), np.uint8)
    for i, (d, _, _) in enumerate(mechs):
        H[i, d] = 1
    obs = np.array([o & 1 for _, o, _ in mechs], np.uint8)
    w = np.array([np.log((1 - p) / p) for _, _, p in mechs])
    subsets = np.array(list(itertools.product([0, 1], repeat=m)), np.uint8)
    syn = (subsets @ H) % 2
    par = (subsets @ obs) % 2
    wt = subsets @ w
    key = syn @ (1 << np.arange(nd))
    best = {}
    for k, p, x in zip(key, par, wt):
        b = best.get(k)
        if b is None or x < b[0] - 1e-9:
            best[k] = (x, {int(p)})
        elif abs(x - b[0]) <= 1e-9:
            b[1].add(int(p))
    keys = [k for k, (x, ps) in best.items() if len(ps) == 1]
    syns = np.array([[(k >> j) & 1 for j in range(nd)] for k in keys], np.uint8)
    pred = dec.decode_observables(syns)[:, 0]
    want = np.array([next(iter(best[k][1])) for k in keys], np.uint8)
    assert len(keys) > 100
    assert np.array_equal(pred, want)


@pytest.mark.parametrize("d,rounds", [(3, 3), (5, 5)])
def test_unambiguous_graphlike_single_faults_are_corrected(d, rounds):
    c = circuits.rotated_surface_code(d, rounds, 1e-3)
    dec = decoders.MatchingDecoder(c.ops, c.det_bits, c.obs_bits)
    mechs = dec.dem.mechanisms()
    variants = defaultdict(set)
    for dd, o, _ in mechs:
        variants[tuple(dd)].add(o)
    syn = np.zeros((len(mechs), dec.ndet), np.uint8)
    for i, (dd, _, _) in enumerate(mechs):
        syn[i, dd] = 1
    pred = dec.decode_observables(syn)[:, 0]
    checked = 0
    for i, (dd, o, _) in enumerate(mechs):
        if len(dd) <= 2 and len(variants[tuple(dd)]) == 1:
            assert pred[i] == (o & 1), (i, dd, o)
            checked += 1
    assert checked > 0.3 * len(mechs)
    assert dec.dropped == 0


def oracle_ler(d, rounds, p, shots, seed):
    c = circuits.rotated_surface_code(d, rounds, p)
    dec = decoders.MatchingDecoder(c.ops, c.det_bits, c.obs_bits)
    prog = dec.dem.program
    _, meas = oracle.pf_run(prog.ops, prog.args, prog.nq, prog.nbits, shots, seed=seed,
                            nthreads=4)
    syn = np.ascontiguousarray(meas[np.asarray(c.det_bits) - 1].T)
    obs = meas[np.asarray(c.obs_bits) - 1].T
    return float(np.mean(np.any(dec.decode_observables(syn) != obs, axis=1)))


def test_logical_error_rate_grows_with_p():
    lers = [oracle_ler(3, 3, p, 20000, seed=11) for p in (2e-3, 6e-3, 1.8e-2)]
    assert lers[0] < lers[1] < lers[2]
    assert lers[0] < 0.02 and lers[2] > 0.03


def test_batchdecode_contract():
    c = circuits.rotated_surface_code(3, 3, 5e-3)
    dec = decoders.MatchingDecoder(c.ops, c.det_bits, c.obs_bits)
    prog = dec.dem.program
    _, meas = oracle.pf_run(prog.ops, prog.args, prog.nq, prog.nbits, 3000, seed=2)
    syn = np.ascontiguousarray(meas[np.asarray(c.det_bits) - 1].T)
    obs = np.ascontiguousarray(meas[np.asarray(c.obs_bits) - 1].T)
    guesses = decoders.batchdecode(dec, syn)
    F = dec.faults_matrix()
    H = decoders.parity_checks(dec)
    assert guesses.shape == (3000, dec.dem.nmech)
    assert H.shape == (dec.ndet, dec.dem.nmech)
    # the guessed mechanisms reproduce the syndrome and the decoded observables
    assert np.array_equal((guesses.astype(np.int64) @ H.T) % 2, syn)
    assert np.array_equal((guesses.astype(np.int64) @ F.T) % 2, dec.decode_observables(syn))
    ler = decoders.evaluate_guesses(obs, guesses, F)
    assert ler == pytest.approx(float(np.mean(np.any(dec.decode_observables(syn) != obs, 1))))
    assert np.array_equal(decoders.decode(dec, syn[7]), guesses[7])
    with pytest.raises(qcf.ArgumentError, match="wrong dimensions"):
        dec.batchdecode(syn[:, :-1])


def test_packed_detectors_decode_like_bytes():
    c = circuits.rotated_surface_code(3, 3, 8e-3)
    dec = decoders.MatchingDecoder(c.ops, c.det_bits, c.obs_bits)
    prog = dec.dem.program
    shots = 1000
    _, meas = oracle.pf_run(prog.ops, prog.args, prog.nq, prog.nbits, shots, seed=9)
    det = meas[np.asarray(c.det_bits) - 1]
    packed = oracle.pack_shots(det)
    assert np.array_equal(dec.decode_packed(packed, shots),
                          dec.decode_observables(np.ascontiguousarray(det.T)))
