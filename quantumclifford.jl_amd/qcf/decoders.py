"""Syndrome decoders over the circuit-level error model of a frame program.

Mirrors the `AbstractSyndromeDecoder` contract of the reference
(src/ecc/decoder_pipeline.jl:1-24): `parity_checks(d)`, `decode(d, syndrome)`,
`batchdecode(d, syndromes)` and `evaluate_guesses(measured_faults, guesses,
faults_matrix)` (:172-189).  The "error" a guess names is a set of error
mechanisms of the detector error model (columns of `parity_checks`), and the
faults matrix maps mechanisms to logical observables, so `evaluate_guesses`
works unchanged.  Matching runs in libqcf (host C++, threads); the reference
delegates it to PyMatching
(ext/QuantumCliffordPyQDecodersExt/QuantumCliffordPyQDecodersExt.jl:83-130).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .frames import Program, pftrajectories


def _np_ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


class DetectorErrorModel:
    """Error mechanisms (detector set, observable mask, rate) of a program,
    read off by pushing every single fault through the compiled circuit."""

    def __init__(self, program, det_bits, obs_bits):
        self.program = program if isinstance(program, Program) else Program(program)
        self.det_bits = np.asarray(det_bits, dtype=np.uint32)
        self.obs_bits = np.asarray(obs_bits, dtype=np.uint32)
        h = C.c_void_p()
        _lib.call("qcf_dem_build", self.program.handle, _np_ptr(self.det_bits),
                  len(self.det_bits), _np_ptr(self.obs_bits), len(self.obs_bits),
                  C.byref(h))
        self.handle = h
        nm, nd, no = C.c_size_t(), C.c_uint32(), C.c_uint32()
        _lib.call("qcf_dem_size", h, C.byref(nm), C.byref(nd), C.byref(no))
        self.nmech, self.ndet, self.nobs = nm.value, nd.value, no.value

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib is not None and _lib._lib is not None:
            _lib._lib.qcf_dem_free(h)
        self.handle = None

    def mechanism(self, i: int):
        buf = np.zeros(max(1, self.ndet), dtype=np.uint32)
        n, obs, p = C.c_size_t(), C.c_uint64(), C.c_double()
        _lib.call("qcf_dem_mechanism", self.handle, i, _np_ptr(buf), len(buf), C.byref(n),
                  C.byref(obs), C.byref(p))
        return buf[: n.value].copy(), int(obs.value), float(p.value)

    def mechanisms(self):
        return [self.mechanism(i) for i in range(self.nmech)]

    def check_matrix(self) -> np.ndarray:
        """Detectors x mechanisms over GF(2)."""
        H = np.zeros((self.ndet, self.nmech), dtype=np.uint8)
        for i in range(self.nmech):
            d, _, _ = self.mechanism(i)
            H[d, i] = 1
        return H

    def faults_matrix(self) -> np.ndarray:
        """Observables x mechanisms over GF(2)."""
        F = np.zeros((self.nobs, self.nmech), dtype=np.uint8)
        for i in range(self.nmech):
            _, o, _ = self.mechanism(i)
            for j in range(self.nobs):
                F[j, i] = (o >> j) & 1
        return F


class AbstractSyndromeDecoder:
    """src/ecc/decoder_pipeline.jl:1-8."""


class MatchingDecoder(AbstractSyndromeDecoder):
    """Minimum-weight perfect matching over a circuit's detector error model.

    `MatchingDecoder(circuit, det_bits, obs_bits)`; det_bits / obs_bits are the
    1-based record bits of the detectors and observables (e.g.
    `circuits.rotated_surface_code(...).det_bits`)."""

    def __init__(self, circuit, det_bits, obs_bits, nthreads: int = 0):
        self.dem = DetectorErrorModel(circuit, det_bits, obs_bits)
        h = C.c_void_p()
        _lib.call("qcf_matcher_create", self.dem.handle, C.byref(h))
        self.handle = h
        ne, dr = C.c_size_t(), C.c_uint32()
        _lib.call("qcf_matcher_info", h, C.byref(ne), C.byref(dr))
        self.nedges, self.dropped = ne.value, dr.value
        self.nthreads = nthreads
        self._H = None
        self._F = None

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib is not None and _lib._lib is not None:
            _lib._lib.qcf_matcher_free(h)
        self.handle = None

    @property
    def ndet(self):
        return self.dem.ndet

    @property
    def nobs(self):
        return self.dem.nobs

    def parity_checks(self) -> np.ndarray:
        if self._H is None:
            self._H = self.dem.check_matrix()
        return self._H

    def faults_matrix(self) -> np.ndarray:
        if self._F is None:
            self._F = self.dem.faults_matrix()
        return self._F

    def _syn(self, syndromes) -> np.ndarray:
        s = np.ascontiguousarray(np.asarray(syndromes, dtype=np.uint8))
        if s.ndim == 1:
            s = s[None, :]
        if s.shape[1] != self.ndet:
            raise _lib.ArgumentError(
                f"The syndromes given to `batchdecode` have the wrong dimensions. The syndrome "
                f"length is {s.shape[1]} while it should be {self.ndet}")
        return s

    def decode_observables(self, syndromes) -> np.ndarray:
        """Predicted observable flips, shots x nobs (uint8)."""
        s = self._syn(syndromes)
        out = np.zeros((s.shape[0], self.nobs), dtype=np.uint8)
        _lib.call("qcf_matcher_decode", self.handle, _np_ptr(s), s.shape[0], _np_ptr(out),
                  None, self.nthreads)
        return out

    def batchdecode(self, syndromes) -> np.ndarray:
        """Guessed mechanisms, shots x nmech (src/ecc/decoder_pipeline.jl:12-24)."""
        s = self._syn(syndromes)
        obs = np.zeros((s.shape[0], self.nobs), dtype=np.uint8)
        guesses = np.zeros((s.shape[0], self.dem.nmech), dtype=np.uint8)
        _lib.call("qcf_matcher_decode", self.handle, _np_ptr(s), s.shape[0], _np_ptr(obs),
                  _np_ptr(guesses), self.nthreads)
        return guesses

    def decode(self, syndrome) -> np.ndarray:
        return self.batchdecode(syndrome)[0]

    def decode_packed(self, dets: np.ndarray, shots: int) -> np.ndarray:
        """Detector planes packed 64 shots per word (ndet x ld uint64)."""
        d = np.ascontiguousarray(dets, dtype=np.uint64)
        if d.ndim != 2 or d.shape[0] != self.ndet:
            raise _lib.DimensionMismatch(f"expected {self.ndet} detector planes, got {d.shape}")
        out = np.zeros((shots, self.nobs), dtype=np.uint8)
        _lib.call("qcf_matcher_decode_packed", self.handle, _np_ptr(d), d.shape[1], shots,
                  _np_ptr(out), self.nthreads)
        return out


def parity_checks(d):
    return d.parity_checks()


def decode(d, syndrome):
    return d.decode(syndrome)


def batchdecode(d, syndromes):
    return d.batchdecode(syndromes)


def evaluate_guesses(measured_faults, guesses, faults_matrix) -> float:
    """Fraction of samples whose guessed logical flips differ from the measured
    ones in any row of the faults matrix (src/ecc/decoder_pipeline.jl:172-189)."""
    g = np.asarray(guesses, dtype=np.uint8)
    F = np.asarray(faults_matrix, dtype=np.uint8)
    m = np.asarray(measured_faults, dtype=np.uint8).reshape(g.shape[0], -1)
    pred = (g.astype(np.int64) @ F.T.astype(np.int64)) % 2
    return float(np.mean(np.any(pred != m, axis=1)))


def sample_detectors(circuit, det_bits, obs_bits, shots: int, seed: int):
    """Sample on the GPU; return (detector planes ndet x ld, observable planes
    nobs x ld) on the host, packed 64 shots per uint64 word."""
    f = pftrajectories(circuit, shots, seed=seed, keep_frame=False)
    rec = f.record_host()
    return (rec[np.asarray(det_bits) - 1], rec[np.asarray(obs_bits) - 1])


def unpack_planes(planes: np.ndarray, shots: int) -> np.ndarray:
    """Planes x words uint64 -> shots x planes uint8."""
    b = np.unpackbits(planes.view(np.uint8).reshape(planes.shape[0], -1), axis=1,
                      bitorder="little")[:, :shots]
    return np.ascontiguousarray(b.T)


def logical_error_rate(decoder: MatchingDecoder, circuit, shots: int, seed: int = 0):
    """Sample `shots` on the GPU, decode the detectors, and return the
    fraction whose predicted observables differ from the sampled ones."""
    dets, obs = sample_detectors(circuit, decoder.dem.det_bits, decoder.dem.obs_bits, shots, seed)
    pred = decoder.decode_packed(dets, shots)
    actual = unpack_planes(obs, shots)
    return float(np.mean(np.any(pred != actual, axis=1)))
