"""Device-resident Pauli operators and tableaux with the reference's
`mul_left!` / `mul_right!` family (src/mul_leftright.jl).

Storage follows the reference exactly: a PauliOperator holds `xz` =
[x words ; z words] of ceil(n/64) UInt64 each (src/pauli_operator.jl:42-56)
and a one-byte phase (0:+1 1:+i 2:-1 3:-i, src/QuantumClifford.jl:118-121);
a Tableau keeps rows contiguous (`fastrow`), row m at xzs[m*ld : m*ld+2W].
Words live in torch int64 tensors on the GPU (same bits as UInt64).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import DimensionMismatch, call

PHASE_STR = {0: "+ ", 1: "+i", 2: "- ", 3: "-i"}
_PHASEDICT = {"": 0, "+": 0, "i": 1, "+i": 1, "-": 2, "-i": 3}


def _device(device=None):
    if device is not None:
        return torch.device(device)
    return torch.device("cuda", torch.cuda.current_device())


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t: torch.Tensor, offset_elems: int = 0) -> C.c_void_p:
    return C.c_void_p(t.data_ptr() + offset_elems * t.element_size())


def words_for(n: int) -> int:
    return (n + 63) // 64


def pack_bits(bits: np.ndarray, nwords: int) -> np.ndarray:
    """Pack a 0/1 vector into little-endian UInt64 words (bit i -> word i//64)."""
    out = np.zeros(nwords * 64, dtype=np.uint8)
    out[: len(bits)] = np.asarray(bits, dtype=np.uint8)
    return np.packbits(out.reshape(-1, 8), axis=1, bitorder="little").reshape(-1).view(np.uint64)


def unpack_bits(words: np.ndarray, n: int) -> np.ndarray:
    w = np.ascontiguousarray(words, dtype=np.uint64)
    return np.unpackbits(w.view(np.uint8), bitorder="little")[:n]


class PauliOperator:
    """PauliOperator(phase, nqubits, xz) on the GPU."""

    def __init__(self, phase: int, nqubits: int, xz, device=None):
        dev = _device(device)
        xz_np = np.ascontiguousarray(np.asarray(xz, dtype=np.uint64))
        self.nqubits = int(nqubits)
        self.xz = torch.from_numpy(xz_np.view(np.int64).copy()).to(dev)
        self.phase = torch.tensor([int(phase) & 3], dtype=torch.uint8, device=dev)

    @property
    def nwords(self) -> int:
        return self.xz.numel() // 2

    @classmethod
    def from_bits(cls, phase, x, z, device=None):
        n = len(x)
        w = words_for(n)
        return cls(phase, n, np.concatenate([pack_bits(x, w), pack_bits(z, w)]), device)

    @classmethod
    def zero(cls, n: int, device=None):
        return cls(0, n, np.zeros(2 * words_for(n), np.uint64), device)

    # host views ---------------------------------------------------------
    def xz_host(self) -> np.ndarray:
        return self.xz.cpu().numpy().view(np.uint64).copy()

    def phase_value(self) -> int:
        return int(self.phase.item())

    def xbits(self) -> np.ndarray:
        return unpack_bits(self.xz_host()[: self.nwords], self.nqubits)

    def zbits(self) -> np.ndarray:
        return unpack_bits(self.xz_host()[self.nwords:], self.nqubits)

    def copy(self) -> "PauliOperator":
        p = PauliOperator.__new__(PauliOperator)
        p.nqubits = self.nqubits
        p.xz = self.xz.clone()
        p.phase = self.phase.clone()
        return p

    def __getitem__(self, i: int):
        return bool(self.xbits()[i - 1]), bool(self.zbits()[i - 1])

    def __repr__(self):
        s = "".join("_XZY"[int(x) | (int(z) << 1)] for x, z in zip(self.xbits(), self.zbits()))
        return PHASE_STR[self.phase_value()] + s

    def __eq__(self, other):
        return (isinstance(other, PauliOperator) and self.nqubits == other.nqubits
                and self.phase_value() == other.phase_value()
                and np.array_equal(self.xz_host(), other.xz_host()))

    def __mul__(self, other: "PauliOperator") -> "PauliOperator":
        # l * r == mul_left!(copy(r), l) (src/QuantumClifford.jl:862)
        return mul_left_(other.copy(), self)


def P(s: str, device=None) -> PauliOperator:
    """The P"..." string macro (src/pauli_operator.jl:97-105)."""
    letters = [c for c in s if c in "_IZXY"]
    phase = _PHASEDICT["".join(c for c in s if c not in "_IZXY").strip()]
    return PauliOperator.from_bits(phase, [c in "XY" for c in letters],
                                   [c in "ZY" for c in letters], device)


def mul_left_(r: PauliOperator, l: PauliOperator, phases: bool = True) -> PauliOperator:
    """mul_left!(r, l; phases): r := l*r (src/mul_leftright.jl:137-142)."""
    if l.nqubits != r.nqubits:
        raise DimensionMismatch("The two Pauli operators should have the same length!")
    call("qcf_mul_left", _ptr(r.xz), _ptr(r.phase), _ptr(l.xz), _ptr(l.phase), r.nwords,
         int(bool(phases)), _stream())
    return r


def mul_right_(l: PauliOperator, r: PauliOperator, phases: bool = True) -> PauliOperator:
    """mul_right!(l, r; phases): l := l*r (src/mul_leftright.jl:144-149)."""
    if l.nqubits != r.nqubits:
        raise DimensionMismatch("The two Pauli operators should have the same length!")
    call("qcf_mul_right", _ptr(l.xz), _ptr(l.phase), _ptr(r.xz), _ptr(r.phase), l.nwords,
         int(bool(phases)), _stream())
    return l


def mul_left_vec_(r: torch.Tensor, l: torch.Tensor, phases: bool = True,
                  right: bool = False) -> int:
    """mul_left!(r::AbstractVector, l::AbstractVector)::UInt8 on [x;z] word tensors
    (src/mul_leftright.jl:123-131); returns the product phase."""
    if r.numel() != l.numel() or r.numel() % 2:
        raise DimensionMismatch("operands must have equal, even word counts")
    out = torch.zeros(1, dtype=torch.uint8, device=r.device)
    call("qcf_mul_vec", _ptr(r), _ptr(l), r.numel() // 2, int(bool(phases)), int(bool(right)),
         _ptr(out), _stream())
    return int(out.item())


class Tableau:
    """Rows of Pauli operators, row-contiguous (`fastrow`), phases one byte each
    (src/QuantumClifford.jl:145-281)."""

    def __init__(self, phases, nqubits: int, xzs, device=None):
        dev = _device(device)
        xzs = np.ascontiguousarray(np.asarray(xzs, dtype=np.uint64))
        self.nrows = xzs.shape[0]
        self.nqubits = int(nqubits)
        self.ld = xzs.shape[1]
        self.xzs = torch.from_numpy(xzs.view(np.int64).copy()).to(dev)
        self.phases = torch.from_numpy(np.asarray(phases, dtype=np.uint8).copy()).to(dev)

    @property
    def nwords(self) -> int:
        return self.ld // 2

    def xzs_host(self) -> np.ndarray:
        return self.xzs.cpu().numpy().view(np.uint64).reshape(self.nrows, self.ld).copy()

    def phases_host(self) -> np.ndarray:
        return self.phases.cpu().numpy().copy()

    def __len__(self):
        return self.nrows

    def row_strings(self) -> list:
        """Rows in the reference's text form, e.g. '+ XZ_Y'."""
        xzs, ph = self.xzs_host(), self.phases_host()
        W = self.nwords
        out = []
        for m in range(self.nrows):
            xs = unpack_bits(xzs[m, :W], self.nqubits)
            zs = unpack_bits(xzs[m, W:2 * W], self.nqubits)
            out.append(PHASE_STR[int(ph[m]) & 3] +
                       "".join("_XZY"[int(x) | (int(z) << 1)] for x, z in zip(xs, zs)))
        return out

    def __repr__(self):
        return "\n".join(self.row_strings())

    def copy(self) -> "Tableau":
        t = Tableau.__new__(Tableau)
        t.nrows, t.nqubits, t.ld = self.nrows, self.nqubits, self.ld
        t.xzs, t.phases = self.xzs.clone(), self.phases.clone()
        return t

    def __eq__(self, other):
        return (isinstance(other, Tableau) and self.nqubits == other.nqubits
                and self.nrows == other.nrows
                and np.array_equal(self.phases_host(), other.phases_host())
                and np.array_equal(self.xzs_host()[:, :2 * self.nwords],
                                   other.xzs_host()[:, :2 * other.nwords]))


def S(text: str, device=None) -> Tableau:
    """The S"..." string macro (src/stabilizer.jl): one Pauli per line, each with
    an optional phase prefix (+, -, i, +i, -i) and letters _ I X Y Z."""
    lines = [ln.strip() for ln in text.strip().splitlines() if ln.strip()]
    rows = []
    for ln in lines:
        letters = [c for c in ln if c in "_IZXY"]
        phase = _PHASEDICT["".join(c for c in ln if c not in "_IZXY").replace(" ", "")]
        rows.append((phase, letters))
    n = len(rows[0][1])
    if any(len(r[1]) != n for r in rows):
        raise DimensionMismatch("all rows of S\"...\" need the same number of qubits")
    W = words_for(n)
    xzs = np.zeros((len(rows), 2 * W), np.uint64)
    for m, (_, letters) in enumerate(rows):
        xzs[m, :W] = pack_bits([c in "XY" for c in letters], W)
        xzs[m, W:] = pack_bits([c in "ZY" for c in letters], W)
    return Tableau([r[0] for r in rows], n, xzs, device)


def canonicalize_(s: Tableau, phases: bool = True, ranks: bool = False):
    """canonicalize!(state; phases, ranks) (src/canonicalization.jl:97-135) in
    place on the device.  With ranks=True returns (s, rx, rz) like the
    reference's (state, rx-1, i-1): pivots of the X stage and of both stages."""
    out = torch.zeros(2, dtype=torch.int32, device=s.xzs.device) if ranks else None
    call("qcf_canonicalize", _ptr(s.xzs), _ptr(s.phases), s.ld, s.nwords, s.nrows, s.nqubits,
         int(bool(phases)), _ptr(out) if ranks else None, _stream())
    if ranks:
        rx, rz = (int(v) for v in out.cpu().tolist())
        return s, rx, rz
    return s


def mul_left_row_(s: Tableau, m: int, i: int, phases: bool = True) -> Tableau:
    """mul_left!(s::Tableau, m, i): row m := row i * row m (src/mul_leftright.jl:177-181).
    Rows are 1-based as in the reference."""
    m0, i0 = m - 1, i - 1
    call("qcf_mul_left", _ptr(s.xzs, m0 * s.ld), _ptr(s.phases, m0), _ptr(s.xzs, i0 * s.ld),
         _ptr(s.phases, i0), s.nwords, int(bool(phases)), _stream())
    return s


def mul_left_tableau_pauli_(s: Tableau, p: PauliOperator, phases: bool = True) -> Tableau:
    """mul_left!(s::Tableau, p::PauliOperator): every row := p * row
    (src/mul_leftright.jl:221-227)."""
    if p.nwords != s.nwords:
        raise DimensionMismatch("Pauli and tableau widths differ")
    call("qcf_mul_rows", _ptr(s.xzs), _ptr(s.phases), s.ld, s.nwords, s.nrows, _ptr(p.xz),
         _ptr(p.phase), C.c_size_t(2**64 - 1), None, int(bool(phases)), 0, _stream())
    return s


def mul_left_sweep_(s: Tableau, pivot: int, row_mask: torch.Tensor | None = None,
                    phases: bool = True) -> Tableau:
    """The pivot-row sweep of canonicalize! (src/canonicalization.jl:105-109): for every
    row m != pivot (selected by row_mask when given), mul_left!(s, m, pivot).  1-based."""
    p0 = pivot - 1
    mask = None if row_mask is None else _ptr(row_mask)
    call("qcf_mul_rows", _ptr(s.xzs), _ptr(s.phases), s.ld, s.nwords, s.nrows,
         _ptr(s.xzs, p0 * s.ld), _ptr(s.phases, p0), p0, mask, int(bool(phases)), 0, _stream())
    return s


def mul_right_row_(s: Tableau, m: int, i: int, phases: bool = True) -> Tableau:
    """mul_right!(s::Tableau, m, i): row m := row m * row i (src/mul_leftright.jl:189-193)."""
    m0, i0 = m - 1, i - 1
    call("qcf_mul_right", _ptr(s.xzs, m0 * s.ld), _ptr(s.phases, m0), _ptr(s.xzs, i0 * s.ld),
         _ptr(s.phases, i0), s.nwords, int(bool(phases)), _stream())
    return s


def _check_widths(a_nwords: int, b_nwords: int):
    if a_nwords != b_nwords:
        raise DimensionMismatch("The two Pauli operators should have the same length!")


def mul_left_tableau_row_(s: Tableau, m: int, t: Tableau, i: int, phases: bool = True) -> Tableau:
    """mul_left!(s::Tableau, m, t::Tableau, i): row m of s := row i of t * row m of s
    (src/mul_leftright.jl:171-175).  As in the reference, the phase added is
    s.phases[i] (of the first tableau), not t.phases[i]."""
    _check_widths(s.nwords, t.nwords)
    m0, i0 = m - 1, i - 1
    call("qcf_mul_left", _ptr(s.xzs, m0 * s.ld), _ptr(s.phases, m0), _ptr(t.xzs, i0 * t.ld),
         _ptr(s.phases, i0), s.nwords, int(bool(phases)), _stream())
    return s


def mul_right_tableau_row_(s: Tableau, m: int, t: Tableau, i: int, phases: bool = True) -> Tableau:
    """mul_right!(s::Tableau, m, t::Tableau, i): row m of s := row m of s * row i of t
    (src/mul_leftright.jl:183-187), adding s.phases[i] like the left form."""
    _check_widths(s.nwords, t.nwords)
    m0, i0 = m - 1, i - 1
    call("qcf_mul_right", _ptr(s.xzs, m0 * s.ld), _ptr(s.phases, m0), _ptr(t.xzs, i0 * t.ld),
         _ptr(s.phases, i0), s.nwords, int(bool(phases)), _stream())
    return s


def mul_left_pauli_row_(r: PauliOperator, l: Tableau, i: int, phases: bool = True) -> PauliOperator:
    """mul_left!(r::PauliOperator, l::Tableau, i): r := row i of l * r
    (src/mul_leftright.jl:151-155)."""
    _check_widths(r.nwords, l.nwords)
    i0 = i - 1
    call("qcf_mul_left", _ptr(r.xz), _ptr(r.phase), _ptr(l.xzs, i0 * l.ld), _ptr(l.phases, i0),
         r.nwords, int(bool(phases)), _stream())
    return r


def mul_right_pauli_row_(l: PauliOperator, r: Tableau, i: int, phases: bool = True) -> PauliOperator:
    """mul_right!(l::PauliOperator, r::Tableau, i): l := l * row i of r
    (src/mul_leftright.jl:159-163)."""
    _check_widths(l.nwords, r.nwords)
    i0 = i - 1
    call("qcf_mul_right", _ptr(l.xz), _ptr(l.phase), _ptr(r.xzs, i0 * r.ld), _ptr(r.phases, i0),
         l.nwords, int(bool(phases)), _stream())
    return l


def mul_right_tableau_pauli_(s: Tableau, p: PauliOperator, phases: bool = True) -> Tableau:
    """mul_right!(s::Tableau, p::PauliOperator): every row := row * p
    (src/mul_leftright.jl:234-240)."""
    _check_widths(s.nwords, p.nwords)
    call("qcf_mul_rows", _ptr(s.xzs), _ptr(s.phases), s.ld, s.nwords, s.nrows, _ptr(p.xz),
         _ptr(p.phase), C.c_size_t(2**64 - 1), None, int(bool(phases)), 1, _stream())
    return s


def mul_left_destabilizer_(t: Tableau, i: int, j: int, phases: bool = True) -> Tableau:
    """mul_left!(s::Destabilizer, i, j) on its 2n-row tableau (destabilizers
    1..n, then stabilizers n+1..2n; src/mul_leftright.jl:200-206): destabilizer
    j := destabilizer i * destabilizer j without phases (the indices flip to
    keep the commutation relations), then stabilizer i := stabilizer j *
    stabilizer i."""
    n = t.nrows // 2
    mul_left_row_(t, j, i, phases=False)
    return mul_left_row_(t, i + n, j + n, phases=phases)
