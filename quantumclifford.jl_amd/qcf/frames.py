"""Pauli frames on the GPU: PauliFrame, pftrajectories, measurements
(src/pauli_frames.jl), backed by the libqcf frame sampler.

Frames are shot-packed (64 shots per UInt64 word, plane-major): `frame` holds
2*nq planes (x of qubits 1..nq, then z), `record` holds one plane per
classical bit.  `measurements(f)` returns the reference's Matrix{Bool}
(shots x bits) view.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np
import torch

from . import _lib
from ._lib import OP_DTYPE, DimensionMismatch, ProgramInfo, call
from .ops import Operation


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return C.c_void_p(0 if t is None else t.data_ptr())


class _SeedSource:
    """Stands in for Julia's global RNG: hands out a fresh 64-bit seed per call."""

    def __init__(self, seed: int = 0x5EED):
        self.rng = np.random.default_rng(seed)

    def next(self) -> int:
        return int(self.rng.integers(0, 2**63, dtype=np.int64))


_SEEDS = _SeedSource()


def seed_(seed: int) -> None:
    """Reseed the source of per-call seeds (Random.seed! for the native sampler)."""
    global _SEEDS
    _SEEDS = _SeedSource(seed)


def lower_circuit(circuit: Sequence[Operation], nqubits: int | None = None):
    """Lower reference ops to (qcf_op records, args).  `nqubits` is the frame's
    qubit count (needed by NoiseOpAll); by default the largest affected qubit."""
    if nqubits is None:
        nqubits = max([max(op.affectedqubits(), default=1) for op in circuit] + [1])
    recs, args = [], []

    def emit(kind, q1=0, q2=0, bit=0, arg_off=0, arg_len=0, p=(0.0, 0.0, 0.0)):
        recs.append((kind, q1, q2, bit, arg_off, arg_len, p))

    for op in circuit:
        op.lower(emit, args, nqubits)
    ops = np.zeros(len(recs), dtype=OP_DTYPE)
    for i, r in enumerate(recs):
        ops[i] = r
    return ops, np.asarray(args, dtype=np.uint32)


class Program:
    """A compiled circuit (qcf_program_build)."""

    def __init__(self, circuit: Sequence[Operation], nqubits: int = 0, nbits: int = 0):
        ops, args = lower_circuit(circuit, nqubits or None)
        self._ops, self._args = ops, args  # keep alive for the oracle
        h = C.c_void_p()
        call("qcf_program_build", ops.ctypes.data_as(C.c_void_p), len(ops),
             args.ctypes.data_as(C.c_void_p) if len(args) else None, len(args),
             int(nqubits), int(nbits), C.byref(h))
        self.handle = h
        info = ProgramInfo()
        call("qcf_program_get_info", h, C.byref(info))
        self.info = info

    @property
    def ops(self):
        return self._ops

    @property
    def args(self):
        return self._args

    @property
    def nq(self) -> int:
        return self.info.nqubits

    @property
    def nbits(self) -> int:
        return self.info.nbits

    @property
    def nreplay(self) -> int:
        return self.info.nreplay

    @property
    def bitops_per_shot(self) -> float:
        return self.info.bitops_per_shot

    def gauge_planes(self):
        """Per replay plane: 0 noise, 1 gauge draw a fresh record-only launch
        makes, 2 gauge draw it leaves out (qcf_program_gauge_planes)."""
        out = np.zeros(max(1, self.nreplay), dtype=np.uint8)
        _lib.call("qcf_program_gauge_planes", self.handle, out.ctypes.data, out.size)
        return out[: self.nreplay]

    def __del__(self):
        h = getattr(self, "handle", None)
        if h and h.value and _lib._lib is not None:
            _lib._lib.qcf_program_free(h)
            self.handle = None


def words(shots: int) -> int:
    """Words that hold `shots` shot bits: ceil(shots/64)."""
    return (shots + 63) // 64


def plane_words(shots: int) -> int:
    """Plane stride in words used for frames, records and masks: ceil(shots/2048)*32,
    at least 32, so that every 2048-shot tile of a plane is one aligned 256-byte
    segment (qcf_plane_words)."""
    return max(1, (int(shots) + 2047) // 2048) * 32


class PauliFrame:
    """PauliFrame(frames, qubits, measurements) (src/pauli_frames.jl:9-37)."""

    def __init__(self, frames: int, qubits: int, measurements: int, *, device=None,
                 init: bool = True, _uninitialized: bool = False):
        self.shots = int(frames)
        self.nq = int(qubits) + 1  # hidden ancilla for PauliMeasurement
        self.nbits = int(measurements)
        self.ld = plane_words(self.shots)
        dev = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        # a fresh launch defines every word (unwritten record columns are
        # zeroed, every plane is stored, padding lanes store zeros), so
        # pftrajectories skips the fill
        alloc = torch.empty if _uninitialized else torch.zeros
        self.frame = alloc(2 * self.nq * self.ld, dtype=torch.int64, device=dev)
        self.record = alloc(max(1, self.nbits) * self.ld, dtype=torch.int64, device=dev)
        # a view (view(frame, r)) shares its parent's planes: its shots start at
        # word w0 of every plane and at absolute shot `offset`
        self.offset, self.w0, self.is_view = 0, 0, False
        if init:
            initZ_(self)

    def nqubits(self) -> int:
        return self.nq - 1

    def __len__(self) -> int:
        return self.shots

    def copy(self) -> "PauliFrame":
        """A frame of its own with this frame's shots (a view copies its range)."""
        f = PauliFrame.__new__(PauliFrame)
        f.shots, f.nq, f.nbits, f.ld = self.shots, self.nq, self.nbits, self.ld
        f.offset, f.w0, f.is_view = self.offset, 0, False
        if self.is_view:
            f.ld = plane_words(self.shots)
            f.frame = torch.zeros(2 * f.nq * f.ld, dtype=torch.int64, device=self.frame.device)
            f.record = torch.zeros(max(1, f.nbits) * f.ld, dtype=torch.int64,
                                   device=self.frame.device)
            nw = words(self.shots)
            f.frame.view(2 * f.nq, f.ld)[:, :nw].copy_(self._planes(self.frame, 2 * self.nq))
            f.record.view(-1, f.ld)[:, :nw].copy_(self._planes(self.record, max(1, self.nbits)))
        else:
            f.frame = self.frame.clone()
            f.record = self.record.clone()
        return f

    def _planes(self, t: torch.Tensor, n: int) -> torch.Tensor:
        """(n, words) view of this frame's shots in a plane buffer of the frame."""
        nw = words(self.shots)
        if not self.is_view:
            return t.view(n, self.ld)[:, :nw]
        base = t._qcf_base  # the parent's whole buffer
        return base.view(n, self.ld)[:, self.w0:self.w0 + nw]

    # host views for tests ---------------------------------------------
    def frame_host(self) -> np.ndarray:
        if self.is_view:
            return self._planes(self.frame, 2 * self.nq).cpu().numpy().view(np.uint64)
        return self.frame.cpu().numpy().view(np.uint64).reshape(2 * self.nq, self.ld)

    def record_host(self) -> np.ndarray:
        if self.is_view:
            return self._planes(self.record, max(1, self.nbits)).cpu().numpy().view(
                np.uint64)[: self.nbits]
        return self.record.cpu().numpy().view(np.uint64).reshape(-1, self.ld)[: self.nbits]


def view(frame: PauliFrame, r) -> PauliFrame:
    """view(frame, r) (src/pauli_frames.jl:17): the frames r (a 0-based range or
    slice of shots) of `frame`, sharing its memory.  r must start on a 64-shot
    word and end on one or at the last shot; a native pftrajectories on the
    view keys its draws on the view's absolute shots (the offset of r within
    the parent), so r must then also start on a multiple of 2048 shots."""
    if isinstance(r, slice):
        if r.step not in (None, 1):
            raise _lib.ArgumentError("view(frame, r) needs a contiguous range of shots")
        start, stop, _ = r.indices(frame.shots)
    else:
        if getattr(r, "step", 1) != 1:
            raise _lib.ArgumentError("view(frame, r) needs a contiguous range of shots")
        start, stop = r.start, r.stop
    if not 0 <= start <= stop <= frame.shots:
        raise _lib.BoundsError(f"view range {start}:{stop} outside 0:{frame.shots}")
    if start % 64 or (stop % 64 and stop != frame.shots):
        raise _lib.ArgumentError(
            "view(frame, r): r must start on a multiple of 64 shots and end on one or at "
            f"the last shot (got {start}:{stop})")
    v = PauliFrame.__new__(PauliFrame)
    v.shots, v.nq, v.nbits, v.ld = stop - start, frame.nq, frame.nbits, frame.ld
    v.offset, v.w0, v.is_view = frame.offset + start, frame.w0 + start // 64, True
    fb = frame.frame._qcf_base if frame.is_view else frame.frame
    rb = frame.record._qcf_base if frame.is_view else frame.record
    v.frame = fb[v.w0:]
    v.frame._qcf_base = fb
    v.record = rb[v.w0:]
    v.record._qcf_base = rb
    return v


def initZ_(frame: PauliFrame, seed: int | None = None) -> PauliFrame:
    """initZ!: x = 0, z uniformly random on every frame qubit (src/pauli_frames.jl:47-56)."""
    prog = Program([], nqubits=frame.nq - 1, nbits=frame.nbits)
    # initZ! leaves the record alone
    record = torch.empty(max(1, frame.nbits) * frame.ld, dtype=torch.int64,
                         device=frame.frame.device)
    call("qcf_pf_sample", prog.handle, frame.shots, frame.offset,
         _SEEDS.next() if seed is None else seed, _ptr(frame.frame), _ptr(record), None, frame.ld,
         1 | (4 if frame.is_view else 0), _stream())
    return frame


def pftrajectories_state(state: PauliFrame, circuit, *, seed: int | None = None,
                         masks: torch.Tensor | None = None) -> PauliFrame:
    """pftrajectories(state::PauliFrame, circuit) (src/pauli_frames.jl:234-239)."""
    prog = circuit if isinstance(circuit, Program) else Program(
        circuit, nqubits=state.nq - 1, nbits=state.nbits)
    if prog.nq != state.nq or prog.nbits > state.nbits:
        raise _lib.BoundsError(
            f"circuit needs {prog.nq} frame qubits / {prog.nbits} bits; frame has "
            f"{state.nq} / {state.nbits}")
    call("qcf_pf_sample", prog.handle, state.shots, state.offset,
         _SEEDS.next() if seed is None else seed, _ptr(state.frame), _ptr(state.record),
         _ptr(masks), state.ld, 4 if state.is_view else 0, _stream())
    return state


def pftrajectories(circuit, trajectories: int = 5000, threads: bool = True, *,
                   seed: int | None = None, masks: torch.Tensor | None = None,
                   export_masks: torch.Tensor | None = None, keep_frame: bool = True,
                   device=None, shot_offset: int = 0) -> PauliFrame:
    """pftrajectories(circuit; trajectories, threads) (src/pauli_frames.jl:193-227):
    a fresh frame sized from the circuit, initZ!, then the circuit.  `threads` is
    accepted for signature compatibility (the GPU runs all shots in parallel).
    `masks` replays pre-drawn randomness; `export_masks` records the native draws.
    `shot_offset` (a multiple of 2048) places these shots in a longer run: the
    draws are those that shots shot_offset.. of one run with this seed make."""
    if isinstance(circuit, PauliFrame):  # pftrajectories(state, circuit)
        return pftrajectories_state(circuit, trajectories, seed=seed, masks=masks)
    prog = circuit if isinstance(circuit, Program) else Program(circuit)
    f = PauliFrame(trajectories, prog.nq - 1, prog.nbits, device=device, init=False,
                   _uninitialized=True)
    if not keep_frame:
        f.frame = torch.empty(0, dtype=torch.int64, device=f.record.device)
    flags = 1
    mptr = masks
    if export_masks is not None:
        flags |= 2
        mptr = export_masks
    f.offset = int(shot_offset)
    call("qcf_pf_sample", prog.handle, f.shots, f.offset, _SEEDS.next() if seed is None else seed,
         _ptr(f.frame) if keep_frame else None, _ptr(f.record), _ptr(mptr), f.ld, flags,
         _stream())
    return f


def apply_(frame: PauliFrame, op: Operation, *, seed: int | None = None) -> PauliFrame:
    """apply!(frame::PauliFrame, op) for any supported operation
    (src/pauli_frames.jl:58-134): a one-operation program on the device."""
    return pftrajectories_state(frame, [op], seed=seed)


def applynoise_(frame: PauliFrame, noise, i: int, *, seed: int | None = None) -> PauliFrame:
    """applynoise!(frame::PauliFrame, noise, i) (src/pauli_frames.jl:136-174):
    the noise channel on qubit i of every frame."""
    from .ops import NoiseOp
    return pftrajectories_state(frame, [NoiseOp(noise, (i,))], seed=seed)


def alloc_masks(prog: Program, shots: int, device=None) -> torch.Tensor:
    dev = torch.device(device) if device is not None else torch.device(
        "cuda", torch.cuda.current_device())
    return torch.zeros(max(1, prog.nreplay) * plane_words(shots), dtype=torch.int64, device=dev)


def _relative_measurements(frame: PauliFrame) -> torch.Tensor:
    out = torch.empty(max(1, frame.nbits) * max(1, frame.shots), dtype=torch.uint8,
                      device=frame.record.device)
    call("qcf_record_to_bool", _ptr(frame.record), frame.ld, frame.shots, frame.nbits,
         _ptr(out), 0, _stream())
    m = out[: frame.nbits * frame.shots].view(frame.nbits, frame.shots).t()
    return m.bool()


def measurements(*args) -> torch.Tensor:
    """measurements(frame): shots x bits Bool matrix of flips relative to the
    reference measurements, column-major like the reference's Matrix{Bool}
    (src/pauli_frames.jl:280).  measurements(register, frame): the absolute
    outcomes, the register's measurement bits XOR each shot's flips
    (src/pauli_frames.jl:289); `register` is the noiseless reference run's bit
    vector (the reference's bitview(register)), one Bool per classical bit."""
    if len(args) == 1:
        return _relative_measurements(args[0])
    if len(args) != 2:
        raise TypeError("measurements(frame) or measurements(register, frame)")
    register, frame = args
    bits = torch.as_tensor(np.asarray(register, dtype=bool), device=frame.record.device)
    if bits.numel() != frame.nbits:
        raise DimensionMismatch(
            f"register holds {bits.numel()} bits, the frame records {frame.nbits}")
    return _relative_measurements(frame) ^ bits.view(1, -1)


pfmeasurements = measurements  # deprecated alias (src/pauli_frames.jl:291-293)


def frame_to_reference(frame: PauliFrame) -> torch.Tensor:
    """The reference `fastcolumn` xzs: shots x 2W UInt64, returned as (2W, shots)."""
    W = (frame.nq + 63) // 64
    out = torch.empty(2 * W * max(1, frame.shots), dtype=torch.int64, device=frame.frame.device)
    call("qcf_frame_to_reference", _ptr(frame.frame), frame.ld, frame.shots, frame.nq,
         _ptr(out), 0, _stream())
    return out.view(2 * W, frame.shots)


def frame_from_reference(xzs: torch.Tensor, frame: PauliFrame) -> PauliFrame:
    call("qcf_frame_from_reference", _ptr(xzs.contiguous()), 0, frame.shots, frame.nq,
         _ptr(frame.frame), frame.ld, _stream())
    return frame
