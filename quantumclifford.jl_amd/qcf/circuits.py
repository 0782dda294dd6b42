"""Memory-experiment circuits written in the reference's operation vocabulary.

The reference ships no rotated-surface-code generator; these build the
circuits the benchmarks name, with detectors and the observable expressed as
ClassicalXOR stores (src/misc_ops.jl:198-205) after the measurement record.
"""
from __future__ import annotations

from dataclasses import dataclass, field

from .ops import (ClassicalXOR, Depolarize2, PauliError, sCNOT, sHadamard, sMRZ, sMZ)


@dataclass
class MemoryCircuit:
    ops: list
    nqubits: int
    nmeas: int                      # measurement bits 1..nmeas
    detectors: list                 # list of tuples of measurement bits
    observables: list               # list of tuples of measurement bits
    det_bits: list = field(default_factory=list)  # record bits of detectors
    obs_bits: list = field(default_factory=list)  # record bits of observables
    meta: dict = field(default_factory=dict)
    det_pos: list = field(default_factory=list)   # op index each detector follows

    @property
    def nbits(self) -> int:
        return self.nmeas + len(self.detectors) + len(self.observables)


def _append_xors(c: MemoryCircuit) -> MemoryCircuit:
    """Store detectors and observables as ClassicalXOR right after the op that
    produces their last input (so their inputs stay on chip only briefly);
    their bits follow the measurement bits."""
    b = c.nmeas
    inserts = {}
    pos = c.det_pos or [len(c.ops)] * len(c.detectors)
    for d, at in zip(c.detectors, pos):
        b += 1
        inserts.setdefault(at, []).append(ClassicalXOR(d, b))
        c.det_bits.append(b)
    for o in c.observables:
        b += 1
        inserts.setdefault(len(c.ops), []).append(ClassicalXOR(o, b))
        c.obs_bits.append(b)
    out = []
    for i in range(len(c.ops) + 1):
        out.extend(inserts.get(i, ()))
        if i < len(c.ops):
            out.append(c.ops[i])
    c.ops = out
    return c


def repetition_code(d: int = 3, rounds: int = 3, p: float = 1e-3) -> MemoryCircuit:
    """Distance-d repetition-code memory (data 1..d, ancillas d+1..2d-1): each round
    CNOT data->ancilla with PauliError on both qubits after every CNOT, X-flip
    before every measurement, sMRZ on the ancillas; final sMZ on the data."""
    data = list(range(1, d + 1))
    anc = list(range(d + 1, 2 * d))
    ops, bit = [], 0
    last = {}
    dets, dpos = [], []
    for r in range(rounds):
        for i, a in enumerate(anc):
            for q in (data[i], data[i + 1]):
                ops.append(sCNOT(q, a))
                ops.append(PauliError((q, a), p))
        for a in anc:
            ops.append(PauliError(a, p, 0.0, 0.0))
            bit += 1
            ops.append(sMRZ(a, bit))
            dets.append((bit,) if r == 0 else (bit, last[a]))
            dpos.append(len(ops))
            last[a] = bit
    mdata = {}
    for q in data:
        ops.append(PauliError(q, p, 0.0, 0.0))
        bit += 1
        ops.append(sMZ(q, bit))
        mdata[q] = bit
    for i, a in enumerate(anc):
        dets.append((last[a], mdata[data[i]], mdata[data[i + 1]]))
        dpos.append(len(ops))
    c = MemoryCircuit(ops, 2 * d - 1, bit, dets, [(mdata[data[0]],)],
                      meta=dict(kind="repetition", d=d, rounds=rounds, p=p), det_pos=dpos)
    return _append_xors(c)


def rotated_surface_code_layout(d: int):
    """Data qubits at odd (x, y) in [1, 2d-1]^2; measure qubits at even coordinates
    in [0, 2d]^2, keeping the d^2-1 that form the rotated code (X-type where
    (x/2 + y/2) is odd).  Returns (data, xmeas, zmeas, index) with 1-based
    qubit indices: data first, then measure qubits."""
    data = [(2 * i + 1, 2 * j + 1) for j in range(d) for i in range(d)]
    xm, zm = [], []
    for j in range(d + 1):
        for i in range(d + 1):
            on_b1 = i == 0 or i == d
            on_b2 = j == 0 or j == d
            parity = (i % 2) != (j % 2)
            if on_b1 and parity:
                continue
            if on_b2 and not parity:
                continue
            (xm if parity else zm).append((2 * i, 2 * j))
    index = {}
    for k, q in enumerate(data + xm + zm):
        index[q] = k + 1
    return data, xm, zm, index


def rotated_surface_code(d: int = 5, rounds: int | None = None, p: float = 1e-3,
                         noise: bool = True) -> MemoryCircuit:
    """Rotated surface-code memory-Z experiment: reset all qubits; each round
    H on X-type measure qubits, four CNOT layers in the N-shaped order (X-type
    measure qubits as controls, Z-type as targets), H again, measure-and-reset
    the measure qubits; finally measure the data.  Circuit-level noise at p:
    two-qubit depolarizing after each CNOT, single-qubit depolarizing after each
    H, X-flip after each reset and before each measurement.  Detectors: Z-type
    checks in round 1, every check against the previous round afterwards, and
    Z-type checks against the final data parity; observable: Z on one column of
    data (the logical Z)."""
    rounds = d if rounds is None else rounds
    data, xm, zm, idx = rotated_surface_code_layout(d)
    meas = xm + zm
    dset = set(data)
    # N-shaped orders (avoid hook errors along the logical direction)
    x_order = [(1, 1), (1, -1), (-1, 1), (-1, -1)]
    z_order = [(1, 1), (-1, 1), (1, -1), (-1, -1)]
    ops = []

    def dep1(qs):
        if noise:
            for q in qs:
                ops.append(PauliError(q, p))

    def xflip(qs):
        if noise:
            for q in qs:
                ops.append(PauliError(q, p, 0.0, 0.0))

    all_q = [idx[q] for q in data + meas]
    for q in all_q:
        ops.append(sMRZ(q, 0))
    xflip(all_q)
    bit = 0
    last = {}
    dets, dpos = [], []
    xq = [idx[q] for q in xm]
    mq = [idx[q] for q in meas]
    for r in range(rounds):
        for q in xq:
            ops.append(sHadamard(q))
        dep1(xq)
        for layer in range(4):
            pairs = []
            for m in meas:
                isx = m in xm
                dx, dy = (x_order if isx else z_order)[layer]
                nb = (m[0] + dx, m[1] + dy)
                if nb not in dset:
                    continue
                c, t = (idx[m], idx[nb]) if isx else (idx[nb], idx[m])
                ops.append(sCNOT(c, t))
                pairs.append((c, t))
            # the layer's CNOTs act on disjoint qubits, so the depolarizing
            # channel after each of them may follow the whole layer
            if noise:
                for c, t in pairs:
                    ops.append(Depolarize2(c, t, p))
        for q in xq:
            ops.append(sHadamard(q))
        dep1(xq)
        xflip(mq)
        cur = {}
        for m in meas:
            bit += 1
            ops.append(sMRZ(idx[m], bit))
            cur[m] = bit
            # each detector right after its measurement: the previous round's
            # value is then the only one that has to stay on chip
            if r == 0:
                if m in zm:
                    dets.append((cur[m],))
                    dpos.append(len(ops))
            else:
                dets.append((cur[m], last[m]))
                dpos.append(len(ops))
        xflip(mq)
        last = cur
    xflip([idx[q] for q in data])
    mdata = {}
    for q in data:
        bit += 1
        ops.append(sMZ(idx[q], bit))
        mdata[q] = bit
    for m in zm:
        nbs = [(m[0] + dx, m[1] + dy) for dx, dy in z_order]
        dets.append((last[m], *[mdata[q] for q in nbs if q in dset]))
        dpos.append(len(ops))
    # logical Z: the data line (row y=1 or column x=1) that overlaps every X-type
    # check an even number of times
    obs_q = None
    for line in ([q for q in data if q[1] == 1], [q for q in data if q[0] == 1]):
        ls = set(line)
        if all(sum((m[0] + dx, m[1] + dy) in ls for dx, dy in x_order) % 2 == 0 for m in xm):
            obs_q = line
            break
    assert obs_q is not None, "no data line commutes with the X checks"
    c = MemoryCircuit(ops, len(all_q), bit, dets, [tuple(mdata[q] for q in obs_q)],
                      meta=dict(kind="rotated_memory_z", d=d, rounds=rounds, p=p,
                                ndata=len(data), nmeas_q=len(meas)), det_pos=dpos)
    return _append_xors(c)


# ---- the reference's own benchmark circuits --------------------------------
def unrotated_surface_code_benchmark(d: int = 11, rounds: int = 12) -> list:
    """The noiseless unrotated surface-code circuit the reference benchmarks its
    GPU frame sampler on (benchmark/surface-code-circuit.jl, included by
    benchmark/benchmark-surface-code-gpu.jl:11; d=11, 12 rounds: 441 qubits,
    17,802 operations, 2,751 measurement bits), rebuilt from its structure.

    Qubits sit on an L x L grid, L = 2d-1, qubit = L*row + col + 1: data where
    row+col is even, X-type measure qubits at (odd row, even col), Z-type at
    (even row, odd col).  Every list below is in ascending qubit order.
      * sMRZ(q, 0) on the data, then on the X-type measure qubits;
      * round 1: H on the X-type measure qubits; CNOT layers in the neighbour
        order left, up, down, right, X-type measure qubit as control, data as
        target; H again; sMZ of the X-type measure qubits;
      * rounds 2..rounds: sMRZ(q, 0) on all measure qubits; H; the same four
        layers over all measure qubits, each layer in ascending measure-qubit
        order (Z-type: data as control, measure qubit as target); H; sMZ of
        all measure qubits, and in the last round of the data after them.
    Measurement bits are numbered in order from 1."""
    L = 2 * d - 1

    def q(r, c):
        return L * r + c + 1

    cells = [(r, c) for r in range(L) for c in range(L)]
    data = [q(r, c) for r, c in cells if (r + c) % 2 == 0]
    xm = [(r, c) for r, c in cells if r % 2 == 1 and c % 2 == 0]
    zm = [(r, c) for r, c in cells if r % 2 == 0 and c % 2 == 1]
    meas = sorted(xm + zm, key=lambda rc: q(*rc))
    xq = [q(*m) for m in xm]
    order = [(0, -1), (-1, 0), (1, 0), (0, 1)]
    ops = [sMRZ(x, 0) for x in data] + [sMRZ(x, 0) for x in xq]
    bit = 0
    for r in range(rounds):
        checks = xm if r == 0 else meas
        if r > 0:
            ops += [sMRZ(q(*m), 0) for m in meas]
        ops += [sHadamard(x) for x in xq]
        for dr, dc in order:
            for m in checks:
                nr, nc = m[0] + dr, m[1] + dc
                if not (0 <= nr < L and 0 <= nc < L):
                    continue
                if m[0] % 2 == 1:
                    ops.append(sCNOT(q(*m), q(nr, nc)))
                else:
                    ops.append(sCNOT(q(nr, nc), q(*m)))
        ops += [sHadamard(x) for x in xq]
        targets = [q(*m) for m in checks] + (data if r == rounds - 1 else [])
        for x in targets:
            bit += 1
            ops.append(sMZ(x, bit))
    return ops


def x_diag_circuit(csize: int) -> list:
    """benchmark/benchmarks.jl:103-110: H on qubits 1..csize, each followed by a
    CNOT onto qubit csize+1, which is then measured into bit 1."""
    ops = []
    for i in range(1, csize + 1):
        ops += [sHadamard(i), sCNOT(i, csize + 1)]
    ops.append(sMZ(csize + 1, 1))
    return ops


def x_diag_circuit_noisy_measurement(csize: int) -> list:
    """benchmark/benchmarks.jl:161-171: per i in 1..csize, PauliError(i, 0.1),
    H(i), CNOT(i -> csize+1), sMZ(csize+1, i), then ClassicalXOR of bits
    1..(i mod 6 + 2) into bit i."""
    ops = []
    for i in range(1, csize + 1):
        ops += [PauliError(i, 0.1), sHadamard(i), sCNOT(i, csize + 1), sMZ(csize + 1, i),
                ClassicalXOR(tuple(range(1, i % 6 + 3)), i)]
    return ops
