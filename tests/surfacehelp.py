"""Stabilizer-algebra invariants of the noiseless unrotated surface-code
benchmark circuit, for records as 0/1 bits or as shot-packed words."""
import numpy as np


def surface_code_check_bits(ops, d):
    """Record-bit bookkeeping of the benchmark circuit: for every measurement
    qubit the bits of its rounds in order, and the final data bits."""
    L = 2 * d - 1
    per_q = {}
    for o in ops:
        if type(o).__name__ == "sMZ":
            per_q.setdefault(o.qubit, []).append(o.bit)

    def rc(q):
        return divmod(q - 1, L)

    xm = {q: b for q, b in per_q.items() if rc(q)[0] % 2 == 1 and rc(q)[1] % 2 == 0}
    zm = {q: b for q, b in per_q.items() if rc(q)[0] % 2 == 0 and rc(q)[1] % 2 == 1}
    data = {q: b[0] for q, b in per_q.items() if sum(rc(q)) % 2 == 0}
    zcheck_data = {}
    for q in zm:
        r, c = rc(q)
        zcheck_data[q] = [L * (r + dr) + (c + dc) + 1 for dr, dc in ((0, -1), (-1, 0), (1, 0), (0, 1))
                          if 0 <= r + dr < L and 0 <= c + dc < L]
    return xm, zm, data, zcheck_data


def check_noiseless_invariants(rec, ops, d):
    """rec: (nbits, shots) 0/1 record flips of the noiseless benchmark circuit.
    Data start in |0>, so every Z check is deterministic from its first
    measurement (round 2) on, every X check repeats its round-1 value, and
    each Z check's last value equals the parity of its final data bits; the
    round-1 X checks are uniformly random."""
    xm, zm, data, zdat = surface_code_check_bits(ops, d)
    for q, bits in zm.items():
        assert not rec[np.array(bits) - 1].any(), q
        par = rec[bits[-1] - 1].copy()
        for dq in zdat[q]:
            par ^= rec[data[dq] - 1]
        assert not par.any(), q
    first = []
    for q, bits in xm.items():
        b = np.array(bits) - 1
        assert not (rec[b[1:]] ^ rec[b[:-1]]).any(), q
        first.append(rec[b[0]])
    return np.array(first)
