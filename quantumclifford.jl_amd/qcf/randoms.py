"""Host-side generators for benchmark and test inputs (src/randoms.jl semantics
for random Paulis; a random-Clifford stabilizer tableau for the row sweep)."""
from __future__ import annotations

import numpy as np


def random_pauli_words(rng: np.random.Generator, n: int):
    """random_pauli(n) (src/randoms.jl:47-52, pad bits cleared as :13-18):
    i.i.d. Bernoulli(1/2) x/z bits packed in UInt64 and a phase in {0,1,2,3}.
    Returns (phase, xz words [x;z])."""
    w = (n + 63) // 64
    xz = rng.integers(0, 2**64, size=2 * w, dtype=np.uint64)
    r = n % 64
    if r:
        m = np.uint64((1 << r) - 1)
        xz[w - 1] &= m
        xz[2 * w - 1] &= m
    return int(rng.integers(4)), xz


def random_clifford_tableau(rng: np.random.Generator, n: int, ngates: int | None = None):
    """Stabilizer tableau of a random Clifford circuit applied to |0...0>:
    start from rows Z_i (x = 0, z = e_i, phase 0) and conjugate every row by
    `ngates` random H / S / CNOT gates with the reference's phase rules
    (src/symbolic_cliffords.jl:105-118, 359-395).  Returns (phases uint8[n],
    xzs uint64[n, 2W]) row-contiguous."""
    w = (n + 63) // 64
    ngates = 4 * n if ngates is None else ngates
    # qubit-major bit matrices (xt[q, row]) so that each gate touches
    # contiguous rows of them
    xt = np.zeros((n, n), np.uint8)
    zt = np.eye(n, dtype=np.uint8)
    ph = np.zeros(n, np.uint8)
    kinds = rng.integers(0, 3, ngates)
    qa = rng.integers(0, n, ngates)
    qb = rng.integers(0, n - 1, ngates) if n > 1 else np.zeros(ngates, np.int64)
    for k, a, b in zip(kinds, qa, qb):
        if k == 0:  # H: (z, x), phase if x&z
            ph = (ph + 2 * (xt[a] & zt[a])) & 3
            xt[a], zt[a] = zt[a].copy(), xt[a].copy()
        elif k == 1:  # S: (x, x^z), phase if x&z
            ph = (ph + 2 * (xt[a] & zt[a])) & 3
            zt[a] ^= xt[a]
        elif n > 1:  # CNOT(a, b)
            b = b + (b >= a)
            x1, z1, x2, z2 = xt[a], zt[a], xt[b], zt[b]
            flip = (x1 & z1 & x2 & z2) | (x1 & z2 & (1 - (z1 | x2)))
            ph = (ph + 2 * flip) & 3
            zt[a] = z1 ^ z2
            xt[b] = x2 ^ x1
    xs = np.zeros((n, w * 64), np.uint8)
    zs = np.zeros((n, w * 64), np.uint8)
    xs[:, :n] = xt.T
    zs[:, :n] = zt.T
    px = np.packbits(xs, axis=1, bitorder="little").view(np.uint64)
    pz = np.packbits(zs, axis=1, bitorder="little").view(np.uint64)
    return ph.astype(np.uint8), np.ascontiguousarray(np.concatenate([px, pz], axis=1))


def random_tableau(rng: np.random.Generator, nrows: int, n: int):
    """random_tableau(nrows, n) (src/randoms.jl:94-105 with the default
    nophase=true): i.i.d. Bernoulli(1/2) x/z bits, pad bits cleared, phases 0.
    Rows need not commute.  Returns (phases uint8[nrows], xzs uint64[nrows, 2W])."""
    w = (n + 63) // 64
    xzs = rng.integers(0, 2**64, size=(nrows, 2 * w), dtype=np.uint64)
    r = n % 64
    if r:
        m = np.uint64((1 << r) - 1)
        xzs[:, w - 1] &= m
        xzs[:, 2 * w - 1] &= m
    return np.zeros(nrows, np.uint8), xzs


def random_stabilizer_layers(rng: np.random.Generator, n: int, layers: int | None = None):
    """A dense random stabilizer state of n qubits (the kind random_stabilizer(n)
    returns, src/randoms.jl): |0...0> (rows Z_i) conjugated by `layers` layers,
    each a random H or S (or nothing) on every qubit followed by CNOTs on a
    random pairing of all qubits, with the reference's phase rules
    (src/symbolic_cliffords.jl:105-118, 359-395).  Random pairings scramble in
    O(log n) layers; the default 2*ceil(log2 n) + 8 leaves rows dense.
    Returns (phases uint8[n], xzs uint64[n, 2W]) row-contiguous."""
    w = (n + 63) // 64
    if layers is None:
        layers = 2 * int(np.ceil(np.log2(max(n, 2)))) + 8
    xt = np.zeros((n, n), np.uint8)   # xt[q, row]
    zt = np.eye(n, dtype=np.uint8)
    ph = np.zeros(n, np.int64)
    for _ in range(layers):
        kind = rng.integers(0, 3, n)  # 0: H, 1: S, 2: nothing
        h = np.nonzero(kind == 0)[0]
        sg = np.nonzero(kind == 1)[0]
        ph = (ph + 2 * (xt[h] & zt[h]).sum(axis=0, dtype=np.int64)) & 3
        xt[h], zt[h] = zt[h].copy(), xt[h].copy()
        ph = (ph + 2 * (xt[sg] & zt[sg]).sum(axis=0, dtype=np.int64)) & 3
        zt[sg] ^= xt[sg]
        perm = rng.permutation(n)
        a, b = perm[: n // 2], perm[n // 2: 2 * (n // 2)]
        x1, z1, x2, z2 = xt[a], zt[a], xt[b], zt[b]
        flip = (x1 & z1 & x2 & z2) | (x1 & z2 & (1 - (z1 | x2)))
        ph = (ph + 2 * flip.sum(axis=0, dtype=np.int64)) & 3
        zt[a] = z1 ^ z2
        xt[b] = x2 ^ x1
    xs = np.zeros((n, w * 64), np.uint8)
    zs = np.zeros((n, w * 64), np.uint8)
    xs[:, :n] = xt.T
    zs[:, :n] = zt.T
    px = np.packbits(xs, axis=1, bitorder="little").view(np.uint64)
    pz = np.packbits(zs, axis=1, bitorder="little").view(np.uint64)
    return ph.astype(np.uint8), np.ascontiguousarray(np.concatenate([px, pz], axis=1))
