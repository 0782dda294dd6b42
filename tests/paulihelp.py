"""Small numpy helpers shared by the Pauli tests (bits <-> UInt64 words)."""
import numpy as np

_PH = {"": 0, "+": 0, "i": 1, "+i": 1, "-": 2, "-i": 3}


def pack(bits, nwords):
    out = np.zeros(nwords * 64, np.uint8)
    out[: len(bits)] = bits
    return np.packbits(out.reshape(-1, 8), axis=1, bitorder="little").reshape(-1).view(np.uint64)


def unpack(words, n):
    return np.unpackbits(np.ascontiguousarray(words, np.uint64).view(np.uint8),
                         bitorder="little")[:n]


def from_bits(phase, xs, zs):
    n = len(xs)
    w = (n + 63) // 64
    return (np.uint8(phase), np.concatenate([pack(xs, w), pack(zs, w)]))


def from_str(s):
    letters = [c for c in s if c in "_IZXY"]
    ph = _PH["".join(c for c in s if c not in "_IZXY").strip()]
    return from_bits(ph, [int(c in "XY") for c in letters], [int(c in "ZY") for c in letters])


def random_pauli(rng, n):
    """random_pauli(n) semantics (src/randoms.jl:47-52): i.i.d. x/z bits,
    phase uniform in 0..3, pad bits cleared."""
    return from_bits(int(rng.integers(4)), rng.integers(0, 2, n), rng.integers(0, 2, n))


def random_words(rng, nwords):
    return rng.integers(0, 2**64, size=2 * nwords, dtype=np.uint64)
