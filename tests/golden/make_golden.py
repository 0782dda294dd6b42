"""Write the golden vectors used by the oracle tests.

pauli_kat.json      -- known answers held by the reference's own tests and
                       docstrings (test/test_paulis.jl:11,26,38-41,
                       src/QuantumClifford.jl:713-720, 772-783,
                       src/pauli_operator.jl:22-28), entered as data.
pauli_brute.json    -- products of Pauli operators computed by brute force:
                       2^n x 2^n matrix products with numpy, decoded back to
                       (phase, x bits, z bits).  Every single-qubit pair with
                       every phase, plus random operators on 2..5 qubits.

Conventions: a Pauli is i^phase * prod_q sigma(x_q, z_q) with sigma(1,0)=X,
sigma(0,1)=Z, sigma(1,1)=Y; phases 0:+1 1:+i 2:-1 3:-i.
"""
import itertools
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
I2 = np.eye(2, dtype=complex)
X = np.array([[0, 1], [1, 0]], dtype=complex)
Z = np.array([[1, 0], [0, -1]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
SIG = {(0, 0): I2, (1, 0): X, (0, 1): Z, (1, 1): Y}


def matrix(phase, xs, zs):
    m = np.array([[1]], dtype=complex)
    for x, z in zip(xs, zs):
        m = np.kron(m, SIG[(x, z)])
    return (1j ** phase) * m


def decode(m, n):
    for xs in itertools.product((0, 1), repeat=n):
        for zs in itertools.product((0, 1), repeat=n):
            b = matrix(0, xs, zs)
            # m = c * b with |c| = 1 and c in {1, i, -1, -i}
            c = np.trace(b.conj().T @ m) / (2 ** n)
            if abs(abs(c) - 1) < 1e-9 and np.allclose(m, c * b):
                ph = {1: 0, 1j: 1, -1: 2, -1j: 3}[complex(round(c.real), round(c.imag))]
                return ph, list(xs), list(zs)
    raise ValueError("not a Pauli")


def main():
    kat = {
        "xz": [
            {"s": "-iXYZ", "xz": [3, 6], "phase": 3},
            {"s": "-IZXY", "xz": [12, 10], "phase": 2},
        ],
        "products": [
            {"l": "X", "r": "Z", "lr": "-iY"},
            {"l": "ZZZ", "r": "XXX", "lr": "-iYYY"},
            {"l": "XX", "r": "ZZ", "lr": "-YY"},
            {"l": "ZZ", "r": "XX", "lr": "-YY"},
        ],
        "prodphase": [
            {"l": "XX", "r": "YY", "phase": 2},
            {"l": "ZZZ", "r": "XXX", "phase": 3},
            {"l": "XXX", "r": "ZZZ", "phase": 1},
        ],
        "comm": [
            {"l": "XX", "r": "YY", "comm": 0},
            {"l": "ZZ", "r": "XX", "comm": 0},
            {"l": "IZ", "r": "XX", "comm": 1},
        ],
    }
    with open(os.path.join(HERE, "pauli_kat.json"), "w") as f:
        json.dump(kat, f, indent=1)

    rng = np.random.default_rng(20240601)
    cases = []
    for pl, pr in itertools.product(range(4), repeat=2):
        for (xl, zl), (xr, zr) in itertools.product(SIG, repeat=2):
            m = matrix(pl, [xl], [zl]) @ matrix(pr, [xr], [zr])
            ph, xs, zs = decode(m, 1)
            cases.append(dict(n=1, l=[pl, [xl], [zl]], r=[pr, [xr], [zr]], lr=[ph, xs, zs]))
    for n in (2, 3, 4, 5):
        for _ in range(40):
            l = [int(rng.integers(4)), rng.integers(0, 2, n).tolist(), rng.integers(0, 2, n).tolist()]
            r = [int(rng.integers(4)), rng.integers(0, 2, n).tolist(), rng.integers(0, 2, n).tolist()]
            m = matrix(*l) @ matrix(*r)
            ph, xs, zs = decode(m, n)
            cases.append(dict(n=n, l=l, r=r, lr=[ph, xs, zs]))
    with open(os.path.join(HERE, "pauli_brute.json"), "w") as f:
        json.dump(cases, f)


if __name__ == "__main__":
    main()
