"""The oracle's canonicalize! (src/canonicalization.jl:97-135) against the
reference's own known answers and against what the mathematics fixes: the
canonical generators (with their signs) are elements of the input's group and
generate all of it (2^n x 2^n matrix products, n <= 4), and the canonical form
depends only on the group, not on the generating set."""
import json
import os

import numpy as np
import pytest

import oracle
from paulihelp import from_str
from qcf import randoms

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "canon_kat.json")
PH_STR = {0: "+ ", 1: "+i", 2: "- ", 3: "-i"}


def tab_from_strings(rows):
    ph, xz = zip(*(from_str(r) for r in rows))
    return np.array(ph, np.uint8), np.ascontiguousarray(np.stack(xz))


def tab_strings(ph, xzs, n):
    W = xzs.shape[1] // 2
    out = []
    for m in range(xzs.shape[0]):
        xs = np.unpackbits(xzs[m, :W].view(np.uint8), bitorder="little")[:n]
        zs = np.unpackbits(xzs[m, W:].view(np.uint8), bitorder="little")[:n]
        out.append(PH_STR[int(ph[m])] + "".join("_XZY"[x | (z << 1)] for x, z in zip(xs, zs)))
    return out


def load_cases():
    with open(GOLDEN) as fh:
        return json.load(fh)["cases"]


@pytest.mark.parametrize("case", load_cases(), ids=lambda c: c["source"])
def test_oracle_canonicalize_known_answers(case):
    ph, xzs = tab_from_strings(case["in"])
    n = len([c for c in case["in"][0] if c in "_IXYZ"])
    ranks = oracle.canonicalize(xzs, ph, n, case["phases"])
    assert tab_strings(ph, xzs, n) == case["out"]
    if "ranks" in case:
        assert list(ranks) == case["ranks"]


_MAT = {0: np.eye(2), 1: np.array([[0, 1], [1, 0]]), 2: np.diag([1, -1]),
        3: np.array([[0, -1j], [1j, 0]])}


def pauli_matrix(ph, xz_row, n):
    W = len(xz_row) // 2
    xs = np.unpackbits(xz_row[:W].view(np.uint8), bitorder="little")[:n]
    zs = np.unpackbits(xz_row[W:].view(np.uint8), bitorder="little")[:n]
    m = np.array([[1j ** int(ph)]])
    for x, z in zip(xs, zs):
        m = np.kron(m, _MAT[int(x) | (int(z) << 1)])
    return m


def group_elements(ph, xzs, n):
    mats = [pauli_matrix(p, r, n) for p, r in zip(ph, xzs)]
    elems = [np.eye(2 ** n, dtype=complex)]
    for g in mats:
        elems = elems + [g @ e for e in elems]
    return elems


def contains(elems, m):
    return any(np.allclose(e, m) for e in elems)


@pytest.mark.parametrize("n", [1, 2, 3, 4])
@pytest.mark.parametrize("seed", range(6))
def test_canonical_rows_generate_the_same_group(n, seed):
    rng = np.random.default_rng(seed)
    ph, xzs = randoms.random_clifford_tableau(rng, n, 6 * n)
    nrows = n if seed % 2 == 0 else max(1, n - 1)
    ph, xzs = ph[:nrows].copy(), xzs[:nrows].copy()
    g_in = group_elements(ph, xzs, n)
    c_ph, c_xz = ph.copy(), xzs.copy()
    oracle.canonicalize(c_xz, c_ph, n)
    for p, r in zip(c_ph, c_xz):
        assert contains(g_in, pauli_matrix(p, r, n))
    g_out = group_elements(c_ph[np.any(c_xz != 0, axis=1)], c_xz[np.any(c_xz != 0, axis=1)], n)
    assert all(contains(g_out, e) for e in g_in)


def recombine(rng, ph, xzs):
    """Another generating set of the same group: random row products (row m :=
    row i * row m with the product's phase) and a random row order."""
    ph, xzs = ph.copy(), xzs.copy()
    r = len(ph)
    for _ in range(4 * r):
        i, m = rng.integers(0, r, 2)
        if i == m:
            continue
        oracle.mul_rows(xzs[m:m + 1], ph[m:m + 1], xzs[i], ph[i:i + 1])
    perm = rng.permutation(r)
    return ph[perm].copy(), np.ascontiguousarray(xzs[perm])


@pytest.mark.parametrize("n", [5, 63, 64, 65, 130])
def test_canonical_form_depends_only_on_the_group(n):
    rng = np.random.default_rng(n)
    ph, xzs = randoms.random_clifford_tableau(rng, n)
    ph2, xzs2 = recombine(rng, ph, xzs)
    assert not np.array_equal(xzs, xzs2)
    r1 = oracle.canonicalize(xzs, ph, n)
    r2 = oracle.canonicalize(xzs2, ph2, n)
    assert r1 == r2 and r1[1] == n
    assert np.array_equal(xzs, xzs2) and np.array_equal(ph, ph2)


def test_dense_generator_gives_dense_valid_stabilizers():
    """random_stabilizer_layers: commuting rows with real phases, full rank, and
    about 3/4 of the entries non-identity (a random stabilizer's density)."""
    n = 200
    ph, xzs = randoms.random_stabilizer_layers(np.random.default_rng(1), n)
    assert set(np.unique(ph).tolist()) <= {0, 2}
    assert all(oracle.comm(xzs[i], xzs[j]) == 0 for i in range(0, n, 7) for j in range(0, n, 5))
    W = xzs.shape[1] // 2
    xs = np.unpackbits(xzs[:, :W].view(np.uint8), axis=1, bitorder="little")[:, :n]
    zs = np.unpackbits(xzs[:, W:].view(np.uint8), axis=1, bitorder="little")[:, :n]
    assert abs(((xs | zs) != 0).mean() - 0.75) < 0.02
    assert oracle.canonicalize(xzs.copy(), ph.copy(), n)[1] == n


def test_phases_off_keeps_phases_and_matches_bits():
    rng = np.random.default_rng(3)
    ph, xzs = randoms.random_clifford_tableau(rng, 70)
    ph_on, xz_on = ph.copy(), xzs.copy()
    ph_off, xz_off = ph.copy(), xzs.copy()
    assert oracle.canonicalize(xz_on, ph_on, 70) == oracle.canonicalize(xz_off, ph_off, 70, False)
    assert np.array_equal(xz_on, xz_off)
    assert np.array_equal(ph_off, ph)
