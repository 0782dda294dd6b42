"""Pins the CPU oracle's Pauli product against the reference's known answers,
brute-force matrix products and algebraic identities (no GPU)."""
import json
import os

import numpy as np
import pytest

import oracle
from paulihelp import from_bits, from_str, random_pauli, random_words

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _mul(l, r):
    lp, lx = l
    rp, rx = r
    rx = rx.copy()
    rph = np.array([rp], np.uint8)
    oracle.mul_left_pauli(rx, rph, lx.copy(), np.array([lp], np.uint8))
    return int(rph[0]), rx


def test_reference_known_answers():
    kat = json.load(open(os.path.join(GOLD, "pauli_kat.json")))
    for e in kat["xz"]:
        ph, xz = from_str(e["s"])
        assert xz.tolist() == e["xz"] and ph == e["phase"]
    for e in kat["products"]:
        ph, xz = _mul(from_str(e["l"]), from_str(e["r"]))
        eph, exz = from_str(e["lr"])
        assert ph == eph and np.array_equal(xz, exz), e
    for e in kat["prodphase"]:
        assert oracle.prodphase(from_str(e["l"])[1], from_str(e["r"])[1]) == e["phase"]
    for e in kat["comm"]:
        assert oracle.comm(from_str(e["l"])[1], from_str(e["r"])[1]) == e["comm"]


def test_brute_force_products():
    cases = json.load(open(os.path.join(GOLD, "pauli_brute.json")))
    for c in cases:
        l = from_bits(*c["l"])
        r = from_bits(*c["r"])
        ph, xz = _mul(l, r)
        eph, exz = from_bits(*c["lr"])
        assert ph == eph and np.array_equal(xz, exz), c


@pytest.mark.parametrize("n", [1, 2, 10, 63, 64, 65, 127, 128, 129])
def test_identities(n):
    """test/test_mul_leftright.jl:2-22 and test/test_paulis.jl:48-66 identities."""
    rng = np.random.default_rng(n)
    for _ in range(20):
        p1, p2 = random_pauli(rng, n), random_pauli(rng, n)
        # product phase of the bit arrays: mul_left! vs the independent prodphase
        r = p2[1].copy()
        s = oracle.mul_left_vec(r, p1[1].copy())
        assert s == oracle.prodphase(p1[1], p2[1])
        # commute <=> the bare product phase is real
        assert (s % 2) == oracle.comm(p1[1], p2[1])
        # mul_left!(p2, p1) == (-1)^comm(p1,p2) * mul_right!(p2, p1)
        a = p2[1].copy()
        aph = np.array([p2[0]], np.uint8)
        oracle.mul_left_pauli(a, aph, p1[1].copy(), np.array([p1[0]], np.uint8))
        b = p2[1].copy()
        bph = np.array([p2[0]], np.uint8)
        oracle.mul_right_pauli(b, bph, p1[1].copy(), np.array([p1[0]], np.uint8))
        assert np.array_equal(a, b)
        assert int(aph[0]) == (int(bph[0]) + 2 * oracle.comm(p1[1], p2[1])) % 4


def test_counter_split_invariance():
    """The phase is a sum of per-bit mod-4 counters, so any split of the words
    gives the same total (what lets the GPU split words across lanes)."""
    rng = np.random.default_rng(7)
    for nw in (1, 5, 33, 100):
        l, r = random_words(rng, nw), random_words(rng, nw)
        whole = oracle.mul_left_vec(r.copy(), l)
        k = nw // 2
        total = 0
        for a, b in ((0, k), (k, nw)):
            if a == b:
                continue
            rr = np.concatenate([r[a:b], r[nw + a:nw + b]])
            ll = np.concatenate([l[a:b], l[nw + a:nw + b]])
            c1 = np.zeros(1, np.uint64)
            c2 = np.zeros(1, np.uint64)
            oracle.lib().oracle_mul_ordered(rr.ctypes.data, ll.ctypes.data, b - a, 1,
                                            c1.ctypes.data, c2.ctypes.data)
            total += int(c1[0]) + 2 * int(c2[0])
        assert total % 4 == whole


def test_self_product_is_identity():
    rng = np.random.default_rng(3)
    p = random_pauli(rng, 77)
    r = p[1].copy()
    assert oracle.mul_left_vec(r, p[1].copy()) == 0
    assert not r.any()


def test_rows_sweep_matches_single_rows():
    rng = np.random.default_rng(11)
    nrows, nw = 9, 3
    xzs = np.stack([random_words(rng, nw) for _ in range(nrows)])
    ph = rng.integers(0, 4, nrows).astype(np.uint8)
    a, aph = xzs.copy(), ph.copy()
    oracle.mul_rows(a, aph, a[4].copy(), aph[4:5].copy(), skip=4)
    for m in range(nrows):
        if m == 4:
            continue
        r = xzs[m].copy()
        s = oracle.mul_left_vec(r, xzs[4].copy())
        assert np.array_equal(a[m], r)
        assert aph[m] == (s + ph[m] + ph[4]) % 4


def test_philox_known_answers():
    """Random123 known-answer vectors for philox4x32-10 (kat_vectors)."""
    kats = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, want in kats:
        assert oracle.philox4x32_10(ctr, key).tolist() == want
