"""mul_left! / mul_right! on the GPU against the oracle: bits and phases
bit-exact (reference tests: test/test_mul_leftright.jl, test/test_paulis.jl,
test/KernelAbstractions/implementation/test_KA_mul_leftright.jl)."""
import json
import os

import numpy as np
import pytest
import torch

import oracle
import qcf
from paulihelp import from_bits, from_str, random_pauli
from qcf import randoms

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def dev_pauli(ph, xz, n):
    return qcf.PauliOperator(ph, n, xz)


def test_known_answers(gpu):
    kat = json.load(open(os.path.join(GOLD, "pauli_kat.json")))
    for e in kat["products"]:
        got = qcf.P(e["l"]) * qcf.P(e["r"])
        assert got == qcf.P(e["lr"]), e
    cases = json.load(open(os.path.join(GOLD, "pauli_brute.json")))
    for c in cases:
        l = qcf.PauliOperator(*from_bits(*c["l"])[:1], c["n"], from_bits(*c["l"])[1])
        r = qcf.PauliOperator(*from_bits(*c["r"])[:1], c["n"], from_bits(*c["r"])[1])
        eph, exz = from_bits(*c["lr"])
        out = qcf.mul_left_(r, l)
        assert out.phase_value() == eph and np.array_equal(out.xz_host(), exz), c


SIZES = [1, 2, 10, 31, 63, 64, 65, 127, 128, 129] + [64 * k for k in (1023, 1024, 1025, 2047, 2048, 2049)]


@pytest.mark.parametrize("n", SIZES)
def test_mul_left_right_random(gpu, n):
    rng = np.random.default_rng(1000 + n)
    for _ in range(5):
        (p1, x1), (p2, x2) = random_pauli(rng, n), random_pauli(rng, n)
        # left
        r = dev_pauli(p2, x2, n)
        qcf.mul_left_(r, dev_pauli(p1, x1, n))
        ox, oph = x2.copy(), np.array([p2], np.uint8)
        oracle.mul_left_pauli(ox, oph, x1.copy(), np.array([p1], np.uint8))
        assert np.array_equal(r.xz_host(), ox) and r.phase_value() == int(oph[0])
        # right
        l = dev_pauli(p2, x2, n)
        qcf.mul_right_(l, dev_pauli(p1, x1, n))
        ox, oph = x2.copy(), np.array([p2], np.uint8)
        oracle.mul_right_pauli(ox, oph, x1.copy(), np.array([p1], np.uint8))
        assert np.array_equal(l.xz_host(), ox) and l.phase_value() == int(oph[0])
        # phases = false is a plain XOR
        r = dev_pauli(p2, x2, n)
        qcf.mul_left_(r, dev_pauli(p1, x1, n), phases=False)
        assert np.array_equal(r.xz_host(), x1 ^ x2) and r.phase_value() == p2


def test_self_alias(gpu):
    rng = np.random.default_rng(5)
    ph, xz = random_pauli(rng, 300)
    p = dev_pauli(ph, xz, 300)
    call = qcf._lib.call
    call("qcf_mul_left", qcf.pauli._ptr(p.xz), qcf.pauli._ptr(p.phase), qcf.pauli._ptr(p.xz),
         qcf.pauli._ptr(p.phase), p.nwords, 1, qcf.pauli._stream())
    assert not p.xz_host().any() and p.phase_value() == (2 * ph) % 4


def test_dimension_mismatch(gpu):
    with pytest.raises(qcf.DimensionMismatch):
        qcf.mul_left_(qcf.P("XX"), qcf.P("XXX"))


@pytest.mark.parametrize("n", [10**6, 10**8, 10**9])
def test_large_operators(gpu, n):
    """The benchmark operand sizes (1 Mqubit and 1 Gqubit) and one between."""
    rng = np.random.default_rng(n % 1000)
    p1, x1 = randoms.random_pauli_words(rng, n)
    p2, x2 = randoms.random_pauli_words(rng, n)
    r = dev_pauli(p2, x2, n)
    qcf.mul_left_(r, dev_pauli(p1, x1, n))
    oph = np.array([p2], np.uint8)
    oracle.mul_left_pauli(x2, oph, x1, np.array([p1], np.uint8))
    assert r.phase_value() == int(oph[0])
    assert np.array_equal(r.xz_host(), x2)


def test_vector_form(gpu):
    rng = np.random.default_rng(9)
    for nw in (1, 3, 16, 1001):
        a = rng.integers(0, 2**64, 2 * nw, dtype=np.uint64)
        b = rng.integers(0, 2**64, 2 * nw, dtype=np.uint64)
        for right in (False, True):
            ta = torch.from_numpy(a.view(np.int64).copy()).cuda()
            tb = torch.from_numpy(b.view(np.int64).copy()).cuda()
            s = qcf.mul_left_vec_(ta, tb, right=right)
            oa = a.copy()
            os_ = oracle.mul_right_vec(oa, b.copy()) if right else oracle.mul_left_vec(oa, b.copy())
            assert s == os_ and np.array_equal(ta.cpu().numpy().view(np.uint64), oa)


@pytest.mark.parametrize("n", [5, 64, 1000, 1500, 10000])
def test_pivot_sweep_and_tableau_pauli(gpu, n):
    """Pivot sweeps (plain and canonicalize!-style masked), tableau x Pauli and
    the single-row form on random-Clifford stabilizer tableaux, including the
    10,000-qubit tableau of the mul_left! benchmark configuration
    (src/canonicalization.jl:105-109, src/mul_leftright.jl:177-181,195-198)."""
    rng = np.random.default_rng(n)
    ph, xzs = randoms.random_clifford_tableau(rng, n, ngates=3 * n)
    t = qcf.Tableau(ph, n, xzs)
    pivot = int(rng.integers(1, n + 1))
    qcf.mul_left_sweep_(t, pivot)
    oxzs, oph = xzs.copy(), ph.copy()
    oracle.mul_rows(oxzs, oph, oxzs[pivot - 1].copy(), oph[pivot - 1:pivot].copy(), skip=pivot - 1)
    assert np.array_equal(t.xzs_host(), oxzs) and np.array_equal(t.phases_host(), oph)
    # masked (canonicalize!-style: only rows with a bit set in the pivot column)
    t2 = qcf.Tableau(ph, n, xzs)
    mask = rng.integers(0, 2, n).astype(np.uint8)
    qcf.mul_left_sweep_(t2, pivot, row_mask=torch.from_numpy(mask).cuda())
    oxzs, oph = xzs.copy(), ph.copy()
    oracle.mul_rows(oxzs, oph, oxzs[pivot - 1].copy(), oph[pivot - 1:pivot].copy(),
                    skip=pivot - 1, row_mask=mask)
    assert np.array_equal(t2.xzs_host(), oxzs) and np.array_equal(t2.phases_host(), oph)
    # mul_left!(s::Tableau, p::PauliOperator)
    pp, pxz = randoms.random_pauli_words(rng, n)
    t3 = qcf.Tableau(ph, n, xzs)
    qcf.mul_left_tableau_pauli_(t3, qcf.PauliOperator(pp, n, pxz))
    oxzs, oph = xzs.copy(), ph.copy()
    oracle.mul_rows(oxzs, oph, pxz.copy(), np.array([pp], np.uint8))
    assert np.array_equal(t3.xzs_host(), oxzs) and np.array_equal(t3.phases_host(), oph)
    # single row form mul_left!(s, m, i)
    t4 = qcf.Tableau(ph, n, xzs)
    m = 1 if pivot != 1 else 2
    if n > 1:
        qcf.mul_left_row_(t4, m, pivot)
        r = xzs[m - 1].copy()
        s = oracle.mul_left_vec(r, xzs[pivot - 1].copy())
        assert np.array_equal(t4.xzs_host()[m - 1], r)
        assert t4.phases_host()[m - 1] == (s + ph[m - 1] + ph[pivot - 1]) % 4


@pytest.mark.parametrize("per_cu", ["1", "2", "3", "7"])
def test_result_independent_of_launch_shape(gpu, per_cu, monkeypatch):
    """Like the reference's random block and batch sizes for its device kernels
    (test/KernelAbstractions/implementation/definitions.jl:8-13): the product
    and phase do not depend on how many workgroups split the words."""
    monkeypatch.setenv("QCF_MUL_BLOCKS_PER_CU", per_cu)
    rng = np.random.default_rng(77)
    for n in (64 * 2049, 3_000_001):
        (p1, x1), (p2, x2) = random_pauli(rng, n), random_pauli(rng, n)
        r = dev_pauli(p2, x2, n)
        qcf.mul_left_(r, dev_pauli(p1, x1, n))
        ox, oph = x2.copy(), np.array([p2], np.uint8)
        oracle.mul_left_pauli(ox, oph, x1.copy(), np.array([p1], np.uint8))
        assert np.array_equal(r.xz_host(), ox) and r.phase_value() == int(oph[0])


def _tab(rng, nrows, n):
    ph, xzs = randoms.random_tableau(rng, nrows, n)
    return rng.integers(0, 4, nrows).astype(np.uint8), xzs


@pytest.mark.parametrize("n", [63, 130, 4097])
def test_tableau_row_forms(gpu, n):
    """The row forms of src/mul_leftright.jl:151-193 and :234-240 against the
    oracle's Pauli products on the same rows, including the quirk of the
    two-tableau forms (the phase added is s.phases[i], not t.phases[i])."""
    rng = np.random.default_rng(n)
    sph, sxz = _tab(rng, 6, n)
    tph, txz = _tab(rng, 6, n)
    pph, pxz = random_pauli(rng, n)
    W = sxz.shape[1] // 2

    def row_op(dst_xz, dst_ph, src_xz, src_ph, right):
        x, ph = dst_xz.copy(), np.array([dst_ph], np.uint8)
        (oracle.mul_right_pauli if right else oracle.mul_left_pauli)(
            x, ph, src_xz.copy(), np.array([src_ph], np.uint8))
        return x, int(ph[0])

    for right in (False, True):
        # s[m] := t[i] * s[m] (or s[m] * t[i]) with s.phases[i]
        s, t = qcf.Tableau(sph, n, sxz), qcf.Tableau(tph, n, txz)
        (qcf.mul_right_tableau_row_ if right else qcf.mul_left_tableau_row_)(s, 2, t, 5)
        ex, eph = row_op(sxz[1], sph[1], txz[4], sph[4], right)
        assert np.array_equal(s.xzs_host()[1], ex) and int(s.phases_host()[1]) == eph
        assert np.array_equal(np.delete(s.xzs_host(), 1, axis=0), np.delete(sxz, 1, axis=0))
        # s[m] := s[i] * s[m] (or s[m] * s[i])
        s = qcf.Tableau(sph, n, sxz)
        (qcf.mul_right_row_ if right else qcf.mul_left_row_)(s, 3, 6)
        ex, eph = row_op(sxz[2], sph[2], sxz[5], sph[5], right)
        assert np.array_equal(s.xzs_host()[2], ex) and int(s.phases_host()[2]) == eph
        # Pauli := row i * Pauli (or Pauli * row i), with the row's phase
        p = qcf.PauliOperator(pph, n, pxz)
        (qcf.mul_right_pauli_row_ if right else qcf.mul_left_pauli_row_)(p, t, 4)
        ex, eph = row_op(pxz, pph, txz[3], tph[3], right)
        assert np.array_equal(p.xz_host(), ex) and p.phase_value() == eph
    # every row := row * p
    s = qcf.Tableau(sph, n, sxz)
    qcf.mul_right_tableau_pauli_(s, qcf.PauliOperator(pph, n, pxz))
    for m in range(6):
        ex, eph = row_op(sxz[m], sph[m], pxz, pph, True)
        assert np.array_equal(s.xzs_host()[m], ex) and int(s.phases_host()[m]) == eph
    with pytest.raises(qcf.DimensionMismatch):
        qcf.mul_left_pauli_row_(qcf.PauliOperator(0, n + 64, np.zeros(2 * (W + 1), np.uint64)), t, 1)


def test_destabilizer_index_flip(gpu):
    """mul_left!(s::Destabilizer, i, j) (src/mul_leftright.jl:200-206):
    destabilizer j := destabilizer i * destabilizer j without phases, then
    stabilizer i := stabilizer j * stabilizer i with phases."""
    n = 70
    rng = np.random.default_rng(9)
    ph, xzs = _tab(rng, 2 * n, n)
    t = qcf.Tableau(ph, n, xzs)
    i, j = 3, 11
    qcf.mul_left_destabilizer_(t, i, j)
    ex = xzs.copy()
    eph = ph.copy()
    ex[j - 1] ^= xzs[i - 1]  # phases off: a plain XOR, phase byte untouched
    x, p = xzs[i - 1 + n].copy(), np.array([ph[i - 1 + n]], np.uint8)
    oracle.mul_left_pauli(x, p, xzs[j - 1 + n].copy(), np.array([ph[j - 1 + n]], np.uint8))
    ex[i - 1 + n], eph[i - 1 + n] = x, p[0]
    assert np.array_equal(t.xzs_host(), ex) and np.array_equal(t.phases_host(), eph)
