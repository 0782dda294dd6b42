"""GPU canonicalize! against the oracle, bit for bit on rows, phases and ranks
(the reference's test/test_gpu_canonicalization.jl compares the CUDA result
with the CPU one the same way; sizes of test/test_stabcanon.jl:3)."""
import json
import os
import time

import numpy as np
import pytest
import torch

import oracle
import qcf
from qcf import randoms

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "canon_kat.json")


def gpu_canon(ph, xzs, n, phases=True):
    t = qcf.Tableau(ph, n, xzs)
    _, rx, rz = qcf.canonicalize_(t, phases=phases, ranks=True)
    return t.phases_host(), t.xzs_host(), (rx, rz)


def check(ph, xzs, n, phases=True):
    gph, gxz, granks = gpu_canon(ph, xzs, n, phases)
    oph, oxz = ph.copy(), xzs.copy()
    oranks = oracle.canonicalize(oxz, oph, n, phases)
    assert granks == oranks
    assert np.array_equal(gxz, oxz)
    assert np.array_equal(gph, oph)
    return granks


def test_known_answers(gpu):
    with open(GOLDEN) as fh:
        cases = json.load(fh)["cases"]
    for case in cases:
        t = qcf.S("\n".join(case["in"]))
        _, rx, rz = qcf.canonicalize_(t, phases=case["phases"], ranks=True)
        assert t.row_strings() == case["out"], case["source"]
        if "ranks" in case:
            assert [rx, rz] == case["ranks"]


@pytest.mark.parametrize("n", [1, 2, 10, 63, 64, 65, 127, 128, 129])
def test_random_stabilizers(gpu, n):
    rng = np.random.default_rng(n)
    print(f"canonicalize! random stabilizer n={n}", flush=True)
    for nrows in sorted({n, int(rng.integers(n // 3, 2 * n // 3 + 1)) or 1}):
        ph, xzs = randoms.random_clifford_tableau(rng, n)
        ph, xzs = ph[:nrows].copy(), np.ascontiguousarray(xzs[:nrows])
        check(ph, xzs, n)
        check(ph, xzs, n, phases=False)


@pytest.mark.parametrize("n", [64, 333, 1000])
def test_dense_random_stabilizers(gpu, n):
    """Dense random stabilizers (3/4 of the entries non-identity, like the
    reference's random_stabilizer), full and half rank, phases on and off."""
    ph, xzs = randoms.random_stabilizer_layers(np.random.default_rng(n + 7), n)
    print(f"canonicalize! dense random stabilizer n={n}", flush=True)
    check(ph, xzs, n)
    check(ph, xzs, n, phases=False)
    check(ph[: n // 2].copy(), np.ascontiguousarray(xzs[: n // 2]), n)


@pytest.mark.parametrize("nrows,n", [(300, 500), (700, 200), (2048 + 3, 2048 + 3), (1, 70)])
def test_random_tableaux_ragged(gpu, nrows, n):
    """random_tableau(nrows, n): rows need not commute; more or fewer rows than
    qubits, rows around the 1024-per-register boundary of the panel kernel."""
    rng = np.random.default_rng(nrows + n)
    ph, xzs = randoms.random_tableau(rng, nrows, n)
    ph = rng.integers(0, 4, nrows).astype(np.uint8)
    print(f"canonicalize! random tableau {nrows}x{n}", flush=True)
    check(ph, xzs, n)


def test_zero_and_duplicate_rows(gpu):
    rng = np.random.default_rng(5)
    ph, xzs = randoms.random_clifford_tableau(rng, 90)
    xzs[10] = 0
    xzs[20] = xzs[30]
    ph[20] = ph[30]
    check(ph, xzs, 90)


def test_6000_qubits_parity_and_time(gpu):
    """test/test_gpu_canonicalization.jl:19-29: Stabilizer(random_tableau(6000,
    6000)) canonicalizes in under 1900 ms on the GPU (the CPU takes 1900-2400 ms
    according to that test); here also bit-exact against the oracle."""
    n = 6000
    ph, xzs = randoms.random_tableau(np.random.default_rng(6000), n, n)
    print("canonicalize! 6000 x 6000", flush=True)
    t = qcf.Tableau(ph, n, xzs)
    qcf.canonicalize_(t.copy())  # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    qcf.canonicalize_(t)
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0)
    print(f"  {ms:.2f} ms", flush=True)
    assert ms < 1900
    oph, oxz = ph.copy(), xzs.copy()
    oracle.canonicalize(oxz, oph, n)
    assert np.array_equal(t.xzs_host(), oxz)
    assert np.array_equal(t.phases_host(), oph)


def test_rejects_bad_shapes(gpu):
    t = qcf.Tableau(np.zeros(3, np.uint8), 70, np.zeros((3, 4), np.uint64))
    t.nqubits = 200  # more qubits than two words hold
    with pytest.raises(qcf.DimensionMismatch):
        qcf.canonicalize_(t)
