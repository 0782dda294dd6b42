"""qcf_canonicalize_noalloc (the entry point that allocates nothing: its
workspace is a fixed device array) against qcf_canonicalize through the
C ABI: identical rows, phases and ranks on two calls in a row, the second on
a smaller tableau (the array keeps the first call's contents), and a clear
error for a tableau whose workspace exceeds the array."""
import ctypes as C

import numpy as np
import pytest
import torch

import qcf
from qcf import _lib, randoms

pytestmark = pytest.mark.gpu


def run(entry, ph, xzs, n):
    t = qcf.Tableau(ph, n, xzs)
    ranks = torch.zeros(2, dtype=torch.int32, device="cuda")
    _lib.call(entry, qcf.pauli._ptr(t.xzs), qcf.pauli._ptr(t.phases), t.ld, t.nwords, t.nrows,
              t.nqubits, 1, qcf.pauli._ptr(ranks), qcf.pauli._stream())
    torch.cuda.synchronize()
    return t.xzs_host(), t.phases_host(), ranks.cpu().numpy()


def test_noalloc_canonicalize_equals_allocating_entry_point(gpu):
    cases = [randoms.random_tableau(np.random.default_rng(0), 1500, 1500) + (1500,),
             randoms.random_stabilizer_layers(np.random.default_rng(1), 300) + (300,)]
    for ph, xzs, n in cases:  # the second call is smaller than the first
        a = run("qcf_canonicalize", ph, xzs, n)
        b = run("qcf_canonicalize_noalloc", ph, xzs, n)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        assert np.array_equal(a[2], b[2])
        assert not np.array_equal(a[0], xzs)  # the tableau was eliminated


def test_noalloc_canonicalize_rejects_a_too_large_workspace(gpu):
    n = 16384
    W = 64
    xzs = torch.zeros(n * 2 * W, dtype=torch.int64, device="cuda")
    ph = torch.zeros(n, dtype=torch.uint8, device="cuda")
    with pytest.raises(qcf.ArgumentError, match="workspace"):
        _lib.call("qcf_canonicalize_noalloc", C.c_void_p(xzs.data_ptr()),
                  C.c_void_p(ph.data_ptr()), 2 * W, W, n, 64 * W, 1, None, qcf.pauli._stream())
