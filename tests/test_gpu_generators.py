"""Known-answer tests of the sampler's two 32-bit generators on the device.

* The per-lane noise-event stream is xoshiro128** (Blackman and Vigna,
  "Scrambled linear pseudorandom number generators", ACM TOMS 2021): from the
  state (1, 2, 3, 4) the reference implementation (xoshiro128starstar.c)
  gives the ten outputs below (the same vector the rand_xoshiro crate tests
  against).
* The gauge words scramble a Weyl sequence with the MurmurHash3 finalizer
  fmix32 (Appleby): MurmurHash3_x86_32 of the empty key is fmix32(seed), and
  its published vectors for seeds 1 and 0xFFFFFFFF are 0x514E28B7 and
  0x81F16F39.
* The gauge stream of a 32-shot word takes its start and increment from
  Philox4x32-10, which the oracle computes (pinned by the Random123 known
  answers in test_gpu_frames.py)."""
import numpy as np
import pytest
import torch

import oracle
import qcf
from qcf import _lib

pytestmark = pytest.mark.gpu

XOSHIRO128SS_1234 = [11520, 0, 5927040, 70819200, 2031721883, 1637235492, 1287239034,
                     3734860849, 3729100597, 4258142804]


def _run(name, state, k, *pre):
    st = torch.as_tensor(np.asarray(state, np.uint32).view(np.int32), device="cuda")
    n = st.numel() // (4 if name == "qcf_debug_xoshiro128ss" else 2)
    out = torch.zeros(n * k, dtype=torch.int32, device="cuda")
    import ctypes as C
    _lib.call(name, *pre, C.c_void_p(st.data_ptr()), C.c_void_p(out.data_ptr()), n, k,
              C.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return out.cpu().numpy().view(np.uint32).reshape(n, k)


def test_xoshiro128ss_known_answers(gpu):
    got = _run("qcf_debug_xoshiro128ss", [1, 2, 3, 4], 10)
    assert got[0].tolist() == XOSHIRO128SS_1234


def test_gauge_words_are_murmur3_finalized_weyl_terms(gpu):
    # start 0, increment 1: the first term is fmix32(1); start 0xFFFFFFFE: fmix32(0xFFFFFFFF)
    got = _run("qcf_debug_gauge", [0, 1, 0xFFFFFFFE, 1, 0xFFFFFFFF, 1], 2)
    assert got[0, 0] == 0x514E28B7
    assert got[1, 0] == 0x81F16F39
    assert got[2, 0] == 0  # fmix32(0) = 0: the Weyl term wrapped to zero
    assert got[1, 1] == 0  # 0xFFFFFFFE + 2 wraps to 0


def test_gauge_stream_seeding(gpu):
    """Start and increment of a word's gauge sequence from Philox4x32-10 of
    (word, 0, 0, 0x47415547) under the seed; the increment made odd and, with
    fewer than 12 bit transitions, XORed with 0xAAAAAAAA."""
    import ctypes as C
    seed = 0x0123456789ABCDEF
    words = np.array([0, 1, 2, 77, 1 << 20, 0xFFFFFFFE], np.uint32)
    wd = torch.as_tensor(words.view(np.int32), device="cuda")
    k = 4
    out = torch.zeros(len(words) * k, dtype=torch.int32, device="cuda")
    _lib.call("qcf_debug_gauge_stream", seed, C.c_void_p(wd.data_ptr()),
              C.c_void_p(out.data_ptr()), len(words), k,
              C.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    got = out.cpu().numpy().view(np.uint32).reshape(len(words), k)
    states = []
    for w in words:
        r = oracle.philox4x32_10([int(w), 0, 0, 0x47415547], [seed & 0xFFFFFFFF, seed >> 32])
        g = int(r[1]) | 1
        if bin(g ^ (g >> 1)).count("1") < 12:
            g ^= 0xAAAAAAAA
        states += [int(r[0]), g]
    # the same words from those states through the device Weyl + fmix32 path
    assert np.array_equal(got, _run("qcf_debug_gauge", states, k))
    assert len({int(x) for x in got.ravel()}) == got.size  # no repeats among these
