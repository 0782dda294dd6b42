"""Perceptron decoder on the GPU against the numpy oracle (oracle/mlp.py).

* the matrix-core GEMM (all three operand layouts) equals float64 products
  of the same BF16 inputs up to float32 summation error;
* forward logits, loss and every gradient equal the oracle's BF16-emulating
  computation (relative error bounds stated per tensor: BF16 activations
  rounded from float32 sums can land one BF16 ulp apart from the oracle's
  float64 sums);
* one Adam step equals the oracle's update of the same gradients;
* training on device-sampled repetition-code syndromes lowers the loss and
  decodes about as well as matching.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import qcf
from oracle import mlp as om
from qcf import circuits, decoders, mlp
from qcf.frames import words

pytestmark = pytest.mark.gpu


def bf16_tensor(a):
    """float32 numpy (already BF16-exact) -> torch bfloat16 on the GPU."""
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(torch.bfloat16).cuda()


@pytest.mark.parametrize("layout", [0, 1, 2])
@pytest.mark.parametrize("K", [160, 192])  # 32-wide and 64-wide K stages
def test_matrix_core_gemm_layouts(gpu, layout, K):
    rng = np.random.default_rng(layout + K)
    M, N = 256, 384
    A = om.bf16(rng.normal(0, 1, (M, K)))
    B = om.bf16(rng.normal(0, 1, (K, N)))
    if layout == 0:
        a, b = A, B.T
    elif layout == 1:
        a, b = A, B
    else:
        a, b = A.T, B
    ta, tb = bf16_tensor(a), bf16_tensor(b)
    out = torch.empty((M, N), dtype=torch.float32, device="cuda")
    qcf._lib.call("qcf_debug_gemm_bf16", layout, M, N, K, C.c_void_p(ta.data_ptr()),
                  C.c_void_p(tb.data_ptr()), C.c_void_p(out.data_ptr()), mlp._stream())
    got = out.cpu().numpy()
    ref = A.astype(np.float64) @ B.astype(np.float64)
    bound = 1e-5 * (np.abs(A) @ np.abs(B)) + 1e-6
    assert np.all(np.abs(got - ref) <= bound)


def sample_record(circ, shots, seed):
    prog = qcf.Program(circ.ops)
    ld = words(shots)
    rec = torch.empty(prog.nbits * ld, dtype=torch.int64, device="cuda")
    qcf._lib.call("qcf_pf_sample", prog.handle, shots, 0, seed, None, C.c_void_p(rec.data_ptr()),
                  None, ld, 1, mlp._stream())
    torch.cuda.synchronize()
    return rec, ld


def host_inputs(rec, ld, det_bits, obs_bit, rows, kin):
    planes = rec.view(-1, ld).cpu().numpy().view(np.uint64)
    bits = decoders.unpack_planes(planes, rows)
    X = np.zeros((rows, kin))
    X[:, :len(det_bits)] = bits[:, np.asarray(det_bits) - 1]
    return X, bits[:, obs_bit - 1].astype(np.float64)


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("d,rounds,hidden,batch", [(3, 3, (128, 256, 128), 256),
                                                   (5, 25, (1024, 1024, 512), 8192)])
def test_gradients_match_oracle(gpu, d, rounds, hidden, batch):
    c = circuits.rotated_surface_code(d, rounds, 4e-3)
    dec = mlp.MLPDecoder(c.det_bits, c.obs_bits[0], hidden=hidden, batch=batch,
                         infer_batch=batch, seed=3)
    # the second batch of a two-batch record (shot offset)
    rec, ld = sample_record(c, 2 * batch, seed=17)
    loss, g = dec.gradients(rec, ld, shot0=batch)
    X, y = host_inputs(rec, ld, c.det_bits, c.obs_bits[0], 2 * batch, dec.kin)
    X, y = X[batch:], y[batch:]
    params = dec.params()
    oloss, og = om.loss_and_grads(params, X, y, dec.dims, emulate=True)
    assert loss == pytest.approx(oloss, rel=2e-4)
    _, _, offs, n = om.layout(*dec.dims)
    ends = offs[1:] + [n]
    for k, (a, b) in enumerate(zip(offs, ends)):
        # weight gradients sum over the batch: tight; bias sums likewise
        assert rel(g[a:b], og[a:b]) < 2e-2, (k, rel(g[a:b], og[a:b]))


def test_adam_step_matches_oracle(gpu):
    c = circuits.repetition_code(3, 3, 2e-2)
    dec = mlp.MLPDecoder(c.det_bits, c.obs_bits[0], hidden=(128, 128, 128), batch=512,
                         infer_batch=512, seed=1)
    rec, ld = sample_record(c, 512, seed=5)
    p0 = dec.params()
    _, g = dec.gradients(rec, ld)
    dec.train_step(rec, ld, lr=1e-3)
    p1 = dec.params()
    want, _, _ = om.adam(p0.astype(np.float64), np.zeros_like(p0, np.float64),
                         np.zeros_like(p0, np.float64), g.astype(np.float64), 1, 1e-3)
    assert np.max(np.abs(p1 - want)) < 1e-6


def test_inference_matches_oracle_forward(gpu):
    c = circuits.rotated_surface_code(3, 3, 5e-3)
    dec = mlp.MLPDecoder(c.det_bits, c.obs_bits[0], hidden=(128, 128, 256), batch=256,
                         infer_batch=384, seed=4)
    shots = 1000  # three inference chunks, the last one partial
    rec, ld = sample_record(c, shots, seed=8)
    z = dec.logits(rec, ld, shots).cpu().numpy()
    X, _ = host_inputs(rec, ld, c.det_bits, c.obs_bits[0], shots, dec.kin)
    oz, _ = om.forward(dec.params(), X, dec.dims, emulate=True)
    assert np.max(np.abs(z - oz)) < 2e-2 * max(1.0, np.max(np.abs(oz)))
    assert np.mean(np.sign(z) == np.sign(oz)) > 0.99


def test_training_on_device_samples_learns_repetition_code(gpu):
    def make(p):
        return circuits.repetition_code(5, 3, p)

    c = make(0.03)
    dec = mlp.MLPDecoder(c.det_bits, c.obs_bits[0], hidden=(128, 128, 128), batch=1024,
                         infer_batch=4096, seed=0)
    rec, ld = sample_record(c, 1024, seed=99)
    loss0, _ = dec.gradients(rec, ld)
    mlp.train(dec, make, steps=2000, p_range=(0.02, 0.04), lr=3e-3, seed=1)
    loss1, _ = dec.gradients(rec, ld)
    assert loss1 < 0.3 * loss0
    matcher = decoders.MatchingDecoder(c.ops, c.det_bits, c.obs_bits)
    r0, r2, rall = mlp.evaluate(dec, matcher, c.ops, shots=100000, taus=(0.0, 2.0, 1e30), seed=3,
                                chunk=50000)
    # about 1.0% against 0.65% for matching after 2000 batches of 1024
    assert r0["ler_nn"] < 2.0 * r0["ler_matching"] + 5e-3
    assert r0["fallback_fraction"] == 0.0
    # what the network gains: with matching only on the unsure shots the
    # combination decodes as well as matching while matching sees a minority
    # of the shots
    assert r2["ler_combined"] <= 1.1 * r2["ler_matching"] + 1e-3
    assert r2["fallback_fraction"] < 0.5
    # with every shot handed to the fallback the combination is matching
    assert rall["fallback_fraction"] == 1.0
    assert rall["ler_combined"] == rall["ler_matching"]
