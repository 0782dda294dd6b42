"""Perceptron syndrome decoder trained on device-sampled syndromes.

`MLPDecoder` wraps the HIP decoder of libqcf (qcf_mlp_*): a network
ndet -> h1 -> h2 -> h3 -> 1 (GELU, BF16 weights and activations, FP32
accumulation on the matrix cores, FP32 master weights with Adam) that
predicts the flip of one observable from the detector bits.  It follows the
AbstractSyndromeDecoder shape of the reference (src/ecc/decoder_pipeline.jl:
1-24) through `decode_observables`, and `train` draws every batch with the
frame sampler on the device, so training never leaves the GPU.

`evaluate` decodes sampled shots with the network and hands the shots whose
|logit| < tau to a matching decoder (the fallback), reporting the logical
error rate of the combination.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np
import torch

from . import _lib
from .frames import Program, plane_words


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t: torch.Tensor):
    return C.c_void_p(t.data_ptr())


class MLPDecoder:
    def __init__(self, det_bits, obs_bit: int, hidden=(1024, 1024, 512), batch: int = 8192,
                 infer_batch: int = 65536, seed: int = 0):
        self.det_bits = np.ascontiguousarray(det_bits, dtype=np.uint32)
        self.ndet = len(self.det_bits)
        self.obs_bit = int(obs_bit)
        self.hidden = tuple(int(h) for h in hidden)
        self.batch = int(batch)
        h = C.c_void_p()
        _lib.call("qcf_mlp_create", self.det_bits.ctypes.data_as(C.c_void_p), self.ndet,
                  self.obs_bit, *self.hidden, self.batch, int(infer_batch), C.byref(h))
        self.handle = h
        n = C.c_size_t()
        _lib.call("qcf_mlp_nparams", h, C.byref(n))
        self.nparams = n.value
        self.kin = (self.ndet + 127) // 128 * 128
        self.set_params(self.init_params(seed))
        self._loss = torch.zeros(1, dtype=torch.float32, device="cuda")

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib is not None and _lib._lib is not None:
            _lib._lib.qcf_mlp_free(h)
        self.handle = None

    @property
    def dims(self):
        return (self.ndet, *self.hidden)

    def init_params(self, seed: int) -> np.ndarray:
        """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights and biases (the
        default of a linear layer); padding input columns of W1 are zero."""
        rng = np.random.default_rng(seed)
        h1, h2, h3 = self.hidden
        parts = []
        for fan_in, fan_out in ((self.ndet, h1), (h1, h2), (h2, h3), (h3, 1)):
            k = 1.0 / np.sqrt(fan_in)
            cols = self.kin if fan_in == self.ndet else fan_in
            W = rng.uniform(-k, k, (fan_out, cols)).astype(np.float32)
            if fan_in == self.ndet:
                W[:, self.ndet:] = 0
            b = rng.uniform(-k, k, fan_out).astype(np.float32)
            parts += [W.reshape(-1), b]
        p = np.concatenate(parts)
        assert p.size == self.nparams
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.nparams:
            raise _lib.DimensionMismatch(f"expected {self.nparams} parameters, got {p.size}")
        _lib.call("qcf_mlp_set_params", self.handle, p.ctypes.data_as(C.c_void_p), _stream())

    def params(self) -> np.ndarray:
        out = np.empty(self.nparams, dtype=np.float32)
        _lib.call("qcf_mlp_get_params", self.handle, out.ctypes.data_as(C.c_void_p), _stream())
        return out

    def gradients(self, record: torch.Tensor, ld: int, shot0: int = 0):
        """(loss, flat gradient) of shots shot0 .. shot0+batch-1 of a packed record."""
        g = np.empty(self.nparams, dtype=np.float32)
        loss = np.empty(1, dtype=np.float32)
        _lib.call("qcf_mlp_gradients", self.handle, _ptr(record), ld, shot0,
                  g.ctypes.data_as(C.c_void_p), loss.ctypes.data_as(C.c_void_p), _stream())
        return float(loss[0]), g

    def train_step(self, record: torch.Tensor, ld: int, lr: float = 3e-4, beta1: float = 0.9,
                   beta2: float = 0.999, eps: float = 1e-8, shot0: int = 0) -> None:
        _lib.call("qcf_mlp_train_step", self.handle, _ptr(record), ld, shot0, lr, beta1, beta2,
                  eps, _ptr(self._loss), _stream())

    def logits(self, record: torch.Tensor, ld: int, shots: int) -> torch.Tensor:
        out = torch.empty(max(1, shots), dtype=torch.float32, device=record.device)
        _lib.call("qcf_mlp_infer", self.handle, _ptr(record), ld, shots, _ptr(out), _stream())
        return out[:shots]

    def take_loss(self) -> float:
        """Loss summed since the last call (one batch mean per step)."""
        v = float(self._loss.item())
        self._loss.zero_()
        return v


def train(decoder: MLPDecoder, make_circuit, steps: int, p_range=(5e-4, 5e-3), lr: float = 3e-4,
          seed: int = 0, n_rates: int = 16, log_every: int = 0, log=print,
          batches_per_sample: int = 16):
    """Train on `steps` batches.  The frame sampler draws `batches_per_sample`
    batches per launch on the device from one of `n_rates` programs whose
    noise rates are evenly spaced over p_range, picked uniformly per launch
    (programs that differ only in their rate share one compiled sampling
    kernel); every batch is used once."""
    rng = np.random.default_rng(seed)
    rates = np.linspace(p_range[0], p_range[1], n_rates)
    progs = [Program(make_circuit(float(p)).ops) for p in rates]
    B = decoder.batch
    per = max(1, int(batches_per_sample))
    ld = plane_words(B * per)
    rec = torch.empty(max(1, progs[0].nbits) * ld, dtype=torch.int64, device="cuda")
    st = _stream()
    hist = []
    t0 = time.perf_counter()
    for step in range(steps):
        k = step % per
        if k == 0:
            prog = progs[int(rng.integers(n_rates))]
            _lib.call("qcf_pf_sample", prog.handle, B * per, 0, int(rng.integers(0, 2**62)), None,
                      _ptr(rec), None, ld, 1, st)
        decoder.train_step(rec, ld, lr=lr, shot0=k * B)
        if log_every and (step + 1) % log_every == 0:
            loss = decoder.take_loss() / log_every
            hist.append((step + 1, loss))
            log(f"step {step + 1} loss {loss:.4f} ({time.perf_counter() - t0:.1f} s)")
    torch.cuda.synchronize()
    return dict(steps=steps, seconds=time.perf_counter() - t0, history=hist)


def _bits_of(planes: np.ndarray, sel: np.ndarray) -> np.ndarray:
    """Selected shots of packed planes (planes x words uint64) -> len(sel) x planes uint8."""
    w = planes[:, sel >> 6]
    return np.ascontiguousarray(((w >> (sel & 63).astype(np.uint64)) & np.uint64(1)).T.astype(np.uint8))


def evaluate(decoder: MLPDecoder, matcher, circuit, shots: int, taus=(0.0,), seed: int = 1,
             chunk: int = 1 << 20, tau=None):
    """Logical error rates on `shots` fresh device-sampled shots: the network
    alone, matching alone, and for each tau the network with matching on the
    shots where |logit| < tau; with each decoder's throughput (network on the
    GPU, matching on the host threads) and the combination's.  Returns one
    dict per tau (a single dict when `tau` is given)."""
    if tau is not None:
        taus = (tau,)
    prog = circuit if isinstance(circuit, Program) else Program(circuit)
    det_idx = torch.from_numpy(np.asarray(decoder.det_bits, dtype=np.int64) - 1)
    obs_idx = decoder.obs_bit - 1
    err_nn = err_mw = 0
    err_comb = [0] * len(taus)
    n_fb = [0] * len(taus)
    t_fb = [0.0] * len(taus)
    t_nn = t_mw = 0.0
    done = k = 0
    while done < shots:
        s = min(chunk, shots - done)
        ld = plane_words(s)
        rec = torch.empty(max(1, prog.nbits) * ld, dtype=torch.int64, device="cuda")
        _lib.call("qcf_pf_sample", prog.handle, s, 0, seed * 1000003 + k, None, _ptr(rec), None, ld,
                  1, _stream())
        torch.cuda.synchronize()
        t = time.perf_counter()
        z = decoder.logits(rec, ld, s)
        torch.cuda.synchronize()
        t_nn += time.perf_counter() - t
        planes = rec.view(-1, ld)
        dets = planes[det_idx.to(rec.device)].cpu().numpy().view(np.uint64)
        op = planes[obs_idx].cpu().numpy().view(np.uint64)
        obs = ((op[np.arange(s) >> 6] >> (np.arange(s) & 63).astype(np.uint64)) &
               np.uint64(1)).astype(np.uint8)
        zc = z.cpu().numpy()
        pred_nn = (zc > 0).astype(np.uint8)
        t = time.perf_counter()
        pred_mw = matcher.decode_packed(dets, s)[:, 0]
        t_mw += time.perf_counter() - t
        err_nn += int(np.sum(pred_nn != obs))
        err_mw += int(np.sum(pred_mw != obs))
        for i, ta in enumerate(taus):
            sel = np.flatnonzero(np.abs(zc) < ta)
            comb = pred_nn.copy()
            if len(sel):
                syn = _bits_of(dets, sel)
                t = time.perf_counter()
                comb[sel] = matcher.decode_observables(syn)[:, 0]
                t_fb[i] += time.perf_counter() - t
            n_fb[i] += len(sel)
            err_comb[i] += int(np.sum(comb != obs))
        done += s
        k += 1
    out = []
    for i, ta in enumerate(taus):
        out.append(dict(shots=shots, tau=ta, ler_nn=err_nn / shots, ler_matching=err_mw / shots,
                        ler_combined=err_comb[i] / shots, fallback_fraction=n_fb[i] / shots,
                        nn_shots_per_s=shots / max(t_nn, 1e-12),
                        matching_shots_per_s=shots / max(t_mw, 1e-12),
                        combined_shots_per_s=shots / max(t_nn + t_fb[i], 1e-12)))
    return out[0] if tau is not None else out
