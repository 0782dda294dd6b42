"""Many circuits in one persistent launch (qcf_batch_*), and the autotuner."""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np
import torch

from . import _lib as _LIB
from ._lib import Tune, call
from .frames import Program, plane_words


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class Batch:
    """Fresh-frame sampling of several compiled circuits in one batched launch:
    circuits with a specialised kernel run it (one launch per distinct kernel,
    on helper streams), the rest share ONE persistent interpreter kernel.

    Records live in `self.records[k]` (nbits_k planes of plane_words(shots_k) words).
    `launch()` enqueues the launches (no allocation, capturable); `launch_graph()`
    replays a captured hipGraph of them."""

    def __init__(self, progs: Sequence[Program], shots: Sequence[int], seeds: Sequence[int],
                 tune: Tune | None = None, device=None):
        dev = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        self.progs = list(progs)
        self.shots = [int(s) for s in shots]
        n = len(self.progs)
        self.records = [torch.empty(max(1, p.nbits) * plane_words(s), dtype=torch.int64,
                                    device=dev) for p, s in zip(self.progs, self.shots)]
        hs = (C.c_void_p * n)(*[p.handle.value for p in self.progs])
        sh = (C.c_uint64 * n)(*self.shots)
        sd = (C.c_uint64 * n)(*[int(s) & (2**64 - 1) for s in seeds])
        ms = (C.c_void_p * n)(*[r.data_ptr() for r in self.records])
        ld = (C.c_size_t * n)(*[plane_words(s) for s in self.shots])
        h = C.c_void_p()
        call("qcf_batch_create", hs, n, sh, sd, ms, ld,
             C.byref(tune) if tune is not None else None, C.byref(h))
        self.handle = h

    def launch(self):
        call("qcf_batch_launch", self.handle, _stream())

    def launch_graph(self):
        call("qcf_batch_launch_graph", self.handle, _stream())

    def record_host(self, k: int) -> np.ndarray:
        p, s = self.progs[k], self.shots[k]
        return self.records[k].cpu().numpy().view(np.uint64).reshape(-1, plane_words(s))[: p.nbits]

    def __del__(self):
        h = getattr(self, "handle", None)
        lib_mod = _LIB
        if h and h.value and lib_mod is not None and lib_mod._lib is not None:
            lib_mod._lib.qcf_batch_free(h)
        self.handle = None


def autotune(progs: Sequence[Program], shots: Sequence[int]) -> Tune:
    n = len(progs)
    hs = (C.c_void_p * n)(*[p.handle.value for p in progs])
    sh = (C.c_uint64 * n)(*[int(s) for s in shots])
    t = Tune()
    call("qcf_autotune", hs, n, sh, C.byref(t), _stream())
    return t


def sample_batched(progs, shots, seeds, tune: Tune | None = None, graph: bool = False):
    b = Batch(progs, shots, seeds, tune)
    (b.launch_graph if graph else b.launch)()
    torch.cuda.synchronize()
    return [b.record_host(k) for k in range(len(progs))]
