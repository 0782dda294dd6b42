"""The batched configuration at full size: 512 rotated surface-code memory
circuits (d ~ U{3..15}, rounds = d, p log-uniform in [1e-4, 1e-2], shots
log-uniform in [1e4, 1e6]; 1.15e8 shots, 36 GB of records), sampled in one
autotuned, graph-captured persistent launch, must equal 512 single launches
bit for bit (records compared on the device)."""
import ctypes as C

import numpy as np
import pytest
import torch

import qcf
from qcf import batch, circuits

pytestmark = pytest.mark.gpu


def test_512_circuit_batch_equals_single_launches(gpu):
    rng = np.random.default_rng(0)
    progs, shots, seeds = [], [], []
    for _ in range(512):
        d = int(rng.choice([3, 5, 7, 9, 11, 13, 15]))
        p = float(10 ** rng.uniform(-4, -2))
        progs.append(qcf.Program(circuits.rotated_surface_code(d, d, p).ops))
        shots.append(int(round(10 ** rng.uniform(4, 6))))
        seeds.append(int(rng.integers(0, 2**62)))
    print(f"512 circuits built, {sum(shots)} shots", flush=True)
    tune = batch.autotune(progs, shots)
    b = batch.Batch(progs, shots, seeds, tune)
    b.launch_graph()
    torch.cuda.synchronize()
    print(f"batched launch done (tune {tune.as_list()})", flush=True)
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    bad = []
    for k, (p, s, sd) in enumerate(zip(progs, shots, seeds)):
        ld = qcf.frames.plane_words(s)
        single = torch.empty(max(1, p.nbits) * ld, dtype=torch.int64, device="cuda")
        qcf._lib.call("qcf_pf_sample", p.handle, s, 0, sd, None, C.c_void_p(single.data_ptr()),
                      None, ld, 1, st)
        n = p.nbits * ld
        if not torch.equal(single[:n], b.records[k][:n]):
            bad.append((k, p.nq, s))
    assert not bad, bad[:10]
    # other launch shapes give the same bits: the persistent kernel sharing
    # the tiles of the launches of more than 768 planes (the default), of all whose frames
    # exceed LDS, or of none, and every circuit of the largest frames on the persistent
    # kernel, started beside the specialised launches on one helper stream
    big = max(p.info.nplanes for p in progs)
    for fields in ((0, 1, 0, 0, 2, 0, 0), (0, 1, 0, 0, 2, 0, 1), (0, 1, 0, 0, 2, 0, 0xFFFFFFFF),
                   (32, 1, 0, big - 1, 1, 1, 0xFFFFFFFF)):
        alt = qcf._lib.Tune(*fields)
        if alt.as_list() == tune.as_list():
            continue
        b2 = batch.Batch(progs, shots, seeds, alt)
        b2.launch_graph()
        torch.cuda.synchronize()
        print(f"launch shape {alt.as_list()} done", flush=True)
        for k, p in enumerate(progs):
            n = p.nbits * qcf.frames.plane_words(shots[k])
            if not torch.equal(b.records[k][:n], b2.records[k][:n]):
                bad.append((fields, k, p.nq, shots[k]))
        del b2
        torch.cuda.empty_cache()
    assert not bad, bad[:10]
    del b
    torch.cuda.empty_cache()
