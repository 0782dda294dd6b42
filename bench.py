"""Benchmark of the Pauli-frame sampling hot path on one MI355X.

Headline (`value`): shots/s of bit-exact Pauli-frame sampling of the rotated
surface-code memory-Z circuit at d=5, 25 rounds, circuit-level noise p=1e-3,
10^7 shots per step (native counter-based RNG, record in HBM).  One step =
one qcf_pf_sample launch over all 10^7 shots.  Also reported: frame bit-ops/s
and record GB/s with their roofline fractions, mul_left! on 1-Gqubit
operators (HBM roofline), the batched persistent launch against per-circuit
launches, and the oracle (CPU restatement) timed on the host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-extras]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "quantumclifford.jl_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
# HBM bytes per shot of the d=5 x 25 launch from rocprofv3 --pmc (profiles/
# pf_jit_kernel_d5x25_traffic.md): WRITE_SIZE 1,496,639.5 KiB and FETCH_SIZE
# 217.2 KiB per 10^7-shot dispatch, scaled by what the counters read for a
# 1 GiB copy (WRITE_SIZE 1.0x, FETCH_SIZE 0.5x of the bytes moved).
TRAFFIC_BYTES_PER_SHOT = (1496639.5 * 1.0 + 217.2 * 2.0) * 1024 / 1e7
INT_VALU_PEAK = 256 * 64 * 2.4e9  # int32 lane-ops/s (256 CUs x 64 lanes x 2.4 GHz)
METRIC = "batched vs per-circuit shots/s, d=5 surface 25 rds; NN vs MWPM LER"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_launches(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(steps):
        fn()
    e1.record(s)
    e1.synchronize()
    return e0.elapsed_time(e1) / steps  # ms per launch


def bench_frames(steps, warmup, shots=10**7):
    import torch
    import qcf
    from qcf import circuits
    c = circuits.rotated_surface_code(5, 25, 1e-3)
    prog = qcf.Program(c.ops)
    ld = qcf.frames.plane_words(shots)
    rec = torch.empty(prog.nbits * ld, dtype=torch.int64, device="cuda")
    import ctypes as C
    seed = [0]

    def step():
        seed[0] += 1
        qcf._lib.call("qcf_pf_sample", prog.handle, shots, 0, seed[0], None,
                      C.c_void_p(rec.data_ptr()), None, ld, 1,
                      C.c_void_p(torch.cuda.current_stream().cuda_stream))

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ms = time_launches(step, steps, warmup)
    wall = time.perf_counter() - t0
    shots_s = shots / (ms * 1e-3)
    bitops = prog.bitops_per_shot
    rec_bytes = prog.nbits / 8.0  # record bits written per shot
    return dict(ms=ms, shots_s=shots_s, bitops_per_shot=bitops,
                bitops_s=bitops * shots_s, lane_ops_s=bitops * shots_s / 32,
                rec_gbs=rec_bytes * shots_s / 1e9, bytes_per_launch=rec_bytes * shots,
                nbits=prog.nbits, wall=wall, shots=shots)


def bench_frame_forms(steps, shots=10**7):
    """The headline workload (d=5 x 25 rounds, p=1e-3, 10^7 shots) in the
    three launch forms of the sampler: fresh with the record only (the
    headline), fresh returning the final frame too (PauliFrame + initZ! +
    pftrajectories, src/pauli_frames.jl:193-227), and in place on a frame the
    caller holds (pftrajectories(frame, circuit), src/pauli_frames.jl:234-239);
    device time per launch inside a captured graph of 10."""
    import ctypes as C
    import torch
    import qcf
    from qcf import circuits
    prog = qcf.Program(circuits.rotated_surface_code(5, 25, 1e-3).ops)
    ld = qcf.frames.plane_words(shots)
    rec = torch.empty(prog.nbits * ld, dtype=torch.int64, device="cuda")
    frame = torch.zeros(2 * prog.nq * ld, dtype=torch.int64, device="cuda")
    out = {}
    seed = [7]
    for name, flags, fr in (("fresh_record_only", 1, None), ("fresh_with_frame", 1, frame),
                            ("in_place_with_frame", 0, frame)):
        def step():
            seed[0] += 1
            qcf._lib.call("qcf_pf_sample", prog.handle, shots, 0, seed[0],
                          C.c_void_p(fr.data_ptr()) if fr is not None else None,
                          C.c_void_p(rec.data_ptr()), None, ld, flags,
                          C.c_void_p(torch.cuda.current_stream().cuda_stream))
        ms = time_graph(step, 10, max(2, steps))
        # record written; with a frame, its 2 nq planes written (and read in place)
        nbytes = prog.nbits * shots / 8 + (0 if fr is None else
                                           2 * prog.nq * shots / 8 * (2 if flags == 0 else 1))
        out[name] = dict(ms=ms, shots_s=shots / (ms * 1e-3), gbs=nbytes / (ms * 1e-3) / 1e9)
    return out


def bench_mul(steps, warmup, n=10**9):
    import numpy as np
    import torch
    import qcf
    rng = np.random.default_rng(0)
    p1, x1 = qcf.randoms.random_pauli_words(rng, n)
    p2, x2 = qcf.randoms.random_pauli_words(rng, n)
    l = qcf.PauliOperator(p1, n, x1)
    r = qcf.PauliOperator(p2, n, x2)
    ms = time_launches(lambda: qcf.mul_left_(r, l), steps, warmup)
    nbytes = 6 * (l.nwords) * 8  # read l.x,l.z,r.x,r.z, write r.x,r.z
    return dict(n=n, ms=ms, gbs=nbytes / (ms * 1e-3) / 1e9, bytes=nbytes)


def time_graph(fn, per_graph, replays, warmup=1):
    """Device time of one fn() call: per_graph calls captured in one hipGraph
    (torch.cuda.CUDAGraph on ROCm), the graph replayed `replays` times between
    two events on the replay stream; no Python or launch API in the timed
    region."""
    import torch
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for _ in range(warmup):
            fn()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    # captured on the warm-up stream: per-stream workspaces (global-memory
    # frames) were sized by the warm-up calls and cannot grow under capture
    with torch.cuda.graph(g, stream=s):
        for _ in range(per_graph):
            fn()
    g.replay()
    torch.cuda.synchronize()
    cs = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(cs)
    for _ in range(replays):
        g.replay()
    e1.record(cs)
    e1.synchronize()
    return e0.elapsed_time(e1) / (replays * per_graph)


def bench_mul_config(steps, warmup):
    """The rest of the mul_left! configuration: two random Paulis of 10^6
    qubits (launch-bound), and the pivot sweep (every row m != pivot:
    mul_left!(s, m, pivot), src/canonicalization.jl:105-109) on stabilizer
    tableaux of 1000 and 10000 qubits built by a random Clifford circuit
    (4n H/S/CNOT gates, seed 0) from |0...0>, plain and canonicalize!-style
    (rows with a random half of a mask); algorithmic bytes of a sweep: read
    and write every other row, read the pivot row, phases."""
    import numpy as np
    import torch
    import qcf
    out = {}
    rng = np.random.default_rng(0)
    n = 10**6
    p1, x1 = qcf.randoms.random_pauli_words(rng, n)
    p2, x2 = qcf.randoms.random_pauli_words(rng, n)
    l, r = qcf.PauliOperator(p1, n, x1), qcf.PauliOperator(p2, n, x2)
    ms_call = time_launches(lambda: qcf.mul_left_(r, l), steps * 10, warmup)
    ms = time_graph(lambda: qcf.mul_left_(r, l), 100, max(2, steps // 2))
    out["mul_left_1mq"] = dict(n=n, ms=ms, gbs=6 * l.nwords * 8 / (ms * 1e-3) / 1e9,
                               ms_per_python_call=ms_call,
                               timing="device time per product inside a captured graph of 100")
    for n in (1000, 10000):
        ph, xzs = qcf.randoms.random_clifford_tableau(np.random.default_rng(0), n)
        t = qcf.Tableau(ph, n, xzs)
        W = t.nwords
        nbytes = (n - 1) * 2 * (2 * W * 8) + 2 * W * 8 + 2 * n
        mask = torch.from_numpy((np.random.default_rng(1).integers(0, 2, n)).astype(np.uint8)).cuda()
        # device time per sweep: 100 sweeps (pivots 1..100 in turn, as
        # canonicalize! runs them) per captured graph
        piv = [0]

        def sweep():
            piv[0] = piv[0] % 100 + 1
            qcf.mul_left_sweep_(t, piv[0])

        def sweep_masked():
            piv[0] = piv[0] % 100 + 1
            qcf.mul_left_sweep_(t, piv[0], row_mask=mask)

        ms = time_graph(sweep, 100, max(2, steps // 2))
        ms_m = time_graph(sweep_masked, 100, max(2, steps // 2))
        ms_call = time_launches(lambda: qcf.mul_left_sweep_(t, 1), steps * 5, warmup)
        out[f"sweep_{n}"] = dict(n=n, rows=n, ms=ms, gbs=nbytes / (ms * 1e-3) / 1e9,
                                 frac_hbm=nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                 bytes=nbytes, ms_masked=ms_m,
                                 ms_per_python_call=ms_call,
                                 timing="device time per sweep inside a captured graph of 100")
    return out


def bench_batched(ncirc=512, seed=0, dmax=15):
    """512 rotated memory circuits, d ~ U{3..dmax odd}, rounds = d, p log-uniform
    [1e-4,1e-2], shots log-uniform [1e4,1e6]: one persistent launch vs one launch
    per circuit."""
    import numpy as np
    import torch
    import qcf
    from qcf import batch, circuits
    rng = np.random.default_rng(seed)
    ds = [d for d in (3, 5, 7, 9, 11, 13, 15) if d <= dmax]
    cache = {}
    progs, shots, seeds = [], [], []
    for k in range(ncirc):
        d = int(rng.choice(ds))
        p = float(10 ** rng.uniform(-4, -2))
        progs.append(qcf.Program(circuits.rotated_surface_code(d, d, p).ops))
        shots.append(int(round(10 ** rng.uniform(4, 6))))
        seeds.append(int(rng.integers(0, 2**62)))
    t0 = time.perf_counter()
    tune = batch.autotune(progs, shots)
    tune_s = time.perf_counter() - t0
    total = sum(shots)
    # the library defaults (every circuit with a specialised kernel runs it on
    # two helper streams; the persistent kernel shares the tiles of those whose
    # frames exceed LDS), then the tune
    b0 = batch.Batch(progs, shots, seeds, qcf._lib.Tune())
    ms_default = time_launches(b0.launch_graph, 3, 1)
    del b0
    # the same without the persistent kernel sharing the tiles of the
    # launches whose frames exceed LDS
    t1 = qcf._lib.Tune()
    t1.steal_planes = 0xFFFFFFFF
    b0 = batch.Batch(progs, shots, seeds, t1)
    ms_unshared = time_launches(b0.launch_graph, 3, 1)
    del b0
    b = batch.Batch(progs, shots, seeds, tune)
    ms_b = time_launches(b.launch_graph, 3, 1)
    # per-circuit launches into the same records
    import ctypes as C
    recs = b.records

    def seq():
        st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        for p, s, sd, r in zip(progs, shots, seeds, recs):
            qcf._lib.call("qcf_pf_sample", p.handle, s, 0, sd, None, C.c_void_p(r.data_ptr()),
                          None, qcf.frames.plane_words(s), 1, st)

    ms_s = time_launches(seq, 2, 1)
    return dict(ncirc=ncirc, total_shots=total, ms_batched=ms_b, ms_sequential=ms_s,
                shots_s_batched=total / (ms_b * 1e-3), shots_s_sequential=total / (ms_s * 1e-3),
                ratio=ms_s / ms_b, ms_batched_default_tune=ms_default,
                ms_batched_unshared=ms_unshared,
                tune=dict(zip([f for f, _ in qcf._lib.Tune._fields_], tune.as_list())),
                autotune_s=tune_s, dmax=dmax)


def bench_reference_circuits(steps, warmup):
    """The reference's own frame-benchmark circuits: the noiseless 441-qubit
    unrotated surface code of benchmark/surface-code-circuit.jl (d=11, 12
    rounds, 2751 record bits; benchmark-surface-code-gpu.jl times 2^3..2^19
    trajectories) at 2^19 and 10^7 shots, x_diag_circuit(1000) at 1000
    trajectories (benchmarks.jl q1001_r10000) and the compactification case,
    x_diag_circuit_noisy_measurement(1000) at 1000 trajectories."""
    import ctypes as C
    import torch
    import qcf
    from qcf import circuits
    out = {}
    cases = (("surface_code_d11_2pow19", circuits.unrotated_surface_code_benchmark(11, 12), 2**19),
             ("surface_code_d11_1e7", circuits.unrotated_surface_code_benchmark(11, 12), 10**7),
             ("x_diag_1000_1000shots", circuits.x_diag_circuit(1000), 1000),
             ("x_diag_noisy_measurement_1000_1000shots",
              circuits.x_diag_circuit_noisy_measurement(1000), 1000))
    for name, ops, shots in cases:
        prog = qcf.Program(ops)
        ld = qcf.frames.plane_words(shots)
        rec = torch.empty(prog.nbits * ld, dtype=torch.int64, device="cuda")
        seed = [100]

        def step():
            seed[0] += 1
            qcf._lib.call("qcf_pf_sample", prog.handle, shots, 0, seed[0], None,
                          C.c_void_p(rec.data_ptr()), None, ld, 1,
                          C.c_void_p(torch.cuda.current_stream().cuda_stream))

        # device time per launch: 20 launches captured in one graph (at 2^19
        # shots and below a launch from Python takes longer than the kernel)
        ms_call = time_launches(step, steps, warmup)
        ms = time_graph(step, 20, max(2, steps))
        sps = shots / (ms * 1e-3)
        out[name] = dict(shots=shots, ms=ms, shots_s=sps, nqubits=prog.nq - 1,
                         record_bits=prog.nbits, bitops_s=prog.bitops_per_shot * sps,
                         rec_gbs=prog.nbits / 8 * sps / 1e9, ms_per_python_call=ms_call,
                         timing="device time per launch inside a captured graph of 20")
        log(f"  {name}: {out[name]}")
        del rec
    return out


def bench_canonicalize(steps, warmup):
    """canonicalize! on the device: a dense random 1000-qubit stabilizer (the
    reference's README.md:107-115 benchmark, 8.5 ms on one Zen4 core; rows
    3/4 non-identity, from 28 layers of random gates and CNOT pairings) and
    Stabilizer(random_tableau(6000, 6000)) (test/test_gpu_canonicalization.jl:
    19-29, under 1900 ms asserted; 1900-2400 ms on the CPU per that test)."""
    import numpy as np
    import qcf
    out = {}
    cases = (("stabilizer_1000", qcf.randoms.random_stabilizer_layers(np.random.default_rng(0), 1000),
              1000),
             ("random_tableau_6000", qcf.randoms.random_tableau(np.random.default_rng(0), 6000, 6000),
              6000))
    for name, (ph, xzs), n in cases:
        src = qcf.Tableau(ph, n, xzs)
        work = src.copy()

        def step():
            work.xzs.copy_(src.xzs)
            work.phases.copy_(src.phases)
            qcf.canonicalize_(work)

        ms = time_launches(step, steps, warmup)
        out[name] = dict(n=n, rows=len(ph), ms=ms)
        log(f"  {name}: {out[name]}")
    return out


def host_cpu():
    """CPU model, physical cores and hardware threads of the host (lscpu)."""
    import subprocess
    info = {"model": None, "physical_cores": None, "threads": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = {}
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
        info["model"] = kv.get("Model name")
        cps, sockets = kv.get("Core(s) per socket"), kv.get("Socket(s)")
        if cps and sockets:
            info["physical_cores"] = int(cps) * int(sockets)
    except Exception as e:  # reported, not hidden
        info["error"] = repr(e)
    return info


def cpu_baseline(target_s=12.0):
    """The oracle (C restatement, OpenMP over shots like the reference's
    Threads.@threads split) on the d=5 x 25 circuit with per-shot RNG draws."""
    import numpy as np
    import oracle
    import qcf
    from qcf import circuits
    c = circuits.rotated_surface_code(5, 25, 1e-3)
    ops, args = qcf.lower_circuit(c.ops, c.nqubits)
    nq, nbits = c.nqubits + 1, c.nbits
    host = host_cpu()
    # the GPU box grants one GPU's share of its host: OMP_NUM_THREADS (16)
    cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    oracle.pf_run(ops, args, nq, nbits, 4000, None, True, seed=1, nthreads=cores)
    # chunks of 10^6 shots (the oracle keeps one byte per record bit) until
    # about target_s seconds of CPU work
    chunk, shots, dt, k = 1_000_000, 0, 0.0, 0
    while dt < target_s and k < 30:
        t0 = time.perf_counter()
        oracle.pf_run(ops, args, nq, nbits, chunk, None, True, seed=2 + k, nthreads=cores)
        dt += time.perf_counter() - t0
        shots += chunk
        k += 1
    # mul_left! (the reference's is single-threaded): the oracle's scalar
    # restatement on one host core, two random 10^8-qubit operators
    n = 10**8
    a_ = np.random.default_rng(2).integers(0, 2**64, 2 * (n // 64), dtype=np.uint64)
    b_ = np.random.default_rng(3).integers(0, 2**64, 2 * (n // 64), dtype=np.uint64)
    oracle.mul_left_vec(a_, b_)
    t0 = time.perf_counter()
    reps = 10
    for _ in range(reps):
        oracle.mul_left_vec(a_, b_)
    dtm = (time.perf_counter() - t0) / reps
    mul = dict(n=n, ms=dtm * 1e3, gbs=6 * (n // 64) * 8 / dtm / 1e9, cores=1, kind="port")
    # canonicalize! (single-threaded in the reference): the oracle on one core
    canon = {}
    for name, (ph, xzs), nq in (
            ("stabilizer_1000", qcf.randoms.random_stabilizer_layers(np.random.default_rng(0), 1000),
             1000),
            ("random_tableau_6000", qcf.randoms.random_tableau(np.random.default_rng(0), 6000, 6000),
             6000)):
        t0 = time.perf_counter()
        oracle.canonicalize(xzs, ph, nq)
        canon[name] = dict(ms=1e3 * (time.perf_counter() - t0), cores=1, kind="port")
    return dict(value=shots / dt, unit="shots/s", cores=cores, kind="port",
                sample=f"{shots} shots of the d=5 x 25-round circuit (p=1e-3) in chunks of "
                       f"{chunk}, oracle restatement with OpenMP over {cores} threads, "
                       f"{dt:.1f} s", host_cpu=host,
                threads_note="OpenMP threads = OMP_NUM_THREADS, the host share one GPU is "
                             "granted on the GPU box (Julia is not installed there, so the "
                             "reference itself is not measured)",
                mul_left=mul, canonicalize=canon)


def bench_mlp_train(steps: int):
    """The decoder's training step alone (for profiling): `steps` batches of
    8192 device-sampled d=5 x 25 shots after 32 warm-up batches; ms per step
    including the sampling of every 16th batch, and the matrix-core rate."""
    from qcf import circuits, mlp

    def make(p):
        return circuits.rotated_surface_code(5, 25, p)

    c = make(1e-3)
    dec = mlp.MLPDecoder(c.det_bits, c.obs_bits[0], batch=8192, infer_batch=65536, seed=0)
    mlp.train(dec, make, steps=32, p_range=(5e-4, 5e-3), lr=3e-4, seed=1)
    tr = mlp.train(dec, make, steps=steps, p_range=(5e-4, 5e-3), lr=3e-4, seed=2)
    layers = ((dec.kin, 1024), (1024, 1024), (1024, 512))
    flops_step = 2 * 8192 * (2 * sum(a * b for a, b in layers) +
                             sum(a * b for a, b in layers[1:]))
    ms = tr["seconds"] * 1e3 / steps
    return dict(steps=steps, ms_per_step=ms, tflops=flops_step / (ms * 1e-3) / 1e12,
                gflop_per_step=flops_step / 1e9)


def bench_gemm(M=8192, N=1024, K=1024, reps=50):
    """The decoder's GEMM kernel alone on a plain product (fp32 output, no
    epilogue work): C = A B^T with A (M x K) and B (N x K) in BF16, timed
    over `reps` launches (for counter runs)."""
    import ctypes as C
    import torch
    from qcf import _lib
    g = torch.Generator(device="cuda").manual_seed(0)
    A = (torch.randn(M, K, device="cuda", generator=g) * 0.1).to(torch.bfloat16)
    B = (torch.randn(N, K, device="cuda", generator=g) * 0.1).to(torch.bfloat16)
    Cm = torch.empty(M, N, device="cuda")
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)

    def go():
        _lib.call("qcf_debug_gemm_bf16", 0, M, N, K, C.c_void_p(A.data_ptr()),
                  C.c_void_p(B.data_ptr()), C.c_void_p(Cm.data_ptr()), st)

    ms = time_launches(go, reps, 3)
    err = float((Cm - A.float() @ B.float().t()).abs().max())
    return dict(M=M, N=N, K=K, us=ms * 1e3, tflops=2 * M * N * K / (ms * 1e-3) / 1e12,
                max_abs_err_vs_fp32=err)


def bench_mlp(steps: int, eval_shots: int):
    """Perceptron decoder (600 -> 1024 -> 1024 -> 512 -> 1, BF16 on the matrix
    cores) trained for `steps` batches of 8192 device-sampled d=5 x 25 shots
    at p evenly spread over [5e-4, 5e-3]; then, on `eval_shots` fresh shots at
    each p in {1e-3, 2e-3, 3e-3}, logical error rates against the matching
    decoder, alone and with the fallback |logit| < tau -> matching (five
    thresholds at 1e-3, three at the higher rates, where matching is slower)."""
    from qcf import circuits, decoders, mlp

    def make(p):
        return circuits.rotated_surface_code(5, 25, p)

    c = make(1e-3)
    dec = mlp.MLPDecoder(c.det_bits, c.obs_bits[0], batch=8192, infer_batch=65536, seed=0)
    tr = mlp.train(dec, make, steps=steps, p_range=(5e-4, 5e-3), lr=3e-4, seed=1)
    # matrix-core work per batch: forward and weight gradients over all three
    # hidden layers, input gradients of layers 2 and 3 (the 512 -> 1 head is
    # not on the matrix cores)
    layers = ((dec.kin, 1024), (1024, 1024), (1024, 512))
    flops_step = 2 * 8192 * (2 * sum(a * b for a, b in layers) +
                             sum(a * b for a, b in layers[1:]))
    log(f"  trained {steps} batches in {tr['seconds']:.1f} s")
    evals = {}
    for p, taus in ((1e-3, (0.0, 1.0, 1.5, 2.0, 4.0)), (2e-3, (0.0, 1.5, 2.0)),
                    (3e-3, (0.0, 1.5, 2.0))):
        n = eval_shots
        cp = make(p)
        matcher = decoders.MatchingDecoder(cp.ops, cp.det_bits, cp.obs_bits)
        rows = []
        # chunks of 2^20 shots, one progress line each
        done = 0
        while done < n:
            s = min(1 << 20, n - done)
            rows.append(mlp.evaluate(dec, matcher, cp.ops, shots=s, taus=taus, seed=7 + done,
                                     chunk=s))
            done += s
            log(f"  p={p}: {done} shots evaluated")
        merged = []
        for i in range(len(rows[0])):
            tot = sum(r[i]["shots"] for r in rows)
            m = {k: sum(r[i][k] * r[i]["shots"] for r in rows) / tot
                 for k in ("ler_nn", "ler_matching", "ler_combined", "fallback_fraction")}
            for k in ("nn_shots_per_s", "matching_shots_per_s", "combined_shots_per_s"):
                m[k] = tot / sum(r[i]["shots"] / r[i][k] for r in rows)
            m["tau"] = rows[0][i]["tau"]
            m["shots"] = tot
            merged.append(m)
        evals[str(p)] = merged
    return dict(train_steps=steps, train_shots=steps * 8192, train_s=tr["seconds"],
                ms_per_step=1e3 * tr["seconds"] / steps,
                train_tflops=flops_step * steps / tr["seconds"] / 1e12,
                evals=evals)


def headline_ratios(extras):
    """The two comparisons the metric names: batched over per-circuit shots/s
    (512 circuits), and the perceptron decoder's logical error rate over
    matching's at p = 1e-3 (with the |logit| fallback closest to 1.05x at or
    above 10x matching's shots/s, and alone)."""
    out = {"batched_vs_per_circuit_shots_s": None, "nn_vs_mwpm_ler_p1e-3": None,
           "nn_alone_vs_mwpm_ler_p1e-3": None}
    b = extras.get("batched") or {}
    if "ratio" in b:
        out["batched_vs_per_circuit_shots_s"] = b["ratio"]
    m = (extras.get("mlp_decoder") or {}).get("evals", {}).get("0.001")
    if m:
        alone = [r for r in m if r["tau"] == 0.0]
        if alone:
            out["nn_alone_vs_mwpm_ler_p1e-3"] = alone[0]["ler_nn"] / alone[0]["ler_matching"]
        ok = [r for r in m if r["combined_shots_per_s"] >= 10 * r["matching_shots_per_s"]]
        if ok:
            best = min(ok, key=lambda r: r["ler_combined"] / r["ler_matching"])
            out["nn_vs_mwpm_ler_p1e-3"] = best["ler_combined"] / best["ler_matching"]
            out["nn_fallback_tau"] = best["tau"]
            out["nn_shots_s_vs_mwpm"] = best["combined_shots_per_s"] / best["matching_shots_per_s"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # the GPU reaches its steady clocks after about 0.1 s of this load: with 3
    # warm-up launches the same kernel measured 0.533-0.539 ms per step, with
    # 300 0.475-0.476 ms
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--shots", type=int, default=10**7)
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batched-dmax", type=int, default=15)
    ap.add_argument("--mlp-steps", type=int, default=24414)  # 2e8 shots of 8192
    ap.add_argument("--mlp-eval-shots", type=int, default=10**7)
    ap.add_argument("--only", choices=("reference_circuits", "canonicalize", "mul_left_1gq",
                                       "mul_config", "batched", "mlp_train", "gemm",
                                       "frame_forms"),
                    help="run one extra alone and print its JSON (for profiling)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}); the frame sampler shards "
              "shots across GPUs without communication, but no multi-GPU run is part of "
              "this build.", file=sys.stderr)
        sys.exit(2)
    import torch
    torch.cuda.set_device(0)
    if a.only:
        fn = {"reference_circuits": lambda: bench_reference_circuits(5, 1),
              "canonicalize": lambda: bench_canonicalize(5, 1),
              "mul_left_1gq": lambda: bench_mul(10, 2),
              "mul_config": lambda: bench_mul_config(10, 2),
              "batched": lambda: bench_batched(dmax=a.batched_dmax),
              "mlp_train": lambda: bench_mlp_train(min(a.mlp_steps, 2000)),
              "gemm": lambda: bench_gemm(reps=a.steps if a.steps < 200 else 50),
              "frame_forms": lambda: bench_frame_forms(5)}[a.only]
        print(json.dumps({a.only: fn()}), flush=True)
        return
    log("frames: d=5 x 25 rounds, p=1e-3 ...")
    fr = bench_frames(a.steps, a.warmup, a.shots)
    log(f"  {fr['ms']:.3f} ms/step, {fr['shots_s']:.3e} shots/s")
    extras = {}
    if not a.no_extras:
        log("launch forms of the d=5 x 25 workload ...")
        try:
            extras["frame_forms"] = bench_frame_forms(5, a.shots)
            log(f"  {extras['frame_forms']}")
        except Exception as e:  # reported, not hidden
            extras["frame_forms"] = {"error": repr(e)}
        log("mul_left! 1 Gqubit ...")
        try:
            extras["mul_left_1gq"] = bench_mul(10, 2)
            log(f"  {extras['mul_left_1gq']}")
        except Exception as e:  # reported, not hidden
            extras["mul_left_1gq"] = {"error": repr(e)}
        log("mul_left! 10^6 qubits and tableau sweeps (1000, 10000 qubits) ...")
        try:
            extras["mul_config"] = bench_mul_config(10, 2)
            log(f"  {extras['mul_config']}")
        except Exception as e:  # reported, not hidden
            extras["mul_config"] = {"error": repr(e)}
        log("the reference's benchmark circuits ...")
        try:
            extras["reference_circuits"] = bench_reference_circuits(5, 1)
        except Exception as e:  # reported, not hidden
            extras["reference_circuits"] = {"error": repr(e)}
        log("canonicalize! ...")
        try:
            extras["canonicalize"] = bench_canonicalize(5, 1)
        except Exception as e:  # reported, not hidden
            extras["canonicalize"] = {"error": repr(e)}
        log("batched 512 circuits ...")
        try:
            extras["batched"] = bench_batched(dmax=a.batched_dmax)
            log(f"  {extras['batched']}")
        except Exception as e:
            extras["batched"] = {"error": repr(e)}
        if a.mlp_steps > 0:
            log(f"perceptron decoder, {a.mlp_steps} training batches ...")
            try:
                extras["mlp_decoder"] = bench_mlp(a.mlp_steps, a.mlp_eval_shots)
                log(f"  {extras['mlp_decoder']}")
            except Exception as e:
                extras["mlp_decoder"] = {"error": repr(e)}
    cpu = None
    if not a.no_cpu:
        log("cpu baseline (oracle) ...")
        cpu = cpu_baseline()
        log(f"  {cpu}")
    achieved = fr["rec_gbs"]
    line = {
        "metric": METRIC,
        "value": fr["shots_s"],
        "unit": "shots/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": fr["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "rotated surface-code memory-Z d=5, 25 rounds, p=1e-3, "
                               "10^7 shots per step, native RNG, record only",
                   "shots_per_step": fr["shots"], "record_bits_per_shot": fr["nbits"],
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": TRAFFIC_BYTES_PER_SHOT * fr["shots"]},
        "int_valu": {"bitops_per_shot": fr["bitops_per_shot"], "bitops_s": fr["bitops_s"],
                     "lane_ops_s": fr["lane_ops_s"], "peak_lane_ops_s": INT_VALU_PEAK,
                     "frac": fr["lane_ops_s"] / INT_VALU_PEAK},
        "ratios": headline_ratios(extras),
        "extras": extras,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
