"""Host side of the per-circuit kernels (no GPU needed): which programs the
split storage mode takes, that its LDS planes fit one compute unit's 160 KB,
and that its source compiles with hiprtc for gfx950."""
import ctypes as C
import re

import pytest

import qcf
from qcf import _lib, circuits

SPLIT = 1 | 4 | 8 | 16  # fresh frame, frame larger than LDS, 32-bit record offsets, split


def source(prog, variant):
    n = C.c_size_t()
    _lib.call("qcf_debug_jit_source", prog.handle, variant, None, 0, C.byref(n))
    buf = C.create_string_buffer(n.value + 1)
    _lib.call("qcf_debug_jit_source", prog.handle, variant, buf, n.value + 1, C.byref(n))
    return buf.value.decode()


def lds_planes(src):
    """Largest LDS plane index the kernel addresses, plus one."""
    hi = 0
    for w, off in re.findall(r"L\[lb(\d+) \+ (\d+)u\]", src):
        hi = max(hi, int(w) * 256 + int(off) // 64 + 1)
    for off in re.findall(r"L\[(\d+)u \+ l\]", src):
        hi = max(hi, int(off) // 64 + 1)
    return hi


def test_split_source_for_the_reference_benchmark_circuit():
    prog = qcf.Program(circuits.unrotated_surface_code_benchmark(11, 12))
    assert prog.info.nplanes == 884 and prog.info.lds_bytes > 160 * 1024
    src = source(prog, SPLIT)
    regs = set(re.findall(r"\br(\d+) = 0u", src))
    n_lds = lds_planes(src)
    assert n_lds <= 640  # 160 KB of 256-byte planes
    assert len(regs) + n_lds >= prog.info.nplanes
    assert "atomicExch" not in src  # noiseless: no slot exchanges


def test_split_source_noisy_rotated_code_uses_slots():
    prog = qcf.Program(circuits.rotated_surface_code(13, 2, 5e-3).ops)
    src = source(prog, SPLIT)
    assert lds_planes(src) <= 640
    assert "atomicExch" in src  # events on register planes pass through slots


def test_split_mode_declines_frames_beyond_registers_and_lds():
    """d=15 needs 1237 planes: more than LDS plus the register budget, so the
    generator declines and launches fall back to global-memory frames."""
    prog = qcf.Program(circuits.rotated_surface_code(15, 15, 1e-3).ops)
    with pytest.raises(_lib.MethodError):
        source(prog, SPLIT)


def test_split_source_compiles():
    prog = qcf.Program(circuits.unrotated_surface_code_benchmark(11, 2))
    n = C.c_size_t()
    _lib.call("qcf_debug_jit_compile", prog.handle, SPLIT, C.byref(n))
    assert n.value > 0


def test_two_wave_split_halves_the_lds():
    """Variant bit 5: split mode within half of a CU's LDS (two waves per CU);
    rotated d=9 (324 qubit planes, one wave per CU in LDS mode) fits it."""
    prog = qcf.Program(circuits.rotated_surface_code(9, 9, 1e-3).ops)
    src = source(prog, 1 | 8 | 16 | 32)
    assert lds_planes(src) <= 320
    big = qcf.Program(circuits.rotated_surface_code(13, 13, 1e-3).ops)
    with pytest.raises(_lib.MethodError):  # 929 planes: too many for 320 in LDS + 400 registers
        source(big, 1 | 4 | 8 | 16 | 32)


def test_half_tile_source_stays_in_lds():
    """Variant bit 6: every plane in LDS at 32 words (one half of the tile's
    shot words), lanes 0..31; rotated d=15 (1237 planes, beyond split mode)
    takes 158 KB.  Every constant LDS address lies inside the planes."""
    prog = qcf.Program(circuits.rotated_surface_code(15, 2, 5e-3).ops)
    src = source(prog, 1 | 4 | 8 | 64)
    hi = 0
    for w, off in re.findall(r"L\[lb(\d+) \+ (\d+)u\]", src):
        hi = max(hi, int(w) * 16384 + int(off) + 32)
    for off in re.findall(r"L\[(\d+)u \+ l\]", src):
        hi = max(hi, int(off) + 32)
    assert 0 < hi <= prog.info.nplanes * 32 <= 160 * 1024 // 4
    assert "if (lh) {" in src and "FlipPart<32> flip{L, hf}" in src
    with pytest.raises(_lib.MethodError):  # 2,004 planes do not fit even at half width
        source(qcf.Program(circuits.x_diag_circuit(1000)), 1 | 4 | 8 | 64)


def test_quarter_tile_source_stays_in_lds():
    """Variant bit 7: planes of 16 words (a quarter of the tile), lanes 0..15,
    four passes: x_diag_circuit(1000) (benchmark/benchmarks.jl:103-110, 2,004
    planes) fits in 128 KB."""
    prog = qcf.Program(circuits.x_diag_circuit(1000))
    src = source(prog, 1 | 4 | 8 | 128)
    hi = 0
    for w, off in re.findall(r"L\[lb(\d+) \+ (\d+)u\]", src):
        hi = max(hi, int(w) * 16384 + int(off) + 16)
    for off in re.findall(r"L\[(\d+)u \+ l\]", src):
        hi = max(hi, int(off) + 16)
    assert 0 < hi <= prog.info.nplanes * 16 <= 160 * 1024 // 4
    assert "const bool lh = l < 16u;" in src and "FlipPart<16> flip{L, hf}" in src


def _part_lds_top(src, pw):
    hi = 0
    for w, off in re.findall(r"L\[lb(\d+) \+ (\d+)u\]", src):
        hi = max(hi, int(w) * 16384 + int(off) + pw)
    for off in re.findall(r"L\[(\d+)u \+ l\]", src):
        hi = max(hi, int(off) + pw)
    return hi


def _reg_planes(src):
    m = re.search(r"^    uint32_t (r\d+ = 0u(?:, r\d+ = 0u)*);$", src, re.M)
    return [int(x) for x in re.findall(r"r(\d+) = 0u", m.group(1))] if m else []


def test_part_tile_kernels_keep_noise_free_planes_in_registers(monkeypatch):
    """x_diag_circuit_noisy_measurement(1000): every qubit 0..999 is measured
    through a noise site and XORed into the target (qubit 1000, planes 2000
    and 2001), and 7 classical slots (planes 2004..2010) go into each of its
    records.  The quarter-tile kernel keeps the target and the slots (no
    noise site names them) in registers and addresses LDS only below plane
    2004; QCF_JIT_PART_REGS=0 puts them back in LDS.  Rotated d=15 (337
    slots) takes at most the default cap of 128, none of them a noise site's."""
    prog = qcf.Program(circuits.x_diag_circuit_noisy_measurement(1000))
    assert prog.info.nplanes == 2 * prog.info.nqubits + 7
    qpl = 2 * prog.info.nqubits
    src = source(prog, 1 | 4 | 8 | 128)
    assert _reg_planes(src) == [2000, 2001] + list(range(2004, 2011))
    assert 0 < _part_lds_top(src, 16) <= qpl * 16
    assert "L[lb1 + 15616u]" not in src  # plane 2000's LDS place
    monkeypatch.setenv("QCF_JIT_PART_REGS", "0")
    src0 = source(prog, 1 | 4 | 8 | 128)
    assert _reg_planes(src0) == []
    assert qpl * 16 < _part_lds_top(src0, 16) <= prog.info.nplanes * 16
    monkeypatch.delenv("QCF_JIT_PART_REGS")
    d15 = qcf.Program(circuits.rotated_surface_code(15, 2, 5e-3).ops)
    s15 = source(d15, 1 | 4 | 8 | 64)
    regs = _reg_planes(s15)
    assert 0 < len(regs) <= 128 and min(regs) >= 2 * d15.info.nqubits
    assert _part_lds_top(s15, 32) <= d15.info.nplanes * 32


def test_large_frames_with_many_noise_groups_stay_on_the_interpreter():
    """1,000 noise sites on one qubit, each between two gates on it, stay
    1,000 separate noise groups, which would compile for minutes: split and
    part-tile kernels decline the program."""
    n = 1000
    ops = []
    for i in range(1, n + 1):
        ops += [qcf.sHadamard(i), qcf.sCNOT(i, n + 1), qcf.ops.PauliError(n + 1, 0.1)]
    ops.append(qcf.sMZ(n + 1, 1))
    prog = qcf.Program(ops)
    for variant in (1 | 4 | 8 | 16, 1 | 4 | 8 | 64, 1 | 4 | 8 | 128):
        with pytest.raises(_lib.MethodError):
            source(prog, variant)


def test_noise_moves_back_past_operations_on_other_qubits():
    """x_diag_circuit_noisy_measurement(1000) (benchmark/benchmarks.jl:161-171):
    each PauliError(i) sits between operations on other qubits only (H(i) is
    a relabelling of planes), so its 1000 sites join one run and the
    quarter-tile kernel makes 16 event passes of at most 63 sites."""
    prog = qcf.Program(circuits.x_diag_circuit_noisy_measurement(1000))
    kernel = source(prog, 1 | 8 | 128).split("qcf_jit_kernel")[-1]
    assert kernel.count("qcfd::group_events") == 16


LINEAR = 1 | 8 | 256  # fresh, record only, 32-bit record offsets, linear mode


def test_linear_source_for_the_reference_benchmark_circuit():
    """Variant bit 8: a noiseless fresh record-only launch keeps no frame; its
    columns are XORs of the kept gauge words (the reference's d=11 benchmark
    circuit: 220 words, 2,928 XORs, 2,751 columns) and the kernel compiles."""
    prog = qcf.Program(circuits.unrotated_surface_code_benchmark(11, 12))
    src = source(prog, LINEAR)
    assert "// linear: 220 gauge words, 2928 XORs, 2751 record columns" in src
    kernel = src.split("qcf_jit_kernel")[-1]
    assert "L[" not in kernel  # no LDS planes
    assert kernel.count("g.next()") == 220
    n = C.c_size_t()
    _lib.call("qcf_debug_jit_compile", prog.handle, LINEAR, C.byref(n))
    assert n.value > 0


def test_linear_mode_needs_a_noiseless_record_only_launch():
    noisy = qcf.Program(circuits.rotated_surface_code(3, 3, 1e-3).ops)
    with pytest.raises(_lib.MethodError):
        source(noisy, LINEAR)
    clean = qcf.Program(circuits.rotated_surface_code(3, 3, 0.0, noise=False).ops)
    source(clean, LINEAR)
    with pytest.raises(_lib.MethodError):  # the frame is returned: no linear mode
        source(clean, LINEAR | 2)


def test_zero_rate_noise_compiles_like_no_noise():
    """Noise ops at rate 0 keep their sites (mask replay needs them), but
    native draws never fire there: the specialised kernels carry no event
    passes or slot exchanges for them, and the record-only launch takes the
    linear kernel with the same XORs as the circuit without noise ops."""
    zero = qcf.Program(circuits.rotated_surface_code(5, 25, 0.0).ops)
    clean = qcf.Program(circuits.rotated_surface_code(5, 25, 0.0, noise=False).ops)
    assert zero.info.nsites > 0 and zero.info.nclasses == 0
    for v in (1 | 8, 2 | 8, 3 | 8):  # fresh record-only, in place + frame, fresh + frame
        kernel = source(zero, v).split("qcf_jit_kernel")[-1]
        assert "group_events" not in kernel and "atomicExch" not in kernel
    lz, lc = source(zero, LINEAR), source(clean, LINEAR)
    assert lz.splitlines()[1] == lc.splitlines()[1] == \
        "// linear: 24 gauge words, 565 XORs, 1226 record columns"


def test_in_place_frame_loads_batched_and_addresses_not_kept():
    """The in-place prologue requests its frame loads in batches before
    storing any (one latency per batch, not per few loads), and the epilogue
    makes the plane stride opaque before its stores, so the compiler does not
    keep every frame address of the loads live through the tile."""
    prog = qcf.Program(circuits.rotated_surface_code(5, 25, 1e-3).ops)
    kernel = source(prog, 2 | 8).split("qcf_jit_kernel")[-1]
    loads = kernel.split("__builtin_amdgcn_sched_barrier(0);")
    assert len(loads) >= 4  # 100 planes in batches of at most 32
    assert loads[0].count("P.frame[") == 32
    last_load = kernel.rindex("= P.frame[")
    barrier = kernel.index('asm volatile("" : "+s"(ld));', last_load)
    assert barrier < kernel.index("P.frame[", barrier + 1)


def test_event_passes_raise_priority_in_lds_mode(monkeypatch):
    """LDS-mode kernels (several waves per SIMD) raise a wave's issue priority
    for the length of each event pass and drop it after; register and split
    kernels keep one priority; QCF_JIT_PASS_PRIO=0 turns it off."""
    lds = qcf.Program(circuits.rotated_surface_code(5, 25, 1e-3).ops)
    kernel = source(lds, 1 | 8).split("qcf_jit_kernel")[-1]
    n_up = kernel.count("__builtin_amdgcn_s_setprio(1);")
    assert n_up > 0 and n_up == kernel.count("__builtin_amdgcn_s_setprio(0);")
    assert kernel.count("qcfd::group_events") == n_up
    for other, v in ((circuits.rotated_surface_code(5, 25, 1e-2).ops, 1 | 8),      # registers
                     (circuits.rotated_surface_code(13, 2, 1e-3).ops, 1 | 8 | 16)):  # split
        assert "s_setprio" not in source(qcf.Program(other), v).split("qcf_jit_kernel")[-1]
    monkeypatch.setenv("QCF_JIT_PASS_PRIO", "0")
    assert "s_setprio" not in source(lds, 1 | 8).split("qcf_jit_kernel")[-1]


def test_lds_mode_kernels_ask_for_ilp_scheduling(monkeypatch):
    """LDS-mode kernels carry a compiler-option line for the max-ilp machine
    scheduler; register and split kernels and the linear kernel do not;
    QCF_JIT_SCHED=0 drops it; the tagged kernel compiles."""
    tag = "// qcf-compile-options: -mllvm -amdgpu-sched-strategy=max-ilp"
    lds = qcf.Program(circuits.rotated_surface_code(5, 25, 1e-3).ops)
    assert tag in source(lds, 1 | 8)
    n = C.c_size_t()
    _lib.call("qcf_debug_jit_compile", lds.handle, 1 | 8, C.byref(n))
    assert n.value > 0
    assert tag not in source(qcf.Program(circuits.rotated_surface_code(5, 25, 1e-2).ops), 1 | 8)
    assert tag not in source(qcf.Program(circuits.rotated_surface_code(13, 2, 1e-3).ops), 1 | 8 | 16)
    clean = qcf.Program(circuits.rotated_surface_code(3, 3, 0.0, noise=False).ops)
    assert tag not in source(clean, LINEAR)
    # two-wave split kernels ask for the low-register-pressure strategy
    minreg = "// qcf-compile-options: -mllvm -amdgpu-sched-strategy=iterative-minreg"
    two = source(qcf.Program(circuits.rotated_surface_code(9, 3, 1e-3).ops), 1 | 8 | 16 | 32)
    assert minreg in two and tag not in two
    monkeypatch.setenv("QCF_JIT_SCHED", "0")
    assert tag not in source(lds, 1 | 8)


def test_split_frames_with_many_register_planes_take_fewer_slots_and_64_bit_records(monkeypatch):
    """Rotated d=13 x 13 (929 planes) would need more than 256 register planes
    beside 80 noise slots: its split kernel takes 24 slots and 64-bit record
    addresses; d=11 x 11 keeps 80 slots and 32-bit record offsets."""
    big = source(qcf.Program(circuits.rotated_surface_code(13, 13, 1e-3).ops), 1 | 8 | 16)
    small = source(qcf.Program(circuits.rotated_surface_code(11, 11, 1e-3).ops), 1 | 8 | 16)
    kb, ks = big.split("qcf_jit_kernel")[-1], small.split("qcf_jit_kernel")[-1]
    assert "(recb +" not in kb and "P.rec[" in kb
    assert "(recb +" in ks and "P.rec[" not in ks
    # its noise groups are cut to fit 24 slots: more event passes than with 80
    monkeypatch.setenv("QCF_JIT_SLOTS", "80")
    k80 = source(qcf.Program(circuits.rotated_surface_code(13, 13, 1e-3).ops),
                 1 | 8 | 16).split("qcf_jit_kernel")[-1]
    assert kb.count("qcfd::group_events") > k80.count("qcfd::group_events")
