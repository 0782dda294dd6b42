"""Shot offsets and frame views.

The native draws are keyed on absolute positions (gauge words per absolute
32-shot word, noise events per absolute 2048-shot tile), so:
* a run split into pieces at multiples of 2048 shots, each launched with its
  shot_offset, reproduces the bits of the one run, on the specialised
  kernels and on the interpreter;
* pftrajectories on view(frame, r) (src/pauli_frames.jl:17) writes only the
  shots of r, and the views of a split frame together equal the whole frame
  run in place (pftrajectories(state, circuit), src/pauli_frames.jl:234-239).
"""
import numpy as np
import pytest
import torch

import oracle
import qcf
from qcf import circuits

pytestmark = pytest.mark.gpu


def bits(planes, shots):
    return oracle.unpack_shots(planes, shots)


@pytest.mark.parametrize("no_jit", ["0", "1"])
def test_split_runs_equal_one_run(gpu, monkeypatch, no_jit):
    monkeypatch.setenv("QCF_NO_JIT", no_jit)
    prog = qcf.Program(circuits.rotated_surface_code(5, 5, 5e-3).ops)
    n, cut, seed = 5 * 2048 + 301, 3 * 2048, 7071
    one = qcf.pftrajectories(prog, n, seed=seed)
    a = qcf.pftrajectories(prog, cut, seed=seed, shot_offset=0)
    b = qcf.pftrajectories(prog, n - cut, seed=seed, shot_offset=cut)
    rec = bits(one.record_host(), n)
    assert np.array_equal(rec[:, :cut], bits(a.record_host(), cut))
    assert np.array_equal(rec[:, cut:], bits(b.record_host(), n - cut))
    fr = bits(one.frame_host(), n)
    assert np.array_equal(fr[:, cut:], bits(b.frame_host(), n - cut))
    assert rec[np.asarray(circuits.rotated_surface_code(5, 5, 5e-3).det_bits) - 1].any()


def test_native_offset_must_be_tile_aligned(gpu):
    prog = qcf.Program(circuits.repetition_code(3, 2, 1e-2).ops)
    with pytest.raises(qcf.ArgumentError):
        qcf.pftrajectories(prog, 100, seed=1, shot_offset=64)


def test_views_of_a_split_frame_equal_the_whole_frame(gpu):
    c = circuits.rotated_surface_code(3, 3, 2e-2)
    prog = qcf.Program(c.ops)
    n, cut = 3 * 2048 + 77, 2048
    base = qcf.PauliFrame(n, prog.nq - 1, prog.nbits, init=False)
    base.frame.random_(-2**62, 2**62)
    base.record.random_(-2**62, 2**62)
    whole, parts = base.copy(), base.copy()
    qcf.pftrajectories(whole, prog, seed=99)
    v1, v2 = qcf.view(parts, slice(0, cut)), qcf.view(parts, slice(cut, n))
    assert (v1.offset, v2.offset, len(v2)) == (0, cut, n - cut)
    qcf.pftrajectories(v1, prog, seed=99)
    # the second view has not run yet: its shots are still the initial ones
    assert np.array_equal(bits(parts.record_host(), n)[:, cut:], bits(base.record_host(), n)[:, cut:])
    qcf.pftrajectories(v2, prog, seed=99)
    assert np.array_equal(bits(parts.frame_host(), n), bits(whole.frame_host(), n))
    assert np.array_equal(bits(parts.record_host(), n), bits(whole.record_host(), n))
    # the view's own accessors see its shots
    assert np.array_equal(bits(v2.record_host(), n - cut), bits(whole.record_host(), n)[:, cut:])
    m = qcf.measurements(v2).cpu().numpy()
    assert np.array_equal(m.T.astype(np.uint8), bits(whole.record_host(), n)[:, cut:])


def test_view_replayed_masks_at_any_word(gpu):
    """Replayed masks need no tile alignment: a view starting on a 64-shot word
    equals the oracle on its shots."""
    rng = np.random.default_rng(5)
    circ = [qcf.sHadamard(1), qcf.sCNOT(1, 2), qcf.PauliError(2, 0.2), qcf.sMRZ(2, 1),
            qcf.sMZ(1, 2)]
    prog = qcf.Program(circ)
    n, start = 1000, 320
    f = qcf.PauliFrame(n, prog.nq - 1, prog.nbits, init=False)
    fb = rng.integers(0, 2, (2 * prog.nq, n)).astype(np.uint8)
    planes = np.zeros((2 * prog.nq, f.ld), np.uint64)
    planes[:, : qcf.frames.words(n)] = oracle.pack_shots(fb)
    f.frame.copy_(torch.from_numpy(planes.view(np.int64).reshape(-1)))
    v = qcf.view(f, slice(start, n))
    mb = rng.integers(0, 2, (prog.nreplay, n - start)).astype(np.uint8)
    mk = np.zeros((prog.nreplay, v.ld), np.uint64)  # one stride for frame, record, masks
    mk[:, : qcf.frames.words(n - start)] = oracle.pack_shots(mb)
    qcf.pftrajectories(v, prog, masks=torch.from_numpy(mk.view(np.int64).reshape(-1)).cuda())
    ops, args = prog.ops, prog.args
    xzs = oracle.frame_to_reference(oracle.pack_shots(fb[:, start:]), prog.nq, n - start)
    _, om = oracle.pf_run(ops, args, prog.nq, prog.nbits, n - start, masks=oracle.pack_shots(mb),
                          fresh=False, xzs=xzs)
    assert np.array_equal(bits(v.record_host(), n - start), om[: prog.nbits])
    # shots before the view are untouched
    assert np.array_equal(bits(f.frame_host(), n)[:, :start], fb[:, :start])


def test_view_range_checks(gpu):
    f = qcf.PauliFrame(1000, 2, 1, init=False)
    with pytest.raises(qcf.ArgumentError):
        qcf.view(f, slice(10, 500))
    with pytest.raises(qcf.ArgumentError):
        qcf.view(f, slice(0, 500))
    with pytest.raises(qcf.BoundsError):
        qcf.view(f, range(0, 2000))
    v = qcf.view(f, slice(64, 1000))
    with pytest.raises(qcf.ArgumentError):  # native draws need a 2048-shot start
        qcf.pftrajectories(v, [qcf.PauliError(1, 0.1)], seed=3)
