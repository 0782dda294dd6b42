/*
 * qcf.h -- C ABI of the MI355X (gfx950) Pauli-frame sampler and bit-packed
 * Pauli product.  Everything here is plain C: device buffers are raw pointers
 * obtained from hipMalloc (or from a ROCArray / torch tensor), streams are
 * hipStream_t passed as void*, and every entry point returns a status code.
 *
 * Reference semantics are those of QuantumClifford.jl v0.11.7; each function
 * names the reference routine it stands in for (paths relative to the
 * reference checkout).
 *
 * Conventions
 *   - Qubits and classical bits are 1-based in qcf_op, exactly as in the
 *     reference operation structs (sCNOT(1,2), sMRZ(4,1), ...).  Bit 0 means
 *     "no classical bit" (reference sMZ(q,0) / sMRZ(q,0)).
 *   - Pauli operators use the reference layout: xz = [x words ; z words], each
 *     half `nwords` UInt64 words, qubit q at word (q-1)/64, bit (q-1)%64
 *     (src/QuantumClifford.jl:994-1001, src/pauli_operator.jl:165).  Phases are
 *     one byte each, 0:+1 1:+i 2:-1 3:-i (src/QuantumClifford.jl:118-121).
 *   - Pauli frames are shot-packed: plane-major, 64 shots per UInt64 word.
 *     A frame buffer holds 2*nq planes (x planes of qubits 1..nq, then z
 *     planes of qubits 1..nq), each `ld` words long; shot s of plane p is bit
 *     s%64 of word p*ld + s/64.  The measurement record uses the same packing:
 *     column b-1 of classical bit b is words [(b-1)*ld, (b-1)*ld + ld).
 *   - Every launching call is stream ordered and asynchronous; the caller owns
 *     all memory passed in.  The library never frees caller memory.
 */
#ifndef QCF_H
#define QCF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (mirroring the reference's exceptions) ---------------- */
#define QCF_OK 0
#define QCF_ERR_DIMENSION 1   /* DimensionMismatch (src/mul_leftright.jl:138)   */
#define QCF_ERR_BOUNDS 2      /* BoundsError (qubit/bit index beyond the frame)  */
#define QCF_ERR_ARGUMENT 3    /* ArgumentError / NoZeroQubit (src/throws.jl:1)   */
#define QCF_ERR_UNSUPPORTED 4 /* MethodError: op has no PauliFrame method        */
#define QCF_ERR_HIP 5         /* HIP runtime failure                             */
#define QCF_ERR_ALLOC 6       /* host or device allocation failure               */

/* Thread-local text of the last error set by any qcf_* call on this thread. */
const char *qcf_last_error(void);

/* Library version string and the offload arch it was built for. */
const char *qcf_version(void);

/* ======================================================================= */
/*  Bit-packed Pauli product (src/mul_leftright.jl)                         */
/* ======================================================================= */

/*
 * r := l * r on two Pauli operators of `nwords` words per half.
 * Replaces mul_left!(r::PauliOperator, l::PauliOperator; phases)
 * (src/mul_leftright.jl:137-142).  With phases != 0:
 *     r_phase[0] = (s + r_phase[0] + l_phase[0]) & 3,
 * where s is the product phase of the bit arrays (src/mul_leftright.jl:123-126).
 * With phases == 0 only the bits are updated (plain XOR).  r and l may not
 * partially overlap; r == l is allowed (the result is the identity with
 * s = 0).  All pointers are device pointers.
 */
int qcf_mul_left(uint64_t *r_xz, uint8_t *r_phase, const uint64_t *l_xz,
                 const uint8_t *l_phase, size_t nwords, int phases,
                 void *stream);

/*
 * l := l * r.  Replaces mul_right!(l::PauliOperator, r::PauliOperator)
 * (src/mul_leftright.jl:144-149): phase fold ((c1 ^ c2<<1) + 2*c1) & 3
 * (src/mul_leftright.jl:128-131), then l_phase = (s + l_phase + r_phase) & 3.
 */
int qcf_mul_right(uint64_t *l_xz, uint8_t *l_phase, const uint64_t *r_xz,
                  const uint8_t *r_phase, size_t nwords, int phases,
                  void *stream);

/*
 * Vector form: r := l * r on raw [x;z] word arrays, writing the 2-bit product
 * phase s to out_phase[0] (device byte).  Replaces
 * mul_left!(r::AbstractVector, l::AbstractVector)::UInt8
 * (src/mul_leftright.jl:123-126); right = 1 selects the mul_right! fold
 * (src/mul_leftright.jl:128-131).  out_phase may be NULL when phases == 0.
 */
int qcf_mul_vec(uint64_t *r_xz, const uint64_t *l_xz, size_t nwords,
                int phases, int right, uint8_t *out_phase, void *stream);

/*
 * Row sweep over a tableau stored row-contiguous (`fastrow`): row m occupies
 * xzs[m*ld .. m*ld + 2*nwords).  For every row m in [0, nrows) with m != skip
 * and (row_mask == NULL || row_mask[m] != 0):
 *     row m := src * row m,   phases[m] = (s_m + phases[m] + src_phase) & 3.
 * `src` is either another row of the same tableau (the pivot sweep of
 * canonicalize!, src/canonicalization.jl:105-109 over
 * mul_left!(s::Tableau, m, i), src/mul_leftright.jl:177-181; pass
 * src = xzs + i*ld, src_phase = phases + i, skip = i) or an independent Pauli
 * (mul_left!(s::Tableau, p::PauliOperator), src/mul_leftright.jl:221-227;
 * pass skip = SIZE_MAX).  `src` must not alias any row that is updated.
 * right = 1 gives the mul_right! variants (src/mul_leftright.jl:234-240).
 */
int qcf_mul_rows(uint64_t *xzs, uint8_t *phases, size_t ld, size_t nwords,
                 size_t nrows, const uint64_t *src, const uint8_t *src_phase,
                 size_t skip, const uint8_t *row_mask, int phases_on,
                 int right, void *stream);

/*
 * canonicalize!(state::Stabilizer; phases, ranks) (src/canonicalization.jl:
 * 97-135) in place on a row-contiguous tableau (row m at xzs + m*ld, words
 * [x 0..nwords) [z nwords..2*nwords), `nqubits` columns, `nrows` rows, phases
 * one byte per row).  Replaces the CUDA extension's gpu_canonicalize!
 * (ext/QuantumCliffordGPUExt/canonicalization.jl:116-166).  phases_on == 0
 * leaves `phases` untouched (the reference then neither updates nor swaps
 * them) and allows phases == NULL.  When `ranks` (device, 2 x uint32) is not
 * NULL it receives the pivots of the X stage and of both stages (the
 * reference's rx-1 and i-1 of ranks=true).  Stream-ordered; the workspace is
 * library-owned (per device, grown on demand) and shared by every call on
 * that device, so calls that overlap on one device race.  nrows <= 16384.
 */
int qcf_canonicalize(uint64_t *xzs, uint8_t *phases, size_t ld, size_t nwords,
                     size_t nrows, size_t nqubits, int phases_on, uint32_t *ranks,
                     void *stream);

/*
 * qcf_canonicalize without any allocation: no reference counterpart (the
 * reference's canonicalize! allocates nothing on the device to begin with).
 * For callers that cannot allocate -- code under hipGraph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  Same arguments, results
 * and errors as qcf_canonicalize; the workspace is a fixed 16 MiB device
 * array that each GPU holds until the process ends (a tableau of nrows rows
 * of W words needs about nrows*(16*W + 13) + 1024*W bytes: 6000 x 6000 fits),
 * and larger tableaux are rejected with QCF_ERR_ARGUMENT.  Every call on one
 * GPU shares that array, so calls that overlap on one GPU (different
 * streams, or host threads) race; calls in one stream are ordered.
 */
int qcf_canonicalize_noalloc(uint64_t *xzs, uint8_t *phases, size_t ld, size_t nwords,
                             size_t nrows, size_t nqubits, int phases_on, uint32_t *ranks,
                             void *stream);

/* ======================================================================= */
/*  Pauli frames (src/pauli_frames.jl)                                      */
/* ======================================================================= */

/* Operation kinds; each names the reference struct it represents. */
enum qcf_op_kind {
  QCF_OP_NOP = 0,
  /* single-qubit symbolic gates, src/symbolic_cliffords.jl:105-118 */
  QCF_OP_ID1 = 1, QCF_OP_HADAMARD, QCF_OP_HADAMARD_XY, QCF_OP_HADAMARD_YZ,
  QCF_OP_PHASE, QCF_OP_INV_PHASE, QCF_OP_X, QCF_OP_Y, QCF_OP_Z,
  QCF_OP_SQRTX, QCF_OP_INV_SQRTX, QCF_OP_SQRTY, QCF_OP_INV_SQRTY,
  QCF_OP_CXYZ, QCF_OP_CZYX,
  /* SingleQubitOperator(q, xx, xz, zx, zz, px, pz),
     src/symbolic_cliffords.jl:168-206: arg[0..3] = xx, xz, zx, zz */
  QCF_OP_SINGLE_QUBIT_OPERATOR = 20,
  /* two-qubit symbolic gates, src/symbolic_cliffords.jl:359-395 */
  QCF_OP_SWAP = 32, QCF_OP_SWAPCX, QCF_OP_INV_SWAPCX, QCF_OP_ISWAP,
  QCF_OP_INV_ISWAP, QCF_OP_CZSWAP, QCF_OP_CXSWAP, QCF_OP_CNOT, QCF_OP_CPHASE,
  QCF_OP_ZCX, QCF_OP_ZCY, QCF_OP_ZCZ, QCF_OP_XCX, QCF_OP_XCY, QCF_OP_XCZ,
  QCF_OP_YCX, QCF_OP_YCY, QCF_OP_YCZ, QCF_OP_ZCRY, QCF_OP_INV_ZCRY,
  QCF_OP_SQRTZZ, QCF_OP_INV_SQRTZZ, QCF_OP_SQRTXX, QCF_OP_INV_SQRTXX,
  QCF_OP_SQRTYY, QCF_OP_INV_SQRTYY,
  /* measurements, src/pauli_frames.jl:74-116 */
  QCF_OP_MZ = 64,   /* sMZ(q, bit)                                     */
  QCF_OP_MRZ,       /* sMRZ(q, bit)                                    */
  QCF_OP_MX,        /* sMX(q, bit)  = H then sMZ, no-op when bit == 0  */
  QCF_OP_MRX,       /* sMRX(q, bit) = H then sMRZ                      */
  /* PauliMeasurement(pauli, bit), src/pauli_frames.jl:118-134:
     arg_off/arg_len index `args` holding one code per qubit 1..len
     (0=I 1=X 2=Z 3=Y) followed by the phase byte. */
  QCF_OP_PAULI_MEASUREMENT,
  /* noise, src/pauli_frames.jl:136-174 and src/noise.jl:26-72 */
  QCF_OP_PAULI_ERROR = 80,  /* UnbiasedUncorrelatedNoise(p[0]) on q1      */
  QCF_OP_PAULI_NOISE,       /* PauliNoise(px=p[0], py=p[1], pz=p[2]) on q1 */
  /* Two-qubit depolarizing on (q1,q2): with probability p[0] one of the 15
     non-identity two-qubit Paulis, uniformly.  No PauliFrame method exists in
     the reference; the distribution is that of
     DepolarizationNoise(two_qubit_error_probability=p[0]) applied jointly
     (src/noise.jl:49-53, 99-129). */
  QCF_OP_DEPOLARIZE2,
  /* ClassicalXOR(bits, store), src/misc_ops.jl:198-205, applied as in
     src/pauli_frames.jl:63-72: args[arg_off .. arg_off+arg_len) are the
     1-based input bits, `bit` is the store bit. */
  QCF_OP_CLASSICAL_XOR = 96,
};

typedef struct qcf_op {
  uint32_t kind;    /* enum qcf_op_kind                              */
  uint32_t q1, q2;  /* 1-based qubits (q2 only for two-qubit ops)    */
  uint32_t bit;     /* 1-based classical bit, 0 = none               */
  uint32_t arg_off; /* offset into the shared `args` array           */
  uint32_t arg_len; /* number of entries of `args` used              */
  double p[3];      /* noise parameters                              */
} qcf_op;

/* Opaque compiled circuit: the op list lowered to the device instruction
 * stream (qubit x/z planes renamed through permutation gates, classical
 * bits assigned to on-chip slots, noise sites grouped into rate classes). */
typedef struct qcf_program qcf_program;

typedef struct qcf_program_info {
  uint32_t nqubits;      /* frame qubits incl. the hidden ancilla (reference n+1) */
  uint32_t nbits;        /* classical bits (record columns)                      */
  uint32_t nplanes;      /* on-chip planes per shot word (2*nqubits + slots)     */
  uint32_t ninstr;       /* device instructions                                  */
  uint32_t nsites;       /* noise sites                                          */
  uint32_t nclasses;     /* rate classes                                         */
  uint32_t nreplay;      /* replay mask planes (RNG draws in program order)      */
  uint32_t lds_bytes;    /* LDS per 2048-shot block                              */
  double bitops_per_shot;/* frame bit-ops per shot (shared accounting)           */
} qcf_program_info;

/*
 * Compile a circuit.  `nqubits`/`nbits` size the frame exactly like
 * PauliFrame(frames, qubits, measurements) (src/pauli_frames.jl:30-37): the
 * frame has nqubits+1 qubits (the extra one serves PauliMeasurement) and
 * nbits record columns.  Pass 0 for either to size from the circuit as
 * _create_pauliframe does (src/pauli_frames.jl:197-201).
 * Errors: QCF_ERR_ARGUMENT (qubit 0, equal qubits in a two-qubit gate, bad
 * probabilities), QCF_ERR_BOUNDS (index beyond the sizes given),
 * QCF_ERR_UNSUPPORTED (no PauliFrame method in the reference, e.g. sMY).
 */
int qcf_program_build(const qcf_op *ops, size_t nops, const uint32_t *args,
                      size_t nargs, uint32_t nqubits, uint32_t nbits,
                      qcf_program **out);
int qcf_program_free(qcf_program *prog);
int qcf_program_get_info(const qcf_program *prog, qcf_program_info *info);

/*
 * Gauge draws of a fresh, record-only launch (QCF_PF_FRESH without a frame
 * buffer).  No counterpart in the reference: it describes what this library
 * leaves out.  out[r] for every replay plane r < info.nreplay: 0 = a noise
 * plane, 1 = a gauge draw (initZ! z word, sMRZ/sMRX z-flip) such a launch
 * makes, 2 = a gauge draw it leaves out because the draws kept before it
 * already span every record pattern it could produce: the record is linear
 * over GF(2) in the gauge bits, so its distribution is unchanged.  A native
 * launch makes no draw there and a mask export leaves that plane as the
 * caller zeroed it; replay applies every plane.  Launches that return the
 * frame draw everything (the frame's gauge part spans more).
 * Errors: QCF_ERR_ARGUMENT, QCF_ERR_DIMENSION (n < info.nreplay).
 */
int qcf_program_gauge_planes(const qcf_program *prog, uint8_t *out, size_t n);

/* qcf_pf_sample flags */
#define QCF_PF_FRESH 1u        /* start from PauliFrame(...) + initZ!: x=0, z random,
                                  record zeroed (pftrajectories(circuit; trajectories)) */
#define QCF_PF_EXPORT_MASKS 2u /* native RNG, and write every drawn mask to `masks`
                                  (which must be zeroed by the caller)            */
#define QCF_PF_VIEW 4u         /* the planes' stride also holds shots outside this
                                  launch (view(frame, r)): write only the launch's own
                                  64-shot words, never the zero padding up to the
                                  stride.  `shots` must then end on a 64-shot word or
                                  at the end of the stride's shots.                 */

/*
 * Run the circuit on `shots` frames: pftrajectories(state::PauliFrame, circuit)
 * (src/pauli_frames.jl:234-239) and, with QCF_PF_FRESH,
 * pftrajectories(circuit; trajectories=shots) (src/pauli_frames.jl:193-227).
 *
 *   frame  : 2*info.nqubits planes of `ld` words (in/out; may be NULL with
 *            QCF_PF_FRESH when the final frame is not wanted)
 *   meas   : info.nbits columns of `ld` words (in/out)
 *   masks  : NULL for the native counter-based RNG; otherwise info.nreplay
 *            planes of `ld` words that replace every random draw in program
 *            order -- the replayed-mask mode.  Planes 0..nqubits-1 are the
 *            initZ! z words of qubits 1..nqubits (read only with
 *            QCF_PF_FRESH); then one plane per sMRZ/sMRX z-flip (and per
 *            PauliMeasurement, which resets its ancilla), two (x flip,
 *            z flip) per single-qubit noise site and four (x1,z1,x2,z2) per
 *            DEPOLARIZE2 site, in op order.  With QCF_PF_EXPORT_MASKS the
 *            native draws are written there instead; a fresh launch without
 *            a frame buffer leaves the planes of the gauge draws it skips
 *            (qcf_program_gauge_planes) as the caller zeroed them.
 *   ld     : words per plane, >= ceil(shots/64); qcf_plane_words(shots)
 *            is the recommended stride
 *   shot_offset : absolute index of this launch's first shot in the caller's
 *            experiment (a multiple of 2048 for native draws; ignored when
 *            masks are replayed).  The buffers hold only this launch's shots.
 *
 * Native draws (no reference counterpart: the reference draws from Julia's
 * task-local Xoshiro, src/pauli_frames.jl:47-56,110-113,136-174):
 *   - gauge words (initZ!, the sMRZ z-flip): per absolute 32-shot word, a
 *     Weyl sequence with its own odd increment, each term scrambled by the
 *     murmur3 32-bit finalizer (SplittableRandom's construction), start and
 *     increment from Philox4x32-10(seed, word);
 *   - noise events: one stream of event positions per absolute 2048-shot
 *     tile along its (site ordinal, shot) sequence, drawn 64 at a time (one
 *     geometric gap per lane from a per-lane xoshiro128** stream seeded by
 *     Philox4x32-10(seed, tile, lane), then a wave prefix sum); per event a
 *     Pauli-choice draw and, only in programs where some site keeps part of
 *     its rate class's events, a thinning draw.
 * So the bits of a shot depend on (seed, its absolute tile) and not on the
 * launch shape, batching or kernel: a run split at multiples of 2048 shots
 * (with shot_offset) reproduces the one-run bits.  A fresh launch without a
 * frame buffer makes only the gauge draws that can reach the record (see
 * qcf_program_gauge_planes): its records have the distribution of a launch
 * that draws all of them, not the same bits.
 * Workspace: launches whose frames live in global memory (frames beyond
 * what the specialised kernels hold on chip, and replay/export launches of
 * frames over 160 KB) use a library-owned workspace per (device, stream):
 * calls on one stream run in order, calls on different streams never share
 * one.  It grows on demand behind a synchronisation of that stream; while
 * the stream is being captured into a graph it cannot grow, and the call
 * returns QCF_ERR_ARGUMENT (make one uncaptured call of the same size
 * first).
 * Errors: QCF_ERR_ARGUMENT for a native shot_offset that is not a multiple
 * of 2048, QCF_ERR_DIMENSION for ld < ceil(shots/64).
 */
int qcf_pf_sample(const qcf_program *prog, uint64_t shots, uint64_t shot_offset, uint64_t seed,
                  uint64_t *frame, uint64_t *meas, uint64_t *masks, size_t ld,
                  uint32_t flags, void *stream);

/*
 * Recommended plane stride in words for `shots` shots: ceil(shots/2048)*32
 * (at least 32).  Every 2048-shot tile of a plane is then one aligned
 * 256-byte segment, so each record store writes whole memory sectors; a
 * stride of exactly ceil(shots/64) words leaves planes misaligned whenever
 * shots is not a multiple of 2048, and the partial sectors at every tile edge
 * cost up to half the record bandwidth.  No reference counterpart (the
 * reference's PauliFrame keeps one Bool per shot, src/pauli_frames.jl:30-37).
 */
size_t qcf_plane_words(uint64_t shots);

/* Launch knobs chosen by the autotuner (0 = library default). */
typedef struct qcf_tune {
  uint32_t waves_per_cu;   /* persistent kernel: resident single-wave workgroups
                              per CU (grid size)                                */
  uint32_t tile_blocks;    /* persistent kernel: 2048-shot tiles per work item    */
  uint32_t lds_cap;        /* persistent kernel: bytes of LDS a circuit's frame may
                              take (0 = all the device allows); larger frames live
                              in global memory, which keeps more waves resident  */
  uint32_t jit_max_planes; /* circuits whose frames have more planes than this
                              run on the persistent kernel instead of their
                              specialised kernel (0 = every circuit that has one
                              runs it)                                           */
  uint32_t streams;        /* helper streams the specialised launches are spread
                              over, 1..4 (0 = QCF_BATCH_STREAMS, default 2)      */
  uint32_t concurrent;     /* 1: the persistent kernel starts beside the
                              specialised launches instead of after them        */
  uint32_t steal_planes;   /* specialised launches whose circuits' frames do not
                              fit in LDS and have more planes than this share
                              their tiles with the persistent kernel, which then
                              runs beside them: both take tiles from one counter
                              per launch (0 = those with more than 768 planes,
                              0xFFFFFFFF = none); the records do not depend on
                              the split                                         */
} qcf_tune;

/*
 * Batched launch (no reference counterpart; the reference runs one circuit
 * per pftrajectories call): `n` compiled circuits, circuit k runs shots[k]
 * fresh frames with seed seeds[k] into meas[k] (its nbits columns of ld[k]
 * words).  qcf_batch_create uploads the per-circuit parameters and the work
 * queue once and compiles the circuits' specialised kernels.
 * qcf_batch_launch first runs the specialised kernels, one launch per
 * distinct kernel, on helper streams forked from `stream` and joined back
 * into it (two by default; QCF_BATCH_STREAMS, 1..4).  It then enqueues
 * ONE persistent kernel for the circuits without one (frames beyond what
 * registers plus LDS hold); its workgroups pull (circuit, tile range) items
 * from a device queue.  By default the specialised launches whose frames
 * exceed LDS and have more than 768 planes (split kernels at one wave per
 * CU, part-tile kernels holding 32 shot words per CU) share their tiles
 * with it: the
 * persistent kernel then starts beside them and, once its own items are
 * done, takes tiles from the same per-launch counters they take them from
 * (qcf_tune.steal_planes).  The last workgroup to leave resets the
 * counters.  The launch allocates and compiles nothing, so it
 * can be captured in a hipGraph; qcf_batch_launch_graph captures it once and
 * replays the graph.  QCF_BATCH_JIT=0 sends every circuit to the persistent
 * kernel.  Results are bit-identical to n qcf_pf_sample(..., QCF_PF_FRESH)
 * calls.
 */
typedef struct qcf_batch qcf_batch;
int qcf_batch_create(const qcf_program *const *progs, size_t n,
                     const uint64_t *shots, const uint64_t *seeds,
                     uint64_t *const *meas, const size_t *ld,
                     const qcf_tune *tune, qcf_batch **out);
int qcf_batch_launch(qcf_batch *batch, void *stream);
int qcf_batch_launch_graph(qcf_batch *batch, void *stream);
int qcf_batch_free(qcf_batch *batch);

/* Blocking convenience: create, launch on `stream`, wait, free. */
int qcf_pf_sample_batched(const qcf_program *const *progs, size_t n,
                          const uint64_t *shots, const uint64_t *seeds,
                          uint64_t *const *meas, const size_t *ld,
                          const qcf_tune *tune, void *stream);

/* Pick qcf_tune for a set of circuits by timing batched launches of them on
 * the current device (on scratch records; the choice never changes results,
 * the RNG being keyed on absolute shots).  Each candidate runs as the graph
 * qcf_batch_launch_graph replays: one warm replay and the median of three
 * timed ones (QCF_AUTOTUNE_LOG=1 prints them).  Candidates: 1 or 2 helper
 * streams; the persistent kernel sharing the tiles of no specialised launch,
 * or of those of the largest one, two or three frame sizes among the ones it
 * can share, at 8 or 16 waves per CU; and, when it has circuits of its own
 * and shares nothing, starting it beside the specialised launches or after
 * them.  The fastest is then timed three more times against the library
 * defaults, alternately; unless it wins on both the median and the slowest
 * of those, the all-zero tune (the defaults) is returned. */
int qcf_autotune(const qcf_program *const *progs, size_t n,
                 const uint64_t *shots, qcf_tune *out, void *stream);

/* Run Philox4x32-10 on the device for n (counter[4], key[2]) records of
 * `in` (6 words each) into `out` (4 words each): exposes the sampler's RNG
 * to known-answer tests. */
int qcf_debug_philox(const uint32_t *in, uint32_t *out, size_t n,
                     void *stream);

/* Source of the kernel specialised to `prog` for a launch variant and its
 * run-time compilation with hiprtc, without loading it on a device
 * (self-check; needs no GPU).  Variant bits:
 *   0  fresh frame;
 *   1  frame stored at the end;
 *   2  frame larger than LDS;
 *   3  record under 4 GB;
 *   4  split storage (registers plus LDS);
 *   5  with 4, half of a CU's LDS per wave;
 *   6  half-tile passes;
 *   7  quarter-tile passes.
 * QCF_ERR_UNSUPPORTED if the generator does not handle the program in that
 * variant. */
int qcf_debug_jit_source(const qcf_program *prog, int variant, char *buf,
                         size_t cap, size_t *needed);
int qcf_debug_jit_compile(const qcf_program *prog, int variant,
                          size_t *code_bytes);

/* The sampler's two 32-bit generators run on the device from given states,
 * `k` outputs per state into out[k * i + j] (self-checks against published
 * known answers):
 *   qcf_debug_xoshiro128ss: the per-lane noise-event stream, xoshiro128**
 *     (Blackman and Vigna) from state[4 i .. 4 i + 3] = s0..s3;
 *   qcf_debug_gauge: the gauge words from (start, increment) =
 *     state[2 i], state[2 i + 1]: output j is fmix32(start + (j + 1) *
 *     increment), fmix32 the MurmurHash3 finalizer;
 *   qcf_debug_gauge_stream: the gauge words of absolute 32-shot word
 *     words[i] under `seed`, as initZ! and the resets draw them: start and
 *     increment from Philox4x32-10(counter (word, 0, 0, 0x47415547), key
 *     (seed low, seed high)) words x and y, the increment made odd and, when
 *     it has fewer than 12 bit transitions, XORed with 0xAAAAAAAA. */
int qcf_debug_xoshiro128ss(const uint32_t *state, uint32_t *out, size_t n, uint32_t k,
                           void *stream);
int qcf_debug_gauge(const uint32_t *state, uint32_t *out, size_t n, uint32_t k, void *stream);
int qcf_debug_gauge_stream(uint64_t seed, const uint32_t *words, uint32_t *out, size_t n,
                           uint32_t k, void *stream);

/* Inclusive prefix sums over each group of 64 words of `in`, computed by the
 * wave-wide scan the sampler uses to place noise events (self-check). */
int qcf_debug_wave_scan(const uint32_t *in, uint32_t *out, size_t nwaves,
                        void *stream);

/* The host-side check that precedes every kernel launch of the library, run
 * for `threads` threads per block against the persistent batch kernel (its
 * __launch_bounds__ is 64), without launching it: QCF_OK, or
 * QCF_ERR_ARGUMENT naming the kernel (self-check). */
int qcf_debug_check_launch(unsigned threads);

/* ---- reference-layout converters (to_gpu / to_cpu counterparts,
 *      ext/QuantumCliffordGPUExt/adapters.jl:7-41) --------------------- */

/*
 * Shot-packed record (nbits columns of ld words) -> Matrix{Bool}(shots, nbits)
 * column-major bytes, as returned by measurements(::PauliFrame)
 * (src/pauli_frames.jl:280).  And the inverse.  `ldr` is the Bool matrix's
 * column stride in elements (0 = shots): a view of a frame's rows r
 * (Base.view(::PauliFrame, r), src/pauli_frames.jl:17) passes the parent's
 * pointer advanced by first(r)-1 and the parent's row count.
 */
int qcf_record_to_bool(const uint64_t *meas, size_t ld, uint64_t shots,
                       uint32_t nbits, uint8_t *out, size_t ldr, void *stream);
int qcf_record_from_bool(const uint8_t *in, size_t ldr, uint64_t shots, uint32_t nbits,
                         uint64_t *meas, size_t ld, void *stream);

/*
 * Shot-packed frame (2*nq planes of ld words) <-> reference frame layout:
 * the `fastcolumn` xzs of PauliFrame.frame, i.e. a shots x 2W column-major
 * UInt64 matrix with W = ceil(nq/64) (src/pauli_frames.jl:32,
 * src/fastmemlayout.jl:47), `ldr` elements between its columns (0 = shots;
 * a view passes its parent's, as above).  Pad bits beyond nq are written as
 * zero by qcf_frame_to_reference and ignored by qcf_frame_from_reference.
 */
int qcf_frame_to_reference(const uint64_t *frame, size_t ld, uint64_t shots,
                           uint32_t nq, uint64_t *xzs, size_t ldr, void *stream);
int qcf_frame_from_reference(const uint64_t *xzs, size_t ldr, uint64_t shots, uint32_t nq,
                             uint64_t *frame, size_t ld, void *stream);

/* ---- matching decoder (host) ------------------------------------------- */

/*
 * Circuit-level detector error model of a compiled program: every noise site
 * and Pauli choice is one fault, pushed through the program with all gauge
 * draws zero; faults with the same detector set and observable mask merge
 * (rates combined as independent flips).  det_bits / obs_bits are 1-based
 * record bits (detectors and observables written by ClassicalXOR or
 * measurements).  Replaces the syndrome/fault bookkeeping that the reference
 * hands to PyMatching (ext/QuantumCliffordPyQDecodersExt/
 * QuantumCliffordPyQDecodersExt.jl:83-130); no reference counterpart for
 * the model itself.  At most 64 observables.
 */
typedef struct qcf_dem qcf_dem;
int qcf_dem_build(const qcf_program *prog, const uint32_t *det_bits, size_t ndet,
                  const uint32_t *obs_bits, size_t nobs, qcf_dem **out);
int qcf_dem_free(qcf_dem *dem);
int qcf_dem_size(const qcf_dem *dem, size_t *nmech, uint32_t *ndet, uint32_t *nobs);
int qcf_dem_mechanism(const qcf_dem *dem, size_t i, uint32_t *dets, size_t cap,
                      size_t *ndets, uint64_t *obs, double *p);

/*
 * Minimum-weight perfect-matching decoder over an error model (the
 * AbstractSyndromeDecoder contract of src/ecc/decoder_pipeline.jl:1-24:
 * decode / batchdecode).  Mechanisms flipping one or two detectors are
 * edges (weight log((1-p)/p)); larger ones are split into existing edges
 * whose observables add up, else dropped (counted by qcf_matcher_info).
 *   qcf_matcher_decode: syndromes shots x ndet bytes (row-major) ->
 *     predicted observable flips shots x nobs bytes, and optionally the
 *     guessed mechanisms shots x nmech bytes (the `guesses` rows that
 *     evaluate_guesses, src/ecc/decoder_pipeline.jl:172-189, multiplies by
 *     the faults matrix).
 *   qcf_matcher_decode_packed: detector bits packed 64 shots per word,
 *     ndet planes of ld words (the record layout of qcf_pf_sample).
 * nthreads <= 0 uses up to 16 host threads.  Host memory only.
 */
typedef struct qcf_matcher qcf_matcher;
int qcf_matcher_create(const qcf_dem *dem, qcf_matcher **out);
int qcf_matcher_free(qcf_matcher *m);
int qcf_matcher_info(const qcf_matcher *m, size_t *nedges, uint32_t *dropped);
int qcf_matcher_decode(const qcf_matcher *m, const uint8_t *syndromes, uint64_t shots,
                       uint8_t *obs_out, uint8_t *mech_out, int nthreads);
int qcf_matcher_decode_packed(const qcf_matcher *m, const uint64_t *dets, size_t ld,
                              uint64_t shots, uint8_t *obs_out, int nthreads);

/* Maximum-weight matching of a dense graph (n x n int64 weights, <= 0 = no
 * edge; mate[i] = -1 if unmatched).  Test entry for the blossom solver. */
int qcf_debug_max_weight_matching(int n, const int64_t *w, int32_t *mate);

/* ---- neural decoder (device) ------------------------------------------ */

/*
 * Perceptron decoder ndet -> h1 -> h2 -> h3 -> 1 (GELU; BF16 weights and
 * activations, FP32 accumulation on the matrix cores, FP32 master weights
 * with Adam).  Inputs are gathered straight from a packed record (the layout
 * qcf_pf_sample writes): det_bits / obs_bit are 1-based record bits.  No
 * reference counterpart; it serves the AbstractSyndromeDecoder contract
 * (src/ecc/decoder_pipeline.jl:1-24) from the host mirror.
 *   widths and batches: multiples of 128; the input is zero-padded to a
 *     multiple of 128.
 *   parameters (fp32, flat): W1[h1][kin] b1 W2[h2][h1] b2 W3[h3][h2] b3
 *     w4[h3] b4, kin = ndet rounded up to 128.  set_params resets Adam.
 *   train_step: gathers shots shot0 .. shot0+batch-1 of `rec` (shot0 a
 *     multiple of 64), forward, binary cross-entropy on the logit, backward,
 *     one Adam step; adds the batch mean loss to *loss_accum (device float,
 *     may be NULL).
 *   gradients: the same forward/backward without the update, copied to host.
 *   infer: logits (device float[shots]) for all shots of `rec`.
 * All device pointers; `stream` is a hipStream_t (NULL = default stream).
 */
typedef struct qcf_mlp qcf_mlp;
int qcf_mlp_create(const uint32_t *det_bits, uint32_t ndet, uint32_t obs_bit, uint32_t h1,
                   uint32_t h2, uint32_t h3, uint32_t batch, uint32_t infer_batch,
                   qcf_mlp **out);
int qcf_mlp_free(qcf_mlp *m);
int qcf_mlp_nparams(const qcf_mlp *m, size_t *n);
int qcf_mlp_set_params(qcf_mlp *m, const float *host, void *stream);
int qcf_mlp_get_params(const qcf_mlp *m, float *host, void *stream);
int qcf_mlp_train_step(qcf_mlp *m, const uint64_t *rec, size_t ld, uint64_t shot0, float lr,
                       float beta1, float beta2, float eps, float *loss_accum, void *stream);
int qcf_mlp_gradients(qcf_mlp *m, const uint64_t *rec, size_t ld, uint64_t shot0,
                      float *grad_host, float *loss_host, void *stream);
int qcf_mlp_infer(qcf_mlp *m, const uint64_t *rec, size_t ld, uint64_t shots, float *logits,
                  void *stream);

/* BF16 GEMM of the decoder (FP32 out C[M][N]): layout 0 = A[M][K] B[N][K],
 * 1 = A[M][K] B[K][N], 2 = A[K][M] B[K][N].  M, N multiples of 128, K of 32.
 * Test entry for the matrix-core kernel. */
int qcf_debug_gemm_bf16(int layout, int M, int N, int K, const void *A, const void *B,
                        float *C, void *stream);

/* ---- shared accounting ------------------------------------------------- */

/* Frame bit-ops per shot of a compiled circuit (CNOT 2, H 2, S 1, SWAP 4,
 * DEPOLARIZE1 2, DEPOLARIZE2 4, X-flip 1, sMZ 1, sMRZ 3, XOR of k bits k-1). */
double qcf_program_bitops(const qcf_program *prog);

#ifdef __cplusplus
}
#endif
#endif /* QCF_H */
