// Pauli-frame sampler kernel for gfx950.
//
// One 64-lane wave owns a tile of 2048 shots: lane l holds 32 shots as one
// 32-bit word of every plane (x and z bit of every frame qubit, plus
// classical bits that are read again).  The planes live in LDS for the whole
// circuit, so HBM carries only the measurement record.  The wave walks the
// device program (src/pauli_frames.jl:234-239, one apply! per op) streamed
// 64 instructions at a time: each lane loads one instruction with a vector
// load and the wave reads instruction k back with v_readlane, so the program
// fetch never waits on the LDS counter.
//
// Randomness (native mode) -- keyed only on (seed, absolute shot position):
//   * gauge words (initZ!, the sMRZ z-flip, src/pauli_frames.jl:47-56,110-113):
//     per lane, a Weyl sequence with a murmur3-finalized output, start and
//     increment from Philox4x32-10(seed, absolute 32-shot word);
//   * noise events: event positions along the (site ordinal, shot) sequence
//     of the wave's 2048-shot tile are drawn 64 at a time, one geometric gap
//     per lane from that lane's xoshiro128** stream, seeded once per tile by
//     Philox4x32-10(seed, absolute tile, lane), and a wave prefix sum.  At a
//     site the lanes holding events that fall in this site flip the target
//     shot's bits with LDS atomics.
// Replay mode reads every draw from caller-supplied mask planes instead.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>

#include "qcf_common.hpp"
#include "qcf_device.h"
#include "qcf_jit.hpp"
#include "qcf_program.hpp"

namespace qcf {
namespace {

enum Mode : int { kNative = 0, kReplay = 1, kExport = 2 };

struct KParams {
  const uint4 *prog;   // 2 uint4 per instruction
  uint32_t ninstr;
  uint32_t nblocks;    // 2048-shot tiles (last tile index + 1)
  uint32_t blk0;       // first tile of this launch
  uint32_t tile0;      // absolute index of tile 0 (shot_offset / 2048): keys the RNG
  uint32_t nw32;       // valid 32-bit shot words
  uint32_t nwbuf;      // 32-bit words of a plane the last tile may store (<= stride)
  uint32_t lastmask;   // valid bits of word nw32-1
  uint64_t ld32;       // 32-bit words per plane of frame / record / masks
  uint64_t seed;
  uint32_t *frame;
  uint32_t *rec;
  uint32_t *masks;
  uint32_t body_rbase; // replay plane of the first body draw
  uint32_t *scratch;   // global plane storage (NULL: planes in LDS)
  uint32_t nplanes;    // planes per tile
  uint32_t sstride;    // 32-bit words of global plane storage per wave slot
  uint32_t nclasses;
  float cls_c[kMaxClasses];
  uint32_t thin;       // some site thins its class's events (qcf_program::thin)
};

using qcfd::EvState;
using qcfd::readlane;

template <int MODE>
__device__ __forceinline__ void export_flip(const KParams &P, uint32_t plane,
                                            uint32_t gword, uint32_t bit) {
  if (MODE != kExport) return;
  if (gword >= P.nw32) return;
  if (gword == P.nw32 - 1 && !(bit & P.lastmask)) return;
  atomicXor(&P.masks[(uint64_t)plane * P.ld32 + gword], bit);
}

// Plane storage of a tile: LDS (kPlLds, the fast path) or, for circuits whose
// frame does not fit on chip, a per-wave region of global memory (kPlAtomic).
// In global memory every access is an agent-scope atomic: bit flips that
// other lanes of the wave write (noise events) are never hidden by a stale L1
// line, and an XOR is one fire-and-forget atomic at L2.  (Plain cached
// accesses, valid for programs without device-drawn events, measured slower:
// d=11 x 12 rounds, 2^19 shots, 12.5 vs 7.3 ms -- the XOR becomes a load on
// the critical path.  Narrow tiles -- fewer lanes per wave, more waves for
// launches with few tiles -- measured slower too: 7.3 ms at 64 lanes, 9.4 /
// 9.8 / 10.6 ms at 32 / 16 / 8.)
constexpr int kPlLds = 0, kPlAtomic = 1;
template <int GL>
struct Planes {
  uint32_t *base;
  __device__ __forceinline__ uint32_t rd(uint32_t p, uint32_t l) const {
    if (GL == kPlAtomic)
      return __hip_atomic_load(base + p * 64u + l, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return base[p * 64u + l];
  }
  __device__ __forceinline__ void wr(uint32_t p, uint32_t l, uint32_t v) const {
    if (GL == kPlAtomic)
      __hip_atomic_store(base + p * 64u + l, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
      base[p * 64u + l] = v;
  }
  __device__ __forceinline__ void xr(uint32_t p, uint32_t l, uint32_t v) const {
    if (GL == kPlAtomic)
      __hip_atomic_fetch_xor(base + p * 64u + l, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
      atomicXor(base + p * 64u + l, v);
  }
};

// vmcnt(0) only (expcnt and lgkmcnt left at their maxima): waits for the
// program prefetch and the record stores once per 64-instruction chunk.
#define QCF_WAIT_VMEM() __builtin_amdgcn_s_waitcnt(0x0F70)

// Run the whole program on one 2048-shot tile.
template <int MODE, int GL>
__device__ __forceinline__ void run_tile(const KParams &P, uint32_t blk,
                                         const Planes<GL> &F, uint32_t lane) {
  const uint4 *__restrict__ prog4 = P.prog;
  const uint32_t gw = blk * 64u + lane;
  // lanes inside the plane stride store (zeros past the shots)
  const bool live = gw < P.nwbuf;
  const uint32_t lmask = gw >= P.nw32 ? 0u : (gw == P.nw32 - 1 ? P.lastmask : 0xFFFFFFFFu);
  const uint64_t goff = gw;
  qcfd::Gauge gauge{0, 1};
  if (MODE != kReplay) gauge = qcfd::gauge_stream(P.seed, gw + P.tile0 * 64u);
  EvState ev0 = qcfd::ev_init(P.seed, blk + P.tile0, lane);
  if (MODE != kReplay && P.nclasses > 0) ev0 = qcfd::gen_batch(ev0, P.cls_c[0], P.thin);
  // Every global load in the instruction loop is waited for on the spot, so
  // no load is ever pending where control flow merges (which would make the
  // compiler wait for all outstanding record stores there).
  auto gload = [&](const uint32_t *base, uint32_t plane) -> uint32_t {
    const uint32_t v = live ? base[(uint64_t)plane * P.ld32 + goff] : 0u;
    QCF_WAIT_VMEM();
    return v;
  };
  auto draw = [&](uint32_t rplane) -> uint32_t {
    if (MODE == kReplay) return gload(P.masks, rplane);
    const uint32_t v = gauge.next();
    if (MODE == kExport && live) P.masks[(uint64_t)rplane * P.ld32 + goff] = v & lmask;
    return v;
  };
  auto RD = [&](uint32_t p) -> uint32_t { return F.rd(p, lane); };
  auto WR = [&](uint32_t p, uint32_t v) { F.wr(p, lane, v); };
  auto XR = [&](uint32_t p, uint32_t v) { F.xr(p, lane, v); };
  uint32_t acc = 0;
  // lanes past the end re-read the last instruction (never executed)
  const uint32_t last = P.ninstr - 1;
  uint4 nA = prog4[2 * min(lane, last)];
  uint4 nB = prog4[2 * min(lane, last) + 1];
  for (uint32_t base = 0; base < P.ninstr; base += 64u) {
    QCF_WAIT_VMEM();
    const uint4 cA = nA, cB = nB;
    const uint32_t nb = base + 64u + lane;
    nA = prog4[2 * min(nb, last)];
    nB = prog4[2 * min(nb, last) + 1];
    const uint32_t cnt = min(64u, P.ninstr - base);
    for (uint32_t k = 0; k < cnt; ++k) {
      const uint32_t opw = readlane(cA.x, k);
      const uint32_t op = opw & 0xFFu;
      // hot operations first (plane XOR bundles, noise groups, measure-reset)
      if (op == I_XORN) {
        const uint32_t n = opw >> 8;
        uint32_t w[7];
        w[0] = readlane(cA.y, k);
        w[1] = n > 1 ? readlane(cA.z, k) : 0u;
        w[2] = n > 2 ? readlane(cA.w, k) : 0u;
        w[3] = n > 3 ? readlane(cB.x, k) : 0u;
        w[4] = n > 4 ? readlane(cB.y, k) : 0u;
        w[5] = n > 5 ? readlane(cB.z, k) : 0u;
        w[6] = n > 6 ? readlane(cB.w, k) : 0u;
        uint32_t v[7];
        switch (n) {  // all sources are read before any target is written
          case 7: v[6] = RD(w[6] >> 16); [[fallthrough]];
          case 6: v[5] = RD(w[5] >> 16); [[fallthrough]];
          case 5: v[4] = RD(w[4] >> 16); [[fallthrough]];
          case 4: v[3] = RD(w[3] >> 16); [[fallthrough]];
          case 3: v[2] = RD(w[2] >> 16); [[fallthrough]];
          case 2: v[1] = RD(w[1] >> 16); [[fallthrough]];
          default: v[0] = RD(w[0] >> 16);
        }
        switch (n) {
          case 7: XR(w[6] & 0xFFFFu, v[6]); [[fallthrough]];
          case 6: XR(w[5] & 0xFFFFu, v[5]); [[fallthrough]];
          case 5: XR(w[4] & 0xFFFFu, v[4]); [[fallthrough]];
          case 4: XR(w[3] & 0xFFFFu, v[3]); [[fallthrough]];
          case 3: XR(w[2] & 0xFFFFu, v[2]); [[fallthrough]];
          case 2: XR(w[1] & 0xFFFFu, v[1]); [[fallthrough]];
          default: XR(w[0] & 0xFFFFu, v[0]);
        }
        continue;
      }
      if (op == I_NOISEG) {
        const uint32_t G = opw >> 8;
        if (MODE == kReplay) {
          for (uint32_t i = 1; i <= G; ++i) {
            const uint32_t pab = readlane(cA.y, k + i), pcd = readlane(cA.z, k + i);
            const uint32_t rb = P.body_rbase + readlane(cA.w, k + i);
            const uint32_t m0 = draw(rb), m1 = draw(rb + 1);
            XR(pab & 0xFFFFu, m0);
            XR(pab >> 16, m1);
            if (pcd != 0u) {
              const uint32_t m2 = draw(rb + 2), m3 = draw(rb + 3);
              XR(pcd & 0xFFFFu, m2);
              XR(pcd >> 16, m3);
            }
          }
        } else if (P.nclasses > 0) {
          auto flip = [&](uint32_t plane, uint32_t tl, uint32_t bit, uint32_t j, uint32_t rs) {
            F.xr(plane, tl, bit);
            export_flip<MODE>(P, P.body_rbase + rs + j, blk * 64u + tl, bit);
          };
          qcfd::group_events(ev0, readlane(cA.y, k), G, k + 1, cA.y, cA.z, cB.x, cB.y, cA.w,
                             P.cls_c[0], P.thin, flip);
        }
        k += G;
        continue;
      }
      const uint32_t a = readlane(cA.y, k);
      if (op == I_MR || op == I_MRX) {
        const uint32_t b = readlane(cA.z, k), c = readlane(cA.w, k);
        const uint32_t col = readlane(cB.x, k), keep = readlane(cB.y, k);
        const uint32_t v = RD(a);
        if (col != kNone && live) P.rec[(uint64_t)col * P.ld32 + goff] = v & lmask;
        if (op == I_MRX) {
          const uint32_t other = readlane(cB.z, k), dcol = readlane(cB.w, k);
          const uint32_t det = v ^ RD(other);
          if (live) P.rec[(uint64_t)dcol * P.ld32 + goff] = det & lmask;
        }
        if (keep != kNone) WR(keep, v);
        WR(a, 0u);
        if (MODE == kReplay || !(opw & kNoGauge)) XR(b, draw(P.body_rbase + c));
        continue;
      }
      const uint32_t b = readlane(cA.z, k);
      if (op == I_XSTORE2) {
        const uint32_t c = readlane(cA.w, k);
        const uint32_t v = RD(a) ^ RD(b);
        if (live) P.rec[(uint64_t)c * P.ld32 + goff] = v & lmask;
        const uint32_t keep = readlane(cB.x, k);
        if (keep != kNone) WR(keep, v);
        continue;
      }
      const uint32_t c = readlane(cA.w, k);
      switch (op) {
        case I_MEAS: {
          const uint32_t v = RD(a);
          if (live) P.rec[(uint64_t)b * P.ld32 + goff] = v & lmask;
          if (c != kNone) WR(c, v);
          break;
        }
        case I_RESETZ: {
          WR(a, 0u);
          if (MODE == kReplay || !(opw & kNoGauge)) XR(b, draw(P.body_rbase + c));
          break;
        }
        case I_ACC_LDS: acc = RD(a); break;
        case I_ACCX_LDS: acc ^= RD(a); break;
        case I_ACC_REC: acc = gload(P.rec, a); break;
        case I_ACCX_REC: acc ^= gload(P.rec, a); break;
        case I_ACC_STORE:
          if (live) P.rec[(uint64_t)a * P.ld32 + goff] = acc & lmask;
          if (b != kNone) WR(b, acc);
          break;
        case I_REC_ZERO:
          if (live) P.rec[(uint64_t)a * P.ld32 + goff] = 0u;
          break;
        case I_INIT: {
          WR(a, 0u);
          WR(b, (MODE == kReplay || !(opw & kNoGauge)) ? draw(c) : 0u);
          break;
        }
        case I_LOAD:
          WR(a, gload(P.frame, b));
          break;
        case I_STORE:
          if (live) P.frame[(uint64_t)b * P.ld32 + goff] = RD(a) & lmask;
          break;
        case I_ZERO: WR(a, 0u); break;
        case I_LIN1: {
          const uint32_t x = RD(a), z = RD(b);
          const uint32_t nx = ((c & 1u) ? x : 0u) ^ ((c & 2u) ? z : 0u);
          const uint32_t nz = ((c & 4u) ? x : 0u) ^ ((c & 8u) ? z : 0u);
          WR(a, nx);
          WR(b, nz);
          break;
        }
        default: break;
      }
    }
  }
}

template <int MODE, int GL>
__global__ __launch_bounds__(64) void pf_kernel(KParams P) {
  extern __shared__ uint32_t lds[];
  const uint32_t lane = threadIdx.x;
  Planes<GL> F;
  F.base = GL != kPlLds ? P.scratch + (uint64_t)blockIdx.x * P.sstride : lds;
  for (uint32_t blk = P.blk0 + blockIdx.x; blk < P.nblocks; blk += gridDim.x)
    run_tile<MODE, GL>(P, blk, F, lane);
}

// Persistent batched kernel: workgroups pull (circuit, tile range) items from
// a device queue until it is drained, then take single tiles of the shared
// specialised launches (those running beside it with a tile counter of their
// own), largest group first.  Every workgroup of this kernel and of the
// sharing launches adds one to queue[1] when it leaves; the last one resets
// the counters, so consecutive launches (and graph replays) need no clearing.
struct WorkItem {
  uint32_t circuit, blk0, nblk, pad;
};

struct BatchArgs {
  const KParams *params;
  const WorkItem *items;     // own items
  uint32_t nitems;
  uint32_t nshared;          // shared groups
  uint32_t *queue;           // [0] next own item, [1] workgroups done, [2 + g] next tile of group g
  const WorkItem *shared;    // tiles of every shared group, in the group's tile order
  const uint32_t *shared_first;  // group g owns shared[shared_first[g] .. shared_first[g + 1])
  uint32_t total;            // workgroups that count into queue[1]
  uint32_t nclear;           // counters the last one resets
};

__device__ __forceinline__ void batch_items(const KParams *params, const WorkItem &w,
                                            uint32_t *lds, uint32_t lane) {
  const KParams P = params[w.circuit];
  if (P.scratch) {
    Planes<kPlAtomic> F{P.scratch + (uint64_t)blockIdx.x * P.sstride};
    for (uint32_t b = 0; b < w.nblk; ++b) run_tile<kNative, kPlAtomic>(P, w.blk0 + b, F, lane);
  } else {
    Planes<kPlLds> F{lds};
    for (uint32_t b = 0; b < w.nblk; ++b) run_tile<kNative, kPlLds>(P, w.blk0 + b, F, lane);
  }
}

__global__ __launch_bounds__(64) void pf_batch_kernel(BatchArgs A) {
  extern __shared__ uint32_t lds[];
  const uint32_t lane = threadIdx.x;
  while (true) {
    uint32_t it = 0;
    if (lane == 0) it = atomicAdd(&A.queue[0], 1u);
    it = readlane(it, 0);
    if (it >= A.nitems) break;
    batch_items(A.params, A.items[it], lds, lane);
  }
  for (uint32_t g = 0; g < A.nshared; ++g) {
    const uint32_t f0 = A.shared_first[g], n = A.shared_first[g + 1] - f0;
    while (true) {
      uint32_t t = 0;
      if (lane == 0) t = atomicAdd(&A.queue[2 + g], 1u);
      t = readlane(t, 0);
      if (t >= n) break;
      batch_items(A.params, A.shared[f0 + t], lds, lane);
    }
  }
  if (lane == 0) {
    __threadfence();
    if (atomicAdd(&A.queue[1], 1u) == A.total - 1u)
      for (uint32_t i = 0; i < A.nclear; ++i) atomicExch(&A.queue[i], 0u);
  }
}

// ---- Philox self-check kernel (exposes the device RNG to tests)
__global__ void philox_kernel(const uint32_t *in, uint32_t *out, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t *c = in + 6 * i;
  const qcfd::u4 r = qcfd::philox(qcfd::u4{c[0], c[1], c[2], c[3]}, c[4], c[5]);
  out[4 * i + 0] = r.x;
  out[4 * i + 1] = r.y;
  out[4 * i + 2] = r.z;
  out[4 * i + 3] = r.w;
}

// ---- generator self-check kernels: the per-lane event stream (xoshiro128**)
// and the gauge words (Weyl sequence + murmur3 finalizer), each run from a
// given state for `k` outputs
__global__ void xoshiro_kernel(const uint32_t *state, uint32_t *out, size_t n, uint32_t k) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t *s = state + 4 * i;
  qcfd::Xoshiro x{s[0], s[1], s[2], s[3]};
  for (uint32_t j = 0; j < k; ++j) out[(size_t)k * i + j] = x.next();
}

__global__ void gauge_kernel(const uint32_t *state, uint32_t *out, size_t n, uint32_t k) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  qcfd::Gauge g{state[2 * i], state[2 * i + 1]};
  for (uint32_t j = 0; j < k; ++j) out[(size_t)k * i + j] = g.next();
}

__global__ void gauge_stream_kernel(uint64_t seed, const uint32_t *words, uint32_t *out,
                                    size_t n, uint32_t k) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  qcfd::Gauge g = qcfd::gauge_stream(seed, words[i]);
  for (uint32_t j = 0; j < k; ++j) out[(size_t)k * i + j] = g.next();
}

// ---- wave scan self-check kernel (one wave per 64 inputs)
__global__ __launch_bounds__(64) void scan_kernel(const uint32_t *in, uint32_t *out) {
  const uint32_t i = blockIdx.x * 64u + threadIdx.x;
  out[i] = qcfd::wave_inclusive_scan(in[i]);
}

// ---- reference-layout converters
// The reference-side matrices are column-major with `ldr` elements between
// columns (ldr > shots for a view of a larger frame's rows).
__global__ void record_to_bool_kernel(const uint64_t *meas, size_t ld,
                                      uint64_t shots, uint32_t nbits,
                                      uint8_t *out, size_t ldr) {
  const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t b = blockIdx.y;
  if (s >= shots || b >= nbits) return;
  out[(uint64_t)b * ldr + s] = (uint8_t)((meas[(uint64_t)b * ld + s / 64] >> (s % 64)) & 1u);
}

__global__ void record_from_bool_kernel(const uint8_t *in, size_t ldr, uint64_t shots,
                                        uint32_t nbits, uint64_t *meas,
                                        size_t ld) {
  const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t b = blockIdx.y;
  if (w >= ld || b >= nbits) return;
  uint64_t v = 0;
  for (int i = 0; i < 64; ++i) {
    const uint64_t s = w * 64 + i;
    if (s < shots && in[(uint64_t)b * ldr + s]) v |= 1ull << i;
  }
  meas[(uint64_t)b * ld + w] = v;
}

// Reference frame layout: shots x 2W column-major UInt64, column j < W holds
// x bits of qubits 64j+1..64j+64, column W+j the z bits.
__global__ void frame_to_reference_kernel(const uint64_t *frame, size_t ld,
                                          uint64_t shots, uint32_t nq,
                                          uint64_t *xzs, size_t ldr) {
  const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t j = blockIdx.y;  // 0 .. 2W-1
  const uint32_t W = (nq + 63) / 64;
  if (s >= shots) return;
  const uint32_t half = j / W, jj = j % W;
  uint64_t v = 0;
  for (uint32_t i = 0; i < 64; ++i) {
    const uint32_t q = jj * 64 + i;
    if (q >= nq) break;
    const uint64_t plane = (uint64_t)half * nq + q;
    v |= ((frame[plane * ld + s / 64] >> (s % 64)) & 1ull) << i;
  }
  xzs[(uint64_t)j * ldr + s] = v;
}

__global__ void frame_from_reference_kernel(const uint64_t *xzs, size_t ldr, uint64_t shots,
                                            uint32_t nq, uint64_t *frame,
                                            size_t ld) {
  const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t plane = blockIdx.y;  // 0 .. 2nq-1
  if (w >= ld) return;
  const uint32_t half = plane / nq, q = plane % nq;
  const uint32_t W = (nq + 63) / 64;
  const uint32_t j = half * W + q / 64;
  uint64_t v = 0;
  for (int i = 0; i < 64; ++i) {
    const uint64_t s = w * 64 + i;
    if (s < shots) v |= ((xzs[(uint64_t)j * ldr + s] >> (q % 64)) & 1ull) << i;
  }
  frame[(uint64_t)plane * ld + w] = v;
}

// ---- host side
int ensure_device_copy(qcf_program *P, int variant, const uint4 **prog,
                       uint32_t *ninstr) {
  int dev = 0;
  QCF_HIP_CHECK(hipGetDevice(&dev));
  static std::mutex mu;
  std::lock_guard<std::mutex> lock(mu);
  qcf_program::DeviceCopy *dc = nullptr;
  for (auto &d : P->dev)
    if (d.device == dev) dc = &d;
  if (!dc) {
    P->dev.emplace_back();
    dc = &P->dev.back();
    dc->device = dev;
  }
  if (!dc->prog[variant]) {
    const std::vector<Instr> code = variant_code(P, variant);
    void *d = nullptr;
    QCF_HIP_CHECK(hipMalloc(&d, code.size() * sizeof(Instr)));
    QCF_HIP_CHECK(hipMemcpy(d, code.data(), code.size() * sizeof(Instr), hipMemcpyHostToDevice));
    dc->prog[variant] = d;
    dc->ninstr[variant] = (uint32_t)code.size();
  }
  *prog = (const uint4 *)dc->prog[variant];
  *ninstr = dc->ninstr[variant];
  return QCF_OK;
}

int max_lds_per_block(int *out) {
  int dev = 0;
  QCF_HIP_CHECK(hipGetDevice(&dev));
  int v = 0;
  QCF_HIP_CHECK(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  *out = v;
  return QCF_OK;
}

int allow_lds(const void *fn, uint32_t lds_bytes) {
  if (lds_bytes <= 64u * 1024u) return QCF_OK;
  QCF_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds_bytes));
  return QCF_OK;
}

template <int MODE>
int launch_pf(const KParams &kp, uint32_t lds_bytes, uint32_t grid, hipStream_t s) {
  if (kp.scratch) {
    QCF_LAUNCH((pf_kernel<MODE, kPlAtomic>), dim3(grid), dim3(64), 0, s, kp);
  } else {
    QCF_TRY(allow_lds((const void *)pf_kernel<MODE, kPlLds>, lds_bytes));
    QCF_LAUNCH((pf_kernel<MODE, kPlLds>), dim3(grid), dim3(64), lds_bytes, s, kp);
  }
  QCF_HIP_CHECK(hipGetLastError());
  return QCF_OK;
}

// Frames too large for LDS live in a per-wave region of a library-owned
// device workspace, one per stream (single-circuit launches; batches own
// theirs).
// (d=11 unrotated surface code, 884 planes, 10^7 shots: 8, 16, 24, 32 waves
// per CU took 36.0, 29.4, 24.0, 23.9 ms; the waves wait on L2/MALL latency,
// so more of them help until occupancy stops them)
constexpr uint32_t kGlobalWavesPerCU = 24;
constexpr uint32_t kMaxBatchWavesPerCU = 32;
constexpr uint32_t kGlobalJitWavesPerCU = 16;  // specialised kernels, planes in global memory
// Frames in global memory use the workspace of the launch's stream.
int frame_scratch(hipStream_t stream, size_t bytes, uint32_t **out) {
  void *p = nullptr;
  QCF_TRY(stream_scratch(kScratchFrames, stream, bytes, &p));
  *out = (uint32_t *)p;
  return QCF_OK;
}

// Waves per CU for frames in global memory (QCF_GLOBAL_WAVES_PER_CU, 1..32;
// 0 = the built-in default of each path).
uint32_t global_waves_per_cu() {
  const char *e = getenv("QCF_GLOBAL_WAVES_PER_CU");
  if (!e) return 0;
  const int v = atoi(e);
  return v >= 1 && v <= 32 ? (uint32_t)v : 0u;
}

bool force_global_frames() {
  const char *e = getenv("QCF_FORCE_GLOBAL_FRAMES");
  return e && e[0] == '1';
}

}  // namespace

// Which kernel a native launch of kp would run if specialised, with its
// arguments and grid (nothing is launched).  Frames too large for LDS take
// the split kernel: the planes used most in registers, the rest in LDS, one
// wave per CU.  With QCF_JIT_GLOBAL=1 (or when the split kernel cannot be
// built) they keep every plane in global memory: the specialised kernel
// there is opt-in (global_planes: the caller provides the plane storage),
// since both it and the interpreter are bound by L2 latency and it measured
// no faster for d >= 9 while costing seconds of compilation.  Returns an
// error (QCF_ERR_UNSUPPORTED) when the launch has no specialised kernel.
int jit_plan(qcf_program *P, const KParams &kp, uint64_t shots, int variant, bool global,
             uint32_t waves_per_cu, size_t ld, JitPlan *out) {
  const bool split_ok = !force_global_frames() && !jit_global_enabled() && jit_split_enabled();
  const bool try_split = global && split_ok;
  if (jit_disabled() || (global && !try_split && !jit_global_enabled()))
    return set_error(QCF_ERR_UNSUPPORTED, "no specialised kernel for this launch");
  DeviceState *st;
  QCF_TRY(device_state(&st));
  JitKernelInfo jk;
  // records under 4 GB take 32-bit byte offsets from one base
  const int v = variant | (global ? 4 : 0) | ((uint64_t)std::max<uint32_t>(P->nbits, 1u) * ld *
                                                        8ull < (1ull << 32) ? 8 : 0);
  // Frames whose qubit planes take more than half of a CU's LDS would run one
  // wave per CU; split mode with half of the LDS each (variant bit 5) lets two
  // waves share a CU when the rest fits the register budget (rotated d=9 x 9,
  // p=1e-3, 10^7 shots: 2.0 ms instead of 3.6 ms in LDS mode).
  // Frames beyond split mode that fit in LDS at half or quarter width (up to
  // 1280 / 2560 planes) run part-tile passes (variant bits 6 / 7): lanes
  // 0..31 (0..15) hold one half (quarter) of the tile's shot words, two
  // (four) passes per tile.
  const bool try_two = split_ok && (global || 2u * P->nq * 256u > 80u * 1024u);
  int js = QCF_ERR_UNSUPPORTED;
  // a noiseless fresh launch that returns only the record: the linear kernel
  // (variant bit 8), which keeps no frame, whatever the frame's size
  bool linear = false;
  if (variant == 1 && P->linear_ok && jit_linear_enabled()) {
    js = jit_get(P, 1 | (v & 8) | 256, &jk);
    if (js == QCF_OK) linear = true;
    else clear_error();
  }
  if (js != QCF_OK && try_two) {
    js = jit_get(P, v | 16 | 32, &jk);
    if (js != QCF_OK) clear_error();
  }
  uint32_t parts = 1;  // passes per tile of the kernel taken
  if (js != QCF_OK && try_split) {
    js = jit_get(P, v | 16, &jk);
    if (js != QCF_OK) {
      clear_error();
      js = jit_get(P, v | 64, &jk);
      if (js == QCF_OK) parts = 2;
    }
    if (js != QCF_OK) {
      clear_error();
      js = jit_get(P, v | 128, &jk);  // quarter tiles: up to 2560 planes
      if (js == QCF_OK) parts = 4;
    }
  }
  if (js != QCF_OK && !try_split) js = jit_get(P, v, &jk);
  QCF_TRY(js);
  const bool gjit = global && !try_split && !linear;  // specialised kernel on global planes
  JitParams jp{};
  jp.rec = kp.rec;
  jp.frame = kp.frame;
  jp.seed = kp.seed;
  jp.ld32 = kp.ld32;
  jp.nblocks = kp.nblocks;
  jp.nw32 = kp.nw32;
  jp.nwbuf = kp.nwbuf;
  jp.tile0 = kp.tile0;
  jp.lastmask = kp.lastmask;
  jp.c = kp.cls_c[0];
  jp.has_class = kp.nclasses > 0 ? 1u : 0u;
  jp.nfull = (uint32_t)(shots / kBlockShots);
  uint32_t jwpc = waves_per_cu ? waves_per_cu : jk.blocks_per_cu;
  if (gjit)
    jwpc = global_waves_per_cu() ? global_waves_per_cu()
                                 : std::min<uint32_t>(jwpc, kGlobalJitWavesPerCU);
  uint64_t g1 = std::min<uint64_t>(kp.nblocks, (uint64_t)st->cu_count * jwpc);
  const uint32_t nx = jit_xcds();
  if (parts > 1 && kp.nblocks < (uint64_t)st->cu_count * jwpc) {
    // fewer tiles than resident workgroups: the parts of a tile run on
    // different workgroups at once (pset of them per tile); the tile loop
    // then sees gridDim / pset workgroups, a whole number of XCD rounds
    uint64_t per = std::min<uint64_t>(kp.nblocks, (uint64_t)st->cu_count * jwpc / parts);
    if (nx > 1 && per >= nx) {
      per -= per % nx;
      jp.xcds = nx;
    }
    jp.pset = parts;
    g1 = std::max<uint64_t>(per, 1) * parts;
  } else if (!gjit && jit_full_grid()) {
    // one workgroup per tile, rounded up to whole XCD rounds (the extra
    // workgroups find no tile and return)
    g1 = kp.nblocks;
    if (nx > 1 && g1 >= nx) {
      g1 += (nx - g1 % nx) % nx;
      jp.xcds = nx;
    }
  } else if (nx > 1 && g1 >= nx) {
    g1 -= g1 % nx;
    jp.xcds = nx;
  }
  out->function = jk.function;
  out->jp = jp;
  out->grid = (uint32_t)g1;
  out->lds_bytes = jk.lds_bytes;
  out->blocks_per_cu = jk.blocks_per_cu;
  out->key = jk.key;
  out->global_planes = gjit;
  return QCF_OK;
}

// Shared with the batched launcher.
int pf_launch_one(const qcf_program *prog, uint64_t shots, uint64_t shot_offset, uint64_t seed,
                  uint64_t *frame, uint64_t *meas, uint64_t *masks, size_t ld,
                  uint32_t flags, uint32_t waves_per_cu, hipStream_t stream) {
  qcf_program *P = const_cast<qcf_program *>(prog);
  if (shots == 0) return QCF_OK;
  if (ld < (shots + 63) / 64)
    return set_error(QCF_ERR_DIMENSION, "pftrajectories: plane stride %zu words < %llu shots",
                     ld, (unsigned long long)shots);
  const bool fresh = flags & QCF_PF_FRESH;
  const bool exp = flags & QCF_PF_EXPORT_MASKS;
  if (!fresh && !frame)
    return set_error(QCF_ERR_ARGUMENT, "pftrajectories(state, circuit) needs the frame buffer");
  if (!meas && P->nbits > 0)
    return set_error(QCF_ERR_ARGUMENT, "pftrajectories: null measurement buffer");
  if (exp && !masks)
    return set_error(QCF_ERR_ARGUMENT, "pftrajectories: mask export needs a mask buffer");
  if (shots > (uint64_t)0xFFFFFFFFu * 16)
    return set_error(QCF_ERR_UNSUPPORTED, "pftrajectories: too many shots in one call");
  if (!masks || exp) {  // the counter-based RNG is keyed on absolute tiles
    if (shot_offset % kBlockShots)
      return set_error(QCF_ERR_ARGUMENT,
                       "pftrajectories: shot_offset %llu is not a multiple of %u shots",
                       (unsigned long long)shot_offset, kBlockShots);
    if ((shot_offset + shots + 31) / 32 > 0xFFFFFFFFull)
      return set_error(QCF_ERR_UNSUPPORTED, "pftrajectories: shot_offset + shots beyond 2^37");
  }
  int lds_max = 0;
  QCF_TRY(max_lds_per_block(&lds_max));
  const uint32_t lds_bytes = P->nplanes * 256u;
  const bool global = (int)lds_bytes > lds_max || force_global_frames();
  const int variant = (fresh ? 1 : 0) | (frame ? 2 : 0);
  KParams kp{};
  QCF_TRY(ensure_device_copy(P, variant, &kp.prog, &kp.ninstr));
  const uint64_t nw32 = (shots + 31) / 32;
  kp.nw32 = (uint32_t)nw32;
  kp.nblocks = (uint32_t)((nw32 + 63) / 64);
  // lanes of the last tile that lie inside the plane stride store zeros past
  // the shots, so a padded stride leaves defined (zero) padding
  // a view shares its plane stride with other shots: only its own words
  kp.nwbuf = (flags & QCF_PF_VIEW)
                 ? (uint32_t)(2 * ((shots + 63) / 64))
                 : (uint32_t)std::min<uint64_t>(2 * (uint64_t)ld, 64ull * kp.nblocks);
  kp.tile0 = (!masks || exp) ? (uint32_t)(shot_offset / kBlockShots) : 0u;
  const uint32_t tail = (uint32_t)(shots % 32);
  kp.lastmask = tail ? ((1u << tail) - 1u) : 0xFFFFFFFFu;
  kp.ld32 = 2 * (uint64_t)ld;
  kp.seed = seed;
  kp.frame = (uint32_t *)frame;
  kp.rec = (uint32_t *)meas;
  kp.masks = (uint32_t *)masks;
  kp.body_rbase = P->ninit_rs;
  kp.nclasses = (uint32_t)P->classes.size();
  for (uint32_t c = 0; c < kMaxClasses; ++c)
    kp.cls_c[c] = c < P->classes.size() ? P->classes[c].c : 0.f;
  kp.thin = P->thin ? 1u : 0u;
  kp.nplanes = P->nplanes;
  DeviceState *st;
  QCF_TRY(device_state(&st));
  uint32_t wpc = waves_per_cu;
  if (wpc == 0)
    wpc = global ? (global_waves_per_cu() ? global_waves_per_cu() : kGlobalWavesPerCU)
                 : std::max<uint32_t>(1, std::min<uint32_t>(16, (uint32_t)lds_max / std::max(1u, lds_bytes)));
  const uint64_t grid = std::min<uint64_t>(kp.nblocks, (uint64_t)st->cu_count * wpc);
  // Native launches run the kernel specialised to this program when it has
  // one (the partial last tile included); the interpreter is the general
  // path.  Both draw the same events for the same shots.
  if (!masks && !exp) {
    JitPlan pl;
    int js = jit_plan(P, kp, shots, variant, global, waves_per_cu, ld, &pl);
    if (js == QCF_OK && pl.global_planes) {
      js = frame_scratch(stream, pl.grid * (size_t)P->nplanes * 256u, &pl.jp.scratch);
      pl.jp.sstride = (uint64_t)P->nplanes * 64u;
    }
    if (js == QCF_OK) js = jit_launch(pl.function, pl.jp, pl.grid, pl.lds_bytes, stream);
    if (js == QCF_OK) return QCF_OK;
    const char *req = getenv("QCF_REQUIRE_JIT");
    if (req && req[0] == '1') return js;
    clear_error();
  }
  if (global) {
    QCF_TRY(frame_scratch(stream, grid * (size_t)P->nplanes * 256u, &kp.scratch));
    kp.sstride = P->nplanes * 64u;
  }
  if (masks && !exp) return launch_pf<kReplay>(kp, lds_bytes, (uint32_t)grid, stream);
  if (exp) return launch_pf<kExport>(kp, lds_bytes, (uint32_t)grid, stream);
  return launch_pf<kNative>(kp, lds_bytes, (uint32_t)grid, stream);
}

}  // namespace qcf

using namespace qcf;

extern "C" size_t qcf_plane_words(uint64_t shots) {
  return (size_t)std::max<uint64_t>(1, (shots + 2047) / 2048) * 32;
}

extern "C" int qcf_pf_sample(const qcf_program *prog, uint64_t shots, uint64_t shot_offset,
                             uint64_t seed, uint64_t *frame, uint64_t *meas,
                             uint64_t *masks, size_t ld, uint32_t flags,
                             void *stream) {
  clear_error();
  if (!prog) return set_error(QCF_ERR_ARGUMENT, "null program");
  return pf_launch_one(prog, shots, shot_offset, seed, frame, meas, masks, ld, flags, 0,
                       (hipStream_t)stream);
}

extern "C" int qcf_debug_philox(const uint32_t *in, uint32_t *out, size_t n,
                                void *stream) {
  clear_error();
  if (n == 0) return QCF_OK;
  QCF_LAUNCH(philox_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, in, out, n);
  QCF_HIP_CHECK(hipGetLastError());
  return QCF_OK;
}

extern "C" int qcf_debug_xoshiro128ss(const uint32_t *state, uint32_t *out, size_t n,
                                      uint32_t k, void *stream) {
  clear_error();
  if (n == 0 || k == 0) return QCF_OK;
  QCF_LAUNCH(xoshiro_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
             (hipStream_t)stream, state, out, n, k);
  QCF_HIP_CHECK(hipGetLastError());
  return QCF_OK;
}

extern "C" int qcf_debug_gauge(const uint32_t *state, uint32_t *out, size_t n, uint32_t k,
                               void *stream) {
  clear_error();
  if (n == 0 || k == 0) return QCF_OK;
  QCF_LAUNCH(gauge_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
             (hipStream_t)stream, state, out, n, k);
  QCF_HIP_CHECK(hipGetLastError());
  return QCF_OK;
}

extern "C" int qcf_debug_gauge_stream(uint64_t seed, const uint32_t *words, uint32_t *out,
                                      size_t n, uint32_t k, void *stream) {
  clear_error();
  if (n == 0 || k == 0) return QCF_OK;
  QCF_LAUNCH(gauge_stream_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
             (hipStream_t)stream, seed, words, out, n, k);
  QCF_HIP_CHECK(hipGetLastError());
  return QCF_OK;
}

extern "C" int qcf_debug_wave_scan(const uint32_t *in, uint32_t *out, size_t nwaves,
                                   void *stream) {
  clear_error();
  if (nwaves == 0) return QCF_OK;
  QCF_LAUNCH(scan_kernel, dim3((unsigned)nwaves), dim3(64), 0, (hipStream_t)stream, in,
                     out);
  QCF_HIP_CHECK(hipGetLastError());
  return QCF_OK;
}

extern "C" int qcf_debug_check_launch(unsigned threads) {
  clear_error();
  return check_block((const void *)pf_batch_kernel, threads, "pf_batch_kernel");
}

// Column stride of a reference-side matrix: 0 means `shots` (a whole matrix).
int ref_stride(size_t *ldr, uint64_t shots) {
  if (*ldr == 0) *ldr = shots;
  if (*ldr < shots)
    return set_error(QCF_ERR_DIMENSION, "reference column stride %zu < %llu shots", *ldr,
                     (unsigned long long)shots);
  return QCF_OK;
}

extern "C" int qcf_record_to_bool(const uint64_t *meas, size_t ld, uint64_t shots,
                                  uint32_t nbits, uint8_t *out, size_t ldr, void *stream) {
  clear_error();
  if (shots == 0 || nbits == 0) return QCF_OK;
  if (ld < (shots + 63) / 64) return set_error(QCF_ERR_DIMENSION, "record stride too short");
  QCF_TRY(ref_stride(&ldr, shots));
  dim3 grid((unsigned)((shots + 255) / 256), nbits);
  QCF_LAUNCH(record_to_bool_kernel, grid, dim3(256), 0, (hipStream_t)stream, meas, ld,
                     shots, nbits, out, ldr);
  QCF_HIP_CHECK(hipGetLastError());
  return QCF_OK;
}

extern "C" int qcf_record_from_bool(const uint8_t *in, size_t ldr, uint64_t shots,
                                    uint32_t nbits, uint64_t *meas, size_t ld, void *stream) {
  clear_error();
  if (shots == 0 || nbits == 0) return QCF_OK;
  if (ld < (shots + 63) / 64) return set_error(QCF_ERR_DIMENSION, "record stride too short");
  QCF_TRY(ref_stride(&ldr, shots));
  dim3 grid((unsigned)((ld + 255) / 256), nbits);
  QCF_LAUNCH(record_from_bool_kernel, grid, dim3(256), 0, (hipStream_t)stream, in, ldr,
                     shots, nbits, meas, ld);
  QCF_HIP_CHECK(hipGetLastError());
  return QCF_OK;
}

extern "C" int qcf_frame_to_reference(const uint64_t *frame, size_t ld, uint64_t shots,
                                      uint32_t nq, uint64_t *xzs, size_t ldr, void *stream) {
  clear_error();
  if (shots == 0 || nq == 0) return QCF_OK;
  if (ld < (shots + 63) / 64) return set_error(QCF_ERR_DIMENSION, "frame stride too short");
  QCF_TRY(ref_stride(&ldr, shots));
  const uint32_t W = (nq + 63) / 64;
  dim3 grid((unsigned)((shots + 255) / 256), 2 * W);
  QCF_LAUNCH(frame_to_reference_kernel, grid, dim3(256), 0, (hipStream_t)stream,
                     frame, ld, shots, nq, xzs, ldr);
  QCF_HIP_CHECK(hipGetLastError());
  return QCF_OK;
}

extern "C" int qcf_frame_from_reference(const uint64_t *xzs, size_t ldr, uint64_t shots,
                                        uint32_t nq, uint64_t *frame, size_t ld, void *stream) {
  clear_error();
  if (shots == 0 || nq == 0) return QCF_OK;
  if (ld < (shots + 63) / 64) return set_error(QCF_ERR_DIMENSION, "frame stride too short");
  QCF_TRY(ref_stride(&ldr, shots));
  dim3 grid((unsigned)((ld + 255) / 256), 2 * nq);
  QCF_LAUNCH(frame_from_reference_kernel, grid, dim3(256), 0, (hipStream_t)stream,
                     xzs, ldr, shots, nq, frame, ld);
  QCF_HIP_CHECK(hipGetLastError());
  return QCF_OK;
}

// ======================================================================
// Batched persistent launch + autotuner
// ======================================================================
// Circuits that have a specialised kernel (frames in LDS, split between
// registers and LDS, or part tiles) run it, spread over helper streams forked
// from the caller's stream; the persistent interpreter kernel takes the rest (frames
// beyond split mode) on the caller's stream.  Both draw identical bits.
constexpr int kBatchStreams = 4;
struct qcf_batch {
  int device = -1;
  size_t n = 0;
  std::vector<qcf::KParams> params;
  std::vector<uint8_t> on_jit;          // circuit k runs its specialised kernel
  std::vector<qcf::JitPlan> jit;        // one launch per kernel, costliest first
  std::vector<double> jit_work;         // estimated cost of each launch
  std::vector<int> jit_stream;          // helper stream of each launch
  std::vector<uint32_t> jit_planes;     // frame planes of each launch's circuits
  std::vector<uint8_t> jit_stealable;   // the persistent kernel can run its tiles
  std::vector<std::vector<qcf::WorkItem>> jit_tiles;  // each launch's tiles, in its order
  std::vector<int> jit_share;           // shared group of each launch, or -1
  uint32_t nshare = 0;                  // launches sharing tiles with the persistent kernel
  qcf::WorkItem *d_shared = nullptr;    // their tiles, group after group
  uint32_t *d_shared_first = nullptr;   // group g: d_shared[first[g] .. first[g + 1])
  int nstreams = 2;                     // helper streams in use
  bool concurrent = false;              // persistent kernel beside the specialised ones
  void *d_jit = nullptr;                // per-circuit parameters and tile offsets
  std::vector<qcf::WorkItem> items;
  qcf::KParams *d_params = nullptr;
  qcf::WorkItem *d_items = nullptr;
  uint32_t *d_queue = nullptr;
  uint32_t *d_frames = nullptr;
  uint32_t lds_bytes = 0;
  uint32_t grid = 0;
  hipStream_t helpers[kBatchStreams] = {};
  hipEvent_t fork = nullptr, join[kBatchStreams] = {};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~qcf_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (int h = 0; h < kBatchStreams; ++h) {
      if (helpers[h]) (void)hipStreamDestroy(helpers[h]);
      if (join[h]) (void)hipEventDestroy(join[h]);
    }
    if (fork) (void)hipEventDestroy(fork);
    if (d_jit) (void)hipFree(d_jit);
    if (d_params) (void)hipFree(d_params);
    if (d_items) (void)hipFree(d_items);
    if (d_shared) (void)hipFree(d_shared);
    if (d_shared_first) (void)hipFree(d_shared_first);
    if (d_queue) (void)hipFree(d_queue);
    if (d_frames) (void)hipFree(d_frames);
  }
};

namespace qcf {
namespace {

int build_kparams(const qcf_program *prog, uint64_t shots, uint64_t seed,
                  uint64_t *meas, size_t ld, KParams *kp) {
  qcf_program *P = const_cast<qcf_program *>(prog);
  if (ld < (shots + 63) / 64)
    return set_error(QCF_ERR_DIMENSION, "batched pftrajectories: stride %zu < %llu shots", ld,
                     (unsigned long long)shots);
  if (!meas && P->nbits > 0)
    return set_error(QCF_ERR_ARGUMENT, "batched pftrajectories: null record");
  *kp = KParams{};
  QCF_TRY(ensure_device_copy(P, 1, &kp->prog, &kp->ninstr));
  const uint64_t nw32 = (shots + 31) / 32;
  kp->nplanes = P->nplanes;
  kp->nw32 = (uint32_t)nw32;
  kp->nblocks = (uint32_t)((nw32 + 63) / 64);
  kp->nwbuf = (uint32_t)std::min<uint64_t>(2 * (uint64_t)ld, 64ull * kp->nblocks);
  const uint32_t tail = (uint32_t)(shots % 32);
  kp->lastmask = tail ? ((1u << tail) - 1u) : 0xFFFFFFFFu;
  kp->ld32 = 2 * (uint64_t)ld;
  kp->seed = seed;
  kp->rec = (uint32_t *)meas;
  kp->body_rbase = P->ninit_rs;
  kp->nclasses = (uint32_t)P->classes.size();
  for (uint32_t c = 0; c < kMaxClasses; ++c)
    kp->cls_c[c] = c < P->classes.size() ? P->classes[c].c : 0.f;
  kp->thin = P->thin ? 1u : 0u;
  return QCF_OK;
}

void make_items(qcf_batch *B, uint32_t tile_blocks) {
  B->items.clear();
  std::vector<std::pair<double, WorkItem>> tmp;
  for (size_t k = 0; k < B->n; ++k) {
    if (B->on_jit[k]) continue;
    const KParams &kp = B->params[k];
    for (uint32_t b0 = 0; b0 < kp.nblocks; b0 += tile_blocks) {
      const uint32_t nb = std::min(tile_blocks, kp.nblocks - b0);
      tmp.push_back({(double)kp.ninstr * nb, WorkItem{(uint32_t)k, b0, nb, 0}});
    }
  }
  // longest items first so the queue drains evenly
  std::stable_sort(tmp.begin(), tmp.end(),
                   [](const auto &a, const auto &b) { return a.first > b.first; });
  for (auto &t : tmp) B->items.push_back(t.second);
}

int upload_items(qcf_batch *B) {
  if (B->d_items) QCF_HIP_CHECK(hipFree(B->d_items));
  B->d_items = nullptr;
  const size_t n = std::max<size_t>(1, B->items.size());
  QCF_HIP_CHECK(hipMalloc(&B->d_items, n * sizeof(WorkItem)));
  if (!B->items.empty())
    QCF_HIP_CHECK(hipMemcpy(B->d_items, B->items.data(), B->items.size() * sizeof(WorkItem),
                            hipMemcpyHostToDevice));
  return QCF_OK;
}

// Specialised launches whose circuits the persistent kernel can run (frames
// in its global slots) and whose frames have more than `planes` planes share
// their tiles with it (0: more than kStealPlanes; 0xFFFFFFFF: none): both
// kernels take tile indices from one counter per launch.  The tiles are the
// same whichever kernel runs them (every draw is keyed on the tile), so the
// records do not depend on the split.
//
// It pays only where the specialised kernel is slow per tile: a persistent
// wave that takes a tile holds it for far longer (it interprets the program
// on global-memory planes, beside thousands of others), so the launch ends
// up to one such tile after the specialised kernel would have.  On the
// 512-circuit memory batch (d = 3..15, 1.6e7 to 2e7 shots per distance),
// sharing took d = 15 (1237 planes) from 108 to 55 ms at 8 persistent waves
// per CU, but d = 13 (929 planes) from 46 to 48-53 ms and d = 11 (665) from
// 14 to 26-30 ms with the kernels of that time.  Since the d = 13 split
// kernels took 24 noise slots and 64-bit record addresses, sharing d = 13 as
// well measured best for the whole batch: 78.0-80.9 ms against 89.2-95.6
// with d = 15 alone and 78.7-79.1 with d = 11 too (8 waves per CU, one or
// two helper streams).
constexpr uint32_t kStealPlanes = 768;
constexpr uint32_t kStealWavesPerCU = 8;  // persistent grid when it shares tiles
int share_tiles(qcf_batch *B, uint32_t planes) {
  if (planes == 0) planes = kStealPlanes;
  B->jit_share.assign(B->jit.size(), -1);
  std::vector<WorkItem> shared;
  std::vector<uint32_t> first{0u};
  uint32_t g = 0;
  for (size_t i = 0; i < B->jit.size(); ++i) {
    if (!B->jit_stealable[i] || planes == 0xFFFFFFFFu || B->jit_planes[i] <= planes) continue;
    B->jit_share[i] = (int)g++;
    shared.insert(shared.end(), B->jit_tiles[i].begin(), B->jit_tiles[i].end());
    first.push_back((uint32_t)shared.size());
  }
  B->nshare = g;
  if (B->d_shared) QCF_HIP_CHECK(hipFree(B->d_shared));
  if (B->d_shared_first) QCF_HIP_CHECK(hipFree(B->d_shared_first));
  B->d_shared = nullptr;
  B->d_shared_first = nullptr;
  if (!g) return QCF_OK;
  QCF_HIP_CHECK(hipMalloc(&B->d_shared, shared.size() * sizeof(WorkItem)));
  QCF_HIP_CHECK(hipMalloc(&B->d_shared_first, first.size() * sizeof(uint32_t)));
  QCF_HIP_CHECK(hipMemcpy(B->d_shared, shared.data(), shared.size() * sizeof(WorkItem),
                          hipMemcpyHostToDevice));
  QCF_HIP_CHECK(hipMemcpy(B->d_shared_first, first.data(), first.size() * sizeof(uint32_t),
                          hipMemcpyHostToDevice));
  return QCF_OK;
}

void assign_streams(qcf_batch *B);
int batch_streams();
bool batch_concurrent();

int set_tune(qcf_batch *B, const qcf_tune *tune) {
  DeviceState *st;
  QCF_TRY(device_state(&st));
  B->nstreams = tune && tune->streams ? std::max(1, std::min(kBatchStreams, (int)tune->streams))
                                      : batch_streams();
  B->concurrent = tune && tune->concurrent ? true : batch_concurrent();
  assign_streams(B);
  uint32_t wpc = tune ? tune->waves_per_cu : 0;
  uint32_t tb = tune ? tune->tile_blocks : 0;
  const uint32_t lds_cap = std::max<uint32_t>(1, (160u * 1024u) / std::max(1u, B->lds_bytes));
  QCF_TRY(share_tiles(B, tune ? tune->steal_planes : 0u));
  if (wpc == 0) wpc = std::min<uint32_t>(B->nshare ? kStealWavesPerCU : 16u, lds_cap);
  wpc = std::max<uint32_t>(1, std::min(std::min(wpc, lds_cap), kMaxBatchWavesPerCU));
  if (tb == 0) tb = 1;
  make_items(B, tb);
  QCF_TRY(upload_items(B));
  uint64_t work = B->items.size();
  for (size_t i = 0; i < B->jit.size(); ++i)
    if (B->jit_share[i] >= 0) work += B->jit_tiles[i].size();
  B->grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)st->cu_count * wpc, work));
  if (B->exec) {
    (void)hipGraphExecDestroy(B->exec);
    B->exec = nullptr;
  }
  if (B->graph) {
    (void)hipGraphDestroy(B->graph);
    B->graph = nullptr;
  }
  return QCF_OK;
}

// QCF_BATCH_JIT=0 keeps every circuit on the persistent interpreter kernel.
bool batch_jit_enabled() {
  const char *e = getenv("QCF_BATCH_JIT");
  return !(e && e[0] == '0');
}

// Helper streams the specialised launches are spread over (QCF_BATCH_STREAMS,
// 1..4; default 2: the 512-circuit batch took 181, 157 and 162 ms on 1, 2
// and 4 -- the per-distance launches each fill the GPU, so two overlap the
// tail of one with the start of the next without splitting CUs further).
int batch_streams() {
  const char *e = getenv("QCF_BATCH_STREAMS");
  const int v = e ? atoi(e) : 2;
  return std::max(1, std::min(kBatchStreams, v));
}

// QCF_BATCH_JIT_ORDER=concurrent starts the persistent kernel beside the
// specialised launches instead of after them.
bool batch_concurrent() {
  const char *e = getenv("QCF_BATCH_JIT_ORDER");
  return e && e[0] == 'c';
}

// Specialised launches of the circuits that have a kernel.  Circuits whose
// kernels have the same source (the same code: for memory experiments, the
// same distance at any noise rate) share ONE launch over all their tiles --
// the kernel finds each tile's circuit in a table of per-circuit parameters
// -- so a distance with many small circuits still fills the GPU.  Launches go
// costliest first, each to the helper stream with the least work so far.
// No allocation and no compilation happen at launch time (both are done
// here), so the launch sequence can be captured into a graph.
// Costliest launch first, each to the helper stream with the least work so far.
void assign_streams(qcf_batch *B) {
  B->jit_stream.assign(B->jit.size(), 0);
  double load[kBatchStreams] = {};
  for (size_t i = 0; i < B->jit.size(); ++i) {
    int h = 0;
    for (int j = 1; j < B->nstreams; ++j)
      if (load[j] < load[h]) h = j;
    load[h] += B->jit_work[i];
    B->jit_stream[i] = h;
  }
}

int plan_jit(qcf_batch *B) {
  std::vector<JitPlan> plans(B->jit);  // indexed by circuit
  B->jit.clear();
  B->jit_stream.clear();
  std::vector<uint64_t> keys;
  std::vector<std::vector<size_t>> members;
  for (size_t k = 0; k < B->n; ++k) {
    if (!B->on_jit[k]) continue;
    size_t g = 0;
    while (g < keys.size() && keys[g] != plans[k].key) ++g;
    if (g == keys.size()) {
      keys.push_back(plans[k].key);
      members.emplace_back();
    }
    members[g].push_back(k);
  }
  if (keys.empty()) return QCF_OK;
  // device table: every member's parameters, then every group's tile offsets
  size_t nparams = 0, nfirst = 0;
  for (const auto &m : members) nparams += m.size(), nfirst += m.size() + 1;
  const size_t pbytes = (nparams * sizeof(JitParams) + 255) / 256 * 256;
  std::vector<char> host(pbytes + nfirst * sizeof(uint32_t), 0);
  JitParams *hp = (JitParams *)host.data();
  uint32_t *hf = (uint32_t *)(host.data() + pbytes);
  QCF_HIP_CHECK(hipMalloc(&B->d_jit, host.size()));
  char *dbase = (char *)B->d_jit;
  DeviceState *st;
  QCF_TRY(device_state(&st));
  struct Launch {
    double work;
    JitPlan plan;
    uint32_t planes;
    bool stealable;
    std::vector<WorkItem> tiles;
  };
  std::vector<Launch> launches;
  size_t pi = 0, fi = 0;
  for (const auto &m : members) {
    const size_t p0 = pi, f0 = fi;
    uint32_t tiles = 0;
    double work = 0.0;
    for (size_t k : m) {
      hp[pi] = plans[k].jp;
      hp[pi].ncirc = 0;
      hp[pi].xcds = 0;
      hf[fi++] = tiles;
      tiles += B->params[k].nblocks;
      work += (double)B->params[k].ninstr * B->params[k].nblocks;
      ++pi;
    }
    hf[fi++] = tiles;
    JitPlan gp = plans[m[0]];
    gp.jp.multi = (const JitParams *)(dbase + p0 * sizeof(JitParams));
    gp.jp.first = (const uint32_t *)(dbase + pbytes + f0 * sizeof(uint32_t));
    gp.jp.ncirc = (uint32_t)m.size();
    gp.jp.pset = 0;  // every workgroup runs all parts of its tiles
    gp.jp.ntiles = tiles;
    uint64_t g1 = std::min<uint64_t>(tiles, (uint64_t)st->cu_count * std::max(1u, gp.blocks_per_cu));
    const uint32_t nx = jit_xcds();
    gp.jp.xcds = 0;
    if (nx > 1 && g1 >= nx) {
      g1 -= g1 % nx;
      gp.jp.xcds = nx;
    }
    gp.grid = (uint32_t)std::max<uint64_t>(1, g1);
    Launch L{work / gp.grid, gp, B->params[m[0]].nplanes, true, {}};
    for (size_t k : m) {
      L.stealable = L.stealable && B->params[k].scratch != nullptr;
      for (uint32_t b = 0; b < B->params[k].nblocks; ++b)
        L.tiles.push_back(WorkItem{(uint32_t)k, b, 1u, 0u});
    }
    if (!L.stealable) L.tiles.clear();
    launches.push_back(std::move(L));
  }
  QCF_HIP_CHECK(hipMemcpy(B->d_jit, host.data(), host.size(), hipMemcpyHostToDevice));
  std::stable_sort(launches.begin(), launches.end(),
                   [](const auto &a, const auto &b) { return a.work > b.work; });
  for (auto &l : launches) {
    B->jit.push_back(l.plan);
    B->jit_work.push_back(l.work);
    B->jit_planes.push_back(l.planes);
    B->jit_stealable.push_back(l.stealable ? 1 : 0);
    B->jit_tiles.push_back(std::move(l.tiles));
  }
  B->jit_share.assign(B->jit.size(), -1);
  assign_streams(B);
  for (int h = 0; h < kBatchStreams; ++h) {
    QCF_HIP_CHECK(hipStreamCreateWithFlags(&B->helpers[h], hipStreamNonBlocking));
    QCF_HIP_CHECK(hipEventCreateWithFlags(&B->join[h], hipEventDisableTiming));
  }
  QCF_HIP_CHECK(hipEventCreateWithFlags(&B->fork, hipEventDisableTiming));
  return QCF_OK;
}

int batch_launch(qcf_batch *B, hipStream_t s) {
  // shared tiles need the persistent kernel running beside the launches
  const bool conc = B->concurrent || B->nshare;
  uint32_t total = B->grid;  // workgroups that count into the done counter
  for (size_t i = 0; i < B->jit.size(); ++i)
    if (B->jit_share[i] >= 0) total += B->jit[i].grid;
  auto persistent = [&]() -> int {
    if (B->items.empty() && !B->nshare) return QCF_OK;
    QCF_TRY(allow_lds((const void *)pf_batch_kernel, B->lds_bytes));
    const BatchArgs A{B->d_params, B->d_items,        (uint32_t)B->items.size(), B->nshare,
                      B->d_queue,  B->d_shared,       B->d_shared_first,         total,
                      2u + B->nshare};
    QCF_LAUNCH(pf_batch_kernel, dim3(B->grid), dim3(64), B->lds_bytes, s, A);
    QCF_HIP_CHECK(hipGetLastError());
    return QCF_OK;
  };
  if (B->jit.empty()) return persistent();
  QCF_HIP_CHECK(hipEventRecord(B->fork, s));
  for (int h = 0; h < kBatchStreams; ++h) QCF_HIP_CHECK(hipStreamWaitEvent(B->helpers[h], B->fork, 0));
  for (size_t i = 0; i < B->jit.size(); ++i) {
    const JitPlan &pl = B->jit[i];
    JitParams jp = pl.jp;
    if (B->jit_share[i] >= 0) {
      jp.steal = B->d_queue + 2 + B->jit_share[i];
      jp.sync = B->d_queue;
      jp.sync_total = total;
      jp.sync_words = 2u + B->nshare;
    }
    QCF_TRY(jit_launch(pl.function, jp, pl.grid, pl.lds_bytes, B->helpers[B->jit_stream[i]]));
  }
  if (conc) QCF_TRY(persistent());
  for (int h = 0; h < kBatchStreams; ++h) {
    QCF_HIP_CHECK(hipEventRecord(B->join[h], B->helpers[h]));
    QCF_HIP_CHECK(hipStreamWaitEvent(s, B->join[h], 0));
  }
  if (!conc) QCF_TRY(persistent());
  return QCF_OK;
}

}  // namespace
}  // namespace qcf

extern "C" int qcf_batch_create(const qcf_program *const *progs, size_t n,
                                const uint64_t *shots, const uint64_t *seeds,
                                uint64_t *const *meas, const size_t *ld,
                                const qcf_tune *tune, qcf_batch **out) {
  clear_error();
  if (!out) return set_error(QCF_ERR_ARGUMENT, "null output pointer");
  *out = nullptr;
  if (n && (!progs || !shots || !seeds || !meas || !ld))
    return set_error(QCF_ERR_ARGUMENT, "null batch argument");
  auto *B = new (std::nothrow) qcf_batch();
  if (!B) return set_error(QCF_ERR_ALLOC, "out of host memory");
  B->n = n;
  QCF_HIP_CHECK(hipGetDevice(&B->device));
  int lds_max = 0;
  if (int st = max_lds_per_block(&lds_max); st != QCF_OK) {
    delete B;
    return st;
  }
  B->params.resize(n);
  B->on_jit.assign(n, 0);
  B->jit.assign(n, JitPlan{});
  const bool use_jit = batch_jit_enabled() && !force_global_frames();
  for (size_t k = 0; k < n; ++k) {
    if (!progs[k]) {
      delete B;
      return set_error(QCF_ERR_ARGUMENT, "null program %zu", k);
    }
    if (int st = build_kparams(progs[k], shots[k], seeds[k], meas[k], ld[k], &B->params[k]);
        st != QCF_OK) {
      delete B;
      return st;
    }
    const KParams &kp = B->params[k];
    if (!use_jit || kp.nblocks == 0) continue;
    // compiled here (once per distinct kernel source), launched later
    qcf_program *P = const_cast<qcf_program *>(progs[k]);
    if (tune && tune->jit_max_planes && P->nplanes > tune->jit_max_planes) continue;
    const bool big = (int)(P->nplanes * 256u) > lds_max;
    JitPlan pl;
    if (jit_plan(P, kp, shots[k], 1, big, 0, ld[k], &pl) == QCF_OK && !pl.global_planes) {
      B->on_jit[k] = 1;
      B->jit[k] = pl;
    } else {
      clear_error();
    }
  }
  B->nstreams = tune && tune->streams ? std::max(1, std::min(kBatchStreams, (int)tune->streams))
                                      : batch_streams();
  // circuits whose frame does not fit in LDS keep their planes in a
  // batch-owned global buffer, one slot of the largest such frame per wave:
  // those of the persistent kernel, and those of specialised launches whose
  // tiles the persistent kernel may share
  const bool force_gl = force_global_frames();
  if (tune && tune->lds_cap) lds_max = std::min<int>(lds_max, (int)tune->lds_cap);
  auto global_frame = [&](size_t k) {
    return (int)(progs[k]->nplanes * 256u) > lds_max || (force_gl && !B->on_jit[k]);
  };
  uint32_t gl_planes = 0;
  for (size_t k = 0; k < n; ++k) {
    if (global_frame(k))
      gl_planes = std::max(gl_planes, progs[k]->nplanes);
    else if (!B->on_jit[k])
      B->lds_bytes = std::max(B->lds_bytes, progs[k]->nplanes * 256u);
  }
  if (gl_planes) {
    DeviceState *st;
    if (int r = device_state(&st); r != QCF_OK) {
      delete B;
      return r;
    }
    const size_t slots = (size_t)st->cu_count * kMaxBatchWavesPerCU;  // the largest grid
    if (hipMalloc(&B->d_frames, slots * gl_planes * 256u) != hipSuccess) {
      delete B;
      return set_error(QCF_ERR_ALLOC, "batch frame buffer (%zu bytes)", slots * gl_planes * 256u);
    }
    for (size_t k = 0; k < n; ++k) {
      if (global_frame(k)) {
        B->params[k].scratch = B->d_frames;
        B->params[k].sstride = gl_planes * 64u;
      }
    }
  }
  if (int st = plan_jit(B); st != QCF_OK) {
    delete B;
    return st;
  }
  hipError_t e = hipMalloc(&B->d_params, std::max<size_t>(1, n) * sizeof(KParams));
  if (e == hipSuccess && n)
    e = hipMemcpy(B->d_params, B->params.data(), n * sizeof(KParams), hipMemcpyHostToDevice);
  // [0] next own item, [1] workgroups done, [2 + g] next tile of shared group g
  const size_t nq = 2 + B->jit.size();
  if (e == hipSuccess) e = hipMalloc(&B->d_queue, nq * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMemset(B->d_queue, 0, nq * sizeof(uint32_t));
  if (e != hipSuccess) {
    delete B;
    return set_error(QCF_ERR_HIP, "batch setup: %s", hipGetErrorString(e));
  }
  if (int st = set_tune(B, tune); st != QCF_OK) {
    delete B;
    return st;
  }
  *out = B;
  return QCF_OK;
}

extern "C" int qcf_batch_launch(qcf_batch *B, void *stream) {
  clear_error();
  if (!B) return set_error(QCF_ERR_ARGUMENT, "null batch");
  return batch_launch(B, (hipStream_t)stream);
}

extern "C" int qcf_batch_launch_graph(qcf_batch *B, void *stream) {
  clear_error();
  if (!B) return set_error(QCF_ERR_ARGUMENT, "null batch");
  hipStream_t s = (hipStream_t)stream;
  if (!B->exec) {
    hipStream_t cap;
    QCF_HIP_CHECK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    QCF_HIP_CHECK(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
    const int st = batch_launch(B, cap);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(cap, &g);
    (void)hipStreamDestroy(cap);
    if (st != QCF_OK) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    if (e != hipSuccess) return set_error(QCF_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
    B->graph = g;
    QCF_HIP_CHECK(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
  }
  QCF_HIP_CHECK(hipGraphLaunch(B->exec, s));
  return QCF_OK;
}

extern "C" int qcf_batch_free(qcf_batch *B) {
  delete B;
  return QCF_OK;
}

extern "C" int qcf_pf_sample_batched(const qcf_program *const *progs, size_t n,
                                     const uint64_t *shots, const uint64_t *seeds,
                                     uint64_t *const *meas, const size_t *ld,
                                     const qcf_tune *tune, void *stream) {
  qcf_batch *B = nullptr;
  QCF_TRY(qcf_batch_create(progs, n, shots, seeds, meas, ld, tune, &B));
  int st = qcf_batch_launch(B, stream);
  if (st == QCF_OK) {
    const hipError_t e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) st = set_error(QCF_ERR_HIP, "batch sync: %s", hipGetErrorString(e));
  }
  qcf_batch_free(B);
  return st;
}

// QCF_AUTOTUNE_LOG=1 prints every candidate's time to stderr.
static bool autotune_log() {
  const char *e = getenv("QCF_AUTOTUNE_LOG");
  return e && e[0] == '1';
}

extern "C" int qcf_autotune(const qcf_program *const *progs, size_t n,
                            const uint64_t *shots, qcf_tune *out, void *stream) {
  clear_error();
  if (!out) return set_error(QCF_ERR_ARGUMENT, "null output pointer");
  *out = qcf_tune{0, 0, 0, 0, 0, 0, 0};
  if (n == 0) return QCF_OK;
  hipStream_t s = (hipStream_t)stream;
  // scratch records so the caller's buffers are untouched
  std::vector<uint64_t *> bufs(n, nullptr);
  std::vector<size_t> lds(n);
  std::vector<uint64_t> seeds(n, 0);
  int st = QCF_OK;
  for (size_t k = 0; k < n && st == QCF_OK; ++k) {
    lds[k] = qcf_plane_words(shots[k]);
    const size_t bytes = std::max<size_t>(1, progs[k]->nbits) * lds[k] * 8;
    if (hipMalloc(&bufs[k], bytes) != hipSuccess) st = set_error(QCF_ERR_ALLOC, "autotune scratch");
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (st == QCF_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    st = set_error(QCF_ERR_HIP, "autotune events");
  float best = 1e30f;
  qcf_tune best_t{0, 1, 0, 0, 0, 0, 0};
  // Candidates run as the captured graph that qcf_batch_launch_graph replays
  // (captured once per candidate): when the persistent kernel shares tiles
  // with a specialised launch, the split between them depends on which
  // starts first, and plain launches and graph replays start them in a
  // different order -- a shape timed at 119 ms as plain launches replayed at
  // 161 ms.  One warm replay, then the median of three timed ones.
  auto time_tune = [&](qcf_batch *B, const qcf_tune &t, float *ms) -> int {
    QCF_TRY(qcf::set_tune(B, &t));
    QCF_TRY(qcf_batch_launch_graph(B, s));  // capture + warm
    float v[3];
    for (float &x : v) {
      if (hipEventRecord(e0, s) != hipSuccess) return set_error(QCF_ERR_HIP, "event");
      QCF_TRY(qcf_batch_launch_graph(B, s));
      if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
        return set_error(QCF_ERR_HIP, "event");
      (void)hipEventElapsedTime(&x, e0, e1);
    }
    std::sort(v, v + 3);
    *ms = v[1];
    if (autotune_log())
      fprintf(stderr, "qcf_autotune: wpc %u streams %u concurrent %u steal_planes %u: %.2f ms\n",
              t.waves_per_cu, t.streams, t.concurrent, t.steal_planes, *ms);
    return QCF_OK;
  };
  auto time_one = [&](qcf_batch *B, const qcf_tune &t) -> int {
    float ms = 0.f;
    QCF_TRY(time_tune(B, t, &ms));
    if (ms < best) {
      best = ms;
      best_t = t;
    }
    return QCF_OK;
  };
  // Every circuit that has a specialised kernel runs it, on 1 or 2 helper
  // streams; the persistent kernel shares the tiles of no launch or of those
  // whose frames have more planes than one of the three largest plane counts
  // among the launches it can share (the largest only, the two largest, the
  // three largest), at 8 or 16 waves per CU; when it has circuits of its own
  // and shares nothing, it starts after the launches or beside them.
  qcf_batch *B = nullptr;
  const qcf_tune tb{0, 1, 0, 0, 0, 0, 0xFFFFFFFFu};
  if (st == QCF_OK)
    st = qcf_batch_create(progs, n, shots, seeds.data(), bufs.data(), lds.data(), &tb, &B);
  if (st == QCF_OK) {
    std::vector<uint32_t> sizes;
    for (size_t i = 0; i < B->jit.size(); ++i)
      if (B->jit_stealable[i]) sizes.push_back(B->jit_planes[i]);
    std::sort(sizes.begin(), sizes.end());
    sizes.erase(std::unique(sizes.begin(), sizes.end()), sizes.end());
    std::vector<uint32_t> steals{0xFFFFFFFFu};
    for (size_t j = 0; j < 3 && j < sizes.size(); ++j)
      steals.push_back(sizes[sizes.size() - 1 - j] - 1u);
    const bool own = !B->items.empty();
    for (uint32_t streams : {1u, 2u})
      for (uint32_t steal : steals) {
        const bool shares = steal != 0xFFFFFFFFu;
        for (uint32_t conc : {0u, 1u}) {
          if (conc && (!own || B->jit.empty() || shares)) continue;
          for (uint32_t wpc : {8u, 16u}) {
            if (!own && !shares && wpc != 8u) continue;  // no persistent kernel
            if (st == QCF_OK) st = time_one(B, qcf_tune{wpc, 1, 0, 0, streams, conc, steal});
          }
        }
      }
    // The pick must beat the library defaults (the all-zero tune) when both
    // are timed again, alternately, three times each; otherwise the defaults
    // are returned.  A shape whose time swings between runs loses here.
    const qcf_tune def{0, 0, 0, 0, 0, 0, 0};
    float tb_[3], td_[3];
    for (int r = 0; r < 3 && st == QCF_OK; ++r) {
      st = time_tune(B, best_t, &tb_[r]);
      if (st == QCF_OK) st = time_tune(B, def, &td_[r]);
    }
    if (st == QCF_OK) {
      std::sort(tb_, tb_ + 3);
      std::sort(td_, td_ + 3);
      if (autotune_log())
        fprintf(stderr, "qcf_autotune: pick %.2f ms (worst %.2f), defaults %.2f ms (worst %.2f)\n",
                tb_[1], tb_[2], td_[1], td_[2]);
      if (!(tb_[1] < td_[1] && tb_[2] < td_[2])) best_t = def;
    }
  }
  if (B) qcf_batch_free(B);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  for (auto *p : bufs)
    if (p) (void)hipFree(p);
  if (st == QCF_OK) *out = best_t;
  return st;
}
