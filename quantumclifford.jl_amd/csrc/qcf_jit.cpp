// Per-circuit frame-sampling kernels compiled at run time with hiprtc.
//
// The interpreter in qcf_frames.hip decodes every device instruction on the
// wave (v_readlane, branches, address arithmetic), which the instruction
// counters show is where its time goes.  For the native-RNG launches this
// file instead writes the program out as straight-line HIP source -- plane
// XORs become ds_read/ds_xor with constant LDS offsets, noise groups call the
// shared event pass with their site tables in constant memory -- compiles it
// once per program and variant, and launches that kernel.  The RNG, event
// pass and gauge streams come from qcf_device.h, the same code the
// interpreter runs, so both produce the same bits for the same seed.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <array>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "qcf_common.hpp"
#include "qcf_jit.hpp"
#include "qcf_program.hpp"

namespace qcf {
namespace {
const char kDeviceSrc[] =
#include "qcf_device_src.inc"
    ;

// Kernel parameters; the generated source declares the same struct.
const char kKernelPrelude[] = R"QCFJIT(
struct KJ {
  uint32_t *rec;
  uint32_t *frame;
  uint64_t seed;
  uint64_t ld32;
  uint32_t nblocks, nw32, nwbuf, lastmask;
  float c;
  uint32_t has_class, nfull, xcds;
  uint32_t *scratch;  // per-wave planes (global mode)
  uint64_t sstride;   // words per wave
  // several circuits with this kernel in one launch (ncirc > 0): circuit i
  // has parameters multi[i] and owns tiles [first[i], first[i+1]) of ntiles
  const KJ *multi;
  const uint32_t *first;
  uint32_t ncirc, ntiles;
  uint32_t tile0;     // absolute index of tile 0: keys the RNG
  uint32_t *steal;    // shared tiles: the next tile index (null: dealt statically)
  uint32_t *sync;     // shared tiles: [1] workgroups done, cleared by the last
  uint32_t sync_total, sync_words;
  uint32_t pset;      // part-tile kernels: workgroups i, i+pset, ... take part i % pset of each tile
};
struct FlipLDS {
  uint32_t *L;
  __device__ __forceinline__ void operator()(uint32_t p, uint32_t tl, uint32_t bit, uint32_t,
                                             uint32_t) const {
    atomicXor(&L[p * 64u + tl], bit);
  }
};
// part-tile passes: planes of PW words (32: halves, 16: quarters); a pass
// keeps the flips that land in its part of the tile
template <unsigned PW>
struct FlipPart {
  uint32_t *L;
  uint32_t part;
  __device__ __forceinline__ void operator()(uint32_t p, uint32_t tl, uint32_t bit, uint32_t,
                                             uint32_t) const {
    if (tl / PW == part) atomicXor(&L[p * PW + (tl % PW)], bit);
  }
};
// planes in global memory: every access at L2 (agent scope) so that the
// event flips other lanes make are seen by the next read
__device__ __forceinline__ uint32_t gld(const uint32_t *a) {
  return __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void gst(uint32_t *a, uint32_t v) {
  __hip_atomic_store(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void gxr(uint32_t *a, uint32_t v) {
  __hip_atomic_fetch_xor(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
struct FlipG {
  uint32_t *F;
  __device__ __forceinline__ void operator()(uint32_t p, uint32_t tl, uint32_t bit, uint32_t,
                                             uint32_t) const {
    gxr(F + p * 64u + tl, bit);
  }
};
)QCFJIT";

// Fields of an instruction that name record columns (shifted from one loop
// iteration to the next), and fields the specialised kernel ignores (replay
// plane numbers).  Everything else must repeat exactly.
struct Roles {
  uint8_t col;     // bit k: word k (a=0 .. g=6) is a record column
  uint8_t ignore;  // bit k: word k is not used
};

Roles roles(uint32_t op) {
  switch (op & 0xFFu) {
    case I_MEAS: return {1u << 1, 0};
    case I_XSTORE2: return {1u << 2, 0};
    case I_RESETZ: return {0, 1u << 2};
    case I_ACC_REC:
    case I_ACCX_REC:
    case I_ACC_STORE:
    case I_REC_ZERO: return {1u << 0, 0};
    case I_INIT: return {0, 1u << 2};
    case I_MR: return {1u << 3, (1u << 2) | (1u << 5) | (1u << 6)};
    case I_MRX: return {(1u << 3) | (1u << 6), 1u << 2};
    case I_SITE: return {0, (1u << 2) | (1u << 5) | (1u << 6)};
    case I_NOISEG: return {0, 0x7Fu};  // the ordinal (word a) is checked separately
    default: return {0, 0};
  }
}

// Does b repeat a with record columns shifted by dcol and (for a noise group)
// the first ordinal shifted by dord?  dcol / dord are fixed by the first pair
// that determines them (*have_col / *have_ord).
bool repeats(const Instr &a, const Instr &b, int64_t *dcol, bool *have_col, int64_t *dord,
             bool *have_ord) {
  if (a.op != b.op) return false;
  const Roles r = roles(a.op);
  const uint32_t *wa = &a.a, *wb = &b.a;
  for (int k = 0; k < 7; ++k) {
    const uint32_t bit = 1u << k;
    if (r.ignore & bit) continue;
    if (r.col & bit) {
      if ((wa[k] == kNone) != (wb[k] == kNone)) return false;
      if (wa[k] == kNone) continue;
      const int64_t d = (int64_t)wb[k] - (int64_t)wa[k];
      if (!*have_col) {
        *dcol = d;
        *have_col = true;
      } else if (d != *dcol) {
        return false;
      }
      continue;
    }
    if (wa[k] != wb[k]) return false;
  }
  if ((a.op & 0xFFu) == I_NOISEG) {
    const int64_t d = (int64_t)b.a - (int64_t)a.a;
    if (!*have_ord) {
      *dord = d;
      *have_ord = true;
    } else if (d != *dord) {
      return false;
    }
  }
  return true;
}

uint64_t shape(const Instr &in) {
  const Roles r = roles(in.op);
  uint64_t h = 1469598103934665603ull ^ in.op;
  const uint32_t *w = &in.a;
  for (int k = 0; k < 7; ++k) {
    const uint32_t bit = 1u << k;
    const uint32_t v = (r.ignore & bit) ? 0u : (r.col & bit) ? (w[k] == kNone ? 1u : 2u) : w[k];
    h = (h ^ v) * 1099511628211ull;
  }
  return h;
}

}  // namespace

// Repeated stretch of the instruction stream (the rounds of a memory
// experiment): code[s, s + R*L) is R copies of code[s, s + L), copy k with
// record columns + k*dcol and noise ordinals + k*dord.
struct Loop {
  size_t s = 0, L = 0, R = 0;
  int64_t dcol = 0, dord = 0;
};

Loop find_loop(const std::vector<Instr> &code) {
  Loop best;
  const size_t n = code.size();
  if (n < 64) return best;
  std::vector<uint64_t> h(n);
  for (size_t i = 0; i < n; ++i) h[i] = shape(code[i]);
  // candidate periods: distances from an anchor in the middle to the next
  // instructions of the same shape
  std::vector<size_t> cands;
  for (size_t anchor : {n / 2, n / 3, (2 * n) / 3}) {
    size_t found = 0;
    for (size_t j = anchor + 1; j < n && found < 6; ++j)
      if (h[j] == h[anchor]) {
        cands.push_back(j - anchor);
        ++found;
      }
  }
  std::sort(cands.begin(), cands.end());
  cands.erase(std::unique(cands.begin(), cands.end()), cands.end());
  for (size_t L : cands) {
    if (L < 8 || 2 * L > n) continue;
    // longest run of i with code[i + L] repeating code[i] under one shift
    size_t i = 0;
    while (i + L < n) {
      if (h[i] != h[i + L]) {
        ++i;
        continue;
      }
      int64_t dcol = 0, dord = 0;
      bool hc = false, ho = false;
      size_t j = i;
      while (j + L < n && h[j] == h[j + L] && repeats(code[j], code[j + L], &dcol, &hc, &dord, &ho))
        ++j;
      const size_t run = j - i;
      const size_t R = (run + L) / L;
      if (R >= 2 && (R - 1) * L > (best.R > 1 ? (best.R - 1) * best.L : 0)) {
        best.s = i;
        best.L = L;
        best.R = R;
        best.dcol = hc ? dcol : 0;
        best.dord = ho ? dord : 0;
      }
      i = j > i ? j : i + 1;
    }
  }
  return best;
}

// The instruction stream a specialised kernel runs: the variant's code
// without the chunk padding, and without the noise groups when no site has a
// nonzero rate (a specialised kernel draws natively, so such sites never
// fire: their event passes and, in register mode, their slot exchanges
// would be dead work).
std::vector<Instr> kernel_code(const qcf_program *P, int variant) {
  const bool silent = P->classes.empty();
  std::vector<Instr> code;
  const std::vector<Instr> all = variant_code(P, variant);
  for (size_t i = 0; i < all.size(); ++i) {
    const uint32_t op = all[i].op & 0xFFu;
    if (op == I_END) continue;
    if (silent && op == I_NOISEG) {
      i += all[i].op >> 8;  // the group's site entries
      continue;
    }
    code.push_back(all[i]);
  }
  return code;
}

// Instructions a specialised kernel spells out: the stream with its repeated
// stretch written once.
size_t emitted_size(const qcf_program *P, int variant) {
  if (variant & 256) {  // linear mode: one store per column, one XOR per kept draw used
    size_t n = P->ngauge_live + P->nbits;
    for (const auto &r : P->rec_gauge) n += r.size();
    return n;
  }
  const std::vector<Instr> code = kernel_code(P, variant);
  const Loop lp = jit_no_loops() ? Loop{} : find_loop(code);
  return lp.R >= 2 ? code.size() - (lp.R - 1) * lp.L : code.size();
}

// Write the kernel source of one variant; returns false for programs the
// generator does not handle (the interpreter runs those).
//
// The stream is first folded: a stretch that repeats (the rounds of a memory
// experiment) is emitted once inside a loop over its copies, so the kernel
// stays small enough for the instruction cache.
//
// Two storage modes.  Programs with at most jit_reg_planes() planes keep every
// plane in a register (r<k>): gates are plain VALU XORs and the only LDS is a
// small buffer of noise slots -- the event pass XORs each event's bit into
// the slots of its site, and the site then takes its slots back with one
// exchange-with-zero per plane (so the buffer is zero again for the next
// group).  Longer programs keep the planes in LDS (L[p*64+l]) as the
// interpreter does and the events XOR straight into them.
// Split mode: LDS of one workgroup (gfx950: 160 KB) and the most planes it
// moves to registers (the rest of the 512 a wave may hold is left to the
// kernel's own values).
constexpr uint32_t kSplitLdsBytes = 160u * 1024u;
constexpr uint32_t kSplitMaxRegs = 400u;
// Noise groups (event passes) a split or part-tile kernel may emit.
constexpr size_t kMaxLargeFrameGroups = 256;

// Pieces of every specialised kernel: the loop over the tiles a workgroup
// takes (dealt out per XCD, or drawn from a shared counter; in a
// multi-circuit launch each tile looks up its circuit), the per-tile
// locals (record addressing, gauge stream), and the end of the kernel.
const char *const kTileLoopHead = R"QCFJIT(
  const uint32_t ntiles = P0.ncirc ? P0.ntiles : P0.nblocks;
  const uint32_t nx = P0.xcds ? P0.xcds : 1u;
  const uint32_t per = gx / nx, xi = bx % nx, wi = bx / nx;
  const uint32_t span = (ntiles + nx - 1u) / nx;
  for (uint32_t t = wi;; t += per) {
    uint32_t tile;
    if (P0.steal) {  // shared with the persistent batch kernel: take the next tile
      uint32_t v = 0u;
      if (l == 0u) v = atomicAdd(P0.steal, 1u);
      tile = __builtin_amdgcn_readfirstlane(v);
    } else {
      if (t >= span) break;
      tile = xi * span + t;
    }
    if (tile >= ntiles) break;
    // the circuit of this tile in a multi-circuit launch: the last i with
    // first[i] <= tile (uniform, so scalar loads, which never wait behind
    // the vector record stores)
    KJ P = P0;
    uint32_t blk = tile;
    if (P0.ncirc) {
      // the table is read-only during the launch: constant address space
      typedef __attribute__((address_space(4))) const uint32_t cu32;
      typedef __attribute__((address_space(4))) const KJ cKJ;
      const cu32 *const first = (const cu32 *)P0.first;
      uint32_t lo = 0u, hi = P0.ncirc;
      while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (first[mid] <= tile) lo = mid;
        else hi = mid;
      }
      const cKJ *const Q = (const cKJ *)P0.multi + lo;
      P.rec = Q->rec;
      P.frame = Q->frame;
      P.seed = Q->seed;
      P.ld32 = Q->ld32;
      P.nw32 = Q->nw32;
      P.nwbuf = Q->nwbuf;
      P.lastmask = Q->lastmask;
      P.c = Q->c;
      P.has_class = Q->has_class;
      P.nfull = Q->nfull;
      P.tile0 = Q->tile0;
      blk = tile - first[lo];
    }
)QCFJIT";
const char *const kTileLocals = R"QCFJIT(
    // an opaque per-tile copy of the plane stride keeps the compiler from
    // hoisting every record address out of the tile loop
    uint64_t ld = P.ld32;
    asm volatile("" : "+s"(ld));
    uint32_t ld4 = (uint32_t)ld * 4u;  // plane stride in bytes (rec32 kernels)
    asm volatile("" : "+s"(ld4));
    char *const recb = (char *)P.rec;
    const uint32_t gw4 = gw * 4u;
    (void)ld4; (void)recb; (void)gw4;
    qcfd::Gauge g = qcfd::gauge_stream(P.seed, gw + P.tile0 * 64u);
)QCFJIT";
const char *const kTileLoopTail = R"QCFJIT(  }
  if (P0.sync && l == 0u) {  // the last workgroup of the sharing launches clears the counters
    __threadfence();
    if (atomicAdd(&P0.sync[1], 1u) == P0.sync_total - 1u)
      for (uint32_t i = 0; i < P0.sync_words; ++i)
        __hip_atomic_store(&P0.sync[i], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}
)QCFJIT";

// Linear mode (variant bit 8): a noiseless program run fresh, returning only
// the record.  Its record is then linear over GF(2) in the gauge draws it
// makes (the kept ones, see mark_dead_gauges): column b is the XOR of the
// kept draws in rec_gauge[b].  The kernel draws those words from the same
// per-word gauge stream in the same order as the frame kernels and the
// interpreter, and stores every column's XOR -- no frame at all, so the same
// bits in a fraction of the work (the reference's d=11 benchmark circuit: 220
// words and 2,928 XORs for 2,751 columns per 32-shot word, where the frame
// runs 17,802 operations over 884 planes).
bool generate_linear(const qcf_program *P, int variant, std::string *out, uint32_t *lds_bytes) {
  if (!P->linear_ok || (variant & 3) != 1 || P->rec_gauge.size() != P->nbits) return false;
  const bool rec32 = (variant & 8) != 0;
  const bool nt = jit_nt_stores();
  size_t nnz = 0;
  for (const auto &r : P->rec_gauge) nnz += r.size();
  std::ostringstream body;
  for (uint32_t k = 0; k < P->ngauge_live; ++k)
    body << "      const uint32_t G" << k << " = g.next();\n";
  for (uint32_t col = 0; col < P->nbits; ++col) {
    std::string v;
    for (uint32_t x : P->rec_gauge[col]) v += (v.empty() ? "G" : " ^ G") + std::to_string(x);
    if (v.empty()) v = "0u";
    const std::string m =
        rec32 ? "*(uint32_t *)(recb + (uint32_t)(" + std::to_string(col) + "u * ld4 + gw4))"
              : "P.rec[(uint64_t)" + std::to_string(col) + "u * ld + gw]";
    const std::string val = "(" + v + ") & lm";
    body << "      " << (nt ? "__builtin_nontemporal_store(" + val + ", &" + m + ");"
                           : m + " = " + val + ";")
         << "\n";
  }
  std::ostringstream src;
  src << "#define QCF_RTC 1\n// linear: " << P->ngauge_live << " gauge words, " << nnz
      << " XORs, " << P->nbits << " record columns\n";
  src << kDeviceSrc << kKernelPrelude;
  src << "extern \"C\" __global__ void __launch_bounds__(64) qcf_jit_kernel(KJ P0) {\n"
      << "  const uint32_t l = threadIdx.x, bx = blockIdx.x, gx = gridDim.x;\n";
  src << kTileLoopHead;
  src << "    const uint32_t gw = blk * 64u + l;\n";
  src << kTileLocals;
  // lanes inside the plane stride store (zeros past the shots)
  src << "    const bool live = gw < P.nwbuf;\n"
      << "    const uint32_t lm = gw >= P.nw32 ? 0u : (gw == P.nw32 - 1u ? P.lastmask : "
         "0xFFFFFFFFu);\n"
      << "    if (live) {\n" << body.str() << "    }\n";
  src << kTileLoopTail;
  *out = src.str();
  if (lds_bytes) *lds_bytes = 0;
  return true;
}

bool generate(const qcf_program *P, int variant, std::string *out, uint32_t *lds_bytes) {
  // the event stream's thinning draw, a constant of the kernel
  if (variant & 256) return generate_linear(P, variant, out, lds_bytes);
  const char *const thin = P->thin ? "1u" : "0u";
  const std::vector<Instr> code = kernel_code(P, variant);
  Loop lp = jit_no_loops() ? Loop{} : find_loop(code);
  if (lp.R < 2) lp = Loop{0, 0, 0, 0, 0};
  // emitted ranges: [0, s), loop body [s, s+L), [s+R*L, n)
  struct Range {
    size_t b, e;
    bool loop;
  };
  std::vector<Range> ranges;
  if (lp.R >= 2) {
    ranges = {{0, lp.s, false}, {lp.s, lp.s + lp.L, true}, {lp.s + lp.R * lp.L, code.size(), false}};
  } else {
    ranges = {{0, code.size(), false}};
  }
  // frames too large for LDS (variant bit 2): with bit 4 the planes used most
  // stay in registers and the rest in LDS (split mode), else every plane in
  // global memory; fitting frames: registers for small programs, else LDS
  // variant bits 6 / 7: part-tile passes -- every plane in LDS at 128 / 64
  // bytes (the shot words of one half / quarter of the tile), two / four
  // passes per tile, for frames beyond split mode that fit in LDS that way
  const bool halfm = (variant & (64 | 128)) != 0;
  const uint32_t pw = (variant & 128) ? 16u : (variant & 64) ? 32u : 64u;  // words per LDS plane
  if (halfm && P->nplanes * pw * 4u > kSplitLdsBytes) return false;
  const bool split = !halfm && (variant & 16) != 0;
  const bool glob = !split && !halfm && (variant & 4) != 0;
  // variant bit 3: the record spans less than 4 GB, so a record address is
  // one 64-bit base plus a 32-bit byte offset (one scalar multiply per
  // store instead of a 64-bit multiply-add)
  bool rec32 = (variant & 8) != 0;
  // registers pay an LDS exchange per noisy plane per noise pass; LDS planes
  // pay LDS traffic per gate but take events for free, and measured faster
  // until events get dense (d=5 x 25: 0.62 vs 0.76 ms at p=1e-3, 2.1 vs
  // 1.64 ms at p=1e-2), so registers are used for small programs with many
  // events per tile
  const double ev_per_tile =
      P->classes.empty() ? 0.0 : (double)P->sites.size() * kBlockShots * P->classes[0].p;
  // noiseless programs take no exchanges at all and gain the occupancy of
  // registers (d=5 x 25: 0.295 vs 0.395 ms)
  const bool regs = !glob && !split && !halfm && P->nplanes <= jit_reg_planes() &&
                    (P->classes.empty() || ev_per_tile >= jit_reg_events() ||
                     2u * P->nq * 256u > 40u * 1024u);
  const bool nt = jit_nt_stores();
  // LDS mode runs several waves per SIMD, and a wave in an event pass (a
  // chain of dependent draws, a wave-wide scan and permutes) issues ahead of
  // the others while it lasts: d=5 x 25, 10^7 shots, p = 1e-4 / 1e-3 / 3e-3:
  // 0.320 / 0.406 / 0.763 ms at one priority, 0.311 / 0.393 / 0.760 with
  // passes raised.  Register, split and part-tile kernels stay at one
  // priority (one wave per SIMD, or no measured gain).
  const int prio_env = jit_pass_prio();
  const bool lds_mode = !glob && !split && !halfm && !regs;
  const int pass_prio = prio_env >= 0 ? prio_env : (lds_mode ? 1 : 0);
  // LDS mode keeps the classical slots (planes 2*nq and up, which no noise
  // site names) in registers: only the qubit planes take LDS, so more waves
  // fit on a CU
  const uint32_t qpl = 2u * P->nq;
  bool hyb = !regs && !glob && !split && !halfm && P->nplanes > qpl;
  for (size_t i = 0; hyb && i < code.size(); ++i) {
    if ((code[i].op & 0xFFu) != I_NOISEG) continue;
    const uint32_t G = code[i].op >> 8;
    for (uint32_t k = 1; k <= G && i + k < code.size(); ++k) {
      const Instr &s = code[i + k];
      const uint32_t pl[4] = {s.a & 0xFFFFu, s.a >> 16, s.b & 0xFFFFu, s.b >> 16};
      for (uint32_t j = 0; j < (s.b ? 4u : 2u); ++j)
        if (pl[j] >= qpl) hyb = false;
    }
    i += G;
  }
  uint32_t kSlots = jit_slots();  // noise slots per event pass (register and split modes)
  // Split mode: sreg[p] marks the planes kept in registers, lidx[p] gives the
  // others their LDS plane (packed), and the noise slots of register planes
  // follow the last LDS plane.  A plane's score is the operations that read
  // or write it (each an LDS access if it stayed in LDS) less its noise sites
  // (each an LDS exchange if it moved to a register); the highest scores take
  // the registers, as many as the LDS budget needs and jit_split_regs() asks.
  std::vector<uint8_t> sreg(P->nplanes, 0);
  std::vector<uint32_t> lidx(P->nplanes, 0);
  uint32_t nlds = 0;
  if (split) {
    std::vector<int64_t> score(P->nplanes, 0);
    bool noisy = false;
    std::vector<uint32_t> pl;
    for (size_t i = 0; i < code.size(); ++i) {
      pl.clear();
      instr_planes(code, i, &pl);
      const bool noise = (code[i].op & 0xFFu) == I_NOISEG;
      for (uint32_t p : pl)
        if (p < P->nplanes) score[p] += noise ? -1 : 1;
      if (noise) {
        noisy = true;
        i += code[i].op >> 8;
      }
    }
    // variant bit 5: half of a CU's LDS per wave, so that two waves share it
    const uint32_t lds_planes = kSplitLdsBytes / 256u / ((variant & 32) ? 2u : 1u);
    // Frames that would need more than 256 register planes beside the
    // default slots take 24 slots instead: the LDS they free holds planes
    // that would otherwise sit in registers (the accumulation registers, past
    // 256), at the cost of event passes split to fit fewer slots.  Rotated
    // d=13, p = 1e-4 / 1e-3 / 1e-2, 4e6 shots: 5.08 / 6.28 / 18.4 ms with 80
    // slots, 3.61 / 4.81 / 17.0 with 24; d=11 (about 110 register planes)
    // and two-wave d=9 measured best at the default.
    if (noisy && !getenv("QCF_JIT_SLOTS") && lds_planes > kSlots &&
        P->nplanes > lds_planes - kSlots && P->nplanes - (lds_planes - kSlots) > 256u)
      kSlots = 24;
    if (noisy && lds_planes <= kSlots) return false;
    const uint32_t budget = lds_planes - (noisy ? kSlots : 0u);
    const uint32_t kmin = P->nplanes > budget ? P->nplanes - budget : 0u;
    const uint32_t k = std::min(P->nplanes, std::max(kmin, jit_split_regs()));
    if (k > kSplitMaxRegs) return false;
    if ((variant & 32) && k > jit_two_wave_regs()) return false;
    // With more than 256 register planes the 64-bit record addressing
    // measured faster than the 32-bit offsets (rotated d=13, 4e6 shots,
    // p = 1e-3: 4.07 against 4.69 ms; every smaller frame the other way
    // round, e.g. d=11 split 2.26 against 2.18 ms, d=5 x 25 0.45 against 0.39)
    if (k > 256u) rec32 = false;
    std::vector<uint32_t> order(P->nplanes);
    for (uint32_t p = 0; p < P->nplanes; ++p) order[p] = p;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t x, uint32_t y) { return score[x] > score[y]; });
    for (uint32_t j = 0; j < k; ++j) sreg[order[j]] = 1;
    for (uint32_t p = 0; p < P->nplanes; ++p)
      if (!sreg[p]) lidx[p] = nlds++;
  }
  // Part-tile kernels: up to jit_part_regs() planes that no noise site names
  // sit in registers, those the most operations touch; the other planes keep
  // their LDS places (part_lds: LDS planes up to the last of them).  Serial
  // chains through such planes then run without an LDS round trip per step:
  // x_diag_circuit_noisy_measurement(1000) XORs every qubit into the target
  // and 7 classical slots into each of its 1000 records.
  std::vector<uint8_t> preg(P->nplanes, 0);
  uint32_t part_lds = P->nplanes;
  if (halfm && jit_part_regs()) {
    std::vector<int64_t> score(P->nplanes, 0);
    std::vector<uint8_t> named(P->nplanes, 0);
    std::vector<uint32_t> pl;
    for (size_t i = 0; i < code.size(); ++i) {
      pl.clear();
      instr_planes(code, i, &pl);
      const bool noise = (code[i].op & 0xFFu) == I_NOISEG;
      for (uint32_t p : pl)
        if (p < P->nplanes) {
          if (noise) named[p] = 1;
          else ++score[p];
        }
      if (noise) i += code[i].op >> 8;
    }
    std::vector<uint32_t> cand;
    for (uint32_t p = 0; p < P->nplanes; ++p)
      if (!named[p] && score[p] >= 2) cand.push_back(p);
    std::stable_sort(cand.begin(), cand.end(),
                     [&](uint32_t x, uint32_t y) { return score[x] > score[y]; });
    if (cand.size() > jit_part_regs()) cand.resize(jit_part_regs());
    for (uint32_t p : cand) preg[p] = 1;
    while (part_lds > 1u && preg[part_lds - 1u]) --part_lds;
  }
  auto in_reg = [&](uint32_t p) {
    return regs || (hyb && p >= qpl) || (split && sreg[p]) || (halfm && preg[p]);
  };
  // LDS plane of an LDS-resident plane
  auto lp_of = [&](uint32_t p) { return split ? lidx[p] : p; };
  const uint32_t slot_base = split ? nlds : 0u;  // first noise slot
  // LDS plane q: ds instructions take a 16-bit byte offset, so planes past
  // the first 64 KB window (256 planes) are addressed from an opaque per-window
  // base (lb<w> = w * 16384 + l); a single base register per window then
  // serves every access in it, where per-access addresses would be hoisted
  // out of the loops and spilled
  const uint32_t per_window = 16384u / pw;  // planes per 64 KB window
  auto plane = [pw, per_window](uint32_t q) {
    const uint32_t w = q / per_window, off = (q % per_window) * pw;
    if (w == 0) return "L[" + std::to_string(off) + "u + l]";
    return "L[lb" + std::to_string(w) + " + " + std::to_string(off) + "u]";
  };
  const size_t split_every = jit_split_barrier();
  constexpr size_t kBarrierOps = 30;  // operands per compiler barrier
  constexpr size_t kEntryRegs = 48;  // site-entry registers a kernel may hold
  struct Group {
    uint32_t site0, count, ord0, slots;
  };
  std::vector<std::array<uint32_t, 4>> sw;  // site table words (w0..w3) per site
  std::vector<Group> groups;
  std::vector<size_t> range_groups;      // first group of each range
  std::vector<const Instr *> sites;
  std::vector<uint32_t> site_slot;  // first slot of each site (register mode)
  uint32_t max_slots = 0;
  for (const Range &rg : ranges) {
    range_groups.push_back(groups.size());
    for (size_t i = rg.b; i < rg.e; ++i) {
      const Instr &in = code[i];
      if ((in.op & 0xFFu) != I_NOISEG) continue;
      const uint32_t G = in.op >> 8;
      for (uint32_t k = 0; k < G;) {
        Group g{(uint32_t)sites.size(), 0u, in.a + k, 0u};
        while (k < G) {
          const Instr &s = code[i + 1 + k];
          const uint32_t np = s.b ? 4u : 2u;
          const uint32_t spl[4] = {s.a & 0xFFFFu, s.a >> 16, s.b & 0xFFFFu, s.b >> 16};
          uint32_t need = regs ? np : 0u;
          if (split)
            for (uint32_t j = 0; j < np; ++j) need += sreg[spl[j]];
          if ((regs || split) && g.count > 0 && g.slots + need > kSlots) break;
          if (regs) {
            const uint32_t b0 = g.slots;
            sw.push_back({b0 | (b0 + 1u) << 16, s.b ? ((b0 + 2u) | (b0 + 3u) << 16) : 0u, s.d, s.e});
          } else if (split) {
            // LDS plane, or (register planes) the next slot after the LDS planes
            uint32_t ix[4] = {0u, 0u, 0u, 0u}, nx = nlds + g.slots;
            for (uint32_t j = 0; j < np; ++j) ix[j] = sreg[spl[j]] ? nx++ : lidx[spl[j]];
            sw.push_back({ix[0] | ix[1] << 16, s.b ? (ix[2] | ix[3] << 16) : 0u, s.d, s.e});
          } else {
            sw.push_back({s.a, s.b, s.d, s.e});
          }
          sites.push_back(&s);
          site_slot.push_back(g.slots);
          g.slots += need;
          ++g.count;
          ++k;
        }
        max_slots = std::max(max_slots, g.slots);
        if (g.count > 64u) return false;  // one site per lane
        groups.push_back(g);
      }
      i += G;
    }
  }
  range_groups.push_back(groups.size());
  // Frames beyond LDS: programs with many separate noise groups (each an
  // inlined event pass) compile for minutes -- noisy x_diag_circuit(1000):
  // 1,003 passes, 218 s of hiprtc -- while the interpreter needs none; the
  // memory experiments emit at most about 100
  if ((split || halfm) && groups.size() > kMaxLargeFrameGroups) return false;
  const uint32_t nsite = (uint32_t)sites.size();
  // Site entries held in registers for the whole kernel: lane i of a group's
  // registers holds its site i.  Planes pack into one register when every
  // plane index fits 8 bits; thresholds shared by a whole group become
  // constants.  Loaded once before the first tile, so no load in the tile
  // loop waits behind the record stores (loads and stores share one counter).
  bool p8 = true;
  for (const auto &w : sw)
    for (uint32_t v : {w[0] & 0xFFFFu, w[0] >> 16, w[1] & 0xFFFFu, w[1] >> 16})
      if (v > 0xFFu) p8 = false;
  struct GDesc {
    bool two, one, w2l, w3l;
    uint32_t w2c, w3c, ra, rb, r2, r3;
  };
  std::vector<GDesc> gd(groups.size());
  std::vector<std::array<uint32_t, 64>> ereg;  // entry registers, 64 lanes each
  auto new_reg = [&](auto word_of, const Group &g) {
    std::array<uint32_t, 64> r{};
    for (uint32_t l = 0; l < 64; ++l)
      r[l] = word_of(sw[g.site0 + std::min(l, g.count - 1u)]);
    ereg.push_back(r);
    return (uint32_t)ereg.size() - 1u;
  };
  for (size_t k = 0; k < groups.size(); ++k) {
    const Group &g = groups[k];
    GDesc d{true, true, false, false, 0u, 0u, kNone, kNone, kNone, kNone};
    bool w2set = false;
    for (uint32_t j = 0; j < g.count; ++j) {
      const auto &w = sw[g.site0 + j];
      (w[1] ? d.one : d.two) = false;
      if (!w[1]) {  // thresholds matter for one-qubit sites only
        if (!w2set) d.w2c = w[2], w2set = true;
        else if (w[2] != d.w2c) d.w2l = true;
      }
      if (j == 0) d.w3c = w[3];
      else if (w[3] != d.w3c) d.w3l = true;
    }
    if (p8) {
      d.ra = new_reg([](const std::array<uint32_t, 4> &w) {
        return (w[0] & 0xFFu) | ((w[0] >> 16) & 0xFFu) << 8 | (w[1] & 0xFFu) << 16 |
               ((w[1] >> 16) & 0xFFu) << 24;
      }, g);
    } else {
      d.ra = new_reg([](const std::array<uint32_t, 4> &w) { return w[0]; }, g);
      d.rb = new_reg([](const std::array<uint32_t, 4> &w) { return w[1]; }, g);
    }
    if (d.w2l) d.r2 = new_reg([](const std::array<uint32_t, 4> &w) { return w[2]; }, g);
    if (d.w3l) d.r3 = new_reg([](const std::array<uint32_t, 4> &w) { return w[3]; }, g);
    gd[k] = d;
  }
  const bool preload = !groups.empty() && ereg.size() <= kEntryRegs;
  std::ostringstream tab;
  if (preload) {
    for (const auto &r : ereg)
      for (uint32_t l = 0; l < 64; ++l) tab << r[l] << "u,";
  } else {
    for (const auto &w : sw) tab << w[0] << "u," << w[1] << "u," << w[2] << "u," << w[3] << "u,\n";
  }
  const uint32_t ntab = preload ? (uint32_t)ereg.size() * 64u : 4u * nsite;
  auto gaddr = [](uint32_t p) { return "F + (" + std::to_string(p * 64u) + "u + l)"; };
  // plane access: read expression, write and XOR statements
  auto PL = [&](uint32_t p) -> std::string {
    if (in_reg(p)) return "r" + std::to_string(p);
    if (glob) return "gld(" + gaddr(p) + ")";
    return plane(lp_of(p));
  };
  auto WR = [&](uint32_t p, const std::string &v) -> std::string {
    if (glob) return "gst(" + gaddr(p) + ", " + v + ");";
    return PL(p) + " = " + v + ";";
  };
  auto XOR = [&](uint32_t p, const std::string &v) -> std::string {
    if (in_reg(p)) return PL(p) + " ^= " + v + ";";
    if (glob) return "gxr(" + gaddr(p) + ", " + v + ");";
    return "atomicXor(&" + plane(lp_of(p)) + ", " + v + ");";
  };
  auto slot = [&](uint32_t k) { return plane(slot_base + k); };
  // Read hoisting (LDS-resident planes): hz maps a plane to the register that
  // holds the value read at the start of the current run of instructions.
  // Not in split mode: its registers already hold the most-used planes, and
  // up to 64 more live values there cost more than the overlapped reads gain
  // (the d=11 benchmark circuit, 2^19 shots: 0.240 ms without, 0.332 with).
  const uint32_t hoist_cap = (glob || split) ? 0u : jit_hoist();
  std::unordered_map<uint32_t, std::string> hz;
  auto RD = [&](uint32_t p) -> std::string {
    const auto it = hz.find(p);
    return it != hz.end() ? it->second : PL(p);
  };
  // planes an instruction reads and writes; false for instructions that end
  // a run (noise groups)
  auto plane_rw = [](const Instr &in, std::vector<uint32_t> *rd, std::vector<uint32_t> *wr) {
    auto R = [&](uint32_t p) { if (p != kNone) rd->push_back(p); };
    auto W = [&](uint32_t p) { if (p != kNone) wr->push_back(p); };
    switch (in.op & 0xFFu) {
      case I_XORN: {
        const uint32_t *w = &in.a;
        for (uint32_t k = 0; k < (in.op >> 8); ++k) R(w[k] >> 16), W(w[k] & 0xFFFFu);
        return true;
      }
      case I_MR: R(in.a), W(in.a), W(in.b), W(in.e); return true;
      case I_MRX: R(in.a), R(in.f), W(in.a), W(in.b), W(in.e); return true;
      case I_MEAS: R(in.a), W(in.c); return true;
      case I_XSTORE2: R(in.a), R(in.b), W(in.d); return true;
      case I_RESETZ:
      case I_INIT: W(in.a), W(in.b); return true;
      case I_ACC_LDS:
      case I_ACCX_LDS:
      case I_STORE: R(in.a); return true;
      case I_ACC_STORE: W(in.b); return true;
      case I_LIN1: R(in.a), R(in.b), W(in.a), W(in.b); return true;
      case I_ZERO:
      case I_LOAD: W(in.a); return true;
      case I_REC_ZERO:
      case I_ACC_REC:
      case I_ACCX_REC: return true;
      default: return false;
    }
  };
  // The run starting at code[i] (ending before `end`): the longest stretch of
  // hoistable instructions in which no LDS plane is read after a write of it,
  // within hoist_cap planes.  Emits the reads and fills hz; returns the end.
  auto open_run = [&](size_t i, size_t end, std::ostringstream &body) -> size_t {
    std::vector<uint32_t> order, written, rd, wr;
    size_t j = i;
    for (; j < end; ++j) {
      rd.clear();
      wr.clear();
      if (!plane_rw(code[j], &rd, &wr)) break;
      bool ok = true;
      std::vector<uint32_t> add;
      for (uint32_t p : rd) {
        if (p >= P->nplanes || in_reg(p)) continue;
        if (std::find(written.begin(), written.end(), p) != written.end()) ok = false;
        else if (std::find(order.begin(), order.end(), p) == order.end() &&
                 std::find(add.begin(), add.end(), p) == add.end())
          add.push_back(p);
      }
      if (!ok || order.size() + add.size() > hoist_cap) break;
      order.insert(order.end(), add.begin(), add.end());
      for (uint32_t p : wr)
        if (p < P->nplanes && !in_reg(p)) written.push_back(p);
    }
    if (j == i || order.empty()) return i;
    body << "    {";
    for (uint32_t p : order) {
      const std::string h = "h" + std::to_string(p);
      body << " const uint32_t " << h << " = " << PL(p) << ";";
      hz[p] = h;
    }
    body << "\n";
    return j;
  };
  // event passes over register-held site entries: template arguments, call
  // arguments, and the permute of the first batch (prefetch_r)
  const bool prefetch_on = jit_prefetch();
  auto q = [](uint32_t r) { return "q" + std::to_string(r); };
  auto pass_tmpl = [&](size_t gi) {
    const GDesc &d = gd[gi];
    std::ostringstream o;
    o << "<" << d.two << ", " << d.one << ", " << d.w2l << ", " << d.w3l << ", " << p8 << ">";
    return o.str();
  };
  auto pass_args = [&](size_t gi, bool in_loop) {
    const GDesc &d = gd[gi];
    std::ostringstream o;
    o << "ev, " << groups[gi].ord0 << (in_loop ? "u + ob, " : "u, ") << groups[gi].count << "u, "
      << q(d.ra) << ", " << (p8 ? "0u" : q(d.rb)) << ", "
      << (d.w2l ? q(d.r2) : std::to_string(d.w2c) + "u") << ", "
      << (d.w3l ? q(d.r3) : std::to_string(d.w3c) + "u");
    return o.str();
  };
  auto emit_prefetch = [&](size_t gi, bool in_loop, std::ostringstream &body) {
    body << "    const qcfd::PrefR<" << p8 << "> pf" << gi << " = qcfd::prefetch_r" << pass_tmpl(gi)
         << "(" << pass_args(gi, in_loop) << ");\n";
  };
  auto load_group = [&](size_t g, const char *pfx) {
    std::ostringstream o;
    const uint32_t off = groups[g].site0, G = groups[g].count;
    o << "    { const uint32_t *q = QS + 4u * (" << off << "u + (l < " << G << "u ? l : " << G - 1
      << "u));\n      " << pfx << "0 = q[0]; " << pfx << "1 = q[1]; " << pfx << "2 = q[2]; "
      << pfx << "3 = q[3]; }\n";
    return o.str();
  };
  auto emit = [&](bool tail, std::ostringstream &body) -> bool {
  if (regs || hyb || split || halfm) {
    const char *sep = "    uint32_t ";
    for (uint32_t p = 0; p < P->nplanes; ++p)
      if (in_reg(p)) {
        body << sep << "r" << p << " = 0u";
        sep = ", ";
      }
    if (*sep == ',') body << ";\n";
  }
  if (!groups.empty() && !preload) body << "    uint32_t t0 = 0u, t1 = 0u, t2 = 0u, t3 = 0u;\n";
  for (size_t ri = 0; ri < ranges.size(); ++ri) {
    const Range &rg = ranges[ri];
    if (rg.b == rg.e) continue;
    const bool in_loop = rg.loop;
    auto at = [&](const char *arr, uint32_t col, bool shifted) {
      if (rec32 && std::string(arr) == "P.rec")
        return "*(uint32_t *)(recb + (uint32_t)((" + std::to_string(col) +
               (shifted ? "u + cb" : "u") + ") * ld4 + gw4))";
      return std::string(arr) + "[(uint64_t)(" + std::to_string(col) +
             (shifted ? "u + cb" : "u") + ") * ld + gw]";
    };
    // the partial tile: lanes past the record read zeros and store nothing,
    // bits past the last shot are stored as zeros
    auto st_rec = [&](uint32_t col, const std::string &v) {
      const std::string m = at("P.rec", col, in_loop);
      const std::string val = tail ? "(" + v + ") & lm" : v;
      const std::string st = nt ? "__builtin_nontemporal_store(" + val + ", &" + m + ");"
                                : m + " = " + val + ";";
      return tail ? "if (live) " + st : st;
    };
    auto ld_rec = [&](uint32_t col) {
      const std::string m = at("P.rec", col, in_loop);
      return tail ? "(live ? " + m + " : 0u)" : m;
    };
    if (in_loop)
      body << "    #pragma nounroll\n    for (uint32_t it = 0; it < " << lp.R
           << "u; ++it) {\n    const uint32_t cb = it * " << (uint32_t)lp.dcol
           << "u, ob = it * " << (uint32_t)lp.dord << "u;\n    (void)cb; (void)ob;\n";
    size_t gi = range_groups[ri];
    const size_t gend = range_groups[ri + 1];
    if (gi < gend && !preload) body << load_group(gi, "t");
    bool prefetched = false;  // the pass of group gi has its prefetch emitted
    if (gi < gend && preload && prefetch_on) {
      emit_prefetch(gi, in_loop, body);
      prefetched = true;
    }
    bool lh_open = false;  // half mode: inside a run guarded by "lanes 0..31"
    size_t run_end = rg.b;  // end of the current read-hoisting run
    bool run_open = false;
    for (size_t i = rg.b; i < rg.e; ++i) {
      const Instr &in = code[i];
      const uint32_t op = in.op & 0xFFu;
      const uint32_t a = in.a, b = in.b, c = in.c;
      // half mode: only lanes 0..31 hold planes; the event passes (noise
      // groups) need every lane of the wave
      if (halfm && (op == I_NOISEG) == lh_open) {
        body << (lh_open ? "    }\n" : "    if (lh) {\n");
        lh_open = !lh_open;
      }
      // split mode: every so many operations all register planes pass a
      // compiler barrier, so the scheduler cannot stretch live ranges (early
      // LDS reads, reassociated XOR chains) over a whole straight-line stretch
      if (split && split_every && i > rg.b && (i - rg.b) % split_every == 0) {
        std::vector<uint32_t> rp;
        if (jit_split_touched()) {
          std::vector<uint32_t> pl;
          for (size_t j = i - split_every; j < i; ++j) instr_planes(code, j, &pl);
          for (uint32_t p : pl)
            if (p < P->nplanes && sreg[p] && std::find(rp.begin(), rp.end(), p) == rp.end())
              rp.push_back(p);
        } else {
          for (uint32_t p = 0; p < P->nplanes; ++p)
            if (sreg[p]) rp.push_back(p);
        }
        for (size_t k0 = 0; k0 < rp.size(); k0 += kBarrierOps) {
          body << "    asm volatile(\"\" :";
          for (size_t k = k0; k < std::min(rp.size(), k0 + kBarrierOps); ++k)
            body << (k > k0 ? ", " : " ") << "\"+v\"(r" << rp[k] << ")";
          body << ");\n";
        }
      }
      if (hoist_cap && !run_open) {
        run_end = open_run(i, rg.e, body);
        run_open = run_end > i;
      }
      switch (op) {
        case I_XORN: {
          const uint32_t n = in.op >> 8;
          const uint32_t *w = &in.a;
          body << "    {";
          for (uint32_t k = 0; k < n; ++k)
            body << " const uint32_t v" << k << " = " << RD(w[k] >> 16) << ";";
          for (uint32_t k = 0; k < n; ++k)
            body << " " << XOR(w[k] & 0xFFFFu, "v" + std::to_string(k));
          body << " }\n";
          break;
        }
        case I_NOISEG: {
          const uint32_t G = in.op >> 8;
          for (uint32_t done = 0; done < G; ++gi) {
            const Group &g = groups[gi];
            body << "    {\n";
            if (pass_prio) body << "    __builtin_amdgcn_s_setprio(" << pass_prio << ");\n";
            if (preload) {
              if (!prefetched) emit_prefetch(gi, in_loop, body);
              body << "    if (P.has_class) qcfd::group_events_r" << pass_tmpl(gi) << "("
                   << pass_args(gi, in_loop) << ", P.c, " << thin << ", pf" << gi << ", flip);\n";
              prefetched = false;
            } else {
              if (gi + 1 < gend)
                body << "    uint32_t n0, n1, n2, n3;\n" << load_group(gi + 1, "n");
              body << "    if (P.has_class) qcfd::group_events(ev, " << g.ord0
                   << (in_loop ? "u + ob, " : "u, ") << g.count
                   << "u, 0u, t0, t1, t2, t3, 0u, P.c, " << thin << ", flip);\n";
              if (gi + 1 < gend) body << "    t0 = n0; t1 = n1; t2 = n2; t3 = n3;\n";
            }
            if (regs || split) {
              std::vector<uint32_t> touched;
              for (uint32_t k = 0; k < g.count; ++k) {
                const Instr &s = *sites[g.site0 + k];
                uint32_t s0 = site_slot[g.site0 + k];
                const uint32_t pl[4] = {s.a & 0xFFFFu, s.a >> 16, s.b & 0xFFFFu, s.b >> 16};
                const uint32_t np = s.b ? 4u : 2u;
                body << "   ";
                for (uint32_t j = 0; j < np; ++j) {
                  if (!in_reg(pl[j])) continue;  // split mode: events went straight to LDS
                  body << " " << PL(pl[j]) << " ^= atomicExch(&" << slot(s0++) << ", 0u);";
                  if (std::find(touched.begin(), touched.end(), pl[j]) == touched.end())
                    touched.push_back(pl[j]);
                }
                body << "\n";
              }
              // one barrier per pass keeps the compiler from reassociating XOR
              // chains across the noise (which stretches live ranges past the
              // register file) while leaving the exchanges free to overlap
              for (size_t k0 = 0; k0 < touched.size(); k0 += kBarrierOps) {
                body << "    asm volatile(\"\" :";
                for (size_t k = k0; k < std::min(touched.size(), k0 + kBarrierOps); ++k)
                  body << (k > k0 ? ", " : " ") << "\"+v\"(" << PL(touched[k]) << ")";
                body << ");\n";
              }
            }
            if (pass_prio) body << "    __builtin_amdgcn_s_setprio(0);\n";
            body << "    }\n";
            done += g.count;
            // the next pass's first permute goes here, ahead of the gates
            // between the two passes
            if (preload && prefetch_on && gi + 1 < gend) {
              emit_prefetch(gi + 1, in_loop, body);
              prefetched = true;
            }
          }
          i += G;
          break;
        }
        case I_MR:
        case I_MRX: {
          const uint32_t col = in.d, keep = in.e;
          body << "    { const uint32_t v = " << RD(a) << ";";
          if (col != kNone) body << " " << st_rec(col, "v");
          if (op == I_MRX) body << " " << st_rec(in.g, "v ^ " + RD(in.f));
          if (keep != kNone) body << " " << WR(keep, "v");
          body << " " << WR(a, "0u");
          if (!(in.op & kNoGauge)) body << " " << XOR(b, "g.next()");
          body << " }\n";
          break;
        }
        case I_XSTORE2:
          body << "    { const uint32_t v = " << RD(a) << " ^ " << RD(b) << "; "
               << st_rec(c, "v");
          if (in.d != kNone) body << " " << WR(in.d, "v");
          body << " }\n";
          break;
        case I_MEAS:
          body << "    { const uint32_t v = " << RD(a) << "; " << st_rec(b, "v");
          if (c != kNone) body << " " << WR(c, "v");
          body << " }\n";
          break;
        case I_RESETZ:
          body << "    " << WR(a, "0u");
          if (!(in.op & kNoGauge)) body << " " << XOR(b, "g.next()");
          body << "\n";
          break;
        case I_ACC_LDS:
          body << "    acc = " << RD(a) << ";\n";
          break;
        case I_ACCX_LDS:
          body << "    acc ^= " << RD(a) << ";\n";
          break;
        case I_ACC_REC:
          body << "    acc = " << ld_rec(a) << ";\n";
          break;
        case I_ACCX_REC:
          body << "    acc ^= " << ld_rec(a) << ";\n";
          break;
        case I_ACC_STORE:
          body << "    " << st_rec(a, "acc");
          if (b != kNone) body << " " << WR(b, "acc");
          body << "\n";
          break;
        case I_REC_ZERO:
          body << "    " << st_rec(a, "0u") << "\n";
          break;
        case I_INIT:
          body << "    " << WR(a, "0u") << " "
               << WR(b, (in.op & kNoGauge) ? "0u" : "g.next()") << "\n";
          break;
        case I_LOAD: {
          // a run of frame loads (the in-place prologue): up to 32 requested
          // before any of them is stored, so that their latencies overlap
          // (left to itself the scheduler waited after every few, 25 round
          // trips per tile at d=5)
          size_t j = i;
          while (j < rg.e && (code[j].op & 0xFFu) == I_LOAD && j - i < 32 &&
                 (!run_open || j < run_end))
            ++j;
          body << "    {\n";
          for (size_t k = i; k < j; ++k) {
            const std::string m = at("P.frame", code[k].b, false);
            body << "    const uint32_t f" << k - i << " = "
                 << (tail ? "live ? " + m + " : 0u" : m) << ";\n";
          }
          body << "    __builtin_amdgcn_sched_barrier(0);\n";
          for (size_t k = i; k < j; ++k)
            body << "    " << WR(code[k].a, "f" + std::to_string(k - i)) << "\n";
          body << "    }\n";
          i = j - 1;
          break;
        }
        case I_STORE:
          // the first store of the epilogue makes the plane stride opaque
          // again: otherwise the compiler keeps the frame addresses of the
          // in-place loads live through the whole tile for these stores
          // (d=5 x 25 in place: 302 VGPRs and four waves per CU, not six)
          if (i == rg.b || (code[i - 1].op & 0xFFu) != I_STORE)
            body << "    asm volatile(\"\" : \"+s\"(ld));\n";
          body << "    " << (tail ? "if (live) " : "") << at("P.frame", b, false) << " = " << RD(a)
               << (tail ? " & lm;\n" : ";\n");
          break;
        case I_ZERO:
          body << "    " << WR(a, "0u") << "\n";
          break;
        case I_LIN1:
          body << "    { const uint32_t x = " << RD(a) << ", z = " << RD(b) << "; "
               << WR(a, std::string((c & 1u) ? "x" : "0u") + " ^ " + ((c & 2u) ? "z" : "0u"))
               << " "
               << WR(b, std::string((c & 4u) ? "x" : "0u") + " ^ " + ((c & 8u) ? "z" : "0u"))
               << " }\n";
          break;
        default:
          return false;
      }
      if (run_open && i + 1 >= run_end) {
        body << "    }\n";
        hz.clear();
        run_open = false;
      }
    }
    if (lh_open) body << "    }\n";
    if (in_loop) body << "    }\n";
  }
    return true;
  };
  std::ostringstream full, part;
  if ((!halfm && !emit(false, full)) || !emit(true, part)) return false;
  std::ostringstream src;
  src << "#define QCF_RTC 1\n";
  // Machine-scheduler strategy by storage mode: LDS-mode and part-tile
  // kernels are scheduled for instruction-level parallelism (max-ilp; d=5 x
  // 25, p = 1e-3, 10^7 shots: 0.393 -> 0.383 ms; half tiles, rotated d=15,
  // 2e6 shots: 5.36 -> 5.25 ms), two-wave split kernels for low register
  // pressure (iterative-minreg; rotated d=9, p = 1e-3: 1.88 -> 1.74 ms).
  // Register and one-wave split kernels keep the default, which measured
  // about as fast or faster there than max-ilp, max-memory-clause,
  // iterative-minreg and iterative-maxocc.
  if (jit_sched_tags() && (lds_mode || halfm))
    src << kCompileOptTag << "-mllvm -amdgpu-sched-strategy=max-ilp\n";
  if (jit_sched_tags() && split && (variant & 32))
    src << kCompileOptTag << "-mllvm -amdgpu-sched-strategy=iterative-minreg\n";
  src << "// folded: prologue " << lp.s << ", loop " << lp.R << " x " << lp.L << ", of "
      << code.size() << " instructions\n";
  src << kDeviceSrc << kKernelPrelude;
  src << "__constant__ uint32_t QS[" << std::max<uint32_t>(4, ntab) << "] = {\n"
      << (nsite ? tab.str() : std::string("0u,0u,0u,0u")) << "};\n";
  src << "extern \"C\" __global__ void __launch_bounds__(64) qcf_jit_kernel(KJ P0) {\n"
      << "  extern __shared__ uint32_t L[];\n"
      << "  const uint32_t l = threadIdx.x;\n";
  // part-tile kernels of a single launch may spread the parts of a tile over
  // pset workgroups (workgroup i takes part i % pset), so that launches of a
  // few tiles use more than one CU per tile
  if (halfm)
    src << "  const uint32_t pset = P0.pset > 1u && !P0.ncirc && !P0.steal ? P0.pset : 1u;\n"
           "  const uint32_t hsel = blockIdx.x % pset;\n"
           "  const uint32_t bx = blockIdx.x / pset, gx = gridDim.x / pset;\n";
  else
    src << "  const uint32_t bx = blockIdx.x, gx = gridDim.x;\n";
  {
    const uint32_t nl = glob    ? 0u
                        : split ? nlds + max_slots
                        : regs  ? max_slots
                        : hyb   ? qpl
                        : halfm ? part_lds
                                : P->nplanes;
    for (uint32_t w = 1; w * per_window < nl; ++w)
      src << "  uint32_t lb" << w << " = " << w * 16384u << "u + l;\n  asm volatile(\"\" : \"+v\"(lb"
          << w << "));\n";
  }
  src
      << (glob ? "  uint32_t *const F = P0.scratch + (uint64_t)blockIdx.x * P0.sstride;\n"
                 "  (void)L;\n  const FlipG flip{F};\n"
               : halfm ? "" : "  const FlipLDS flip{L};\n");
  if (preload) {
    // entry registers, made opaque so that they stay in registers
    for (size_t r = 0; r < ereg.size(); ++r)
      src << "  uint32_t q" << r << " = QS[" << r * 64 << "u + l];\n";
    for (size_t r0 = 0; r0 < ereg.size(); r0 += kBarrierOps) {
      src << "  asm volatile(\"\" :";
      for (size_t r = r0; r < std::min(ereg.size(), r0 + kBarrierOps); ++r)
        src << (r > r0 ? ", " : " ") << "\"+v\"(q" << r << ")";
      src << ");\n";
    }
  }
  if (regs || (split && max_slots))  // the slot buffer starts at zero and every pass leaves it at zero
    src << "  for (uint32_t k = " << slot_base << "u; k < " << slot_base + std::max<uint32_t>(max_slots, 1u)
        << "u; ++k) L[k * 64u + l] = 0u;\n";
  // Tiles are dealt out per XCD: with X = P0.xcds > 1 the grid is a multiple of
  // X, workgroup i runs on XCD i mod X, and XCD x takes the x-th contiguous
  // run of tiles, so each XCD's L2 writes whole contiguous stretches of every
  // record column.
  src << kTileLoopHead;
  // part-tile mode: 64/pw passes over the tile, each with the shot words of
  // one part in lanes 0..pw-1 (the event stream is drawn for the whole tile
  // in every pass); passes past the plane stride hold no shots and end it
  if (halfm)
    src << "    for (uint32_t hf = hsel; hf < " << 64u / pw << "u; hf += pset) {\n"
        << "    if (blk * 64u + hf * " << pw << "u >= P.nwbuf) break;\n"
        << "    const bool lh = l < " << pw << "u;\n"
        << "    const uint32_t gw = blk * 64u + hf * " << pw << "u + (l & " << pw - 1 << "u);\n"
        << "    const FlipPart<" << pw << "> flip{L, hf};\n";
  else
    src << "    const uint32_t gw = blk * 64u + l;\n";
  src << kTileLocals;
  src << "    qcfd::EvState ev = qcfd::ev_init(P.seed, blk + P.tile0, l);\n"
         "    if (P.has_class) ev = qcfd::gen_batch(ev, P.c, " << thin << ");\n"
         "    uint32_t acc = 0u;\n    (void)acc;\n";
  if (halfm)
    src << "    const bool live = lh && gw < P.nwbuf;\n"
        << "    const uint32_t lm = gw >= P.nw32 ? 0u : (gw == P.nw32 - 1u ? P.lastmask : "
           "0xFFFFFFFFu);\n"
        << part.str() << "    }\n";
  else
    src << "    if (blk < P.nfull) {  // every lane's 32 shots exist\n" << full.str()
        << "    } else {\n"
        << "    const bool live = gw < P.nwbuf;\n"
        << "    const uint32_t lm = gw >= P.nw32 ? 0u : (gw == P.nw32 - 1u ? P.lastmask : "
           "0xFFFFFFFFu);\n"
        << part.str() << "    }\n";
  src << kTileLoopTail;
  *out = src.str();
  if (lds_bytes)
    *lds_bytes = glob    ? 0u
                 : halfm ? part_lds * pw * 4u
                 : split ? (nlds + max_slots) * 256u
                 : regs  ? std::max<uint32_t>(max_slots, 1u) * 256u
                 : hyb  ? qpl * 256u
                        : P->nplanes * 256u;
  return true;
}

int compile(const std::string &src, std::vector<char> *code, std::string *log) {
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "qcf_jit.hip", 0, nullptr, nullptr) !=
      HIPRTC_SUCCESS)
    return set_error(QCF_ERR_HIP, "hiprtcCreateProgram failed");
  std::vector<const char *> opts = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang",
                                    "-target-feature", "-Xclang", "-packed-fp32-ops"};
  // options the generator asked for (lines of the source starting with the
  // tag), then QCF_JIT_OPTS: extra compiler options separated by spaces
  std::vector<std::string> extra;
  for (size_t at = src.find(kCompileOptTag); at != std::string::npos;
       at = src.find(kCompileOptTag, at + 1)) {
    const size_t b = at + strlen(kCompileOptTag), e = src.find('\n', b);
    std::istringstream is(src.substr(b, e == std::string::npos ? std::string::npos : e - b));
    for (std::string w; is >> w;) extra.push_back(w);
  }
  if (const char *e = getenv("QCF_JIT_OPTS")) {
    std::istringstream is(e);
    for (std::string w; is >> w;) extra.push_back(w);
  }
  for (const std::string &w : extra) opts.push_back(w.c_str());
  const hiprtcResult r = hiprtcCompileProgram(prog, (int)opts.size(), opts.data());
  size_t n = 0;
  hiprtcGetProgramLogSize(prog, &n);
  if (n > 1) {
    log->resize(n);
    hiprtcGetProgramLog(prog, &(*log)[0]);
  }
  if (r != HIPRTC_SUCCESS) {
    hiprtcDestroyProgram(&prog);
    return set_error(QCF_ERR_HIP, "hiprtc compile failed: %.900s", log->c_str());
  }
  hiprtcGetCodeSize(prog, &n);
  code->resize(n);
  hiprtcGetCode(prog, code->data());
  hiprtcDestroyProgram(&prog);
  return QCF_OK;
}

namespace {
std::mutex g_jit_mu;
constexpr size_t kCodeCacheMax = 32;
std::unordered_map<std::string, std::vector<char>> g_code_cache;  // source -> code object
}  // namespace

// Planes go to registers only with at least this many expected noise events
// per 2048-shot tile (QCF_JIT_REG_EVENTS, default 60,000).  d=5 x 25 (3874
// sites), 10^7 shots, LDS against register mode: p = 2e-3 (15.9 k events per
// tile) 0.50 / 0.68 ms, 3e-3 (23.8 k) 0.61 / 0.77, 5e-3 (39.7 k) 0.84 /
// 0.93, 1e-2 (79.3 k) 1.43 / 1.35 -- since the LDS kernels raise the
// priority of event passes and are scheduled for ILP (the threshold was
// 20,000 before).
double jit_reg_events() {
  const char *e = getenv("QCF_JIT_REG_EVENTS");
  return e ? strtod(e, nullptr) : 60000.0;
}

// Noise slots (256 bytes of LDS each) one event pass may use in register
// mode (QCF_JIT_SLOTS, 16..128; default 80: 20 KB, so LDS still lets eight
// waves share a CU, and a layer of 20 two-qubit noise sites is one pass).
uint32_t jit_slots() {
  const char *e = getenv("QCF_JIT_SLOTS");
  const uint32_t v = e ? (uint32_t)strtoul(e, nullptr, 10) : 80u;
  return std::min<uint32_t>(128u, std::max<uint32_t>(16u, v));
}

// Programs with at most this many planes keep them in registers.
uint32_t jit_reg_planes() {
  const char *e = getenv("QCF_JIT_REG_PLANES");
  return e ? (uint32_t)strtoul(e, nullptr, 10) : 192u;
}

// Programs longer than this many device instructions stay on the interpreter:
// compile time grows faster than linearly with the straight-line code.
size_t jit_max_instr() {
  const char *e = getenv("QCF_JIT_MAX_INSTR");
  return e ? (size_t)strtoull(e, nullptr, 10) : (size_t)12000;
}

// Split mode (frames larger than LDS): planes kept in registers, at least
// what the LDS budget needs (QCF_JIT_SPLIT_REGS raises it).
uint32_t jit_split_regs() {
  const char *e = getenv("QCF_JIT_SPLIT_REGS");
  return e ? (uint32_t)strtoul(e, nullptr, 10) : 0u;
}

// Split mode: operations between compiler barriers over the register planes
// (QCF_JIT_SPLIT_BARRIER, 0 = none; the d=11 reference circuit at 2^19 shots
// took 0.30 ms at 16 and 64, 0.33 ms at 32, 0.38 ms at 128).
size_t jit_split_barrier() {
  const char *e = getenv("QCF_JIT_SPLIT_BARRIER");
  return e ? (size_t)strtoull(e, nullptr, 10) : (size_t)16;
}

// A split-mode barrier names only the register planes of the operations
// since the last one (QCF_JIT_SPLIT_TOUCHED=0: every register plane; d=13 x 2
// rounds: 15 s instead of 119 s of compilation, and no scratch spills).
bool jit_split_touched() {
  const char *e = getenv("QCF_JIT_SPLIT_TOUCHED");
  return !(e && e[0] == '0');
}

// Two-wave split mode (half of a CU's LDS per wave) only with at most this
// many register planes (QCF_JIT_TWO_WAVE_REGS, default 256): beyond the
// architectural VGPRs the planes spill into the accumulation registers and
// the second wave costs more than it hides.  10^7 shots, frame returned:
// rotated d=9 (125 / 205 register planes at p = 0 / 1e-3) 0.80 / 2.18 ms
// two-wave against 1.43 / 3.30 one-wave; d=11 at p = 0 (345 planes) 5.89 ms
// two-wave against 3.70 one-wave.
uint32_t jit_two_wave_regs() {
  const char *e = getenv("QCF_JIT_TWO_WAVE_REGS");
  return e ? (uint32_t)strtoul(e, nullptr, 10) : 256u;
}

// QCF_JIT_SCHED=0 compiles every kernel with the default machine scheduler
// (instead of the strategy chosen per storage mode).
bool jit_sched_tags() {
  const char *e = getenv("QCF_JIT_SCHED");
  return !(e && e[0] == '0');
}

// Issue priority of a wave during its event passes (s_setprio), -1 for the
// default: QCF_JIT_PASS_PRIO=0 keeps every wave at priority 0, 1..3 raises
// passes to that level in every storage mode.
int jit_pass_prio() {
  const char *e = getenv("QCF_JIT_PASS_PRIO");
  return e && e[0] >= '0' && e[0] <= '3' ? e[0] - '0' : -1;
}

// Planes a part-tile kernel may keep in registers (QCF_JIT_PART_REGS,
// default 128; 0 keeps every plane in LDS).  Rotated d=15 (337 noise-free
// classical slots), half tiles, 2e6 shots at p=1e-3 / 1e6 at p=1e-4:
// 5.24 / 1.72 ms with none, 5.21 / 1.70 with 32, 5.16 / 1.68 with 128,
// 5.18-5.21 / 1.70-1.71 with 192, 256 or all.
uint32_t jit_part_regs() {
  const char *e = getenv("QCF_JIT_PART_REGS");
  return e ? (uint32_t)strtoul(e, nullptr, 10) : 128u;
}

// QCF_JIT_SPLIT=0 leaves frames larger than LDS to the global-memory paths.
bool jit_split_enabled() {
  const char *e = getenv("QCF_JIT_SPLIT");
  return !(e && e[0] == '0');
}

// QCF_JIT_GLOBAL=1 also specialises programs whose frames live in global memory.
bool jit_global_enabled() {
  const char *e = getenv("QCF_JIT_GLOBAL");
  return e && e[0] == '1';
}

// QCF_JIT_NT=1 writes the record with nontemporal stores.
bool jit_nt_stores() {
  const char *e = getenv("QCF_JIT_NT");
  return e && e[0] == '1';
}

// QCF_JIT_FULL_GRID=1 launches one workgroup per tile (the dispatcher then
// balances tiles over CUs) instead of a grid of resident waves.
bool jit_full_grid() {
  const char *e = getenv("QCF_JIT_FULL_GRID");
  return e && e[0] == '1';
}

// XCDs the grid is dealt over (QCF_XCDS, default 8; 1 = plain grid stride).
uint32_t jit_xcds() {
  const char *e = getenv("QCF_XCDS");
  return e ? (uint32_t)strtoul(e, nullptr, 10) : 8u;
}

// QCF_JIT_NO_LOOPS=1 emits the whole stream straight-line.
bool jit_no_loops() {
  const char *e = getenv("QCF_JIT_NO_LOOPS");
  return e && e[0] == '1';
}

bool jit_disabled() {
  const char *e = getenv("QCF_NO_JIT");
  return e && e[0] == '1';
}

// LDS-resident planes read by a run of gates and measurements go first, all
// of them before the run's first write, as long as no read follows a write
// of the same plane inside the run: one LDS round trip per run instead of
// one per 7-XOR bundle and per measurement.  QCF_JIT_HOIST caps the planes
// one run may hold in registers (0 = off; default 64).
uint32_t jit_hoist() {
  const char *e = getenv("QCF_JIT_HOIST");
  return e ? (uint32_t)strtoul(e, nullptr, 10) : 64u;
}

// QCF_JIT_LINEAR=0 runs noiseless record-only launches on the frame kernels
// instead of the linear one.
bool jit_linear_enabled() {
  const char *e = getenv("QCF_JIT_LINEAR");
  return !(e && e[0] == '0');
}

// QCF_JIT_PREFETCH=0 leaves the first site-entry permute of every noise pass
// at the pass instead of after the previous pass.
bool jit_prefetch() {
  const char *e = getenv("QCF_JIT_PREFETCH");
  return !(e && e[0] == '0');
}

// The values of every knob generate() reads: a kernel compiled under other
// values is not reused.
std::string jit_knob_signature() {
  static const char *const kKnobs[] = {
      "QCF_JIT_REG_EVENTS", "QCF_JIT_SLOTS", "QCF_JIT_REG_PLANES", "QCF_JIT_MAX_INSTR",
      "QCF_JIT_SPLIT_REGS", "QCF_JIT_SPLIT_BARRIER", "QCF_JIT_SPLIT_TOUCHED", "QCF_JIT_NT",
      "QCF_JIT_NO_LOOPS", "QCF_JIT_HOIST", "QCF_JIT_PREFETCH", "QCF_JIT_TWO_WAVE_REGS", "QCF_JIT_PASS_PRIO", "QCF_JIT_OPTS", "QCF_JIT_SCHED", "QCF_JIT_PART_REGS"};
  std::string s;
  for (const char *k : kKnobs) {
    const char *e = getenv(k);
    s += e ? e : "-";
    s += '|';
  }
  return s;
}

int jit_get(qcf_program *P, int variant, JitKernelInfo *info) {
  *info = JitKernelInfo{};
  int dev = 0;
  QCF_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_jit_mu);
  const std::string knobs = jit_knob_signature();
  for (auto &j : P->jit)
    if (j.device == dev && j.variant == variant && j.knobs == knobs) {
      if (j.failed) return set_error(QCF_ERR_UNSUPPORTED, "specialised kernel unavailable");
      info->function = j.function;
      info->lds_bytes = j.lds_bytes;
      info->blocks_per_cu = j.blocks_per_cu;
      info->key = j.key;
      return QCF_OK;
    }
  qcf_program::JitKernel jk;
  jk.device = dev;
  jk.variant = variant;
  jk.knobs = knobs;
  std::string src, log;
  std::vector<char> bin;
  int st = QCF_OK;
  if (emitted_size(P, variant) > jit_max_instr())
    st = set_error(QCF_ERR_UNSUPPORTED, "program longer than QCF_JIT_MAX_INSTR instructions");
  else if (!generate(P, variant, &src, &jk.lds_bytes))
    st = set_error(QCF_ERR_UNSUPPORTED, "op not specialisable");
  if (st == QCF_OK) {
    jk.key = std::hash<std::string>{}(src);
    // programs with the same source (the same circuit built twice) share one
    // compilation
    // (keyed with the extra compiler options, which change the code too)
    const char *xo = getenv("QCF_JIT_OPTS");
    const std::string ckey = src + "\n//" + (xo ? xo : "");
    auto it = g_code_cache.find(ckey);
    if (it != g_code_cache.end()) {
      bin = it->second;
    } else {
      st = compile(src, &bin, &log);
      if (st == QCF_OK) {
        if (g_code_cache.size() >= kCodeCacheMax) g_code_cache.clear();
        g_code_cache.emplace(ckey, bin);
      }
    }
  }
  if (st == QCF_OK) {
    hipModule_t m;
    hipFunction_t f;
    int nb = 0;
    if (hipModuleLoadData(&m, bin.data()) != hipSuccess) {
      st = set_error(QCF_ERR_HIP, "hipModuleLoadData failed");
    } else if (hipModuleGetFunction(&f, m, "qcf_jit_kernel") != hipSuccess) {
      (void)hipModuleUnload(m);
      st = set_error(QCF_ERR_HIP, "hipModuleGetFunction failed");
    } else if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 64, jk.lds_bytes) !=
                   hipSuccess ||
               nb < 1) {
      (void)hipModuleUnload(m);
      st = set_error(QCF_ERR_HIP, "specialised kernel does not fit a compute unit");
    } else {
      jk.module = m;
      jk.function = f;
      jk.blocks_per_cu = (uint32_t)nb;
    }
  }
  jk.failed = st != QCF_OK;
  P->jit.push_back(jk);
  info->function = jk.function;
  info->lds_bytes = jk.lds_bytes;
  info->blocks_per_cu = jk.blocks_per_cu;
  info->key = jk.key;
  return st;
}

int jit_launch(void *function, const JitParams &kp, uint32_t grid, uint32_t lds_bytes,
               hipStream_t stream) {
  JitParams p = kp;
  void *args[] = {&p};
  QCF_TRY(check_block_module((hipFunction_t)function, 64, "qcf_jit_kernel"));
  const hipError_t e = hipModuleLaunchKernel((hipFunction_t)function, grid, 1, 1, 64, 1, 1,
                                             lds_bytes, stream, args, nullptr);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return set_error(QCF_ERR_HIP, "specialised kernel launch: %s", hipGetErrorString(e));
  }
  return QCF_OK;
}

}  // namespace qcf

extern "C" int qcf_debug_jit_source(const qcf_program *prog, int variant, char *buf, size_t cap,
                                    size_t *needed) {
  qcf::clear_error();
  if (!prog) return qcf::set_error(QCF_ERR_ARGUMENT, "null program");
  std::string src;
  if (!qcf::generate(prog, variant & 511, &src, nullptr))
    return qcf::set_error(QCF_ERR_UNSUPPORTED, "op not specialisable");
  if (needed) *needed = src.size() + 1;
  if (buf && cap) {
    const size_t n = std::min(cap - 1, src.size());
    memcpy(buf, src.data(), n);
    buf[n] = 0;
  }
  return QCF_OK;
}

extern "C" int qcf_debug_jit_compile(const qcf_program *prog, int variant, size_t *code_bytes) {
  qcf::clear_error();
  if (!prog) return qcf::set_error(QCF_ERR_ARGUMENT, "null program");
  std::string src, log;
  std::vector<char> bin;
  if (!qcf::generate(prog, variant & 511, &src, nullptr))
    return qcf::set_error(QCF_ERR_UNSUPPORTED, "op not specialisable");
  QCF_TRY(qcf::compile(src, &bin, &log));
  if (code_bytes) *code_bytes = bin.size();
  return QCF_OK;
}
