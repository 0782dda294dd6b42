// Interface of the per-circuit specialised kernels (qcf_jit.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

struct qcf_program;

namespace qcf {

// Must match struct KJ in the generated source.
struct JitParams {
  uint32_t *rec;
  uint32_t *frame;
  uint64_t seed;
  uint64_t ld32;
  uint32_t nblocks, nw32, nwbuf, lastmask;
  float c;
  uint32_t has_class, nfull, xcds;
  uint32_t *scratch;  // per-wave planes when the frame does not fit in LDS
  uint64_t sstride;   // words per wave
  // multi-circuit launch (ncirc > 0): circuit i has parameters multi[i] and
  // tiles [first[i], first[i+1]) of ntiles; a single launch leaves them zero
  const JitParams *multi;
  const uint32_t *first;
  uint32_t ncirc, ntiles;
  uint32_t tile0;     // absolute index of tile 0: keys the RNG
  // shared tiles (steal non-null): workgroups take tile indices from *steal,
  // which the persistent batch kernel draws from too; every workgroup of the
  // launches that share counters adds one to sync[1] when it leaves, and the
  // last of sync_total clears sync[0 .. sync_words)
  uint32_t *steal;
  uint32_t *sync;
  uint32_t sync_total, sync_words;
  // part-tile kernels of a single launch: workgroups i, i + pset, ... take
  // part i % pset of every tile they visit (0 or 1: every part)
  uint32_t pset;
};

// QCF_NO_JIT=1 keeps every launch on the interpreter.
bool jit_disabled();

// Largest plane count kept in registers (QCF_JIT_REG_PLANES, default 192) and
// longest program specialised (QCF_JIT_MAX_INSTR, default 12000).
uint32_t jit_reg_planes();
double jit_reg_events();
bool jit_no_loops();
bool jit_global_enabled();
bool jit_split_enabled();
uint32_t jit_split_regs();
uint32_t jit_two_wave_regs();
int jit_pass_prio();
bool jit_sched_tags();
uint32_t jit_part_regs();
// Source lines starting with this tag carry compiler options for the kernel.
constexpr const char *kCompileOptTag = "// qcf-compile-options: ";
size_t jit_split_barrier();
bool jit_split_touched();
bool jit_nt_stores();
uint32_t jit_xcds();
bool jit_full_grid();
uint32_t jit_slots();
size_t jit_max_instr();
uint32_t jit_hoist();
bool jit_prefetch();
bool jit_linear_enabled();

struct JitKernelInfo {
  void *function = nullptr;     // hipFunction_t
  uint32_t lds_bytes = 0;       // dynamic LDS per wave
  uint32_t blocks_per_cu = 0;   // resident waves per compute unit
  uint64_t key = 0;             // hash of the kernel source: equal keys, equal code
};

// The compiled kernel of `variant` for the current device (compiled on first
// use, cached in the program).  Returns an error if it cannot be built.
int jit_get(qcf_program *P, int variant, JitKernelInfo *info);

// A specialised launch ready to go: kernel, arguments, grid and LDS.
struct JitPlan {
  void *function = nullptr;
  JitParams jp{};
  uint32_t grid = 0;
  uint32_t lds_bytes = 0;
  uint32_t blocks_per_cu = 0;
  uint64_t key = 0;            // circuits with equal keys can share one launch
  bool global_planes = false;  // planes in global memory: jp.scratch still to be set
};

int jit_launch(void *function, const JitParams &p, uint32_t grid, uint32_t lds_bytes,
               hipStream_t stream);

}  // namespace qcf
