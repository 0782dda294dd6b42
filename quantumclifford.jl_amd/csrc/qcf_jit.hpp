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
  uint32_t has_class, pad0, pad1;
};

// QCF_NO_JIT=1 keeps every launch on the interpreter.
bool jit_disabled();

// The compiled kernel of `variant` for the current device (compiled on first
// use, cached in the program).  Returns an error if it cannot be built.
int jit_get(qcf_program *P, int variant, void **function);

int jit_launch(void *function, const JitParams &p, uint32_t grid, uint32_t lds_bytes,
               hipStream_t stream);

}  // namespace qcf
