// Shared internals of libqcf: error reporting, HIP checks, per-device state
// and the counter-based RNG (Philox4x32-10) used by the frame sampler.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/qcf.h"

namespace qcf {

// ---------------------------------------------------------------- errors --
int set_error(int code, const char *fmt, ...);
void clear_error();

#define QCF_HIP_CHECK(expr)                                                  \
  do {                                                                       \
    hipError_t qcf_e_ = (expr);                                              \
    if (qcf_e_ != hipSuccess)                                                \
      return ::qcf::set_error(QCF_ERR_HIP, "%s failed: %s (%s:%d)", #expr,   \
                              hipGetErrorString(qcf_e_), __FILE__, __LINE__); \
  } while (0)

#define QCF_TRY(expr)                                                        \
  do {                                                                       \
    int qcf_s_ = (expr);                                                     \
    if (qcf_s_ != QCF_OK) return qcf_s_;                                     \
  } while (0)

// ------------------------------------------------------------- launches --
// Every kernel launch goes through QCF_LAUNCH, which first checks on the host
// that the block size is within the kernel's __launch_bounds__ (the
// maxThreadsPerBlock of its compiled code, looked up once per kernel): a
// larger block fails to launch, and a kernel whose indexing assumes its
// bound would run out of range if it did.  check_block returns
// QCF_ERR_ARGUMENT naming the kernel.
int check_block(const void *kernel, unsigned threads, const char *name);
// The same for a kernel loaded from a code object (hipModuleLaunchKernel).
int check_block_module(hipFunction_t kernel, unsigned threads, const char *name);

#define QCF_LAUNCH(kernel, grid, block, lds, stream, ...)                              \
  do {                                                                                \
    const dim3 qcf_b_ = (block);                                                      \
    QCF_TRY(::qcf::check_block((const void *)(kernel), qcf_b_.x * qcf_b_.y * qcf_b_.z, \
                               #kernel));                                             \
    hipLaunchKernelGGL(kernel, grid, qcf_b_, lds, stream, __VA_ARGS__);               \
  } while (0)

// Per-device scratch used by reductions that span several workgroups: a ring
// of (sum, done) counter pairs that the last workgroup of a launch resets to
// zero after reading, so no clearing launch is needed between calls.
struct DeviceState {
  uint32_t *counters = nullptr;  // kCounterSlots * 2 words
  uint32_t next_slot = 0;
  int cu_count = 0;
};
constexpr uint32_t kCounterSlots = 4096;
int device_state(DeviceState **out);

// Library-owned device workspace, one per (device, stream, kind): calls on one
// stream run in order, so they never use a workspace at the same time, and
// calls on different streams have different ones.  It grows on demand behind
// a synchronization of that stream only; it cannot grow while the stream is
// being captured into a graph (QCF_ERR_ARGUMENT: make one uncaptured call of
// the same size first).  Workspaces live until the process ends.
enum ScratchKind : int { kScratchFrames = 0, kScratchCanon = 1 };
int stream_scratch(int kind, hipStream_t stream, size_t bytes, void **out);

// ------------------------------------------------------------------ RNG --
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11): 10 rounds of the 4x32 multiply-xorshift round with the
// Weyl key schedule.  Constants are those of the paper.
constexpr uint32_t kPhiloxM0 = 0xD2511F53u;
constexpr uint32_t kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u;
constexpr uint32_t kPhiloxW1 = 0xBB67AE85u;

struct u32x4 {
  uint32_t x, y, z, w;
};

__host__ __device__ inline uint32_t mulhi32(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umulhi(a, b);
#else
  return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
#endif
}

__host__ __device__ inline u32x4 philox4x32_10(u32x4 c, uint32_t k0,
                                               uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = mulhi32(kPhiloxM0, c.x);
    const uint32_t lo0 = kPhiloxM0 * c.x;
    const uint32_t hi1 = mulhi32(kPhiloxM1, c.z);
    const uint32_t lo1 = kPhiloxM1 * c.z;
    c = u32x4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += kPhiloxW0;
    k1 += kPhiloxW1;
  }
  return c;
}

// Stream identifiers (fourth counter word) so that the gauge stream and the
// noise-event stream of the same shot word never share a counter.
constexpr uint32_t kStreamGauge = 0x47415547u;  // "GAUG"
constexpr uint32_t kStreamEvent = 0x45564E54u;  // "EVNT"

// xoshiro128** (Blackman & Vigna, "Scrambled linear pseudorandom number
// generators", 2021): the per-lane sequential generator for gauge words,
// seeded once per shot word from Philox so the stream depends only on
// (seed, shot word).
struct Xoshiro128 {
  uint32_t s0, s1, s2, s3;
  __host__ __device__ inline uint32_t next() {
    const uint32_t r = rotl(s1 * 5u, 7) * 9u;
    const uint32_t t = s1 << 9;
    s2 ^= s0;
    s3 ^= s1;
    s1 ^= s2;
    s0 ^= s3;
    s2 ^= t;
    s3 = rotl(s3, 11);
    return r;
  }
  __host__ __device__ static inline uint32_t rotl(uint32_t x, int k) {
    return (x << k) | (x >> (32 - k));
  }
};

__host__ __device__ inline Xoshiro128 gauge_stream(uint64_t seed,
                                                   uint32_t word) {
  const u32x4 r = philox4x32_10(u32x4{word, 0u, 0u, kStreamGauge},
                                (uint32_t)seed, (uint32_t)(seed >> 32));
  Xoshiro128 g{r.x, r.y, r.z, r.w | 1u};  // never the all-zero state
  return g;
}

}  // namespace qcf
