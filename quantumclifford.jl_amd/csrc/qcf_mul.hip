// Bit-packed Pauli product with phase tracking on gfx950.
//
// Semantics follow src/mul_leftright.jl of QuantumClifford.jl: for r := l*r
// on [x;z] word arrays, per 64-qubit word
//     x1,z1 = l;  x2,z2 = r;  r.x = x1^x2;  r.z = z1^z2
//     ac = (x2 & z1) ^ (x1 & z2)            anticommuting positions
//     d  = r.x ^ r.z ^ (x1 & z2)            set where the factor is -i
// and every bit position keeps a two-bit counter (c1 low, c2 high) that counts
// +i (d=0) or -i (d=1) factors mod 4 (src/mul_leftright.jl:11-31, 75-121).
// The product phase is (sum c1 + 2 sum c2) mod 4 (mul_left!, :123-126) or
// (3 sum c1 + 2 sum c2) mod 4 (mul_right!, :128-131).  Because each counter
// is a mod-4 accumulator, the sum is independent of how words are split among
// lanes, so the kernels split freely.
//
// Roofline: HBM-bound, 48 bytes per 64 qubits (read l.x, l.z, r.x, r.z, write
// r.x, r.z) against ~16 int32 lane-ops.
#include <hip/hip_runtime.h>

#include <atomic>
#include <algorithm>
#include <climits>
#include <cstdlib>

#include "qcf_common.hpp"

namespace qcf {
namespace {

template <typename W>
__device__ __forceinline__ void mul_word(W x1, W z1, W &x2, W &z2, uint64_t &c1,
                                         uint64_t &c2) {
  const uint64_t nx = x1 ^ x2;
  const uint64_t nz = z1 ^ z2;
  const uint64_t x1z2 = x1 & z2;
  const uint64_t ac = (x2 & z1) ^ x1z2;
  c2 ^= (c1 ^ nx ^ nz ^ x1z2) & ac;
  c1 ^= ac;
  x2 = nx;
  z2 = nz;
}

__device__ __forceinline__ uint32_t fold_phase(uint64_t c1, uint64_t c2,
                                               bool right) {
  const uint32_t p1 = __popcll(c1);
  const uint32_t p2 = __popcll(c2);
  return right ? 3u * p1 + 2u * p2 : p1 + 2u * p2;
}

// Sum of v over the workgroup, valid in thread 0.  blockDim.x <= 1024.
__device__ __forceinline__ uint32_t block_sum(uint32_t v) {
  __shared__ uint32_t part[16];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) part[wave] = v;
  __syncthreads();
  uint32_t s = 0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; ++i) s += part[i];
  }
  return s;
}

// Epilogue of a launch whose phase sum spans workgroups: the last workgroup
// to finish folds the total and resets its counter slot to zero.
__device__ __forceinline__ void finish_phase(uint32_t block_total,
                                             uint32_t *slot, uint8_t *out_s,
                                             uint8_t *dest_phase,
                                             const uint8_t *src_phase) {
  if (threadIdx.x != 0) return;
  atomicAdd(&slot[0], block_total);
  __threadfence();
  const uint32_t done = atomicAdd(&slot[1], 1u);
  if (done != gridDim.x - 1) return;
  __threadfence();
  const uint32_t total = atomicAdd(&slot[0], 0u);
  const uint32_t s = total & 3u;
  if (out_s) out_s[0] = (uint8_t)s;
  if (dest_phase) {
    const uint32_t extra = src_phase ? src_phase[0] : 0u;
    dest_phase[0] = (uint8_t)((s + dest_phase[0] + extra) & 3u);
  }
  atomicExch(&slot[0], 0u);
  atomicExch(&slot[1], 0u);
}

// dest := src * dest on [x;z] arrays of nw words per half, two words per
// thread per step through 16-byte loads (requires nw even and 16-byte aligned
// halves; the host falls back to mul_vec_u64 otherwise).  Every word is read
// and written once, so the accesses are nontemporal (no cache allocation):
// at 10^9 qubits that and one workgroup per CU took the launch from 0.166 to
// 0.141 ms.
typedef unsigned long long qcf_u64x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ qcf_u64x2 ld_nt(const qcf_u64x2 *p) {
  return __builtin_nontemporal_load(p);
}
__device__ __forceinline__ void st_nt(qcf_u64x2 *p, qcf_u64x2 v) {
  __builtin_nontemporal_store(v, p);
}
__device__ __forceinline__ void mul_pair(qcf_u64x2 a1, qcf_u64x2 b1, qcf_u64x2 &a2,
                                         qcf_u64x2 &b2, uint64_t &c1, uint64_t &c2) {
  unsigned long long ax = a2.x, ay = a2.y, bx = b2.x, by = b2.y;
  mul_word((unsigned long long)a1.x, (unsigned long long)b1.x, ax, bx, c1, c2);
  mul_word((unsigned long long)a1.y, (unsigned long long)b1.y, ay, by, c1, c2);
  a2 = qcf_u64x2{ax, ay};
  b2 = qcf_u64x2{bx, by};
}

template <bool PHASES>
__global__ __launch_bounds__(256) void mul_vec_x4(
    uint64_t *__restrict__ dest, const uint64_t *__restrict__ src, size_t nw,
    int right, uint32_t *slot, uint8_t *out_s, uint8_t *dest_phase,
    const uint8_t *src_phase) {
  const size_t npairs = nw >> 1;
  const qcf_u64x2 *sx = reinterpret_cast<const qcf_u64x2 *>(src);
  const qcf_u64x2 *sz = reinterpret_cast<const qcf_u64x2 *>(src + nw);
  qcf_u64x2 *dx = reinterpret_cast<qcf_u64x2 *>(dest);
  qcf_u64x2 *dz = reinterpret_cast<qcf_u64x2 *>(dest + nw);
  uint64_t c1 = 0, c2 = 0;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  // two independent pairs per iteration keep 8 loads in flight per thread
  for (; i + stride < npairs; i += 2 * stride) {
    const qcf_u64x2 a1 = ld_nt(sx + i), b1 = ld_nt(sz + i), e1 = ld_nt(sx + i + stride),
                    f1 = ld_nt(sz + i + stride);
    qcf_u64x2 a2 = ld_nt(dx + i), b2 = ld_nt(dz + i), e2 = ld_nt(dx + i + stride),
              f2 = ld_nt(dz + i + stride);
    if (PHASES) {
      mul_pair(a1, b1, a2, b2, c1, c2);
      mul_pair(e1, f1, e2, f2, c1, c2);
    } else {
      a2 ^= a1; b2 ^= b1; e2 ^= e1; f2 ^= f1;
    }
    st_nt(dx + i, a2); st_nt(dz + i, b2); st_nt(dx + i + stride, e2); st_nt(dz + i + stride, f2);
  }
  for (; i < npairs; i += stride) {
    const qcf_u64x2 a1 = ld_nt(sx + i), b1 = ld_nt(sz + i);
    qcf_u64x2 a2 = ld_nt(dx + i), b2 = ld_nt(dz + i);
    if (PHASES)
      mul_pair(a1, b1, a2, b2, c1, c2);
    else {
      a2 ^= a1; b2 ^= b1;
    }
    st_nt(dx + i, a2); st_nt(dz + i, b2);
  }
  if (!PHASES) return;
  const uint32_t total = block_sum(fold_phase(c1, c2, right != 0));
  finish_phase(total, slot, out_s, dest_phase, src_phase);
}

template <bool PHASES>
__global__ __launch_bounds__(256) void mul_vec_u64(
    uint64_t *__restrict__ dest, const uint64_t *__restrict__ src, size_t nw,
    int right, uint32_t *slot, uint8_t *out_s, uint8_t *dest_phase,
    const uint8_t *src_phase) {
  uint64_t c1 = 0, c2 = 0;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nw;
       i += stride) {
    const uint64_t x1 = src[i], z1 = src[nw + i];
    uint64_t x2 = dest[i], z2 = dest[nw + i];
    if (PHASES) {
      mul_word(x1, z1, x2, z2, c1, c2);
    } else {
      x2 ^= x1;
      z2 ^= z1;
    }
    dest[i] = x2;
    dest[nw + i] = z2;
  }
  if (!PHASES) return;
  const uint32_t total = block_sum(fold_phase(c1, c2, right != 0));
  finish_phase(total, slot, out_s, dest_phase, src_phase);
}

// One group of RL lanes per tableau row: row m := src * row m for every row
// m != skip selected by row_mask.  The source row is read by every group and
// stays in L2; no cross-group reduction is needed.
template <int RL, bool PHASES>
__global__ __launch_bounds__(256) void mul_rows_kernel(
    uint64_t *__restrict__ xzs, uint8_t *__restrict__ phases, size_t ld,
    size_t nw, size_t nrows, const uint64_t *__restrict__ src,
    const uint8_t *src_phase, size_t skip, const uint8_t *row_mask,
    int right) {
  const int g = threadIdx.x / RL;
  const int l = threadIdx.x % RL;
  const size_t m = (size_t)blockIdx.x * (blockDim.x / RL) + g;
  bool active = m < nrows && m != skip && (!row_mask || row_mask[m] != 0);
  uint64_t c1 = 0, c2 = 0;
  if (active) {
    uint64_t *row = xzs + m * ld;
    for (size_t k = l; k < nw; k += RL) {
      const uint64_t x1 = src[k], z1 = src[nw + k];
      uint64_t x2 = row[k], z2 = row[nw + k];
      if (PHASES) {
        mul_word(x1, z1, x2, z2, c1, c2);
      } else {
        x2 ^= x1;
        z2 ^= z1;
      }
      row[k] = x2;
      row[nw + k] = z2;
    }
  }
  if (!PHASES) return;
  uint32_t s = fold_phase(c1, c2, right != 0);
#pragma unroll
  for (int off = RL / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, RL);
  if (active && l == 0) {
    const uint32_t extra = src_phase ? src_phase[0] : 0u;
    phases[m] = (uint8_t)((s + phases[m] + extra) & 3u);
  }
}

int launch_vec(uint64_t *dest, const uint64_t *src, size_t nw, int phases,
               int right, uint8_t *out_s, uint8_t *dest_phase,
               const uint8_t *src_phase, hipStream_t stream) {
  if (nw == 0) {
    // empty operators: the product phase is 0
    if (phases && (out_s || dest_phase)) {
      return set_error(QCF_ERR_ARGUMENT,
                       "mul_left!: zero-length operators carry no words");
    }
    return QCF_OK;
  }
  if (!dest || !src)
    return set_error(QCF_ERR_ARGUMENT, "mul_left!: null operand pointer");
  // r and l may coincide but must not partially overlap
  const uint64_t *d0 = dest, *d1 = dest + 2 * nw;
  const uint64_t *s0 = src, *s1 = src + 2 * nw;
  if (d0 != s0 && d0 < s1 && s0 < d1)
    return set_error(QCF_ERR_ARGUMENT,
                     "mul_left!: operands partially overlap in memory");
  DeviceState *st;
  QCF_TRY(device_state(&st));
  uint32_t *slot = nullptr;
  if (phases) {
    const uint32_t k = __atomic_fetch_add(&st->next_slot, 1u, __ATOMIC_RELAXED);
    slot = st->counters + 2 * (k % kCounterSlots);
  }
  const bool aligned = (nw % 2 == 0) && ((uintptr_t)dest % 16 == 0) &&
                       ((uintptr_t)src % 16 == 0);
  const size_t items = aligned ? nw / 2 : nw;
  const int threads = 256;
  size_t blocks = (items + threads - 1) / threads;
  // 256-thread workgroups per CU: at 10^9 qubits 1, 2, 3 took 0.140, 0.162,
  // 0.176 ms (fewer streams in flight keep HBM pages open longer)
  size_t per_cu = 1;
  if (const char *e = getenv("QCF_MUL_BLOCKS_PER_CU")) per_cu = std::max(1, atoi(e));
  const size_t cap = (size_t)st->cu_count * per_cu;
  if (blocks > cap) blocks = cap;
  if (blocks == 0) blocks = 1;
  if (aligned) {
    if (phases)
      QCF_LAUNCH(mul_vec_x4<true>, dim3(blocks), dim3(threads), 0,
                         stream, dest, src, nw, right, slot, out_s, dest_phase,
                         src_phase);
    else
      QCF_LAUNCH(mul_vec_x4<false>, dim3(blocks), dim3(threads), 0,
                         stream, dest, src, nw, right, slot, out_s, dest_phase,
                         src_phase);
  } else {
    if (phases)
      QCF_LAUNCH(mul_vec_u64<true>, dim3(blocks), dim3(threads), 0,
                         stream, dest, src, nw, right, slot, out_s, dest_phase,
                         src_phase);
    else
      QCF_LAUNCH(mul_vec_u64<false>, dim3(blocks), dim3(threads), 0,
                         stream, dest, src, nw, right, slot, out_s, dest_phase,
                         src_phase);
  }
  QCF_HIP_CHECK(hipGetLastError());
  return QCF_OK;
}

template <int RL>
int launch_rows_rl(uint64_t *xzs, uint8_t *phases, size_t ld, size_t nw,
                    size_t nrows, const uint64_t *src,
                    const uint8_t *src_phase, size_t skip,
                    const uint8_t *row_mask, int phases_on, int right,
                    hipStream_t stream) {
  const int threads = 256;
  const size_t rows_per_block = threads / RL;
  const size_t blocks = (nrows + rows_per_block - 1) / rows_per_block;
  if (phases_on)
    QCF_LAUNCH((mul_rows_kernel<RL, true>), dim3(blocks), dim3(threads),
                       0, stream, xzs, phases, ld, nw, nrows, src, src_phase,
                       skip, row_mask, right);
  else
    QCF_LAUNCH((mul_rows_kernel<RL, false>), dim3(blocks),
                       dim3(threads), 0, stream, xzs, phases, ld, nw, nrows,
                       src, src_phase, skip, row_mask, right);
  return QCF_OK;
}

}  // namespace
}  // namespace qcf

using namespace qcf;

extern "C" int qcf_mul_left(uint64_t *r_xz, uint8_t *r_phase,
                            const uint64_t *l_xz, const uint8_t *l_phase,
                            size_t nwords, int phases, void *stream) {
  clear_error();
  if (phases && (!r_phase || !l_phase))
    return set_error(QCF_ERR_ARGUMENT, "mul_left!: null phase pointer");
  return launch_vec(r_xz, l_xz, nwords, phases, 0, nullptr,
                    phases ? r_phase : nullptr, l_phase, (hipStream_t)stream);
}

extern "C" int qcf_mul_right(uint64_t *l_xz, uint8_t *l_phase,
                             const uint64_t *r_xz, const uint8_t *r_phase,
                             size_t nwords, int phases, void *stream) {
  clear_error();
  if (phases && (!r_phase || !l_phase))
    return set_error(QCF_ERR_ARGUMENT, "mul_right!: null phase pointer");
  return launch_vec(l_xz, r_xz, nwords, phases, 1, nullptr,
                    phases ? l_phase : nullptr, r_phase, (hipStream_t)stream);
}

extern "C" int qcf_mul_vec(uint64_t *r_xz, const uint64_t *l_xz,
                           size_t nwords, int phases, int right,
                           uint8_t *out_phase, void *stream) {
  clear_error();
  if (phases && !out_phase)
    return set_error(QCF_ERR_ARGUMENT, "mul_left!: null out_phase pointer");
  return launch_vec(r_xz, l_xz, nwords, phases, right ? 1 : 0, out_phase,
                    nullptr, nullptr, (hipStream_t)stream);
}

extern "C" int qcf_mul_rows(uint64_t *xzs, uint8_t *phases, size_t ld,
                            size_t nwords, size_t nrows, const uint64_t *src,
                            const uint8_t *src_phase, size_t skip,
                            const uint8_t *row_mask, int phases_on, int right,
                            void *stream) {
  clear_error();
  if (nrows == 0 || nwords == 0) return QCF_OK;
  if (!xzs || !src)
    return set_error(QCF_ERR_ARGUMENT, "mul_left!: null tableau pointer");
  if (ld < 2 * nwords)
    return set_error(QCF_ERR_DIMENSION,
                     "mul_left!: row stride %zu is shorter than a row (%zu words)",
                     ld, 2 * nwords);
  if (phases_on && (!phases || !src_phase))
    return set_error(QCF_ERR_ARGUMENT, "mul_left!: null phase pointer");
  // the source must not be one of the rows being rewritten
  const uint64_t *t0 = xzs, *t1 = xzs + (nrows - 1) * ld + 2 * nwords;
  if (src >= t0 && src < t1) {
    const size_t off = (size_t)(src - t0);
    if (off % ld != 0 || off / ld != skip)
      return set_error(QCF_ERR_ARGUMENT,
                       "mul_left!: source row aliases a row that is updated");
  }
  hipStream_t s = (hipStream_t)stream;
  if (nwords <= 16)
    QCF_TRY(launch_rows_rl<16>(xzs, phases, ld, nwords, nrows, src, src_phase, skip,
                               row_mask, phases_on, right, s));
  else if (nwords <= 32)
    QCF_TRY(launch_rows_rl<32>(xzs, phases, ld, nwords, nrows, src, src_phase, skip,
                               row_mask, phases_on, right, s));
  else
    QCF_TRY(launch_rows_rl<64>(xzs, phases, ld, nwords, nrows, src, src_phase, skip,
                               row_mask, phases_on, right, s));
  QCF_HIP_CHECK(hipGetLastError());
  return QCF_OK;
}
