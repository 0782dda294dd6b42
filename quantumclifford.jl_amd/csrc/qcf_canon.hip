// canonicalize! of a stabilizer tableau on gfx950, blocked by 64-column panels.
//
// Semantics: src/canonicalization.jl:97-135 (Garcia-Markov-Cross canonical
// form).  X stage, for each qubit column j: the first row k >= i with an X or
// Y at j is swapped into row i (rowswap!, src/QuantumClifford.jl:899-905;
// phases move only when tracked), then every other row m with an X or Y at j
// becomes row_i * row_m (mul_left!(s, m, i), src/mul_leftright.jl:195-198,
// phases[m] = (s + phases[m] + phases[i]) & 3); i advances.  Z stage: the same
// on Z bits, continuing from the X stage's i.
//
// Blocking.  The pivots of 64 consecutive columns (one x or z word) depend
// only on that word of every row, so one workgroup eliminates the panel word
// of all rows in registers (canon_factor): 64 sequential pivot steps, each a
// workgroup-wide minimum and one broadcast.  It records, per row, the 64-bit
// set of the panel's pivots that multiplied it (in pivot order), the pivot
// rows, and the row order as a slot -> logical-row map (swaps move no data).
// canon_snap then computes, for each pivot k, the full pivot row as it was at
// step k (its original row times the earlier pivot rows in its set); every
// word of that is independent of the others, so one thread per word runs the
// products in order.  canon_apply finally replays each row's set against those
// snapshots: every row is read and written once per panel instead of once per
// pivot.  Phases: each product's phase is a sum of per-position mod-4
// counters, so a lane may accumulate the counters of all products on its
// words and the group reduces once per row.  After the last panel the rows
// move to their logical places (canon_gather, canon_scatter_back).
//
// Roofline per panel: the row pass reads and writes the tableau once
// (2 * nrows * 16 * W bytes); snapshots (64 * 16 * W bytes) are L2-resident.
// The factor step is latency-bound (64 dependent workgroup reductions).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <mutex>
#include <vector>

#include "qcf_common.hpp"

namespace qcf {
namespace {

struct CanonState {
  uint32_t i;     // pivots so far (the next logical pivot row)
  uint32_t npiv;  // pivots of the current panel
  uint32_t rx;    // pivots of the X stage
  uint32_t pad;
};

constexpr int kFactorThreads = 1024;
constexpr uint32_t kMaxCanonRows = 16 * kFactorThreads;  // 16 slots per thread

__device__ __forceinline__ void mul_word_c(uint64_t x1, uint64_t z1, uint64_t &x2, uint64_t &z2,
                                           uint64_t &c1, uint64_t &c2) {
  const uint64_t nx = x1 ^ x2;
  const uint64_t nz = z1 ^ z2;
  const uint64_t x1z2 = x1 & z2;
  const uint64_t ac = (x2 & z1) ^ x1z2;
  c2 ^= (c1 ^ nx ^ nz ^ x1z2) & ac;
  c1 ^= ac;
  x2 = nx;
  z2 = nz;
}

__global__ void canon_init(uint32_t *inv, uint32_t nrows, CanonState *st) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nrows) inv[s] = s;
  if (s == 0) *st = CanonState{0, 0, 0, 0};
}

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(v, off, 64);
    v = o < v ? o : v;
  }
  return v;
}

// One panel: word `c` of every row (x word w, or z word W+w), `nb` valid bits.
// Slot s (physical row) belongs to thread s % 1024, register j = s / 1024.
template <int J>
__global__ __launch_bounds__(kFactorThreads) void canon_factor(
    const uint64_t *__restrict__ xzs, size_t ld, uint32_t nrows, uint32_t c, uint32_t nb,
    uint32_t *__restrict__ inv, uint64_t *__restrict__ mask, uint32_t *__restrict__ piv,
    CanonState *st, int mark_rx) {
  __shared__ unsigned long long s_part[kFactorThreads / 64];
  __shared__ uint32_t s_si;
  __shared__ unsigned long long s_pw;
  const uint32_t t = threadIdx.x;
  uint64_t word[J], m[J];
  uint32_t lid[J];
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const uint32_t s = t + j * kFactorThreads;
    const bool ok = s < nrows;
    word[j] = ok ? xzs[(size_t)s * ld + c] : 0ull;
    lid[j] = ok ? inv[s] : UINT_MAX;
    m[j] = 0;
  }
  const uint32_t i0 = st->i;
  uint32_t i = i0, k = 0;
  for (uint32_t b = 0; b < nb && i < nrows; ++b) {
    // the first logical row >= i with bit b set: minimum of (logical, slot)
    unsigned long long key = ULLONG_MAX;
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const unsigned long long kk = ((unsigned long long)lid[j] << 32) | (t + j * kFactorThreads);
      if (((word[j] >> b) & 1ull) && lid[j] >= i && kk < key) key = kk;
      if (lid[j] == i) s_si = t + j * kFactorThreads;
    }
    key = wave_min_u64(key);
    if ((t & 63) == 0) s_part[t >> 6] = key;
    __syncthreads();
    unsigned long long best = s_part[0];
#pragma unroll
    for (int w = 1; w < kFactorThreads / 64; ++w) best = s_part[w] < best ? s_part[w] : best;
    const uint32_t si = s_si;
    if (best == ULLONG_MAX) {  // no pivot in this column
      __syncthreads();
      continue;
    }
    const uint32_t ps = (uint32_t)best, pl = (uint32_t)(best >> 32);
#pragma unroll
    for (int j = 0; j < J; ++j)
      if (t + j * kFactorThreads == ps) s_pw = word[j];
    __syncthreads();
    const uint64_t pw = s_pw;
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const uint32_t s = t + j * kFactorThreads;
      if (s == ps) {
        lid[j] = i;  // the pivot row moves to logical row i ...
      } else {
        if (s == si) lid[j] = pl;  // ... and row i to the pivot's old place
        if ((word[j] >> b) & 1ull) {
          word[j] ^= pw;
          m[j] |= 1ull << k;
        }
      }
    }
    if (t == 0) piv[k] = ps;
    ++k;
    ++i;
  }
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const uint32_t s = t + j * kFactorThreads;
    if (s < nrows) {
      inv[s] = lid[j];
      mask[s] = m[j];
    }
  }
  __syncthreads();  // every wave has read st->i before it changes
  if (t == 0) {
    st->i = i;
    st->npiv = k;
    if (mark_rx) st->rx = i0;
  }
}

// Snapshots: pivot k's full row at its step = original row times the earlier
// pivots (k' < k) of its set, in order.  One thread per word position.
__global__ __launch_bounds__(1024) void canon_snap(
    const uint64_t *__restrict__ xzs, const uint8_t *__restrict__ phases, size_t ld, size_t W,
    const uint64_t *__restrict__ mask, const uint32_t *__restrict__ piv, const CanonState *st,
    uint64_t *__restrict__ snap, uint32_t *__restrict__ snap_ph, int phases_on) {
  __shared__ uint32_t cnt[64];
  const uint32_t npiv = st->npiv;
  if (npiv == 0) return;  // uniform
  const uint32_t t = threadIdx.x;
  if (t < 64) cnt[t] = 0;
  __syncthreads();
  const size_t row = 2 * W;
  for (size_t w = t; w < W; w += blockDim.x) {
    for (uint32_t k = 0; k < npiv; ++k) {
      const uint32_t ps = piv[k];
      uint64_t x = xzs[(size_t)ps * ld + w], z = xzs[(size_t)ps * ld + W + w];
      uint64_t early = mask[ps] & ((1ull << k) - 1ull);
      uint64_t c1 = 0, c2 = 0;
      while (early) {
        const int k2 = __ffsll((unsigned long long)early) - 1;
        early &= early - 1ull;
        mul_word_c(snap[k2 * row + w], snap[k2 * row + W + w], x, z, c1, c2);
      }
      snap[k * row + w] = x;
      snap[k * row + W + w] = z;
      if (phases_on) {
        const uint32_t f = (__popcll(c1) + 2u * __popcll(c2)) & 3u;
        if (f) atomicAdd(&cnt[k], f);
      }
    }
  }
  __syncthreads();
  if (t == 0 && phases_on) {
    uint32_t ph[64];
    for (uint32_t k = 0; k < npiv; ++k) {
      const uint32_t ps = piv[k];
      uint32_t v = phases[ps] + cnt[k];
      uint64_t early = mask[ps] & ((1ull << k) - 1ull);
      while (early) {
        const int k2 = __ffsll((unsigned long long)early) - 1;
        early &= early - 1ull;
        v += ph[k2];
      }
      ph[k] = v & 3u;
      snap_ph[k] = ph[k];
    }
  }
}

// Every row with a nonempty set: row := P_k * row for each pivot k of its set
// in order, P_k the snapshots.  G lanes per row.
template <int G>
__global__ __launch_bounds__(256) void canon_apply(
    uint64_t *__restrict__ xzs, uint8_t *__restrict__ phases, size_t ld, size_t W,
    uint32_t nrows, const uint64_t *__restrict__ mask, const uint64_t *__restrict__ snap,
    const uint32_t *__restrict__ snap_ph, const CanonState *st, int phases_on) {
  if (st->npiv == 0) return;  // uniform
  constexpr uint32_t kRowsPerBlock = 256 / G;
  const uint32_t lane = threadIdx.x % G;
  const size_t row = 2 * W;
  for (uint32_t r = blockIdx.x * kRowsPerBlock + threadIdx.x / G; r < nrows;
       r += gridDim.x * kRowsPerBlock) {
    const uint64_t m = mask[r];
    uint64_t c1 = 0, c2 = 0;
    if (m) {
      uint64_t *xr = xzs + (size_t)r * ld;
      for (size_t w = lane; w < W; w += G) {
        uint64_t x = xr[w], z = xr[W + w];
        uint64_t mm = m;
        // snapshot words of up to four pivots are loaded before they are used
        while (mm) {
          uint64_t sx[4], sz[4];
          int n = 0;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            if (mm) {
              const int k = __ffsll((unsigned long long)mm) - 1;
              mm &= mm - 1ull;
              sx[q] = snap[k * row + w];
              sz[q] = snap[k * row + W + w];
              n = q + 1;
            }
          }
#pragma unroll
          for (int q = 0; q < 4; ++q)
            if (q < n) mul_word_c(sx[q], sz[q], x, z, c1, c2);
        }
        xr[w] = x;
        xr[W + w] = z;
      }
    }
    if (phases_on) {
      uint32_t f = __popcll(c1) + 2u * __popcll(c2);
#pragma unroll
      for (int off = G / 2; off > 0; off >>= 1) f += __shfl_xor(f, off, G);
      if (lane == 0 && m) {
        uint32_t v = phases[r] + f;
        uint64_t mm = m;
        while (mm) {
          const int k = __ffsll((unsigned long long)mm) - 1;
          mm &= mm - 1ull;
          v += snap_ph[k];
        }
        phases[r] = (uint8_t)(v & 3u);
      }
    }
  }
}

// Rows (and phases, when tracked) to their logical places: tmp[inv[s]] = row s.
__global__ __launch_bounds__(256) void canon_gather(const uint64_t *__restrict__ xzs,
                                                    const uint8_t *__restrict__ phases,
                                                    size_t ld, size_t W, uint32_t nrows,
                                                    const uint32_t *__restrict__ inv,
                                                    uint64_t *__restrict__ tmp,
                                                    uint8_t *__restrict__ tmp_ph) {
  const size_t row = 2 * W;
  const size_t total = (size_t)nrows * row;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    const size_t s = e / row, w = e % row;
    const uint32_t l = inv[s];
    tmp[(size_t)l * row + w] = xzs[s * ld + w];
    if (w == 0 && tmp_ph) tmp_ph[l] = phases[s];
  }
}

__global__ __launch_bounds__(256) void canon_scatter_back(uint64_t *__restrict__ xzs,
                                                          uint8_t *__restrict__ phases,
                                                          size_t ld, size_t W, uint32_t nrows,
                                                          const uint64_t *__restrict__ tmp,
                                                          const uint8_t *__restrict__ tmp_ph) {
  const size_t row = 2 * W;
  const size_t total = (size_t)nrows * row;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    const size_t l = e / row, w = e % row;
    xzs[l * ld + w] = tmp[e];
    if (w == 0 && tmp_ph) phases[l] = tmp_ph[l];
  }
}

__global__ void canon_ranks(const CanonState *st, uint32_t *ranks) {
  if (threadIdx.x == 0) {
    ranks[0] = st ? st->rx : 0u;
    ranks[1] = st ? st->i : 0u;
  }
}

// Workspace per device, grown on demand (owned by the library).
struct CanonScratch {
  void *ptr = nullptr;
  size_t bytes = 0;
};

int canon_scratch(size_t bytes, char **out) {
  static std::mutex mu;
  static std::vector<CanonScratch> per_dev;
  int dev = 0;
  QCF_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  if ((int)per_dev.size() <= dev) per_dev.resize(dev + 1);
  CanonScratch &cs = per_dev[dev];
  if (cs.bytes < bytes) {
    if (cs.ptr) {
      QCF_HIP_CHECK(hipDeviceSynchronize());
      QCF_HIP_CHECK(hipFree(cs.ptr));
      cs.ptr = nullptr;
      cs.bytes = 0;
    }
    QCF_HIP_CHECK(hipMalloc(&cs.ptr, bytes));
    cs.bytes = bytes;
  }
  *out = (char *)cs.ptr;
  return QCF_OK;
}

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

template <int J>
void launch_factor(const uint64_t *xzs, size_t ld, uint32_t nrows, uint32_t c, uint32_t nb,
                   uint32_t *inv, uint64_t *mask, uint32_t *piv, CanonState *st, int mark_rx,
                   hipStream_t stream) {
  hipLaunchKernelGGL(canon_factor<J>, dim3(1), dim3(kFactorThreads), 0, stream, xzs, ld, nrows,
                     c, nb, inv, mask, piv, st, mark_rx);
}

template <int G>
void launch_apply(uint64_t *xzs, uint8_t *phases, size_t ld, size_t W, uint32_t nrows,
                  const uint64_t *mask, const uint64_t *snap, const uint32_t *snap_ph,
                  const CanonState *st, int phases_on, uint32_t grid, hipStream_t stream) {
  hipLaunchKernelGGL(canon_apply<G>, dim3(grid), dim3(256), 0, stream, xzs, phases, ld, W, nrows,
                     mask, snap, snap_ph, st, phases_on);
}

}  // namespace
}  // namespace qcf

using namespace qcf;

extern "C" int qcf_canonicalize(uint64_t *xzs, uint8_t *phases, size_t ld, size_t nwords,
                                size_t nrows, size_t nqubits, int phases_on, uint32_t *ranks,
                                void *stream_) {
  clear_error();
  hipStream_t stream = (hipStream_t)stream_;
  if (nrows == 0) {
    if (ranks) {
      hipLaunchKernelGGL(canon_ranks, dim3(1), dim3(64), 0, stream, nullptr, ranks);
      QCF_HIP_CHECK(hipGetLastError());
    }
    return QCF_OK;
  }
  if (!xzs) return set_error(QCF_ERR_ARGUMENT, "canonicalize!: null tableau");
  if (phases_on && !phases) return set_error(QCF_ERR_ARGUMENT, "canonicalize!: null phases");
  if (nqubits > 64 * nwords)
    return set_error(QCF_ERR_DIMENSION, "canonicalize!: %zu qubits need more than %zu words",
                     nqubits, nwords);
  if (ld < 2 * nwords)
    return set_error(QCF_ERR_DIMENSION, "canonicalize!: row stride %zu < 2*%zu words", ld,
                     nwords);
  if (nrows > kMaxCanonRows)
    return set_error(QCF_ERR_UNSUPPORTED, "canonicalize!: %zu rows exceed the limit of %u", nrows,
                     kMaxCanonRows);
  const size_t W = nwords;
  // workspace: state | inv | mask | piv | snapshot phases | snapshots | gathered rows | phases
  size_t off = 0;
  const size_t o_st = off;
  off = align256(off + sizeof(CanonState));
  const size_t o_inv = off;
  off = align256(off + nrows * sizeof(uint32_t));
  const size_t o_mask = off;
  off = align256(off + nrows * sizeof(uint64_t));
  const size_t o_piv = off;
  off = align256(off + 64 * sizeof(uint32_t));
  const size_t o_sph = off;
  off = align256(off + 64 * sizeof(uint32_t));
  const size_t o_snap = off;
  off = align256(off + 64 * 2 * W * sizeof(uint64_t));
  const size_t o_tmp = off;
  off = align256(off + nrows * 2 * W * sizeof(uint64_t));
  const size_t o_tph = off;
  off = align256(off + nrows);
  char *base = nullptr;
  QCF_TRY(canon_scratch(off, &base));
  CanonState *st = (CanonState *)(base + o_st);
  uint32_t *inv = (uint32_t *)(base + o_inv);
  uint64_t *mask = (uint64_t *)(base + o_mask);
  uint32_t *piv = (uint32_t *)(base + o_piv);
  uint32_t *sph = (uint32_t *)(base + o_sph);
  uint64_t *snap = (uint64_t *)(base + o_snap);
  uint64_t *tmp = (uint64_t *)(base + o_tmp);
  uint8_t *tph = (uint8_t *)(base + o_tph);
  const uint32_t n = (uint32_t)nrows;
  hipLaunchKernelGGL(canon_init, dim3((n + 255) / 256), dim3(256), 0, stream, inv, n, st);
  DeviceState *ds;
  QCF_TRY(device_state(&ds));
  const uint32_t G = W > 32 ? 64 : W > 16 ? 32 : W > 8 ? 16 : 8;
  const uint32_t rows_per_block = 256 / G;
  const uint32_t apply_grid =
      std::max<uint32_t>(1, std::min<uint32_t>((n + rows_per_block - 1) / rows_per_block,
                                               (uint32_t)ds->cu_count * 8));
  const uint32_t J = (n + kFactorThreads - 1) / kFactorThreads;
  const uint32_t snap_threads = (uint32_t)std::min<size_t>(1024, ((W + 63) / 64) * 64);
  for (int zstage = 0; zstage < 2; ++zstage) {
    for (size_t w = 0; w < W; ++w) {
      const size_t first = 64 * w;
      if (first >= nqubits) break;
      const uint32_t nb = (uint32_t)std::min<size_t>(64, nqubits - first);
      const uint32_t c = (uint32_t)(zstage ? W + w : w);
      const int mark = (zstage == 1 && w == 0) ? 1 : 0;
      if (J <= 1)
        launch_factor<1>(xzs, ld, n, c, nb, inv, mask, piv, st, mark, stream);
      else if (J <= 2)
        launch_factor<2>(xzs, ld, n, c, nb, inv, mask, piv, st, mark, stream);
      else if (J <= 4)
        launch_factor<4>(xzs, ld, n, c, nb, inv, mask, piv, st, mark, stream);
      else if (J <= 8)
        launch_factor<8>(xzs, ld, n, c, nb, inv, mask, piv, st, mark, stream);
      else
        launch_factor<16>(xzs, ld, n, c, nb, inv, mask, piv, st, mark, stream);
      hipLaunchKernelGGL(canon_snap, dim3(1), dim3(snap_threads), 0, stream, xzs, phases, ld, W,
                         mask, piv, st, snap, sph, phases_on);
      switch (G) {
        case 64:
          launch_apply<64>(xzs, phases, ld, W, n, mask, snap, sph, st, phases_on, apply_grid, stream);
          break;
        case 32:
          launch_apply<32>(xzs, phases, ld, W, n, mask, snap, sph, st, phases_on, apply_grid, stream);
          break;
        case 16:
          launch_apply<16>(xzs, phases, ld, W, n, mask, snap, sph, st, phases_on, apply_grid, stream);
          break;
        default:
          launch_apply<8>(xzs, phases, ld, W, n, mask, snap, sph, st, phases_on, apply_grid, stream);
      }
    }
  }
  const size_t total = nrows * 2 * W;
  const uint32_t pgrid = (uint32_t)std::min<size_t>((total + 255) / 256, (size_t)ds->cu_count * 16);
  hipLaunchKernelGGL(canon_gather, dim3(pgrid), dim3(256), 0, stream, xzs, phases, ld, W, n, inv,
                     tmp, phases_on ? tph : nullptr);
  hipLaunchKernelGGL(canon_scatter_back, dim3(pgrid), dim3(256), 0, stream, xzs, phases, ld, W, n,
                     tmp, phases_on ? tph : nullptr);
  if (ranks) hipLaunchKernelGGL(canon_ranks, dim3(1), dim3(64), 0, stream, st, ranks);
  QCF_HIP_CHECK(hipGetLastError());
  return QCF_OK;
}
