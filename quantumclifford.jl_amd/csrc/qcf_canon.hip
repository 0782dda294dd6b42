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
// word of that is independent of the others, so one lane per word runs the
// products in order (eight waves share the chain by pivot groups).
// canon_apply finally replays each row's set against those snapshots: every
// row is read and written once per panel instead of once per pivot.  Phases: each product's phase is a sum of per-position mod-4
// counters, so a lane may accumulate the counters of all products on its
// words and the group reduces once per row.  After the last panel the rows
// move to their logical places (canon_gather, canon_scatter_back).
//
// Roofline per panel: the row pass reads and writes the tableau once
// (2 * nrows * 16 * W bytes); snapshots (64 * 16 * W bytes) are L2-resident.
// The elimination (64 dependent workgroup reductions) and the snapshot chain
// are latency-bound and make the critical path.
//
// The workspace (per device, grown on demand) is shared by all calls on a
// device: calls that overlap on one device race.
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
  uint32_t done;  // snapshot workgroups finished (reset by the last one)
  uint64_t ph_lo, ph_hi;  // bit k: low / high bit of snapshot k's phase
};

constexpr uint32_t kMaxCanonRows = 16 * 1024;  // 16 slots per thread of 1024

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
  if (s == 0) *st = CanonState{0, 0, 0, 0, 0, 0};
}

// Minimum over the 64 lanes with DPP moves (an inclusive min-scan whose last
// lane holds the total: row shifts 1, 2, 4, 8, then the row broadcasts).
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
  const int inf = (int)UINT_MAX;
  v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(inf, (int)v, 0x111, 0xf, 0xf, false));
  v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(inf, (int)v, 0x112, 0xf, 0xf, false));
  v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(inf, (int)v, 0x114, 0xf, 0xf, false));
  v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(inf, (int)v, 0x118, 0xf, 0xf, false));
  v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(inf, (int)v, 0x142, 0xa, 0xf, false));
  v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(inf, (int)v, 0x143, 0xc, 0xf, false));
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// One panel: word `c` of every row (x word w, or z word W+w), `nb` valid bits.
// Slot s (physical row) belongs to thread s % NT, register j = s / NT; the
// panel words also live in LDS (s_word[s], dynamic) so that every thread can
// read the pivot's word after the step's one barrier.  Keys are
// (logical row << 14) | slot, so nrows <= 16384.  The step's shared values are
// double-buffered by step parity: a buffer is rewritten two steps later, after
// a barrier that every reader of it has passed.
template <int NT, int J>
__global__ __launch_bounds__(NT) void canon_factor(
    const uint64_t *__restrict__ xzs, size_t ld, uint32_t nrows, uint32_t c, uint32_t nb,
    uint32_t *__restrict__ inv, uint64_t *__restrict__ mask, uint32_t *__restrict__ piv,
    CanonState *st, uint32_t *__restrict__ cnt, int mark_rx) {
  extern __shared__ uint64_t s_word[];
  __shared__ uint32_t s_part[2][NT / 64];
  __shared__ uint32_t s_si[2];
  const uint32_t t = threadIdx.x;
  uint64_t word[J], m[J];
  uint32_t lid[J];
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const uint32_t s = t + j * NT;
    const bool ok = s < nrows;
    word[j] = ok ? xzs[(size_t)s * ld + c] : 0ull;
    lid[j] = ok ? inv[s] : UINT_MAX;
    m[j] = 0;
    if (ok) s_word[s] = word[j];
  }
  if (t < 64) cnt[t] = 0;  // per-pivot phase counters of this panel's snapshot pass
  const uint32_t i0 = st->i;
  uint32_t i = i0, k = 0;
  __syncthreads();
  for (uint32_t b = 0; b < nb && i < nrows; ++b) {
    const uint32_t par = b & 1u;
    // the first logical row >= i with bit b set
    uint32_t key = UINT_MAX;
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const uint32_t s = t + j * NT;
      const uint32_t kk = (lid[j] << 14) | s;
      if (((word[j] >> b) & 1ull) && lid[j] >= i && kk < key) key = kk;
      if (lid[j] == i) s_si[par] = s;
    }
    key = wave_min_u32(key);
    if ((t & 63) == 0) s_part[par][t >> 6] = key;
    __syncthreads();
    uint32_t best = UINT_MAX;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) best = min(best, s_part[par][w]);
    if (best == UINT_MAX) continue;  // no pivot in this column (uniform)
    const uint32_t ps = best & 0x3FFFu, pl = best >> 14;
    const uint32_t si = s_si[par];
    const uint64_t pw = s_word[ps];
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const uint32_t s = t + j * NT;
      if (s == ps) {
        lid[j] = i;  // the pivot row moves to logical row i ...
      } else {
        if (s == si) lid[j] = pl;  // ... and row i to the pivot's old place
        if ((word[j] >> b) & 1ull) {
          word[j] ^= pw;
          s_word[s] = word[j];
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
    const uint32_t s = t + j * NT;
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
// pivots (k' < k) of its set, in order.  One lane per word position, 64 words
// per workgroup of 8 waves; the 64 snapshots of a lane's word stay in LDS
// (s_x/s_z[k*64 + lane], 64 KB).  The pivots go in 8 groups of 8: for group
// g, wave j multiplies pivot 8g+j by the snapshots of earlier groups in its
// set (all final, so the 8 rows proceed in parallel), then wave 0 finishes
// the group's rows with the products inside the group, in pivot order.  That
// is each row's products in increasing pivot order, as the reference applies
// them, on a critical path of about a quarter of the products.  Per-pivot
// phase counts go to cnt[] (zeroed by the factor kernel); the last workgroup
// to finish resolves the snapshot phases.
constexpr uint32_t kSnapWaves = 8;

__device__ __forceinline__ void snap_products(const uint64_t *s_x, const uint64_t *s_z,
                                              uint32_t t, uint64_t early, uint64_t &x,
                                              uint64_t &z, uint64_t &c1, uint64_t &c2) {
  while (early) {
    uint64_t ax[8], az[8];
    int n = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      if (early) {
        const int k2 = __ffsll((unsigned long long)early) - 1;
        early &= early - 1ull;
        ax[q] = s_x[k2 * 64 + t];
        az[q] = s_z[k2 * 64 + t];
        n = q + 1;
      }
    }
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (q < n) mul_word_c(ax[q], az[q], x, z, c1, c2);
  }
}

__global__ __launch_bounds__(64 * kSnapWaves) void canon_snap(
    const uint64_t *__restrict__ xzs, const uint8_t *__restrict__ phases, size_t ld, size_t W,
    const uint64_t *__restrict__ mask, const uint32_t *__restrict__ piv, CanonState *st,
    uint64_t *__restrict__ snap, uint32_t *__restrict__ cnt, int phases_on) {
  extern __shared__ uint64_t s_snap[];  // x: [0, 4096), z: [4096, 8192)
  __shared__ uint64_t l_mask[64];
  __shared__ uint32_t l_piv[64];
  __shared__ uint32_t l_cnt[64];
  __shared__ uint32_t l_last;
  const uint32_t npiv = st->npiv;
  if (npiv == 0) return;  // uniform
  const uint32_t t = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint64_t *s_x = s_snap, *s_z = s_snap + 64 * 64;
  if (threadIdx.x < 64) {
    l_cnt[t] = 0;
    if (t < npiv) {
      l_piv[t] = piv[t];
      l_mask[t] = mask[piv[t]];
    }
  }
  __syncthreads();
  const size_t w = (size_t)blockIdx.x * 64 + t;
  const bool ok = w < W;
  const size_t row = 2 * W;
  // the original pivot rows: all loads in flight at once
  for (uint32_t k = wave; k < npiv; k += kSnapWaves) {
    const size_t ps = l_piv[k];
    s_x[k * 64 + t] = ok ? xzs[ps * ld + w] : 0ull;
    s_z[k * 64 + t] = ok ? xzs[ps * ld + W + w] : 0ull;
  }
  __syncthreads();
  for (uint32_t g0 = 0; g0 < npiv; g0 += kSnapWaves) {
    const uint64_t before = (1ull << g0) - 1ull;  // pivots of earlier groups
    const uint32_t k = g0 + wave;
    if (k < npiv) {
      uint64_t x = s_x[k * 64 + t], z = s_z[k * 64 + t], c1 = 0, c2 = 0;
      snap_products(s_x, s_z, t, l_mask[k] & before, x, z, c1, c2);
      s_x[k * 64 + t] = x;
      s_z[k * 64 + t] = z;
      const uint32_t f = (__popcll(c1) + 2u * __popcll(c2)) & 3u;
      if (phases_on && f) atomicAdd(&l_cnt[k], f);
    }
    __syncthreads();
    if (wave == 0) {
      const uint32_t g1 = min(npiv, g0 + kSnapWaves);
      for (uint32_t kk = g0 + 1; kk < g1; ++kk) {
        const uint64_t inside = l_mask[kk] & ((1ull << kk) - 1ull) & ~before;
        if (!inside) continue;  // uniform
        uint64_t x = s_x[kk * 64 + t], z = s_z[kk * 64 + t], c1 = 0, c2 = 0;
        snap_products(s_x, s_z, t, inside, x, z, c1, c2);
        s_x[kk * 64 + t] = x;
        s_z[kk * 64 + t] = z;
        const uint32_t f = (__popcll(c1) + 2u * __popcll(c2)) & 3u;
        if (phases_on && f) atomicAdd(&l_cnt[kk], f);
      }
    }
    __syncthreads();
  }
  if (ok)
    for (uint32_t k = wave; k < npiv; k += kSnapWaves) {
      snap[k * row + w] = s_x[k * 64 + t];
      snap[k * row + W + w] = s_z[k * 64 + t];
    }
  if (!phases_on) return;  // uniform
  if (threadIdx.x < npiv && l_cnt[threadIdx.x]) atomicAdd(&cnt[threadIdx.x], l_cnt[threadIdx.x]);
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) l_last = atomicAdd(&st->done, 1u) == gridDim.x - 1 ? 1u : 0u;
  __syncthreads();
  if (!l_last) return;  // uniform
  __threadfence();
  // snapshot phases: phases[p_k] + its product counters + the phases of the
  // earlier snapshots in its set, as two bit masks (low and high phase bits)
  if (threadIdx.x < npiv)
    l_cnt[threadIdx.x] = (uint32_t)phases[l_piv[threadIdx.x]] + atomicAdd(&cnt[threadIdx.x], 0u);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t lo = 0, hi = 0;
    for (uint32_t k = 0; k < npiv; ++k) {
      const uint64_t early = l_mask[k] & ((1ull << k) - 1ull);
      const uint32_t v = (l_cnt[k] + __popcll(early & lo) + 2u * __popcll(early & hi)) & 3u;
      lo |= (uint64_t)(v & 1u) << k;
      hi |= (uint64_t)(v >> 1) << k;
    }
    st->ph_lo = lo;
    st->ph_hi = hi;
    atomicExch(&st->done, 0u);
  }
}

// Every row with a nonempty set: row := P_k * row for each pivot k of its set
// in order, P_k the snapshots.  G lanes per row.
template <int G>
__global__ __launch_bounds__(256) void canon_apply(
    uint64_t *__restrict__ xzs, uint8_t *__restrict__ phases, size_t ld, size_t W,
    uint32_t nrows, const uint64_t *__restrict__ mask, const uint64_t *__restrict__ snap,
    const CanonState *st, int phases_on) {
  if (st->npiv == 0) return;  // uniform
  const uint64_t ph_lo = st->ph_lo, ph_hi = st->ph_hi;
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
        // snapshot words of up to eight pivots are loaded before they are used
        while (mm) {
          uint64_t sx[8], sz[8];
          int n = 0;
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            if (mm) {
              const int k = __ffsll((unsigned long long)mm) - 1;
              mm &= mm - 1ull;
              sx[q] = snap[k * row + w];
              sz[q] = snap[k * row + W + w];
              n = q + 1;
            }
          }
#pragma unroll
          for (int q = 0; q < 8; ++q)
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
        const uint32_t v = phases[r] + f + __popcll(m & ph_lo) + 2u * __popcll(m & ph_hi);
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

// Workspace of the launch's stream (library-owned, grown on demand).
int canon_scratch(hipStream_t stream, size_t bytes, char **out) {
  void *p = nullptr;
  QCF_TRY(stream_scratch(kScratchCanon, stream, bytes, &p));
  *out = (char *)p;
  return QCF_OK;
}

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

constexpr uint32_t kSnapLds = 2 * 64 * 64 * sizeof(uint64_t);  // 64 KB, the default limit

template <int NT, int J>
int launch_factor(const uint64_t *xzs, size_t ld, uint32_t nrows, uint32_t c, uint32_t nb,
                  uint32_t *inv, uint64_t *mask, uint32_t *piv, CanonState *st, uint32_t *cnt,
                  int mark_rx, hipStream_t stream) {
  const uint32_t lds = nrows * (uint32_t)sizeof(uint64_t);  // at most 128 KB
  if (lds > 64u * 1024u)
    QCF_HIP_CHECK(hipFuncSetAttribute((const void *)canon_factor<NT, J>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  QCF_LAUNCH((canon_factor<NT, J>), dim3(1), dim3(NT), lds, stream, xzs, ld, nrows, c, nb,
                     inv, mask, piv, st, cnt, mark_rx);
  return QCF_OK;
}

// Threads of the elimination workgroup.  A pivot step is a chain of
// dependent LDS round trips, a DPP minimum and a barrier; 1024 threads (fewest
// rows per thread) measured fastest: 1000 rows 1.62 / 1.47 / 1.47 ms and 6000
// rows - / 25.2 / 22.3 ms for 256 / 512 / 1024 threads.  QCF_CANON_THREADS
// (256, 512 or 1024) overrides the choice.
uint32_t factor_threads(uint32_t nrows) {
  uint32_t nt = 1024;
  if (const char *e = getenv("QCF_CANON_THREADS")) {
    const int v = atoi(e);
    if (v == 256 || v == 512 || v == 1024) nt = (uint32_t)v;
  }
  while (nt < 1024 && (nrows + nt - 1) / nt > 16) nt *= 2;  // at most 16 rows per thread
  return nt;
}

template <int NT>
int dispatch_factor(uint32_t J, const uint64_t *xzs, size_t ld, uint32_t n, uint32_t c,
                    uint32_t nb, uint32_t *inv, uint64_t *mask, uint32_t *piv, CanonState *st,
                    uint32_t *cnt, int mark, hipStream_t stream) {
#define QCF_LAUNCH_FACTOR(JJ) \
  launch_factor<NT, JJ>(xzs, ld, n, c, nb, inv, mask, piv, st, cnt, mark, stream)
  if (J <= 1) return QCF_LAUNCH_FACTOR(1);
  if (J <= 2) return QCF_LAUNCH_FACTOR(2);
  if (J <= 3) return QCF_LAUNCH_FACTOR(3);
  if (J <= 4) return QCF_LAUNCH_FACTOR(4);
  if (J <= 6) return QCF_LAUNCH_FACTOR(6);
  if (J <= 8) return QCF_LAUNCH_FACTOR(8);
  if (J <= 12) return QCF_LAUNCH_FACTOR(12);
  return QCF_LAUNCH_FACTOR(16);
#undef QCF_LAUNCH_FACTOR
}

template <int G>
int launch_apply(uint64_t *xzs, uint8_t *phases, size_t ld, size_t W, uint32_t nrows,
                 const uint64_t *mask, const uint64_t *snap, const CanonState *st,
                 int phases_on, uint32_t grid, hipStream_t stream) {
  QCF_LAUNCH(canon_apply<G>, dim3(grid), dim3(256), 0, stream, xzs, phases, ld, W, nrows,
             mask, snap, st, phases_on);
  return QCF_OK;
}

}  // namespace
}  // namespace qcf

// Workspace of qcf_canonicalize_noalloc: 16 MiB of device memory that every
// GPU the library runs on holds until the process ends.  Host code never
// passes this name to a kernel (it is a host-side stand-in there); it uses the
// device address hipGetSymbolAddress returns for the current device.
constexpr size_t kNoAllocCanonBytes = 16u << 20;
static __device__ __align__(256) char canon_noalloc_scratch[kNoAllocCanonBytes];

namespace qcf {
namespace {

// Checks of both canonicalize! entry points; returns QCF_OK or an error code.
int canon_check(const uint64_t *xzs, const uint8_t *phases, size_t ld, size_t nwords,
                size_t nrows, size_t nqubits, int phases_on) {
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
  return QCF_OK;
}

// Workspace layout: state | inv | mask | piv | phase counters | snapshots |
// gathered rows | phases.  Every part is written by the kernels before it is
// read, so the workspace may hold anything on entry.
struct CanonLayout {
  size_t o_st, o_inv, o_mask, o_piv, o_cnt, o_snap, o_tmp, o_tph, bytes;
};
CanonLayout canon_layout(size_t nrows, size_t W) {
  CanonLayout L{};
  size_t off = 0;
  L.o_st = off;
  off = align256(off + sizeof(CanonState));
  L.o_inv = off;
  off = align256(off + nrows * sizeof(uint32_t));
  L.o_mask = off;
  off = align256(off + nrows * sizeof(uint64_t));
  L.o_piv = off;
  off = align256(off + 64 * sizeof(uint32_t));
  L.o_cnt = off;
  off = align256(off + 64 * sizeof(uint32_t));
  L.o_snap = off;
  off = align256(off + 64 * 2 * W * sizeof(uint64_t));
  L.o_tmp = off;
  off = align256(off + nrows * 2 * W * sizeof(uint64_t));
  L.o_tph = off;
  off = align256(off + nrows);
  L.bytes = off;
  return L;
}

// The elimination itself, on a checked tableau with a workspace at `base`.
int canon_run(uint64_t *xzs, uint8_t *phases, size_t ld, size_t nwords, size_t nrows,
              size_t nqubits, int phases_on, uint32_t *ranks, char *base, const CanonLayout &Lo,
              hipStream_t stream) {
  const size_t W = nwords;
  CanonState *st = (CanonState *)(base + Lo.o_st);
  uint32_t *inv = (uint32_t *)(base + Lo.o_inv);
  uint64_t *mask = (uint64_t *)(base + Lo.o_mask);
  uint32_t *piv = (uint32_t *)(base + Lo.o_piv);
  uint32_t *cnt = (uint32_t *)(base + Lo.o_cnt);
  uint64_t *snap = (uint64_t *)(base + Lo.o_snap);
  uint64_t *tmp = (uint64_t *)(base + Lo.o_tmp);
  uint8_t *tph = (uint8_t *)(base + Lo.o_tph);
  const uint32_t n = (uint32_t)nrows;
  QCF_LAUNCH(canon_init, dim3((n + 255) / 256), dim3(256), 0, stream, inv, n, st);
  int dev = 0, cus = 0;
  QCF_HIP_CHECK(hipGetDevice(&dev));
  QCF_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const uint32_t G = W > 32 ? 64 : W > 16 ? 32 : W > 8 ? 16 : 8;
  const uint32_t rows_per_block = 256 / G;
  const uint32_t apply_grid =
      std::max<uint32_t>(1, std::min<uint32_t>((n + rows_per_block - 1) / rows_per_block,
                                               (uint32_t)cus * 8));
  const uint32_t NT = factor_threads(n);
  const uint32_t J = (n + NT - 1) / NT;
  const uint32_t snap_grid = (uint32_t)((W + 63) / 64);
  for (int zstage = 0; zstage < 2; ++zstage) {
    for (size_t w = 0; w < W; ++w) {
      const size_t first = 64 * w;
      if (first >= nqubits) break;
      const uint32_t nb = (uint32_t)std::min<size_t>(64, nqubits - first);
      const uint32_t c = (uint32_t)(zstage ? W + w : w);
      const int mark = (zstage == 1 && w == 0) ? 1 : 0;
      if (NT == 256)
        QCF_TRY(dispatch_factor<256>(J, xzs, ld, n, c, nb, inv, mask, piv, st, cnt, mark, stream));
      else if (NT == 512)
        QCF_TRY(dispatch_factor<512>(J, xzs, ld, n, c, nb, inv, mask, piv, st, cnt, mark, stream));
      else
        QCF_TRY(dispatch_factor<1024>(J, xzs, ld, n, c, nb, inv, mask, piv, st, cnt, mark, stream));
      QCF_LAUNCH(canon_snap, dim3(snap_grid), dim3(64 * kSnapWaves), kSnapLds, stream,
                         xzs, phases, ld, W, mask, piv, st, snap, cnt, phases_on);
      switch (G) {
        case 64:
          QCF_TRY(launch_apply<64>(xzs, phases, ld, W, n, mask, snap, st, phases_on, apply_grid, stream));
          break;
        case 32:
          QCF_TRY(launch_apply<32>(xzs, phases, ld, W, n, mask, snap, st, phases_on, apply_grid, stream));
          break;
        case 16:
          QCF_TRY(launch_apply<16>(xzs, phases, ld, W, n, mask, snap, st, phases_on, apply_grid, stream));
          break;
        default:
          QCF_TRY(launch_apply<8>(xzs, phases, ld, W, n, mask, snap, st, phases_on, apply_grid, stream));
      }
    }
  }
  const size_t total = nrows * 2 * W;
  const uint32_t pgrid = (uint32_t)std::min<size_t>((total + 255) / 256, (size_t)cus * 16);
  QCF_LAUNCH(canon_gather, dim3(pgrid), dim3(256), 0, stream, xzs, phases, ld, W, n, inv,
                     tmp, phases_on ? tph : nullptr);
  QCF_LAUNCH(canon_scatter_back, dim3(pgrid), dim3(256), 0, stream, xzs, phases, ld, W, n,
                     tmp, phases_on ? tph : nullptr);
  if (ranks) QCF_LAUNCH(canon_ranks, dim3(1), dim3(64), 0, stream, st, ranks);
  QCF_HIP_CHECK(hipGetLastError());
  return QCF_OK;
}

// The no-allocation entry point's workspace: a device array of fixed size,
// one copy per GPU, whose address is looked up once per device.
constexpr int kMaxNoAllocDevices = 64;
std::mutex g_noalloc_mu;
void *g_noalloc_ptr[kMaxNoAllocDevices];

int noalloc_scratch(char **out) {
  int dev = 0;
  QCF_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxNoAllocDevices)
    return set_error(QCF_ERR_UNSUPPORTED, "canonicalize!: device %d beyond %d", dev,
                     kMaxNoAllocDevices);
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(canon_noalloc_scratch));
    if (e != hipSuccess || !p)
      return set_error(QCF_ERR_HIP, "canonicalize!: workspace symbol address: %s",
                       hipGetErrorString(e));
    g_noalloc_ptr[dev] = p;
  }
  *out = (char *)g_noalloc_ptr[dev];
  return QCF_OK;
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
      QCF_LAUNCH(canon_ranks, dim3(1), dim3(64), 0, stream, nullptr, ranks);
      QCF_HIP_CHECK(hipGetLastError());
    }
    return QCF_OK;
  }
  QCF_TRY(canon_check(xzs, phases, ld, nwords, nrows, nqubits, phases_on));
  const CanonLayout Lo = canon_layout(nrows, nwords);
  char *base = nullptr;
  QCF_TRY(canon_scratch(stream, Lo.bytes, &base));
  return canon_run(xzs, phases, ld, nwords, nrows, nqubits, phases_on, ranks, base, Lo, stream);
}

extern "C" int qcf_canonicalize_noalloc(uint64_t *xzs, uint8_t *phases, size_t ld,
                                        size_t nwords, size_t nrows, size_t nqubits,
                                        int phases_on, uint32_t *ranks, void *stream_) {
  clear_error();
  hipStream_t stream = (hipStream_t)stream_;
  if (nrows == 0) {
    if (ranks) {
      QCF_LAUNCH(canon_ranks, dim3(1), dim3(64), 0, stream, nullptr, ranks);
      QCF_HIP_CHECK(hipGetLastError());
    }
    return QCF_OK;
  }
  QCF_TRY(canon_check(xzs, phases, ld, nwords, nrows, nqubits, phases_on));
  const CanonLayout Lo = canon_layout(nrows, nwords);
  if (Lo.bytes > kNoAllocCanonBytes)
    return set_error(QCF_ERR_ARGUMENT,
                     "canonicalize! without allocation: %zu rows of %zu words need %zu bytes of "
                     "workspace, more than the %zu it holds",
                     nrows, nwords, Lo.bytes, kNoAllocCanonBytes);
  char *base = nullptr;
  QCF_TRY(noalloc_scratch(&base));
  return canon_run(xzs, phases, ld, nwords, nrows, nqubits, phases_on, ranks, base, Lo, stream);
}
