// Neural syndrome decoder: a 4-layer perceptron (detectors -> h1 -> h2 -> h3
// -> 1 logit of the observable flip, GELU between layers) trained on
// syndromes that the frame sampler draws on the device.  BF16 weights and
// activations, FP32 accumulation, FP32 master weights with Adam.
//
// No reference counterpart: the reference decodes through
// AbstractSyndromeDecoder (src/ecc/decoder_pipeline.jl:1-24) with table,
// belief-propagation and matching back ends; this decoder plugs into the same
// contract from the host side.
//
// Every dense contraction runs in one hand-written GEMM kernel on the gfx950
// matrix cores (v_mfma_f32_32x32x16_bf16): 128x128 output tiles per
// 256-thread workgroup (four waves, 64x64 each as 2x2 MFMA tiles), K in
// stages of 64 through one LDS buffer with the stage after next prefetched
// into registers, tiles renumbered so that each XCD works on one band of
// rows.
// Operands are read K-contiguous (or, in the debug entry, transposed).
// Epilogues fuse bias + GELU, keeping the GELU derivative for the backward
// pass (forward), the multiplication by that derivative (backward), and
// optionally write a transposed copy of their output through LDS: the
// batch-minor operands of the weight gradients.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/qcf.h"
#include "qcf_common.hpp"

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

namespace qcf {
namespace {

constexpr int kTile = 128;  // output tile edge
constexpr int kXcds = 8;    // accelerator dies of an MI355X
constexpr int kBK = 32;     // K per LDS stage
constexpr int kMaxSplitK = 16;
constexpr bool kWideK = true;  // 64-wide K stages when K allows

enum Epi : int {
  kEpiStoreF32 = 0,   // C (fp32) = acc
  kEpiBiasGelu = 1,   // p = acc + bias: act = gelu(p), dg = gelu'(p) (both bf16)
  kEpiDGelu = 2,      // out = acc * dg_in (bf16)
  kEpiBiasGeluInfer = 3,  // act = gelu(acc + bias) (bf16), no derivative kept
};

struct GemmArgs {
  const bf16 *A, *B;
  int M, N, K, lda, ldb, ldc;
  float *Cf;           // kEpiStoreF32 output
  bf16 *Cb;            // bf16 output
  const float *bias;   // per column
  const bf16 *dg_in;   // kEpiDGelu: gelu' at the layer's pre-activations
  bf16 *CbT;           // optional transposed copy of Cb: CbT[n * M + m]
  bf16 *Cg;            // kEpiBiasGelu: gelu'(p), for the backward pass
};

// GELU(p) = p Phi(p) and GELU'(p) = Phi(p) + p phi(p), phi the standard
// normal density.  erfc comes from Abramowitz & Stegun, Handbook of
// Mathematical Functions, 7.1.26: for x >= 0, erfc(x) = (a1 t + a2 t^2 +
// a3 t^3 + a4 t^4 + a5 t^5) e^{-x^2} with t = 1 / (1 + 0.3275911 x), absolute
// error below 1.5e-7 -- far inside the BF16 rounding of both outputs.  With
// x = |p| / sqrt(2), e^{-x^2} = e^{-p^2/2} is also what phi(p) needs, so both
// outputs cost one exponential and one reciprocal (the exact erff and expf
// took about 80 instructions per element and dominated the forward
// epilogue).
struct GeluPair {
  float act, der;
};
__device__ __forceinline__ GeluPair gelu_pair(float p) {
  const float x = fabsf(p) * 0.70710678118654752f;
  const float t = __builtin_amdgcn_rcpf(fmaf(0.3275911f, x, 1.0f));
  float q = fmaf(t, 1.061405429f, -1.453152027f);
  q = fmaf(t, q, 1.421413741f);
  q = fmaf(t, q, -0.284496736f);
  q = fmaf(t, q, 0.254829592f);
  const float e = __builtin_amdgcn_exp2f(-x * x * 1.4426950408889634f);  // e^{-x^2}
  const float h = 0.5f * t * q * e;      // erfc(x) / 2 = Phi(-|p|)
  const float cdf = p >= 0.0f ? 1.0f - h : h;
  return GeluPair{p * cdf, fmaf(p, 0.39894228040143268f * e, cdf)};
}

// One operand tile (ROWS rows of the output dimension x BK of K) moves from
// global memory to LDS as [row][k].  K-contiguous operands move 16 bytes per
// thread-chunk; transposed ones are read along the output dimension and
// scattered into the [row][k] layout.  ROWS * BK / 2048 chunks of 16 bytes
// per thread, held as one vector of 32-bit words (a plain array of uint4
// ends up in scratch memory).
template <int ROWS, int BK>
using TileRegs = uint32_t __attribute__((ext_vector_type(ROWS * BK / 512)));

template <bool KMAJOR, int ROWS, int BK>
__device__ __forceinline__ void load_tile(const bf16 *G, int ld, int r0, int k0, int t,
                                          TileRegs<ROWS, BK> &regs) {
#pragma unroll
  for (int h = 0; h < ROWS * BK / 2048; ++h) {
    const int c = t + h * 256;
    uint4 v;
    if constexpr (KMAJOR) {
      const int row = c / (BK / 8), kc = (c % (BK / 8)) * 8;
      v = *(const uint4 *)(G + (size_t)(r0 + row) * ld + k0 + kc);
    } else {
      const int k = c / (ROWS / 8), rc = (c % (ROWS / 8)) * 8;
      v = *(const uint4 *)(G + (size_t)(k0 + k) * ld + r0 + rc);
    }
    regs[4 * h + 0] = v.x;
    regs[4 * h + 1] = v.y;
    regs[4 * h + 2] = v.z;
    regs[4 * h + 3] = v.w;
  }
}

template <bool KMAJOR, int ROWS, int BK>
__device__ __forceinline__ void store_tile(bf16 (*S)[BK + 8], int t,
                                           const TileRegs<ROWS, BK> &regs) {
#pragma unroll
  for (int h = 0; h < ROWS * BK / 2048; ++h) {
    const int c = t + h * 256;
    if constexpr (KMAJOR) {
      const int row = c / (BK / 8), kc = (c % (BK / 8)) * 8;
      *(uint4 *)&S[row][kc] =
          uint4{regs[4 * h], regs[4 * h + 1], regs[4 * h + 2], regs[4 * h + 3]};
    } else {
      // the 8 values go to 8 rows
      const int k = c / (ROWS / 8), rc = (c % (ROWS / 8)) * 8;
#pragma unroll
      for (int j = 0; j < 8; ++j)
        S[rc + j][k] =
            __builtin_bit_cast(bf16, (uint16_t)(regs[4 * h + (j >> 1)] >> (16 * (j & 1))));
    }
  }
}

// C[m][n] = sum_k A(m,k) B(k,n) with A(m,k) = A[m*lda+k] (A_KMAJOR) or
// A[k*lda+m], and B(k,n) = B[n*ldb+k] (B_KMAJOR) or B[k*ldb+n].  A TM x TN
// output tile per 256-thread workgroup: four waves in a 2 x 2 grid, each
// (TM/2) x (TN/2) as 32 x 32 MFMA tiles.  Requires M, N multiples of TM, TN
// and K a multiple of BK (checked on the host).
template <bool A_KMAJOR, bool B_KMAJOR, int EPI, int BK, int STAGES, int TM, int TN>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
gemm_bf16(GemmArgs g) {
  // STAGES == 2: double-buffered LDS, one barrier per K stage, the next
  // stage prefetched into registers; STAGES == 1: one LDS buffer (half the
  // LDS, twice the resident workgroups), two barriers per stage, the stage
  // after next prefetched into registers
  constexpr int FM = TM / 64, FN = TN / 64;  // 32 x 32 tiles per wave
  // Output orientation.  The input-gradient epilogue reads the GELU
  // derivative at every output: with a lane per row and four consecutive
  // columns it takes one 8-byte load per four outputs (36.1 -> 26.9 us per
  // launch at M = 8192, N = K = 1024).  The other epilogues store only, and
  // a lane per column writes 64- and 128-byte runs of a row per instruction,
  // which measured faster than 8- and 16-byte pieces of 32 rows (forward
  // 26.0 against 30.0 us, weight gradients 24.5 against 26.3).
  constexpr bool kRowLanes = EPI == kEpiDGelu;
  // operand stages As, Bs; after the K loop the same LDS holds the output
  // tile's transpose for the optional batch-minor copy (rows padded)
  constexpr int kTSp = TM + 8;
  constexpr int kA = TM * (BK + 8), kB = TN * (BK + 8);
  constexpr int kLds = STAGES * (kA + kB) > TN * kTSp ? STAGES * (kA + kB) : TN * kTSp;
  __shared__ __attribute__((aligned(16))) bf16 lds_raw[kLds];
  bf16(*As)[TM][BK + 8] = reinterpret_cast<bf16(*)[TM][BK + 8]>(lds_raw);
  bf16(*Bs)[TN][BK + 8] = reinterpret_cast<bf16(*)[TN][BK + 8]>(lds_raw + STAGES * kA);
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = w >> 1, wn = w & 1;
  // Workgroup i runs on XCD i mod 8, each XCD with its own L2.  Tiles are
  // renumbered so that XCD x takes the x-th contiguous run of row-major tile
  // indices (whole rows of output tiles where the count allows): its L2 then
  // holds one band of A and the B tiles, instead of every XCD streaming all
  // of A.
  const int ntn = (int)gridDim.x, nt = (int)(gridDim.x * gridDim.y);
  int lin = (int)(blockIdx.y * gridDim.x + blockIdx.x);
  if (nt % kXcds == 0) lin = (lin % kXcds) * (nt / kXcds) + lin / kXcds;
  const int m0 = (lin / ntn) * TM, n0 = (lin % ntn) * TN;
  // split K: workgroup z covers K range [z*K/Z, (z+1)*K/Z) and, for the fp32
  // store, writes its partial sum to slice z of the output
  const int kz = g.K / (int)gridDim.z, kbase = (int)blockIdx.z * kz;
  f32x16 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;
  const int KT = kz / BK;
  // the matrix-core work of one K stage from LDS buffer `cur`
  auto mma_stage = [&](int cur) {
#pragma unroll
    for (int kk = 0; kk < BK; kk += 16) {
      bf16x8 a[FM], b[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i)
        a[i] = *(const bf16x8 *)&As[cur][wm * (TM / 2) + i * 32 + (lane & 31)][kk + 8 * (lane >> 5)];
#pragma unroll
      for (int j = 0; j < FN; ++j)
        b[j] = *(const bf16x8 *)&Bs[cur][wn * (TN / 2) + j * 32 + (lane & 31)][kk + 8 * (lane >> 5)];
      // kRowLanes: B as the matrix core's first operand, so the product is
      // the tile's transpose and each lane holds one output row m and runs
      // of four consecutive columns n
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
          acc[i][j] = kRowLanes
                          ? __builtin_amdgcn_mfma_f32_32x32x16_bf16(b[j], a[i], acc[i][j], 0, 0, 0)
                          : __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
  };
  if (STAGES == 1) {
    // One LDS buffer, the global loads two K stages ahead in two register
    // sets: a stage's tiles are requested while the matrix cores work on the
    // two stages before it, so a load has two stages of MFMA time to arrive
    // (one stage ahead, the wave waited on nearly every load).
    // The refills are unconditional (stages past the end reload the last
    // one, never stored): with a load under a branch the compiler's wait
    // counting lost track of which set was in flight across the loop and
    // waited for both sets before every store, one stage ahead after all.
    TileRegs<TM, BK> ra0, ra1;
    TileRegs<TN, BK> rb0, rb1;
    auto k_of = [&](int kt) { return kbase + (kt < KT ? kt : KT - 1) * BK; };
    load_tile<A_KMAJOR, TM, BK>(g.A, g.lda, m0, k_of(0), t, ra0);
    load_tile<B_KMAJOR, TN, BK>(g.B, g.ldb, n0, k_of(0), t, rb0);
    load_tile<A_KMAJOR, TM, BK>(g.A, g.lda, m0, k_of(1), t, ra1);
    load_tile<B_KMAJOR, TN, BK>(g.B, g.ldb, n0, k_of(1), t, rb1);
    auto stage = [&](int kt, TileRegs<TM, BK> &ra, TileRegs<TN, BK> &rb, bool refill) {
      store_tile<A_KMAJOR, TM, BK>(As[0], t, ra);
      store_tile<B_KMAJOR, TN, BK>(Bs[0], t, rb);
      __syncthreads();
      if (refill) {  // the set just stored takes the stage after next
        load_tile<A_KMAJOR, TM, BK>(g.A, g.lda, m0, k_of(kt + 2), t, ra);
        load_tile<B_KMAJOR, TN, BK>(g.B, g.ldb, n0, k_of(kt + 2), t, rb);
      }
      mma_stage(0);
      __syncthreads();
    };
    int kt = 0;
    for (; kt + 1 < KT; kt += 2) {
      stage(kt, ra0, rb0, true);
      stage(kt + 1, ra1, rb1, true);
    }
    if (kt < KT) stage(kt, ra0, rb0, false);
  } else {
    TileRegs<TM, BK> ra;
    TileRegs<TN, BK> rb;
    load_tile<A_KMAJOR, TM, BK>(g.A, g.lda, m0, kbase, t, ra);
    load_tile<B_KMAJOR, TN, BK>(g.B, g.ldb, n0, kbase, t, rb);
    store_tile<A_KMAJOR, TM, BK>(As[0], t, ra);
    store_tile<B_KMAJOR, TN, BK>(Bs[0], t, rb);
    __syncthreads();
    for (int kt = 0; kt < KT; ++kt) {
      const int cur = kt & 1;
      if (kt + 1 < KT) {
        load_tile<A_KMAJOR, TM, BK>(g.A, g.lda, m0, kbase + (kt + 1) * BK, t, ra);
        load_tile<B_KMAJOR, TN, BK>(g.B, g.ldb, n0, kbase + (kt + 1) * BK, t, rb);
      }
      mma_stage(cur);
      if (kt + 1 < KT) {
        store_tile<A_KMAJOR, TM, BK>(As[cur ^ 1], t, ra);
        store_tile<B_KMAJOR, TN, BK>(Bs[cur ^ 1], t, rb);
      }
      __syncthreads();
    }
  }
  if constexpr (kRowLanes) {
    // in each 32 x 32 tile the lane holds row m = (lane & 31) and, for
    // q = 0..3, the four consecutive columns n = 8q + 4 (lane / 32) + u
    // (accumulator element 4q + u)
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int m = m0 + wm * (TM / 2) + i * 32 + (lane & 31);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int nq = n0 + wn * (TN / 2) + j * 32 + 8 * q + 4 * (lane >> 5);
          const size_t o = (size_t)m * g.ldc + nq;
          const uint2 dg = *(const uint2 *)(g.dg_in + o);
          const uint32_t dw[2] = {dg.x, dg.y};
          uint16_t h[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const float du = __builtin_bit_cast(float, (dw[u >> 1] >> (16 * (u & 1))) << 16);
            h[u] = __builtin_bit_cast(uint16_t, (bf16)(acc[i][j][4 * q + u] * du));
          }
          *(uint2 *)(g.Cb + o) =
              uint2{h[0] | ((uint32_t)h[1] << 16), h[2] | ((uint32_t)h[3] << 16)};
          if (g.CbT) {
            // the tile's transpose goes through LDS: T[n - n0][m - m0]
            uint16_t *T = (uint16_t *)lds_raw;
#pragma unroll
            for (int u = 0; u < 4; ++u) T[(size_t)(nq + u - n0) * kTSp + (m - m0)] = h[u];
          }
        }
      }
  } else {
    // in each 32 x 32 tile the lane holds column n = (lane & 31) and rows
    // 8 (r / 4) + 4 (lane / 32) + r % 4 (accumulator element r)
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int n = n0 + wn * (TN / 2) + j * 32 + (lane & 31);
        const float bn = (EPI == kEpiBiasGelu || EPI == kEpiBiasGeluInfer) ? g.bias[n] : 0.0f;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          // r = 4q .. 4q+3 are four consecutive rows m
          const int mq = m0 + wm * (TM / 2) + i * 32 + 8 * q + 4 * (lane >> 5);
          uint16_t h[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int r = 4 * q + u, m = mq + u;
            const size_t o = (size_t)m * g.ldc + n;
            const float v = acc[i][j][r];
            bf16 b = (bf16)0.0f;
            if (EPI == kEpiStoreF32) {
              g.Cf[o + (size_t)blockIdx.z * g.M * g.ldc] = v;
            } else if (EPI == kEpiBiasGelu) {
              const GeluPair gp = gelu_pair(v + bn);
              b = (bf16)gp.act;
              g.Cg[o] = (bf16)gp.der;
            } else {
              b = (bf16)gelu_pair(v + bn).act;
            }
            if (EPI != kEpiStoreF32) g.Cb[o] = b;
            h[u] = __builtin_bit_cast(uint16_t, b);
          }
          if (EPI != kEpiStoreF32 && g.CbT) {
            // the tile's transpose goes through LDS: T[n - n0][m - m0]
            uint16_t *T = (uint16_t *)lds_raw;
            *(uint2 *)(T + (size_t)(n - n0) * kTSp + (mq - m0)) =
                uint2{h[0] | ((uint32_t)h[1] << 16), h[2] | ((uint32_t)h[3] << 16)};
          }
        }
      }
  }
  if (EPI != kEpiStoreF32 && g.CbT) {
    // TN rows n of TM consecutive m: 16 bytes per thread
    __syncthreads();
    const uint16_t *T = (const uint16_t *)lds_raw;
#pragma unroll
    for (int h = 0; h < TM * TN / 8 / 256; ++h) {
      const int c = t + h * 256, row = c / (TM / 8), col = (c % (TM / 8)) * 8;
      *(uint4 *)(g.CbT + (size_t)(n0 + row) * g.M + m0 + col) =
          *(const uint4 *)(T + (size_t)row * kTSp + col);
    }
  }
}

// out[i] = sum_z part[z][i] in a fixed order (split-K reduction).
__global__ void sum_slices(const float *part, uint32_t nz, size_t n, float *out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.0f;
  for (uint32_t z = 0; z < nz; ++z) s += part[z * n + i];
  out[i] = s;
}

// LDS buffers per K stage (QCF_GEMM_STAGES, 1 or 2; one measured faster:
// 588 vs 536 TFLOP/s on the inference forward).
int gemm_stages() {
  static const int v = [] {
    const char *e = getenv("QCF_GEMM_STAGES");
    return e && e[0] == '2' ? 2 : 1;
  }();
  return v;
}

// Workgroups a split-K weight gradient aims at (QCF_SPLITK_WGS, default
// 256, one per CU).  Batch 8192: the three weight-gradient GEMMs took 24.5,
// 23.9 and 35.1 us per launch at targets 512, 256 and 128, and the slice
// sums after them 8.8, 5.6 and 4.9 us (twice the slices, twice the fp32
// partials written and read back).
int splitk_workgroups() {
  static const int v = [] {
    const char *e = getenv("QCF_SPLITK_WGS");
    const int x = e ? atoi(e) : 0;
    return x > 0 ? x : 256;
  }();
  return v;
}

template <bool AKM, bool BKM, int EPI, int TM, int TN>
int launch_tiled(const GemmArgs &gp, bool wide, bool one, int splitk, hipStream_t s) {
  const dim3 grid(gp.N / TN, gp.M / TM, splitk);
  if (wide && one)
    QCF_LAUNCH((gemm_bf16<AKM, BKM, EPI, 64, 1, TM, TN>), grid, dim3(256), 0, s, gp);
  else if (wide)
    QCF_LAUNCH((gemm_bf16<AKM, BKM, EPI, 64, 2, TM, TN>), grid, dim3(256), 0, s, gp);
  else if (one)
    QCF_LAUNCH((gemm_bf16<AKM, BKM, EPI, 32, 1, TM, TN>), grid, dim3(256), 0, s, gp);
  else
    QCF_LAUNCH((gemm_bf16<AKM, BKM, EPI, 32, 2, TM, TN>), grid, dim3(256), 0, s, gp);
  return QCF_OK;
}

// `splitk` > 1 (fp32 store only) runs that many K slices into `part`
// (splitk * M * ldc floats) and sums them into C.
template <bool AKM, bool BKM, int EPI>
int launch_gemm(const GemmArgs &g, hipStream_t s, int splitk = 1, float *part = nullptr) {
  if (g.M % kTile || g.N % kTile || g.K % kBK || g.M <= 0 || g.N <= 0 || g.K <= 0)
    return set_error(QCF_ERR_DIMENSION, "gemm: M=%d N=%d K=%d must be multiples of %d/%d/%d",
                     g.M, g.N, g.K, kTile, kTile, kBK);
  if (splitk > 1 && (EPI != kEpiStoreF32 || !part || g.K % (splitk * kBK)))
    splitk = 1;
  // K stages of 64 where K allows (half the barriers per MFMA), else 32
  const bool wide = kWideK && (g.K % (splitk * 64)) == 0;
  GemmArgs gp = g;
  if (splitk > 1) gp.Cf = part;
  const bool one = gemm_stages() == 1;
  // 128 x 128 tiles: 128 x 256, 256 x 128, 128 x 64 and 64 x 128 measured
  // slower or no faster on the decoder's shapes (M = 8192, N = K = 1024:
  // 577, 539, 574, 547 against 606 TFLOP/s), as did 128-wide K stages (547)
  QCF_TRY((launch_tiled<AKM, BKM, EPI, kTile, kTile>(gp, wide, one, splitk, s)));
  if (splitk > 1) {
    const size_t n = (size_t)g.M * g.ldc;
    QCF_LAUNCH(sum_slices, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, part,
                       (uint32_t)splitk, n, g.Cf);
  }
  QCF_HIP_CHECK(hipGetLastError());
  return QCF_OK;
}

// Syndromes of `rows` shots starting at shot `s0` (a multiple of 64) of a
// packed record: x[m][j] = detector j of shot m (0 past the detectors and past
// the shots), optionally the transpose xT[j][m], and y[m] = observable.  One
// workgroup per 64 shots and 128 columns: the 128 detector words go to LDS
// once, then every thread writes 16-byte runs of x (8 columns of one shot)
// and of xT (8 shots of one column).
__global__ void gather_inputs(const uint64_t *rec, size_t ld, uint64_t s0, uint64_t nshots,
                              const uint32_t *det_cols, uint32_t ndet, uint32_t kin,
                              uint32_t obs_col, float *y, uint16_t *x, uint16_t *xT,
                              uint32_t rows) {
  __shared__ uint64_t wd[128];
  const uint32_t m0 = blockIdx.x * 64, j0 = blockIdx.y * 128;
  const uint64_t first = s0 + m0, wi = first >> 6;
  const uint64_t left = nshots > first ? nshots - first : 0;
  const uint64_t valid = left >= 64 ? ~0ull : ((1ull << left) - 1ull);
  constexpr uint32_t kOne = 0x3F80;  // bf16 1.0
  if (threadIdx.x < 128) {
    const uint32_t j = j0 + threadIdx.x;
    uint64_t w = 0;
    if (j < ndet && valid) w = rec[(size_t)det_cols[j] * ld + wi] & valid;
    wd[threadIdx.x] = w;
  }
  __syncthreads();
  const uint32_t ncol = min(128u, kin - j0);  // kin is a multiple of 128
  for (uint32_t c = threadIdx.x; c < 64 * (ncol / 8); c += blockDim.x) {
    const uint32_t b = c / (ncol / 8), jc = (c % (ncol / 8)) * 8;
    uint32_t v[4];
#pragma unroll
    for (int h = 0; h < 4; ++h)
      v[h] = (((wd[jc + 2 * h] >> b) & 1u) ? kOne : 0u) |
             ((((wd[jc + 2 * h + 1] >> b) & 1u) ? kOne : 0u) << 16);
    *(uint4 *)(x + (size_t)(m0 + b) * kin + j0 + jc) = uint4{v[0], v[1], v[2], v[3]};
  }
  if (xT)
    for (uint32_t c = threadIdx.x; c < ncol * 8; c += blockDim.x) {
      const uint32_t j = c >> 3, b0 = (c & 7) * 8;
      const uint64_t w = wd[j] >> b0;
      uint32_t v[4];
#pragma unroll
      for (int h = 0; h < 4; ++h)
        v[h] = (((w >> (2 * h)) & 1u) ? kOne : 0u) | ((((w >> (2 * h + 1)) & 1u) ? kOne : 0u) << 16);
      *(uint4 *)(xT + (size_t)(j0 + j) * rows + m0 + b0) = uint4{v[0], v[1], v[2], v[3]};
    }
  if (y && blockIdx.y == 0 && threadIdx.x < 64) {
    const uint64_t o = valid ? rec[(size_t)obs_col * ld + wi] & valid : 0;
    y[m0 + threadIdx.x] = (float)((o >> threadIdx.x) & 1u);
  }
}

// dst[c][r] = src[r][c] for a rows x cols bf16 matrix (both multiples of 64).
__global__ void transpose_bf16(const uint16_t *src, uint16_t *dst, uint32_t rows, uint32_t cols) {
  __shared__ uint16_t t[64][66];
  const uint32_t r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
  const uint32_t tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (uint32_t i = ty; i < 64; i += 4) t[i][tx] = src[(size_t)(r0 + i) * cols + c0 + tx];
  __syncthreads();
  for (uint32_t i = ty; i < 64; i += 4) dst[(size_t)(c0 + i) * rows + r0 + tx] = t[tx][i];
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Output layer h3 -> 1: z[m] = a3[m] . w4 + b4, one wave per row, each wave
// kHeadRows rows.  With labels: dz = (sigmoid(z) - y) / rows, and the block's
// share of mean(softplus(z) - y z) goes to lpart[block] (summed by
// head_sums in a fixed order: one atomic per row on one address serialises).
constexpr uint32_t kHeadRows = 2;
__global__ void head_forward(const bf16 *a3, const bf16 *w4, const float *b4, uint32_t h3,
                             uint32_t rows, float *z, const float *y, float *dz, float *lpart,
                             float inv_rows) {
  __shared__ float sh[4];
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  float lsum = 0.0f;
  // 8 consecutive k per lane and step (16-byte loads); h3 is a multiple of 128
  for (uint32_t i = 0; i < kHeadRows; ++i) {
    const uint32_t m = (blockIdx.x * 4 + w) * kHeadRows + i;
    if (m >= rows) break;
    float s = 0.0f;
    for (uint32_t k = lane * 8; k < h3; k += 512) {
      const uint4 a = *(const uint4 *)(a3 + (size_t)m * h3 + k);
      const uint4 b = *(const uint4 *)(w4 + k);
      const uint32_t ua[4] = {a.x, a.y, a.z, a.w}, ub[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
      for (int h = 0; h < 4; ++h)
        s += __builtin_bit_cast(float, ua[h] << 16) * __builtin_bit_cast(float, ub[h] << 16) +
             __builtin_bit_cast(float, ua[h] & 0xFFFF0000u) *
                 __builtin_bit_cast(float, ub[h] & 0xFFFF0000u);
    }
    s = wave_sum(s) + b4[0];
    if (lane == 0) {
      z[m] = s;
      if (y) {
        const float yy = y[m];
        const float sp = s > 0 ? s + log1pf(__expf(-s)) : log1pf(__expf(s));
        const float sg = 1.0f / (1.0f + __expf(-s));
        dz[m] = (sg - yy) * inv_rows;
        lsum += (sp - yy * s) * inv_rows;
      }
    }
  }
  if (!lpart) return;
  if (lane == 0) sh[w] = lsum;
  __syncthreads();
  if (threadIdx.x == 0) lpart[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// d3[m][k] = dz[m] * w4[k] * gelu'(pre3[m][k]) (dg3 holds the derivative),
// written row-major and as the batch-minor copy d3T[k][m], one 64 x 64 tile
// per workgroup through LDS.
__global__ void head_backward(const float *dz, const bf16 *w4, const bf16 *dg3, uint32_t h3,
                              uint32_t rows, bf16 *d3, bf16 *d3T) {
  __shared__ bf16 t[64][66];
  const uint32_t r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
  const uint32_t tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const float wk = (float)w4[c0 + tx];
  for (uint32_t i = ty; i < 64; i += 4) {
    const size_t o = (size_t)(r0 + i) * h3 + c0 + tx;
    const bf16 v = (bf16)(dz[r0 + i] * wk * (float)dg3[o]);
    d3[o] = v;
    t[i][tx] = v;
  }
  __syncthreads();
  for (uint32_t i = ty; i < 64; i += 4) d3T[(size_t)(c0 + i) * rows + r0 + tx] = t[tx][i];
}

// grad b4 = sum of dz over the batch and loss += sum of the head's block
// partials, both in a fixed order.
__global__ void head_sums(const float *dz, uint32_t n, const float *lpart, uint32_t nl,
                          float *gb4, float *loss) {
  __shared__ float sh[2][256];
  float s = 0.0f, l = 0.0f;
  for (uint32_t i = threadIdx.x; i < n; i += 256) s += dz[i];
  if (loss)
    for (uint32_t i = threadIdx.x; i < nl; i += 256) l += lpart[i];
  sh[0][threadIdx.x] = s;
  sh[1][threadIdx.x] = l;
  __syncthreads();
  for (uint32_t o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + o];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    gb4[0] = sh[0][0];
    if (loss) loss[0] += sh[1][0];
  }
}

// out[r] = sum of row r of a rows x cols bf16 matrix (cols a multiple of
// 8), one workgroup per row, fixed order: the bias gradients as row sums
// of the batch-minor gradient copies.
__global__ void row_sums(const bf16 *T, uint32_t cols, float *out) {
  __shared__ float sh[4];
  const uint4 *row = (const uint4 *)(T + (size_t)blockIdx.x * cols);
  float s = 0.0f;
  for (uint32_t c = threadIdx.x; c < cols / 8; c += 256) {
    const uint4 v = row[c];
    const uint32_t u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int h = 0; h < 4; ++h)
      s += __builtin_bit_cast(float, u[h] << 16) + __builtin_bit_cast(float, u[h] & 0xFFFF0000u);
  }
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// Column sums out[n] = sum_m D[m][n] * (scale ? scale[m] : 1), in two fixed
// passes (partials per kColChunk-row chunk, then their sum) so the result
// does not depend on scheduling.
constexpr uint32_t kColChunk = 256;
__global__ void colsum_partial(const bf16 *D, const float *scale, uint32_t rows, uint32_t cols,
                               float *part) {
  const uint32_t n = blockIdx.x * 64 + (threadIdx.x & 63);
  const uint32_t r0 = blockIdx.y * kColChunk;
  __shared__ float sh[4][64];
  float s = 0.0f;
  if (n < cols) {
    const uint32_t r1 = min(rows, r0 + kColChunk);
#pragma unroll 8
    for (uint32_t m = r0 + (threadIdx.x >> 6); m < r1; m += 4)
      s += (float)D[(size_t)m * cols + n] * (scale ? scale[m] : 1.0f);
  }
  sh[threadIdx.x >> 6][threadIdx.x & 63] = s;
  __syncthreads();
  if (threadIdx.x < 64 && n < cols)
    part[(size_t)blockIdx.y * cols + n] = (sh[0][threadIdx.x] + sh[1][threadIdx.x]) +
                                          (sh[2][threadIdx.x] + sh[3][threadIdx.x]);
}

// out[n] = sum over chunks of part[c][n]: 64 columns per workgroup, the four
// waves take every fourth chunk, then a fixed-order sum of the four.
__global__ void colsum_final(const float *part, uint32_t nchunks, uint32_t cols, float *out) {
  __shared__ float sh[4][64];
  const uint32_t n = blockIdx.x * 64 + (threadIdx.x & 63), w = threadIdx.x >> 6;
  float s = 0.0f;
  if (n < cols) {
#pragma unroll 8
    for (uint32_t c = w; c < nchunks; c += 4) s += part[(size_t)c * cols + n];
  }
  sh[w][threadIdx.x & 63] = s;
  __syncthreads();
  if (threadIdx.x < 64 && n < cols)
    out[n] = (sh[0][threadIdx.x] + sh[1][threadIdx.x]) + (sh[2][threadIdx.x] + sh[3][threadIdx.x]);
}

// Adam (Kingma & Ba 2015) on the fp32 master weights; refreshes the bf16 copy.
__global__ void adam_step(float *w, float *mo, float *ve, const float *g, bf16 *wb, size_t n,
                          float lr, float b1, float b2, float eps, float c1, float c2) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float m = b1 * mo[i] + (1.0f - b1) * gi;
  const float v = b2 * ve[i] + (1.0f - b2) * gi * gi;
  mo[i] = m;
  ve[i] = v;
  const float x = w[i] - lr * (m * c1) / (sqrtf(v * c2) + eps);
  w[i] = x;
  wb[i] = (bf16)x;
}

__global__ void to_bf16(const float *w, bf16 *wb, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) wb[i] = (bf16)w[i];
}

}  // namespace
}  // namespace qcf

// Parameter layout (fp32, flat): W1[h1][kin] b1[h1] W2[h2][h1] b2[h2]
// W3[h3][h2] b3[h3] w4[h3] b4[1]; kin = ndet rounded up to 128 (the padding
// columns of W1 see zero inputs).
struct qcf_mlp {
  int device = 0;
  uint32_t ndet = 0, kin = 0, h1 = 0, h2 = 0, h3 = 0, batch = 0, ibatch = 0, obs_col = 0;
  size_t off[8] = {};
  size_t nparam = 0;
  uint64_t step = 0;
  uint32_t *det_cols = nullptr;
  float *w = nullptr, *mo = nullptr, *ve = nullptr, *g = nullptr;
  bf16 *wb = nullptr;
  // activations, sized for max(batch, ibatch) rows
  bf16 *x = nullptr, *a1 = nullptr, *a2 = nullptr, *a3 = nullptr;
  bf16 *d1 = nullptr, *d2 = nullptr, *d3 = nullptr;
  bf16 *xT = nullptr, *a1T = nullptr, *a2T = nullptr, *dT = nullptr;  // batch-minor copies
  bf16 *WT = nullptr;  // transposed weights for the input gradients
  bf16 *dg1 = nullptr, *dg2 = nullptr, *dg3 = nullptr;  // gelu' at the pre-activations
  float *y = nullptr, *z = nullptr, *dz = nullptr, *part = nullptr, *lossbuf = nullptr;
  float *lpart = nullptr;  // per-workgroup loss partials of the head
  float *kpart = nullptr;  // split-K partial sums of the weight gradients
  std::vector<void *> allocs;
  ~qcf_mlp() {
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(device);
    for (void *p : allocs) (void)hipFree(p);
    (void)hipSetDevice(cur);
  }
};

namespace qcf {
namespace {

template <class T>
int dev_alloc(qcf_mlp *M, T **p, size_t n) {
  void *q = nullptr;
  if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess)
    return set_error(QCF_ERR_ALLOC, "hipMalloc of %zu bytes failed", n * sizeof(T));
  M->allocs.push_back(q);
  *p = (T *)q;
  return QCF_OK;
}

int forward(qcf_mlp *M, uint32_t rows, bool train, hipStream_t s) {
  GemmArgs g{};
  const bf16 *W1 = M->wb + M->off[0], *W2 = M->wb + M->off[2], *W3 = M->wb + M->off[4];
  const float *b1 = M->w + M->off[1], *b2 = M->w + M->off[3], *b3 = M->w + M->off[5];
  // training also keeps the batch-minor copies of a1 and a2 that the weight
  // gradients contract (written by the epilogue)
  auto layer = [&](const bf16 *X, int kin, const bf16 *W, const float *b, int nout, bf16 *dg,
                   bf16 *act, bf16 *actT) {
    g = GemmArgs{X, W, (int)rows, nout, kin, kin, kin, nout, nullptr, act, b, nullptr, actT, dg};
    return train ? launch_gemm<true, true, kEpiBiasGelu>(g, s)
                 : launch_gemm<true, true, kEpiBiasGeluInfer>(g, s);
  };
  QCF_TRY(layer(M->x, M->kin, W1, b1, M->h1, M->dg1, M->a1, train ? M->a1T : nullptr));
  QCF_TRY(layer(M->a1, M->h1, W2, b2, M->h2, M->dg2, M->a2, train ? M->a2T : nullptr));
  QCF_TRY(layer(M->a2, M->h2, W3, b3, M->h3, M->dg3, M->a3, nullptr));
  return QCF_OK;
}

int colsum(qcf_mlp *M, const bf16 *D, const float *scale, uint32_t rows, uint32_t cols,
           float *out, hipStream_t s) {
  const uint32_t nch = (rows + kColChunk - 1) / kColChunk;
  QCF_LAUNCH(colsum_partial, dim3((cols + 63) / 64, nch), dim3(256), 0, s, D, scale,
                     rows, cols, M->part);
  QCF_LAUNCH(colsum_final, dim3((cols + 63) / 64), dim3(256), 0, s, M->part, nch, cols,
                     out);
  QCF_HIP_CHECK(hipGetLastError());
  return QCF_OK;
}

int transpose(const bf16 *src, bf16 *dst, uint32_t rows, uint32_t cols, hipStream_t s) {
  QCF_LAUNCH(transpose_bf16, dim3(cols / 64, rows / 64), dim3(256), 0, s,
                     (const uint16_t *)src, (uint16_t *)dst, rows, cols);
  QCF_HIP_CHECK(hipGetLastError());
  return QCF_OK;
}

// Forward + backward of one batch already gathered into x / xT / y;
// gradients in g.  Weight gradients sum over the batch, so they contract two
// batch-minor operands (copies of the activations and of the back-propagated
// pre-activation gradients, written by the kernels that produce them) with
// K-contiguous reads; the bias gradients are row sums of the latter.
int fwd_bwd(qcf_mlp *M, float *loss, hipStream_t s) {
  const uint32_t B = M->batch;
  QCF_TRY(forward(M, B, true, s));
  const bf16 *w4 = M->wb + M->off[6];
  const uint32_t nhead = B / (4 * kHeadRows);
  QCF_LAUNCH(head_forward, dim3(nhead), dim3(256), 0, s, M->a3, w4, M->w + M->off[7],
                     M->h3, B, M->z, M->y, M->dz, M->lpart, 1.0f / (float)B);
  QCF_LAUNCH(head_backward, dim3(M->h3 / 64, B / 64), dim3(256), 0, s, M->dz, w4,
                     M->dg3, M->h3, B, M->d3, M->dT);
  QCF_HIP_CHECK(hipGetLastError());
  float *G = M->g;
  // output layer: grad w4 = a3^T dz, grad b4 = sum dz (and the loss)
  QCF_TRY(colsum(M, M->a3, M->dz, B, M->h3, G + M->off[6], s));
  QCF_LAUNCH(head_sums, dim3(1), dim3(256), 0, s, M->dz, B, M->lpart, nhead,
                     G + M->off[7], loss);
  const bf16 *W2 = M->wb + M->off[2], *W3 = M->wb + M->off[4];
  // one layer back, with dT = d^T already in M->dT: grad W = dT aT^T, grad b
  // = row sums of dT, and (unless first) d_prev = (d W) * dg_prev,
  // whose epilogue also writes d_prev^T into M->dT (every reader of the old
  // contents is earlier on the stream)
  auto back = [&](const bf16 *d, uint32_t hout, const bf16 *aT, uint32_t hin, float *gW,
                  float *gb, const bf16 *W, const bf16 *dg_prev, bf16 *d_prev) -> int {
    GemmArgs g{M->dT, aT, (int)hout, (int)hin, (int)B, (int)B, (int)B, (int)hin,
               gW, nullptr, nullptr, nullptr, nullptr, nullptr};
    // few output tiles, long K: split K so that ~splitk_workgroups() run
    const int tiles = (int)((hout / kTile) * (hin / kTile));
    int sk = 1;
    while (sk * tiles < splitk_workgroups() && (int)B % (2 * sk * 256) == 0 &&
           2 * sk <= kMaxSplitK)
      sk *= 2;
    QCF_TRY((launch_gemm<true, true, kEpiStoreF32>(g, s, sk, M->kpart)));
    QCF_LAUNCH(row_sums, dim3(hout), dim3(256), 0, s, M->dT, B, gb);
    QCF_HIP_CHECK(hipGetLastError());
    if (d_prev) {
      // W^T (hin x hout) so both operands are read K-contiguous
      QCF_TRY(transpose(W, M->WT, hout, hin, s));
      g = GemmArgs{d, M->WT, (int)B, (int)hin, (int)hout, (int)hout, (int)hout, (int)hin,
                   nullptr, d_prev, nullptr, dg_prev, M->dT, nullptr};
      QCF_TRY((launch_gemm<true, true, kEpiDGelu>(g, s)));
    }
    return QCF_OK;
  };
  QCF_TRY(back(M->d3, M->h3, M->a2T, M->h2, G + M->off[4], G + M->off[5], W3, M->dg2, M->d2));
  QCF_TRY(back(M->d2, M->h2, M->a1T, M->h1, G + M->off[2], G + M->off[3], W2, M->dg1, M->d1));
  QCF_TRY(back(M->d1, M->h1, M->xT, M->kin, G + M->off[0], G + M->off[1], nullptr, nullptr,
               nullptr));
  return QCF_OK;
}

int gather(qcf_mlp *M, const uint64_t *rec, size_t ld, uint64_t s0, uint64_t nshots,
           uint32_t rows, bool train, hipStream_t s) {
  if (s0 % 64) return set_error(QCF_ERR_ARGUMENT, "first shot %llu is not a multiple of 64",
                                (unsigned long long)s0);
  // one workgroup per 64 shots and 128 input columns
  QCF_LAUNCH(gather_inputs, dim3(rows / 64, M->kin / 128), dim3(256), 0, s, rec,
                     ld, s0, nshots,
                     M->det_cols, M->ndet, M->kin, M->obs_col, train ? M->y : nullptr,
                     (uint16_t *)M->x, train ? (uint16_t *)M->xT : nullptr, rows);
  QCF_HIP_CHECK(hipGetLastError());
  return QCF_OK;
}

}  // namespace
}  // namespace qcf

using namespace qcf;

extern "C" int qcf_mlp_create(const uint32_t *det_bits, uint32_t ndet, uint32_t obs_bit,
                              uint32_t h1, uint32_t h2, uint32_t h3, uint32_t batch,
                              uint32_t infer_batch, qcf_mlp **out) {
  clear_error();
  if (!out || (ndet && !det_bits)) return set_error(QCF_ERR_ARGUMENT, "null argument");
  *out = nullptr;
  if (ndet == 0 || obs_bit == 0) return set_error(QCF_ERR_ARGUMENT, "need detectors and an observable");
  for (uint32_t v : {h1, h2, h3, batch, infer_batch})
    if (v == 0 || v % 128)
      return set_error(QCF_ERR_ARGUMENT, "layer widths and batches must be multiples of 128 (got %u)", v);
  for (uint32_t j = 0; j < ndet; ++j)
    if (det_bits[j] == 0) return set_error(QCF_ERR_ARGUMENT, "record bits are 1-based");
  qcf_mlp *M = new qcf_mlp();
  (void)hipGetDevice(&M->device);
  M->ndet = ndet;
  M->kin = (ndet + 127) / 128 * 128;
  M->h1 = h1;
  M->h2 = h2;
  M->h3 = h3;
  M->batch = batch;
  M->ibatch = infer_batch;
  M->obs_col = obs_bit - 1;
  size_t o = 0;
  const size_t sizes[8] = {(size_t)h1 * M->kin, h1, (size_t)h2 * h1, h2, (size_t)h3 * h2, h3, h3, 1};
  for (int i = 0; i < 8; ++i) {
    M->off[i] = o;
    o += sizes[i];
  }
  M->nparam = o;
  const size_t R = std::max(batch, infer_batch);
  int st = QCF_OK;
  auto A = [&](auto **p, size_t n) {
    if (st == QCF_OK) st = dev_alloc(M, p, n);
  };
  A(&M->det_cols, ndet);
  A(&M->w, o);
  A(&M->mo, o);
  A(&M->ve, o);
  A(&M->g, o);
  A(&M->wb, o);
  A(&M->x, R * M->kin);
  A(&M->a1, R * h1);
  A(&M->a2, R * h2);
  A(&M->a3, R * h3);
  A(&M->d1, (size_t)batch * h1);
  A(&M->d2, (size_t)batch * h2);
  A(&M->d3, (size_t)batch * h3);
  A(&M->xT, (size_t)batch * M->kin);
  A(&M->a1T, (size_t)batch * h1);
  A(&M->WT, std::max((size_t)h2 * h1, (size_t)h3 * h2));
  A(&M->a2T, (size_t)batch * h2);
  A(&M->dT, (size_t)batch * std::max(std::max(h1, h2), h3));
  A(&M->dg1, (size_t)batch * h1);
  A(&M->dg2, (size_t)batch * h2);
  A(&M->dg3, (size_t)batch * h3);
  A(&M->y, R);
  A(&M->z, R);
  A(&M->dz, R);
  A(&M->lossbuf, 1);
  A(&M->lpart, (size_t)batch / (4 * kHeadRows));
  A(&M->kpart, (size_t)kMaxSplitK * std::max((size_t)h1 * M->kin, std::max((size_t)h2 * h1, (size_t)h3 * h2)));
  A(&M->part, (size_t)((batch + 127) / 128) * std::max(std::max(h1, h2), std::max(h3, M->kin)));
  if (st == QCF_OK) {
    std::vector<uint32_t> cols(ndet);
    for (uint32_t j = 0; j < ndet; ++j) cols[j] = det_bits[j] - 1;
    if (hipMemcpy(M->det_cols, cols.data(), ndet * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(M->w, 0, o * 4) != hipSuccess || hipMemset(M->wb, 0, o * 2) != hipSuccess ||
        hipMemset(M->mo, 0, o * 4) != hipSuccess || hipMemset(M->ve, 0, o * 4) != hipSuccess)
      st = set_error(QCF_ERR_HIP, "initialising the decoder buffers failed");
  }
  if (st != QCF_OK) {
    delete M;
    return st;
  }
  *out = M;
  return QCF_OK;
}

extern "C" int qcf_mlp_free(qcf_mlp *m) {
  delete m;
  return QCF_OK;
}

extern "C" int qcf_mlp_nparams(const qcf_mlp *m, size_t *n) {
  clear_error();
  if (!m || !n) return set_error(QCF_ERR_ARGUMENT, "null argument");
  *n = m->nparam;
  return QCF_OK;
}

extern "C" int qcf_mlp_set_params(qcf_mlp *m, const float *host, void *stream) {
  clear_error();
  if (!m || !host) return set_error(QCF_ERR_ARGUMENT, "null argument");
  hipStream_t s = (hipStream_t)stream;
  QCF_HIP_CHECK(hipMemcpyAsync(m->w, host, m->nparam * 4, hipMemcpyHostToDevice, s));
  QCF_HIP_CHECK(hipMemsetAsync(m->mo, 0, m->nparam * 4, s));
  QCF_HIP_CHECK(hipMemsetAsync(m->ve, 0, m->nparam * 4, s));
  QCF_LAUNCH(to_bf16, dim3((unsigned)((m->nparam + 255) / 256)), dim3(256), 0, s, m->w,
                     m->wb, m->nparam);
  QCF_HIP_CHECK(hipGetLastError());
  QCF_HIP_CHECK(hipStreamSynchronize(s));
  m->step = 0;
  return QCF_OK;
}

extern "C" int qcf_mlp_get_params(const qcf_mlp *m, float *host, void *stream) {
  clear_error();
  if (!m || !host) return set_error(QCF_ERR_ARGUMENT, "null argument");
  QCF_HIP_CHECK(hipMemcpyAsync(host, m->w, m->nparam * 4, hipMemcpyDeviceToHost,
                               (hipStream_t)stream));
  QCF_HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
  return QCF_OK;
}

extern "C" int qcf_mlp_gradients(qcf_mlp *m, const uint64_t *rec, size_t ld, uint64_t shot0,
                                 float *grad_host, float *loss_host, void *stream) {
  clear_error();
  if (!m || !rec || !grad_host || !loss_host) return set_error(QCF_ERR_ARGUMENT, "null argument");
  if (ld * 64 < shot0 + m->batch)
    return set_error(QCF_ERR_DIMENSION, "record holds fewer than shot0 + batch shots");
  hipStream_t s = (hipStream_t)stream;
  QCF_TRY(gather(m, rec, ld, shot0, shot0 + m->batch, m->batch, true, s));
  float *loss = m->lossbuf;
  QCF_HIP_CHECK(hipMemsetAsync(loss, 0, 4, s));
  QCF_TRY(fwd_bwd(m, loss, s));
  QCF_HIP_CHECK(hipMemcpyAsync(grad_host, m->g, m->nparam * 4, hipMemcpyDeviceToHost, s));
  QCF_HIP_CHECK(hipMemcpyAsync(loss_host, loss, 4, hipMemcpyDeviceToHost, s));
  QCF_HIP_CHECK(hipStreamSynchronize(s));
  return QCF_OK;
}

extern "C" int qcf_mlp_train_step(qcf_mlp *m, const uint64_t *rec, size_t ld, uint64_t shot0,
                                  float lr, float beta1, float beta2, float eps,
                                  float *loss_accum, void *stream) {
  clear_error();
  if (!m || !rec) return set_error(QCF_ERR_ARGUMENT, "null argument");
  if (ld * 64 < shot0 + m->batch)
    return set_error(QCF_ERR_DIMENSION, "record holds fewer than shot0 + batch shots");
  hipStream_t s = (hipStream_t)stream;
  QCF_TRY(gather(m, rec, ld, shot0, shot0 + m->batch, m->batch, true, s));
  QCF_TRY(fwd_bwd(m, loss_accum, s));
  m->step += 1;
  const float c1 = 1.0f / (1.0f - std::pow(beta1, (float)m->step));
  const float c2 = 1.0f / (1.0f - std::pow(beta2, (float)m->step));
  QCF_LAUNCH(adam_step, dim3((unsigned)((m->nparam + 255) / 256)), dim3(256), 0, s, m->w,
                     m->mo, m->ve, m->g, m->wb, m->nparam, lr, beta1, beta2, eps, c1, c2);
  QCF_HIP_CHECK(hipGetLastError());
  return QCF_OK;
}

extern "C" int qcf_mlp_infer(qcf_mlp *m, const uint64_t *rec, size_t ld, uint64_t shots,
                             float *logits, void *stream) {
  clear_error();
  if (!m || (shots && (!rec || !logits))) return set_error(QCF_ERR_ARGUMENT, "null argument");
  if (ld * 64 < shots) return set_error(QCF_ERR_DIMENSION, "ld %zu words < %llu shots", ld,
                                         (unsigned long long)shots);
  hipStream_t s = (hipStream_t)stream;
  for (uint64_t s0 = 0; s0 < shots; s0 += m->ibatch) {
    const uint32_t rows = (uint32_t)std::min<uint64_t>(m->ibatch, shots - s0);
    const uint32_t rows128 = (rows + 127) / 128 * 128;
    QCF_TRY(gather(m, rec, ld, s0, shots, rows128, false, s));
    QCF_TRY(forward(m, rows128, false, s));
    QCF_LAUNCH(head_forward, dim3(rows128 / (4 * kHeadRows)), dim3(256), 0, s, m->a3,
                       m->wb + m->off[6], m->w + m->off[7], m->h3, rows, logits + s0,
                       (const float *)nullptr, (float *)nullptr, (float *)nullptr, 0.0f);
    QCF_HIP_CHECK(hipGetLastError());
  }
  return QCF_OK;
}

extern "C" int qcf_debug_gemm_bf16(int layout, int M, int N, int K, const void *A, const void *B,
                                   float *C, void *stream) {
  clear_error();
  GemmArgs g{(const bf16 *)A, (const bf16 *)B, M, N, K, 0, 0, N, C, nullptr, nullptr, nullptr,
             nullptr, nullptr};
  hipStream_t s = (hipStream_t)stream;
  switch (layout) {
    case 0:  // A[M][K], B[N][K]
      g.lda = K;
      g.ldb = K;
      return launch_gemm<true, true, kEpiStoreF32>(g, s);
    case 1:  // A[M][K], B[K][N]
      g.lda = K;
      g.ldb = N;
      return launch_gemm<true, false, kEpiStoreF32>(g, s);
    case 2:  // A[K][M], B[K][N]
      g.lda = M;
      g.ldb = N;
      return launch_gemm<false, false, kEpiStoreF32>(g, s);
    default:
      return set_error(QCF_ERR_ARGUMENT, "layout %d", layout);
  }
}
