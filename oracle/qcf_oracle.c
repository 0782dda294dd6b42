/*
 * qcf_oracle.c -- CPU restatement of the QuantumClifford.jl v0.11.7 hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP library is compared
 * against; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it.  The product (libqcf) never links, loads or calls it.
 *
 * Layouts are the reference's own, not the device library's:
 *   - Pauli operators: xz = [x words ; z words] (src/pauli_operator.jl:42-56,
 *     src/QuantumClifford.jl:994-1001);
 *   - Pauli frames: the `fastcolumn` xzs of PauliFrame.frame, a shots x 2W
 *     column-major UInt64 matrix, W = ceil(nq/64) with nq = qubits + 1
 *     (src/pauli_frames.jl:30-37, src/fastmemlayout.jl:47): element
 *     (shot f, word row j) at xzs[j*shots + f];
 *   - measurements: Matrix{Bool}(shots, nbits), column-major bytes
 *     (src/pauli_frames.jl:33).
 * Every random draw of the reference (initZ!, sMRZ, PauliError/PauliNoise)
 * is either replayed from shot-packed mask planes (the interchange format of
 * qcf_pf_sample) or, with masks == NULL, drawn per shot from xoshiro256++ as
 * the reference does with Julia's default RNG (used only to time a CPU
 * baseline; that stream is not reproducible across thread counts, like the
 * reference's Threads.@threads split, src/pauli_frames.jl:215-222).
 *
 * Parity unpinned: the two-qubit depolarizing channel (the reference has no
 * PauliFrame method for it; only its closed-form rates and the base-4 digit
 * order of src/noise.jl:112-118 check it) and the z pad bits that initZ!
 * randomizes past the last qubit (written as zeros here; no operation reads
 * them).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/qcf.h"

static inline int popc64(uint64_t v) { return __builtin_popcountll(v); }

/* ====================================================================== */
/* mul_left! / mul_right!  (src/mul_leftright.jl)                          */
/* ====================================================================== */

/* _mul_ordered_nonvec! (src/mul_leftright.jl:11-31): r := l*r in place,
 * returns the two counters' popcounts. */
void oracle_mul_ordered(uint64_t *r, const uint64_t *l, size_t len, int phases,
                        uint64_t *rcnt1, uint64_t *rcnt2) {
  if (!phases) {
    for (size_t i = 0; i < 2 * len; ++i) r[i] ^= l[i];
    *rcnt1 = 0;
    *rcnt2 = 0;
    return;
  }
  uint64_t cnt1 = 0, cnt2 = 0;
  for (size_t i = 0; i < len; ++i) {
    const uint64_t x1 = l[i], x2 = r[i], z1 = l[i + len], z2 = r[i + len];
    const uint64_t newx1 = x1 ^ x2, newz1 = z1 ^ z2;
    r[i] = newx1;
    r[i + len] = newz1;
    const uint64_t x1z2 = x1 & z2;
    const uint64_t anti_comm = (x2 & z1) ^ x1z2;
    cnt2 ^= (cnt1 ^ newx1 ^ newz1 ^ x1z2) & anti_comm;
    cnt1 ^= anti_comm;
  }
  *rcnt1 = (uint64_t)popc64(cnt1);
  *rcnt2 = (uint64_t)popc64(cnt2);
}

/* mul_left!(r::AbstractVector, l::AbstractVector)::UInt8 (:123-126) */
uint8_t oracle_mul_left_vec(uint64_t *r, const uint64_t *l, size_t len, int phases) {
  uint64_t c1, c2;
  oracle_mul_ordered(r, l, len, phases, &c1, &c2);
  return (uint8_t)((c1 ^ (c2 << 1)) & 3u);
}

/* mul_right!(l::AbstractVector, r::AbstractVector)::UInt8 (:128-131) */
uint8_t oracle_mul_right_vec(uint64_t *l, const uint64_t *r, size_t len, int phases) {
  uint64_t c1, c2;
  oracle_mul_ordered(l, r, len, phases, &c1, &c2);
  return (uint8_t)(((c1 ^ (c2 << 1)) + c1 * 2) & 3u);
}

/* mul_left!(r::PauliOperator, l::PauliOperator) (:137-142) */
void oracle_mul_left_pauli(uint64_t *r, uint8_t *rphase, const uint64_t *l,
                           const uint8_t *lphase, size_t len, int phases) {
  const uint8_t s = oracle_mul_left_vec(r, l, len, phases);
  if (phases) rphase[0] = (uint8_t)((s + rphase[0] + lphase[0]) & 3u);
}

/* mul_right!(l::PauliOperator, r::PauliOperator) (:144-149) */
void oracle_mul_right_pauli(uint64_t *l, uint8_t *lphase, const uint64_t *r,
                            const uint8_t *rphase, size_t len, int phases) {
  const uint8_t s = oracle_mul_right_vec(l, r, len, phases);
  if (phases) lphase[0] = (uint8_t)((s + rphase[0] + lphase[0]) & 3u);
}

/* Row loops of a fastrow tableau (row m at xzs + m*ld):
 *   mul_left!(s::Tableau, m, i)  (:177-181), swept over m != i as in
 *   canonicalize! (src/canonicalization.jl:105-109), or
 *   mul_left!(s::Tableau, p::PauliOperator) (:221-227) with skip = SIZE_MAX.
 * right != 0 gives mul_right! (:189-193, :234-240). */
void oracle_mul_rows(uint64_t *xzs, uint8_t *phases, size_t ld, size_t len,
                     size_t nrows, const uint64_t *src, const uint8_t *src_phase,
                     size_t skip, const uint8_t *row_mask, int phases_on, int right) {
  for (size_t m = 0; m < nrows; ++m) {
    if (m == skip) continue;
    if (row_mask && !row_mask[m]) continue;
    uint64_t *row = xzs + m * ld;
    const uint8_t s = right ? oracle_mul_right_vec(row, src, len, phases_on)
                            : oracle_mul_left_vec(row, src, len, phases_on);
    if (phases_on) phases[m] = (uint8_t)((s + phases[m] + src_phase[0]) & 3u);
  }
}

/* canonicalize!(state::Stabilizer; phases, ranks) (src/canonicalization.jl:97-135)
 * on a row-contiguous tableau (row m at xzs + m*ld, words [x 0..len); z
 * len..2len)).  X stage: for every qubit column j, the first row k >= i with
 * an X (or Y) at j is swapped into row i (rowswap!, src/QuantumClifford.jl:
 * 899-905: the phases move only when phases are tracked), then every other row
 * m with an X at j becomes row_i * row_m (mul_left!(s, m, i; phases),
 * src/mul_leftright.jl:195-198); i advances.  Z stage: the same on Z bits,
 * continuing from the X stage's i.  ranks[0] = pivots of the X stage,
 * ranks[1] = all pivots (the reference's rx-1, i-1). */
static inline int tab_bit(const uint64_t *row, size_t len, size_t j, int z) {
  return (int)((row[(z ? len : 0) + j / 64] >> (j % 64)) & 1u);
}
static void tab_rowswap(uint64_t *xzs, uint8_t *phases, size_t ld, size_t len, size_t a,
                        size_t b, int phases_on) {
  if (a == b) return;
  if (phases_on) {
    const uint8_t t = phases[a];
    phases[a] = phases[b];
    phases[b] = t;
  }
  for (size_t w = 0; w < 2 * len; ++w) {
    const uint64_t t = xzs[a * ld + w];
    xzs[a * ld + w] = xzs[b * ld + w];
    xzs[b * ld + w] = t;
  }
}
void oracle_canonicalize(uint64_t *xzs, uint8_t *phases, size_t ld, size_t len, size_t nrows,
                         size_t nqubits, int phases_on, uint64_t ranks[2]) {
  size_t i = 0;
  for (int z = 0; z < 2; ++z) {
    for (size_t j = 0; j < nqubits; ++j) {
      size_t k = i;
      while (k < nrows && !tab_bit(xzs + k * ld, len, j, z)) ++k;
      if (k >= nrows) continue;
      tab_rowswap(xzs, phases, ld, len, k, i, phases_on);
      for (size_t m = 0; m < nrows; ++m) {
        if (m == i || !tab_bit(xzs + m * ld, len, j, z)) continue;
        const uint8_t s = oracle_mul_left_vec(xzs + m * ld, xzs + i * ld, len, phases_on);
        if (phases_on) phases[m] = (uint8_t)((s + phases[m] + phases[i]) & 3u);
      }
      ++i;
    }
    if (ranks) ranks[z] = i;
  }
}

/* prodphase(l, r) (src/QuantumClifford.jl:723-732): independent formula */
uint8_t oracle_prodphase(const uint64_t *l, const uint64_t *r, size_t len) {
  int64_t res = 0;
  for (size_t i = 0; i < len; ++i) {
    const uint64_t x1 = l[i], x2 = r[i], z1 = l[i + len], z2 = r[i + len];
    res += popc64((~z2 & x2 & ~x1 & z1) | (z2 & ~x2 & x1 & z1) | (z2 & x2 & x1 & ~z1));
    res -= popc64((z2 & x2 & ~x1 & z1) | (~z2 & x2 & x1 & z1) | (z2 & ~x2 & x1 & ~z1));
  }
  return (uint8_t)((uint64_t)res & 3u);
}

/* comm(l, r) (src/QuantumClifford.jl:793-800) */
uint8_t oracle_comm(const uint64_t *l, const uint64_t *r, size_t len) {
  uint64_t res = 0;
  for (size_t i = 0; i < len; ++i) res ^= (l[i + len] & r[i]) ^ (l[i] & r[i + len]);
  return (uint8_t)(popc64(res) & 1);
}

/* ====================================================================== */
/* Philox4x32-10 (Salmon et al., SC'11) -- pins the device RNG by KAT      */
/* ====================================================================== */
void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* ====================================================================== */
/* Pauli frames (src/pauli_frames.jl)                                      */
/* ====================================================================== */

/* qubit_kernel bit maps, phases dropped (src/symbolic_cliffords.jl:105-118,
 * :168-206 for SingleQubitOperator, :359-395).  Returns 0 for unknown kinds. */
static int gate_kernel(uint32_t kind, const uint32_t *m, uint64_t *x1p, uint64_t *z1p,
                       uint64_t *x2p, uint64_t *z2p) {
  const uint64_t x1 = *x1p, z1 = *z1p, x2 = *x2p, z2 = *z2p;
  uint64_t a = x1, b = z1, c = x2, d = z2;
  switch (kind) {
    /* single qubit: (x, z) -> (a, b) */
    case QCF_OP_ID1: case QCF_OP_X: case QCF_OP_Y: case QCF_OP_Z: break;
    case QCF_OP_HADAMARD: a = z1; b = x1; break;
    case QCF_OP_HADAMARD_XY: a = x1; b = x1 ^ z1; break;
    case QCF_OP_HADAMARD_YZ: a = x1 ^ z1; b = z1; break;
    case QCF_OP_PHASE: case QCF_OP_INV_PHASE: a = x1; b = x1 ^ z1; break;
    case QCF_OP_SQRTX: case QCF_OP_INV_SQRTX: a = x1 ^ z1; b = z1; break;
    case QCF_OP_SQRTY: case QCF_OP_INV_SQRTY: a = z1; b = x1; break;
    case QCF_OP_CXYZ: a = x1 ^ z1; b = x1; break;
    case QCF_OP_CZYX: a = z1; b = x1 ^ z1; break;
    case QCF_OP_SINGLE_QUBIT_OPERATOR: {
      const uint64_t xx = m[0] ? ~0ull : 0, xz = m[1] ? ~0ull : 0;
      const uint64_t zx = m[2] ? ~0ull : 0, zz = m[3] ? ~0ull : 0;
      a = (x1 & xx) ^ (z1 & zx);
      b = (x1 & xz) ^ (z1 & zz);
      break;
    }
    /* two qubit: (x1, z1, x2, z2) -> (a, b, c, d) */
    case QCF_OP_SWAP: a = x2; b = z2; c = x1; d = z1; break;
    case QCF_OP_SWAPCX: a = x2; b = z2 ^ z1; c = x2 ^ x1; d = z1; break;
    case QCF_OP_INV_SWAPCX: a = x2 ^ x1; b = z2; c = x1; d = z2 ^ z1; break;
    case QCF_OP_ISWAP: case QCF_OP_INV_ISWAP:
      a = x2; b = x1 ^ z2 ^ x2; c = x1; d = x1 ^ x2 ^ z1; break;
    case QCF_OP_CZSWAP: a = x2; b = z2 ^ x1; c = x1; d = x2 ^ z1; break;
    case QCF_OP_CXSWAP: a = x2 ^ x1; b = z2; c = x1; d = z2 ^ z1; break;
    case QCF_OP_CNOT: case QCF_OP_ZCX: a = x1; b = z1 ^ z2; c = x2 ^ x1; d = z2; break;
    case QCF_OP_CPHASE: case QCF_OP_ZCZ: a = x1; b = z1 ^ x2; c = x2; d = z2 ^ x1; break;
    case QCF_OP_ZCY: a = x1; b = x2 ^ z1 ^ z2; c = x2 ^ x1; d = z2 ^ x1; break;
    case QCF_OP_XCX: a = z2 ^ x1; b = z1; c = x2 ^ z1; d = z2; break;
    case QCF_OP_XCY: a = x1 ^ x2 ^ z2; b = z1; c = z1 ^ x2; d = z2 ^ z1; break;
    case QCF_OP_XCZ: a = x1 ^ x2; b = z1; c = x2; d = z2 ^ z1; break;
    case QCF_OP_YCX: a = x1 ^ z2; b = z2 ^ z1; c = x1 ^ z1 ^ x2; d = z2; break;
    case QCF_OP_YCY: a = x1 ^ z2 ^ x2; b = z1 ^ x2 ^ z2; c = x1 ^ x2 ^ z1; d = x1 ^ z1 ^ z2; break;
    case QCF_OP_YCZ: a = x1 ^ x2; b = x2 ^ z1; c = x2; d = z2 ^ x1 ^ z1; break;
    case QCF_OP_ZCRY: case QCF_OP_INV_ZCRY:
      a = x1; b = x1 ^ z1 ^ x2 ^ z2; c = x1 ^ x2; d = x1 ^ z2; break;
    case QCF_OP_SQRTZZ: case QCF_OP_INV_SQRTZZ: a = x1; b = x1 ^ x2 ^ z1; c = x2; d = x1 ^ z2 ^ x2; break;
    case QCF_OP_SQRTXX: case QCF_OP_INV_SQRTXX: a = z1 ^ z2 ^ x1; b = z1; c = z1 ^ x2 ^ z2; d = z2; break;
    case QCF_OP_SQRTYY: case QCF_OP_INV_SQRTYY:
      a = z1 ^ x2 ^ z2; b = x1 ^ z2 ^ x2; c = x1 ^ z1 ^ z2; d = x1 ^ x2 ^ z1; break;
    default: return 0;
  }
  *x1p = a; *z1p = b; *x2p = c; *z2p = d;
  return 1;
}

static int is_two_qubit(uint32_t kind) { return kind >= QCF_OP_SWAP && kind <= QCF_OP_INV_SQRTYY; }
static int is_one_qubit(uint32_t kind) {
  return (kind >= QCF_OP_ID1 && kind <= QCF_OP_CZYX) || kind == QCF_OP_SINGLE_QUBIT_OPERATOR;
}

/* xoshiro256++ -- the generator family of Julia's default task-local RNG */
typedef struct { uint64_t s[4]; } xo256;
static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static inline uint64_t xo_next(xo256 *g) {
  const uint64_t r = rotl64(g->s[0] + g->s[3], 23) + g->s[0];
  const uint64_t t = g->s[1] << 17;
  g->s[2] ^= g->s[0]; g->s[3] ^= g->s[1]; g->s[1] ^= g->s[2]; g->s[0] ^= g->s[3];
  g->s[2] ^= t; g->s[3] = rotl64(g->s[3], 45);
  return r;
}
static inline double xo_double(xo256 *g) { return (double)(xo_next(g) >> 11) * 0x1.0p-53; }
static void xo_seed(xo256 *g, uint64_t seed) {
  uint64_t z = seed;
  for (int i = 0; i < 4; ++i) {  /* splitmix64 expansion */
    z += 0x9E3779B97F4A7C15ull;
    uint64_t v = z;
    v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
    v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
    g->s[i] = v ^ (v >> 31);
  }
}

typedef struct {
  uint64_t *xzs;
  uint8_t *meas;
  uint64_t S;       /* total shots (column stride) */
  uint32_t W;       /* words per half */
  const uint64_t *masks;
  size_t ld;
  xo256 rng;
} frame_ctx;

#define XW(ctx, q) ((ctx)->xzs + (size_t)(((q) - 1) / 64) * (ctx)->S)
#define ZW(ctx, q) ((ctx)->xzs + (size_t)((ctx)->W + ((q) - 1) / 64) * (ctx)->S)
#define QMASK(q) (1ull << (((q) - 1) % 64))

static inline uint64_t mask_bit(const frame_ctx *c, uint32_t plane, uint64_t f) {
  return (c->masks[(size_t)plane * c->ld + f / 64] >> (f % 64)) & 1u;
}

/* _apply!(stab, gate; phases=Val(false)) per frame row (symbolic_cliffords.jl:62-74, 293-316) */
static void apply_gate(frame_ctx *c, uint32_t kind, uint32_t q1, uint32_t q2,
                       const uint32_t *m, uint64_t f0, uint64_t f1) {
  uint64_t *xa = XW(c, q1), *za = ZW(c, q1);
  const uint64_t ma = QMASK(q1);
  if (is_two_qubit(kind)) {
    uint64_t *xb = XW(c, q2), *zb = ZW(c, q2);
    const uint64_t mb = QMASK(q2);
    for (uint64_t f = f0; f < f1; ++f) {
      uint64_t x1 = (xa[f] & ma) ? 1 : 0, z1 = (za[f] & ma) ? 1 : 0;
      uint64_t x2 = (xb[f] & mb) ? 1 : 0, z2 = (zb[f] & mb) ? 1 : 0;
      gate_kernel(kind, m, &x1, &z1, &x2, &z2);
      xa[f] = (xa[f] & ~ma) | ((x1 & 1) ? ma : 0);
      za[f] = (za[f] & ~ma) | ((z1 & 1) ? ma : 0);
      xb[f] = (xb[f] & ~mb) | ((x2 & 1) ? mb : 0);
      zb[f] = (zb[f] & ~mb) | ((z2 & 1) ? mb : 0);
    }
  } else {
    for (uint64_t f = f0; f < f1; ++f) {
      uint64_t x1 = (xa[f] & ma) ? 1 : 0, z1 = (za[f] & ma) ? 1 : 0, u = 0, v = 0;
      gate_kernel(kind, m, &x1, &z1, &u, &v);
      xa[f] = (xa[f] & ~ma) | ((x1 & 1) ? ma : 0);
      za[f] = (za[f] & ~ma) | ((z1 & 1) ? ma : 0);
    }
  }
}

/* apply!(f::PauliFrame, op::sMZ) (:85-97) */
static void apply_mz(frame_ctx *c, uint32_t q, uint32_t bit, uint64_t f0, uint64_t f1) {
  if (bit == 0) return;
  const uint64_t *xa = XW(c, q);
  const uint64_t ma = QMASK(q);
  uint8_t *col = c->meas + (size_t)(bit - 1) * c->S;
  for (uint64_t f = f0; f < f1; ++f) col[f] = (xa[f] & ma) != 0;
}

/* apply!(f::PauliFrame, op::sMRZ) (:99-116) */
static void apply_mrz(frame_ctx *c, uint32_t q, uint32_t bit, uint32_t *rs, uint64_t f0,
                      uint64_t f1) {
  uint64_t *xa = XW(c, q), *za = ZW(c, q);
  const uint64_t ma = QMASK(q);
  if (bit != 0) {
    uint8_t *col = c->meas + (size_t)(bit - 1) * c->S;
    for (uint64_t f = f0; f < f1; ++f) col[f] = (xa[f] & ma) != 0;
  }
  const uint32_t plane = (*rs)++;
  for (uint64_t f = f0; f < f1; ++f) {
    xa[f] &= ~ma;
    const int flip = c->masks ? (int)mask_bit(c, plane, f) : (int)(xo_next(&c->rng) >> 63);
    if (flip) za[f] ^= ma;
  }
}

/* applynoise! for UnbiasedUncorrelatedNoise (:136-155) and PauliNoise
 * (:157-174): thresholds t1 < t2 < t3 select X, Z, Y from one rand(). */
static void apply_noise1(frame_ctx *c, uint32_t q, double t1, double t2, double t3,
                         uint32_t *rs, uint64_t f0, uint64_t f1) {
  uint64_t *xa = XW(c, q), *za = ZW(c, q);
  const uint64_t ma = QMASK(q);
  const uint32_t plane = *rs;
  *rs += 2;
  for (uint64_t f = f0; f < f1; ++f) {
    if (c->masks) {
      if (mask_bit(c, plane, f)) xa[f] ^= ma;
      if (mask_bit(c, plane + 1, f)) za[f] ^= ma;
      continue;
    }
    const double r = xo_double(&c->rng);
    if (r < t1) {
      xa[f] ^= ma;
    } else if (r < t2) {
      za[f] ^= ma;
    } else if (r < t3) {
      xa[f] ^= ma;
      za[f] ^= ma;
    }
  }
}

/* Two-qubit depolarizing: with probability e one of the 15 non-identity
 * two-qubit Paulis, digits as _apply_depolarization_pauli! (src/noise.jl:112-118). */
static void apply_dep2(frame_ctx *c, uint32_t q1, uint32_t q2, double e, uint32_t *rs,
                       uint64_t f0, uint64_t f1) {
  uint64_t *xa = XW(c, q1), *za = ZW(c, q1), *xb = XW(c, q2), *zb = ZW(c, q2);
  const uint64_t ma = QMASK(q1), mb = QMASK(q2);
  const uint32_t plane = *rs;
  *rs += 4;
  for (uint64_t f = f0; f < f1; ++f) {
    int d1, d2;
    if (c->masks) {
      if (mask_bit(c, plane, f)) xa[f] ^= ma;
      if (mask_bit(c, plane + 1, f)) za[f] ^= ma;
      if (mask_bit(c, plane + 2, f)) xb[f] ^= mb;
      if (mask_bit(c, plane + 3, f)) zb[f] ^= mb;
      continue;
    }
    if (xo_double(&c->rng) >= e) continue;
    const int code = 1 + (int)(xo_next(&c->rng) % 15);
    d1 = code & 3;
    d2 = code >> 2;
    if (d1 & 1) xa[f] ^= ma;
    if (d1 & 2) za[f] ^= ma;
    if (d2 & 1) xb[f] ^= mb;
    if (d2 & 2) zb[f] ^= mb;
  }
}

/* apply!(frame::PauliFrame, xor::ClassicalXOR) (:63-72) */
static void apply_cxor(frame_ctx *c, const uint32_t *bits, uint32_t n, uint32_t store,
                       uint64_t f0, uint64_t f1) {
  for (uint64_t f = f0; f < f1; ++f) {
    uint8_t v = c->meas[(size_t)(bits[0] - 1) * c->S + f];
    for (uint32_t k = 1; k < n; ++k) v ^= c->meas[(size_t)(bits[k] - 1) * c->S + f];
    c->meas[(size_t)(store - 1) * c->S + f] = v;
  }
}

/* Validates like qcf_program_build; returns a QCF_ status. */
static int check_op(const qcf_op *op, const uint32_t *args, size_t nargs, uint32_t nq,
                    uint32_t nbits) {
  if (is_two_qubit(op->kind) || op->kind == QCF_OP_DEPOLARIZE2) {
    if (op->q1 == 0 || op->q2 == 0) return QCF_ERR_ARGUMENT;
    if (op->q1 == op->q2) return QCF_ERR_ARGUMENT;
    if (op->q1 > nq || op->q2 > nq) return QCF_ERR_BOUNDS;
  } else if (is_one_qubit(op->kind) || (op->kind >= QCF_OP_MZ && op->kind <= QCF_OP_MRX) ||
             op->kind == QCF_OP_PAULI_ERROR || op->kind == QCF_OP_PAULI_NOISE) {
    if (op->q1 == 0) return QCF_ERR_ARGUMENT;
    if (op->q1 > nq) return QCF_ERR_BOUNDS;
  }
  if (op->kind >= QCF_OP_MZ && op->kind <= QCF_OP_PAULI_MEASUREMENT && op->bit > nbits)
    return QCF_ERR_BOUNDS;
  if (op->kind == QCF_OP_CLASSICAL_XOR) {
    if (op->arg_len == 0 || op->bit == 0 || op->bit > nbits) return QCF_ERR_BOUNDS;
    if (op->arg_off + (size_t)op->arg_len > nargs) return QCF_ERR_ARGUMENT;
    for (uint32_t k = 0; k < op->arg_len; ++k)
      if (args[op->arg_off + k] == 0 || args[op->arg_off + k] > nbits) return QCF_ERR_BOUNDS;
  }
  if (op->kind == QCF_OP_PAULI_MEASUREMENT) {
    if (op->arg_len < 1 || op->arg_off + (size_t)op->arg_len > nargs) return QCF_ERR_ARGUMENT;
    if (op->arg_len > nq) return QCF_ERR_BOUNDS; /* ancilla = len(pauli)+1 */
  }
  if (op->kind == QCF_OP_SINGLE_QUBIT_OPERATOR && op->arg_off + (size_t)4 > nargs)
    return QCF_ERR_ARGUMENT;
  return QCF_OK;
}

/* Run ops on shots [f0, f1) of one frame (pftrajectories(state, circuit),
 * src/pauli_frames.jl:234-239). */
static int run_range(frame_ctx *c, const qcf_op *ops, size_t nops, const uint32_t *args,
                     uint32_t nq, uint64_t f0, uint64_t f1) {
  uint32_t rs = nq; /* planes 0..nq-1 are the initZ! draws */
  for (size_t i = 0; i < nops; ++i) {
    const qcf_op *op = &ops[i];
    const uint32_t k = op->kind;
    if (is_one_qubit(k) || is_two_qubit(k)) {
      apply_gate(c, k, op->q1, op->q2, args + op->arg_off, f0, f1);
    } else if (k == QCF_OP_MZ) {
      apply_mz(c, op->q1, op->bit, f0, f1);
    } else if (k == QCF_OP_MRZ) {
      apply_mrz(c, op->q1, op->bit, &rs, f0, f1);
    } else if (k == QCF_OP_MX) { /* (:74-78) */
      if (op->bit == 0) continue;
      apply_gate(c, QCF_OP_HADAMARD, op->q1, 0, NULL, f0, f1);
      apply_mz(c, op->q1, op->bit, f0, f1);
    } else if (k == QCF_OP_MRX) { /* (:80-83) */
      apply_gate(c, QCF_OP_HADAMARD, op->q1, 0, NULL, f0, f1);
      apply_mrz(c, op->q1, op->bit, &rs, f0, f1);
    } else if (k == QCF_OP_PAULI_MEASUREMENT) { /* (:118-134) */
      const uint32_t n = op->arg_len - 1, anc = n + 1;
      for (uint32_t q = 1; q <= n; ++q) {
        const uint32_t code = args[op->arg_off + q - 1] & 3;
        if (code == 1) apply_gate(c, QCF_OP_XCZ, q, anc, NULL, f0, f1);
        else if (code == 2) apply_gate(c, QCF_OP_CNOT, q, anc, NULL, f0, f1);
        else if (code == 3) apply_gate(c, QCF_OP_YCX, q, anc, NULL, f0, f1);
      }
      /* sX on the ancilla for a nonzero phase leaves frame bits unchanged */
      apply_mrz(c, anc, op->bit, &rs, f0, f1);
    } else if (k == QCF_OP_PAULI_ERROR) {
      const double p = op->p[0] / 3;
      apply_noise1(c, op->q1, p, 2 * p, 3 * p, &rs, f0, f1);
    } else if (k == QCF_OP_PAULI_NOISE) {
      const double px = op->p[0], py = op->p[1], pz = op->p[2];
      apply_noise1(c, op->q1, px, px + pz, px + pz + py, &rs, f0, f1);
    } else if (k == QCF_OP_DEPOLARIZE2) {
      apply_dep2(c, op->q1, op->q2, op->p[0], &rs, f0, f1);
    } else if (k == QCF_OP_CLASSICAL_XOR) {
      apply_cxor(c, args + op->arg_off, op->arg_len, op->bit, f0, f1);
    } else if (k != QCF_OP_NOP) {
      return QCF_ERR_UNSUPPORTED;
    }
  }
  return QCF_OK;
}

/*
 * Frame oracle.  nq = frame qubits incl. the hidden ancilla (qubits + 1),
 * xzs = shots x 2W reference frame (W = ceil(nq/64)), meas = shots x nbits
 * bytes.  fresh != 0: PauliFrame(shots, nq-1, nbits) + initZ!
 * (src/pauli_frames.jl:30-56) first, with the initZ! draws taken from mask
 * planes 0..nq-1 (pad bits beyond nq are left zero).  masks == NULL draws
 * from per-thread xoshiro256++ streams seeded from `seed` and the thread.
 */
int oracle_pf_run(const qcf_op *ops, size_t nops, const uint32_t *args, size_t nargs,
                  uint32_t nq, uint32_t nbits, uint64_t shots, uint64_t *xzs, uint8_t *meas,
                  const uint64_t *masks, size_t ld, int fresh, uint64_t seed, int nthreads) {
  for (size_t i = 0; i < nops; ++i) {
    const int st = check_op(&ops[i], args, nargs, nq, nbits);
    if (st != QCF_OK) return st;
    const uint32_t k = ops[i].kind;
    if (!(is_one_qubit(k) || is_two_qubit(k) || (k >= QCF_OP_MZ && k <= QCF_OP_PAULI_MEASUREMENT) ||
          k == QCF_OP_PAULI_ERROR || k == QCF_OP_PAULI_NOISE || k == QCF_OP_DEPOLARIZE2 ||
          k == QCF_OP_CLASSICAL_XOR || k == QCF_OP_NOP))
      return QCF_ERR_UNSUPPORTED;
  }
  const uint32_t W = (nq + 63) / 64;
  if (fresh) {
    memset(xzs, 0, sizeof(uint64_t) * (size_t)2 * W * shots);
    memset(meas, 0, (size_t)nbits * shots);
  }
  if (nthreads < 1) nthreads = 1;
  if ((uint64_t)nthreads > shots / 100 + 1) nthreads = (int)(shots / 100 + 1);
  int status = QCF_OK;
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
#endif
  {
    int t = 0, nt = 1;
#ifdef _OPENMP
    t = omp_get_thread_num();
    nt = omp_get_num_threads();
#endif
    const uint64_t b = shots / (uint64_t)nt;
    const uint64_t f0 = (uint64_t)t * b;
    const uint64_t f1 = t == nt - 1 ? shots : f0 + b;
    frame_ctx c;
    c.xzs = xzs;
    c.meas = meas;
    c.S = shots;
    c.W = W;
    c.masks = masks;
    c.ld = ld;
    xo_seed(&c.rng, seed * 0x9E3779B97F4A7C15ull + (uint64_t)t);
    if (fresh) { /* initZ! (:47-56) */
      for (uint32_t q = 1; q <= nq; ++q) {
        uint64_t *za = ZW(&c, q);
        const uint64_t ma = QMASK(q);
        for (uint64_t f = f0; f < f1; ++f) {
          const int bit = masks ? (int)mask_bit(&c, q - 1, f) : (int)(xo_next(&c.rng) & 1);
          if (bit) za[f] |= ma;
        }
      }
    }
    const int st = run_range(&c, ops, nops, args, nq, f0, f1);
    if (st != QCF_OK) {
#ifdef _OPENMP
#pragma omp critical
#endif
      status = st;
    }
  }
  return status;
}
