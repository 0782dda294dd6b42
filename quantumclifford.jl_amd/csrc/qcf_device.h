// Device code shared by the frame-sampler interpreter (qcf_frames.hip) and the
// per-circuit kernels compiled at run time with hiprtc (qcf_jit.cpp): the
// counter-based RNG, the wave prefix sum and the noise-event pass.  Both
// paths include this one file, so for the same (seed, shot) they draw the
// same events and produce the same bits.  The file has no includes so that
// hiprtc can compile it as is (QCF_RTC is defined there).
#ifndef QCF_DEVICE_H
#define QCF_DEVICE_H

#ifdef QCF_RTC
typedef unsigned int uint32_t;
typedef int int32_t;
typedef unsigned long long uint64_t;
#endif

namespace qcfd {

constexpr uint32_t kTileShots = 2048;  // shots per wave tile: 64 lanes x 32

struct u4 {
  uint32_t x, y, z, w;
};

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): 10 rounds of the 4x32
// multiply-xorshift round with the Weyl key schedule.
__device__ __forceinline__ u4 philox(u4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
    const uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
    c = u4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// Counter word 3 separates the gauge stream from the event stream.
constexpr uint32_t kStreamGauge = 0x47415547u;
constexpr uint32_t kStreamEvent = 0x45564E54u;

// xoshiro128** (Blackman & Vigna 2021), one per lane, seeded from Philox.
struct Xoshiro {
  uint32_t s0, s1, s2, s3;
  __device__ __forceinline__ uint32_t next() {
    const uint32_t m = s1 * 5u;
    const uint32_t r = ((m << 7) | (m >> 25)) * 9u;
    const uint32_t t = s1 << 9;
    s2 ^= s0;
    s3 ^= s1;
    s1 ^= s2;
    s0 ^= s3;
    s2 ^= t;
    s3 = (s3 << 11) | (s3 >> 21);
    return r;
  }
};

// Gauge words of one 32-shot word (initZ! and the sMRZ z-flips): a Weyl
// sequence with its own odd increment, each term scrambled by the murmur3
// 32-bit finalizer -- the construction of SplittableRandom (Steele, Lea,
// Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014),
// 32-bit.  Start and increment come from Philox4x32-10(seed, word); the
// increment is made odd, and one with few bit transitions is flipped as that
// paper's mixGamma does, so no two words walk the same sequence.  A draw
// advances the state by one add: the draws of a tile form no chain of
// dependent state updates (a xoshiro128** stream per word did, and took
// 0.06 ms of the d=5 x 25 launch's 0.53).
struct Gauge {
  uint32_t s, g;
  __device__ __forceinline__ uint32_t next() {
    uint32_t z = (s += g);
    z = (z ^ (z >> 16)) * 0x85EBCA6Bu;
    z = (z ^ (z >> 13)) * 0xC2B2AE35u;
    return z ^ (z >> 16);
  }
};

__device__ __forceinline__ Gauge gauge_stream(uint64_t seed, uint32_t word) {
  const u4 r = philox(u4{word, 0u, 0u, kStreamGauge}, (uint32_t)seed, (uint32_t)(seed >> 32));
  uint32_t g = r.y | 1u;
  if (__popc(g ^ (g >> 1)) < 12) g ^= 0xAAAAAAAAu;
  return Gauge{r.x, g};
}

__device__ __forceinline__ uint32_t readlane(uint32_t v, uint32_t k) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)k);
}

__device__ __forceinline__ uint32_t bperm(uint32_t src_lane, uint32_t v) {
  return (uint32_t)__builtin_amdgcn_ds_bpermute((int)(src_lane << 2), (int)v);
}

// A value barrier: v must be present here, so a load into v is waited for at
// this point rather than wherever the register is next written.
__device__ __forceinline__ void consume(uint32_t &v) { asm volatile("" : "+v"(v)); }

// Inclusive prefix sum over the 64 lanes with DPP moves only: shifts by
// 1, 2, 4, 8 inside each row of 16 lanes, then lane 15 -> next row and lane 31
// -> upper half.  Lanes without a source read 0.
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t x) {
  int v = (int)x;
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return (uint32_t)v;
}

// Noise events of one tile.  Events occur at rate p (one class for the whole
// program) along the sequence (site ordinal, shot in tile): position
// ordinal*2048 + shot.  Lane i holds the i-th event of the current batch of 64.
// Each lane draws its gaps and Pauli choices from its own xoshiro128** stream,
// seeded once per tile from Philox(tile, lane).
struct EvState {
  uint32_t pos;    // event position (per lane)
  uint32_t upau;   // uniform for the Pauli choice
  uint32_t uacc;   // 31-bit uniform for thinning; 0xFFFFFFFF = null event
  uint32_t last;   // position held by lane 63 (wave uniform)
  Xoshiro rng;     // per-lane event stream
};

__device__ __forceinline__ EvState ev_init(uint64_t seed, uint32_t tile, uint32_t lane) {
  const u4 r = philox(u4{tile, lane, 0u, kStreamEvent}, (uint32_t)seed, (uint32_t)(seed >> 32));
  EvState e{0u, 0u, 0xFFFFFFFFu, 0xFFFFFFFFu, Xoshiro{r.x, r.y, r.z, r.w | 1u}};
  return e;
}

// One lane's draws for the next batch: the geometric gap to its event
// (failures before it, floor(log2(u) * c) with c = 1/log2(1-p)), the uniform
// for the Pauli choice and the 31-bit uniform for thinning.  Gaps beyond 2^24
// trials advance without an event (null, uacc = 0xFFFFFFFF; memorylessness
// keeps the process exact).  A program whose sites all accept every event of
// their class or none (thresholds 2^31 or 0) needs no thinning uniform
// (thin = 0): uacc is then 0 for every event, which passes 2^31 and fails 0,
// and the stream advances by two draws per event instead of three.
struct Draw {
  uint32_t incr, upau, uacc;
};
__device__ __forceinline__ Draw draw_event(Xoshiro &rng, float c, uint32_t thin) {
  const uint32_t r0 = rng.next();
  const uint32_t r1 = rng.next();
  const uint32_t r2 = thin ? rng.next() : 0u;
  const float uf = ((float)(r0 >> 8) + 0.5f) * (1.0f / 16777216.0f);
  const float g = __log2f(uf) * c;
  Draw d;
  if (!(g < 16777216.0f)) {
    d.incr = 1u << 24;
    d.uacc = 0xFFFFFFFFu;
  } else {
    d.incr = (uint32_t)g + 1u;
    d.uacc = r2 >> 1;
  }
  d.upau = r1;
  return d;
}

// Draw the next 64 events: one gap per lane and a wave prefix sum.
__device__ __forceinline__ EvState gen_batch(EvState ev, float c, uint32_t thin) {
  const Draw d = draw_event(ev.rng, c, thin);
  ev.pos = ev.last + wave_inclusive_scan(d.incr);
  ev.last = readlane(ev.pos, 63);
  ev.upau = d.upau;
  ev.uacc = d.uacc;
  return ev;
}

// Site entry (one per noise site, 4 words):
//   w0 = x1 | z1 << 16        planes of the (first) qubit
//   w1 = x2 | z2 << 16        planes of the second qubit, 0 for one qubit
//   w2 = t1 | t2 << 16        2^15-scaled thresholds: X if u < t1, Z if
//                             u < t2, else Y (u = top 15 bits of the draw)
//   w3 = accept               2^31-scaled thinning threshold
// Returns the flip mask over (x1, z1, x2, z2).
__device__ __forceinline__ uint32_t event_flips(uint32_t u, uint32_t w1, uint32_t w2) {
  if (w1 == 0u) {
    const uint32_t u15 = u >> 17, t1 = w2 & 0xFFFFu, t2 = w2 >> 16;
    return (u15 < t1 || u15 >= t2 ? 1u : 0u) | (u15 >= t1 ? 2u : 0u);
  }
  // uniform over the 15 non-identity two-qubit Paulis; the first qubit is the
  // least significant base-4 digit, 1=X 2=Z 3=Y
  return __umulhi(u, 15u) + 1u;
}

// Event pass of one noise group: G sites with ordinals ord0 .. ord0+G-1 whose
// entries sit in lanes base .. base+G-1 of (e0..e3); `pay` is carried along
// to the flip callback (the interpreter passes the replay plane there).
// flip(plane, target_lane, bit, j, payload) is called for every flipped plane
// j in 0..3 of an event.  ds_bpermute runs with all lanes active.
template <class Flip>
__device__ __forceinline__ void group_events(EvState &ev, uint32_t ord0, uint32_t G,
                                             uint32_t base, uint32_t e0, uint32_t e1,
                                             uint32_t e2, uint32_t e3, uint32_t pay, float c,
                                             uint32_t thin, Flip flip) {
  const uint32_t lo = ord0 * kTileShots;
  const uint32_t span = G * kTileShots;
  while (true) {
    const uint32_t rel = ev.pos - lo;
    const bool inr = rel < span;
    if (__builtin_amdgcn_ballot_w64(inr && ev.uacc != 0xFFFFFFFFu) != 0) {
      const uint32_t src = base + (inr ? (rel >> 11) : 0u);
      uint32_t w0 = bperm(src, e0), w1 = bperm(src, e1);
      uint32_t w2 = bperm(src, e2), w3 = bperm(src, e3);
      uint32_t py = bperm(src, pay);
      consume(w0);
      consume(w1);
      consume(w2);
      consume(w3);
      consume(py);
      if (inr && ev.uacc < w3) {
        const uint32_t within = rel & (kTileShots - 1u);
        const uint32_t tl = within >> 5;
        const uint32_t bit = 1u << (within & 31u);
        const uint32_t fl = event_flips(ev.upau, w1, w2);
        if (fl & 1u) flip(w0 & 0xFFFFu, tl, bit, 0u, py);
        if (fl & 2u) flip(w0 >> 16, tl, bit, 1u, py);
        if (fl & 4u) flip(w1 & 0xFFFFu, tl, bit, 2u, py);
        if (fl & 8u) flip(w1 >> 16, tl, bit, 3u, py);
      }
    }
    if (ev.last >= lo + span) break;
    ev = gen_batch(ev, c, thin);
  }
}

// Event pass of one noise group whose site entries a specialised kernel holds
// in registers for its whole run: lane i has site i's planes, packed four
// 8-bit fields x1 | z1 << 8 | x2 << 16 | z2 << 24 in `pa` (P8), else the
// 16-bit pairs of w0/w1 in `pa`/`pb`; `w2`/`w3` are per-lane like `pa` when
// W2L/W3L, else the group's common value.  TWO / ONE: every site of the
// group acts on two / on one qubit.  Same events and flips as group_events;
// the flip callback must accept a zero bit (a no-op XOR).
//
// The site entries of the first batch come from prefetch_r, called with the
// same event state earlier in the kernel (before the gates that precede the
// pass): LDS operations complete in order, so that permute is back without
// waiting for the gate XORs issued after it.  Later batches permute here.
template <bool P8>
struct PrefR {
  uint32_t a, b, t, acc;
};

template <bool TWO, bool ONE, bool W2L, bool W3L, bool P8>
__device__ __forceinline__ PrefR<P8> prefetch_r(const EvState &ev, uint32_t ord0, uint32_t G,
                                                uint32_t pa, uint32_t pb, uint32_t w2,
                                                uint32_t w3) {
  const uint32_t rel = ev.pos - ord0 * kTileShots;
  const uint32_t src = rel < G * kTileShots ? (rel >> 11) : 0u;
  PrefR<P8> p;
  p.a = bperm(src, pa);
  p.b = P8 ? 0u : bperm(src, pb);
  p.t = W2L ? bperm(src, w2) : w2;
  p.acc = W3L ? bperm(src, w3) : w3;
  // keep the scheduler from sinking the permutes to their use, behind the
  // gates that follow
  __builtin_amdgcn_sched_barrier(0);
  return p;
}

template <bool TWO, bool ONE, bool W2L, bool W3L, bool P8, class Flip>
__device__ __forceinline__ void flips_r(const EvState &ev, uint32_t rel, uint32_t a, uint32_t b,
                                        uint32_t t, Flip flip) {
  const uint32_t within = rel & (kTileShots - 1u);
  const uint32_t tl = within >> 5;
  const uint32_t bit = 1u << (within & 31u);
  const uint32_t x1 = P8 ? (a & 0xFFu) : (a & 0xFFFFu);
  const uint32_t z1 = P8 ? ((a >> 8) & 0xFFu) : (a >> 16);
  const uint32_t q2 = P8 ? (a >> 16) : b;
  const bool two = TWO ? true : ONE ? false : q2 != 0u;
  uint32_t fl;
  if (two) {
    fl = __umulhi(ev.upau, 15u) + 1u;
  } else {
    const uint32_t u15 = ev.upau >> 17, t1 = t & 0xFFFFu, t2 = t >> 16;
    fl = (u15 < t1 || u15 >= t2 ? 1u : 0u) | (u15 >= t1 ? 2u : 0u);
  }
  // every plane of the site takes an XOR, with zero where the Pauli leaves
  // it alone: no divergent branch per plane
  flip(x1, tl, (fl & 1u) ? bit : 0u, 0u, 0u);
  flip(z1, tl, (fl & 2u) ? bit : 0u, 1u, 0u);
  if (!ONE) {
    const uint32_t x2 = P8 ? (q2 & 0xFFu) : (q2 & 0xFFFFu);
    const uint32_t z2 = P8 ? (q2 >> 8) : (q2 >> 16);
    flip(x2, tl, (fl & 4u) ? bit : 0u, 2u, 0u);
    flip(z2, tl, (fl & 8u) ? bit : 0u, 3u, 0u);
  }
}

template <bool TWO, bool ONE, bool W2L, bool W3L, bool P8, class Flip>
__device__ __forceinline__ void group_events_r(EvState &ev, uint32_t ord0, uint32_t G,
                                               uint32_t pa, uint32_t pb, uint32_t w2,
                                               uint32_t w3, float c, uint32_t thin,
                                               PrefR<P8> pf, Flip flip) {
  const uint32_t lo = ord0 * kTileShots;
  const uint32_t span = G * kTileShots;
  // the prefetched values are used on every path, so no permute is left
  // pending where control flow joins
  consume(pf.a);
  if (!P8) consume(pf.b);
  if (W2L) consume(pf.t);
  if (W3L) consume(pf.acc);
  {
    const uint32_t rel = ev.pos - lo;
    if (rel < span && ev.uacc < pf.acc)
      flips_r<TWO, ONE, W2L, W3L, P8>(ev, rel, pf.a, pf.b, pf.t, flip);
  }
  while (ev.last < lo + span) {
    ev = gen_batch(ev, c, thin);
    const uint32_t rel = ev.pos - lo;
    const bool inr = rel < span;
    if (__builtin_amdgcn_ballot_w64(inr && ev.uacc != 0xFFFFFFFFu) != 0) {
      const uint32_t src = inr ? (rel >> 11) : 0u;
      uint32_t a = bperm(src, pa);
      uint32_t b = P8 ? 0u : bperm(src, pb);
      uint32_t t = W2L ? bperm(src, w2) : w2;
      uint32_t acc = W3L ? bperm(src, w3) : w3;
      // the permuted values are used here, on every path: no LDS result is
      // left pending where the loop branches back, so the loop head (and the
      // code after the pass) does not wait for every outstanding LDS write
      consume(a);
      if (!P8) consume(b);
      if (W2L) consume(t);
      if (W3L) consume(acc);
      if (inr && ev.uacc < acc) flips_r<TWO, ONE, W2L, W3L, P8>(ev, rel, a, b, t, flip);
    }
  }
}

}  // namespace qcfd

#endif  // QCF_DEVICE_H
