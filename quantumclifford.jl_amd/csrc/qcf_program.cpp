// Host compiler: reference operation list -> device plane program.
//
// Lowering rules (reference semantics cited per case):
//   * symbolic Cliffords act on (x,z) bits through a GF(2)-linear map; phases
//     are discarded on frames (src/pauli_frames.jl:58-61 passes
//     phases=Val(false)).  Each map M is realised as P*M = E_k...E_1 with
//     elementary plane XORs E_i and a plane renaming P, found by breadth-first
//     search over GL(n,2), so H, SWAP, ... cost no device work.
//   * sMZ / sMRZ / sMX / sMRX follow src/pauli_frames.jl:74-116 exactly,
//     including "bit == 0 is a no-op" for sMZ/sMX and "reset still happens"
//     for sMRZ/sMRX.
//   * PauliError / PauliNoise sites follow the thresholds of
//     src/pauli_frames.jl:136-174; DEPOLARIZE2 is this library's two-qubit
//     channel (uniform over the 15 non-identity Paulis).
//   * ClassicalXOR follows src/pauli_frames.jl:63-72; bit values that are read
//     again later stay on chip in classical planes.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "qcf_common.hpp"
#include "qcf_program.hpp"

namespace qcf {
namespace {

// GF(2) bit maps of the symbolic gates.  Row i is the mask of inputs (bit0=x1,
// bit1=z1, bit2=x2, bit3=z2) whose XOR gives output i (x1', z1', x2', z2').
// Taken from the qubit_kernel tables of src/symbolic_cliffords.jl:105-118
// (single qubit) and :359-395 (two qubit), phase outputs dropped.
struct GateMap {
  uint32_t kind;
  int nq;
  uint8_t rows[4];
};

constexpr GateMap kGateMaps[] = {
    {QCF_OP_ID1, 1, {1, 2}},         {QCF_OP_X, 1, {1, 2}},
    {QCF_OP_Y, 1, {1, 2}},           {QCF_OP_Z, 1, {1, 2}},
    {QCF_OP_HADAMARD, 1, {2, 1}},    {QCF_OP_HADAMARD_XY, 1, {1, 3}},
    {QCF_OP_HADAMARD_YZ, 1, {3, 2}}, {QCF_OP_PHASE, 1, {1, 3}},
    {QCF_OP_INV_PHASE, 1, {1, 3}},   {QCF_OP_SQRTX, 1, {3, 2}},
    {QCF_OP_INV_SQRTX, 1, {3, 2}},   {QCF_OP_SQRTY, 1, {2, 1}},
    {QCF_OP_INV_SQRTY, 1, {2, 1}},   {QCF_OP_CXYZ, 1, {3, 1}},
    {QCF_OP_CZYX, 1, {2, 3}},
    {QCF_OP_SWAP, 2, {4, 8, 1, 2}},        {QCF_OP_SWAPCX, 2, {4, 10, 5, 2}},
    {QCF_OP_INV_SWAPCX, 2, {5, 8, 1, 10}}, {QCF_OP_ISWAP, 2, {4, 13, 1, 7}},
    {QCF_OP_INV_ISWAP, 2, {4, 13, 1, 7}},  {QCF_OP_CZSWAP, 2, {4, 9, 1, 6}},
    {QCF_OP_CXSWAP, 2, {5, 8, 1, 10}},     {QCF_OP_CNOT, 2, {1, 10, 5, 8}},
    {QCF_OP_CPHASE, 2, {1, 6, 4, 9}},      {QCF_OP_ZCX, 2, {1, 10, 5, 8}},
    {QCF_OP_ZCY, 2, {1, 14, 5, 9}},        {QCF_OP_ZCZ, 2, {1, 6, 4, 9}},
    {QCF_OP_XCX, 2, {9, 2, 6, 8}},         {QCF_OP_XCY, 2, {13, 2, 6, 10}},
    {QCF_OP_XCZ, 2, {5, 2, 4, 10}},        {QCF_OP_YCX, 2, {9, 10, 7, 8}},
    {QCF_OP_YCY, 2, {13, 14, 7, 11}},      {QCF_OP_YCZ, 2, {5, 6, 4, 11}},
    {QCF_OP_ZCRY, 2, {1, 15, 5, 9}},       {QCF_OP_INV_ZCRY, 2, {1, 15, 5, 9}},
    {QCF_OP_SQRTZZ, 2, {1, 7, 4, 13}},     {QCF_OP_INV_SQRTZZ, 2, {1, 7, 4, 13}},
    {QCF_OP_SQRTXX, 2, {11, 2, 14, 8}},    {QCF_OP_INV_SQRTXX, 2, {11, 2, 14, 8}},
    {QCF_OP_SQRTYY, 2, {14, 13, 11, 7}},   {QCF_OP_INV_SQRTYY, 2, {14, 13, 11, 7}},
};

const GateMap *find_gate(uint32_t kind) {
  for (const auto &g : kGateMaps)
    if (g.kind == kind) return &g;
  return nullptr;
}

// ---- decomposition of GF(2) maps into plane XORs + renaming --------------
// A matrix of n<=4 rows is packed 4 bits per row into a uint16.
struct Decomp {
  int len = -1;              // -1: singular (no decomposition)
  uint8_t ops[16][2];        // (dst, src): slot[dst] ^= slot[src]
  uint8_t perm[4];           // output i ends in slot perm_inv...: see apply
};

struct GLTables {
  int n;
  std::vector<int8_t> dist;
  std::vector<uint16_t> parent;
  std::vector<uint8_t> pop;  // op index used to reach a state
  explicit GLTables(int n_) : n(n_), dist(65536, -1), parent(65536), pop(65536) {
    uint16_t id = 0;
    for (int i = 0; i < n; ++i) id |= (uint16_t)(1u << i) << (4 * i);
    std::vector<uint16_t> q{id};
    dist[id] = 0;
    for (size_t h = 0; h < q.size(); ++h) {
      const uint16_t s = q[h];
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
          if (i == j) continue;
          const uint16_t rj = (s >> (4 * j)) & 0xF;
          const uint16_t t = s ^ (uint16_t)(rj << (4 * i));
          if (dist[t] >= 0) continue;
          dist[t] = dist[s] + 1;
          parent[t] = s;
          pop[t] = (uint8_t)(i * 4 + j);
          q.push_back(t);
        }
    }
  }
};

const GLTables &gl_tables(int n) {
  static GLTables t2(2), t4(4);
  return n == 2 ? t2 : t4;
}

// Find E_k..E_1 and a permutation sigma with E_k..E_1 = P_sigma * M, where
// (P_sigma M) row i = M row sigma(i).  After executing the XORs, slot i holds
// logical output sigma(i).
Decomp decompose(const uint8_t *rows, int n) {
  const GLTables &T = gl_tables(n);
  std::array<int, 4> sigma{0, 1, 2, 3}, best_sigma{};
  int best = 1 << 30;
  uint16_t best_state = 0;
  do {
    bool valid = true;
    for (int i = n; i < 4; ++i)
      if (sigma[i] != i) valid = false;
    if (!valid) continue;
    uint16_t s = 0;
    for (int i = 0; i < n; ++i) s |= (uint16_t)(rows[sigma[i]] & 0xF) << (4 * i);
    if (T.dist[s] >= 0 && T.dist[s] < best) {
      best = T.dist[s];
      best_state = s;
      best_sigma = sigma;
    }
  } while (std::next_permutation(sigma.begin(), sigma.end()));
  Decomp d;
  if (best == (1 << 30)) return d;
  d.len = best;
  uint16_t s = best_state;
  for (int k = best - 1; k >= 0; --k) {
    const uint8_t o = T.pop[s];
    d.ops[k][0] = o / 4;
    d.ops[k][1] = o % 4;
    s = T.parent[s];
  }
  for (int i = 0; i < 4; ++i) d.perm[i] = (uint8_t)best_sigma[i];
  return d;
}

const Decomp &gate_decomp(const GateMap &g) {
  static std::mutex mu;
  static std::vector<std::pair<uint32_t, Decomp>> cache;
  std::lock_guard<std::mutex> lock(mu);
  for (auto &e : cache)
    if (e.first == g.kind) return e.second;
  cache.emplace_back(g.kind, decompose(g.rows, 2 * g.nq));
  return cache.back().second;
}

int popcount4(uint32_t v) { return __builtin_popcount(v & 0xF); }

// ---------------------------------------------------------------- compiler -
struct Compiler {
  qcf_program *P;
  uint32_t nq, nbits;
  std::vector<uint32_t> xs, zs;   // plane slot of x / z of qubit q (0-based)
  std::vector<uint32_t> bit_slot; // classical slot holding bit b, or kNone
  std::vector<uint32_t> reads_left;
  std::vector<uint32_t> free_slots;
  uint32_t next_slot;
  uint32_t rs = 0;                // next body replay plane
  std::vector<double> site_pe;    // event rate of each site
  std::vector<uint8_t> written, zero_first;

  Compiler(qcf_program *p, uint32_t nq_, uint32_t nbits_)
      : P(p), nq(nq_), nbits(nbits_), xs(nq_), zs(nq_), bit_slot(nbits_, kNone),
        reads_left(nbits_, 0), next_slot(2 * nq_), written(nbits_, 0),
        zero_first(nbits_, 0) {
    for (uint32_t q = 0; q < nq; ++q) {
      xs[q] = 2 * q;
      zs[q] = 2 * q + 1;
    }
  }

  void emit(uint32_t op, uint32_t a, uint32_t b = 0, uint32_t c = 0) {
    P->body.push_back(Instr{op, a, b, c, 0, 0, 0, 0});
  }
  void emit_xor(uint32_t dst, uint32_t src) {
    emit(I_XORN | (1u << 8), dst | (src << 16));
  }

  uint32_t alloc_slot() {
    if (!free_slots.empty()) {
      uint32_t s = free_slots.back();
      free_slots.pop_back();
      return s;
    }
    return next_slot++;
  }

  // ---- gates
  void gate(const GateMap &g, uint32_t q1, uint32_t q2) {
    uint32_t slot[4] = {xs[q1], zs[q1], 0, 0};
    if (g.nq == 2) {
      slot[2] = xs[q2];
      slot[3] = zs[q2];
    }
    const Decomp &d = gate_decomp(g);
    for (int k = 0; k < d.len; ++k) emit_xor(slot[d.ops[k][0]], slot[d.ops[k][1]]);
    uint32_t ns[4];
    for (int i = 0; i < 2 * g.nq; ++i) ns[d.perm[i]] = slot[i];
    xs[q1] = ns[0];
    zs[q1] = ns[1];
    if (g.nq == 2) {
      xs[q2] = ns[2];
      zs[q2] = ns[3];
    }
    for (int i = 0; i < 2 * g.nq; ++i)
      if (g.rows[i] != (1u << i)) P->bitops += 1;
  }

  int single_qubit_operator(uint32_t q, const uint32_t *m) {
    // x' = (x & xx) ^ (z & zx),  z' = (x & xz) ^ (z & zz)
    // (src/symbolic_cliffords.jl:184-189)
    const uint8_t rows[2] = {(uint8_t)((m[0] ? 1 : 0) | (m[2] ? 2 : 0)),
                             (uint8_t)((m[1] ? 1 : 0) | (m[3] ? 2 : 0))};
    Decomp d = decompose(rows, 2);
    if (d.len < 0) {
      // singular map (not a Clifford, but the reference applies it anyway)
      const uint32_t code = (rows[0] & 1) | ((rows[0] & 2)) | ((rows[1] & 1) << 2) |
                            ((rows[1] & 2) << 2);
      emit(I_LIN1, xs[q], zs[q], code);
      P->bitops += 2;
      return QCF_OK;
    }
    uint32_t slot[2] = {xs[q], zs[q]};
    for (int k = 0; k < d.len; ++k) emit_xor(slot[d.ops[k][0]], slot[d.ops[k][1]]);
    uint32_t ns[2];
    for (int i = 0; i < 2; ++i) ns[d.perm[i]] = slot[i];
    xs[q] = ns[0];
    zs[q] = ns[1];
    for (int i = 0; i < 2; ++i)
      if (rows[i] != (1u << i)) P->bitops += 1;
    return QCF_OK;
  }

  // ---- classical bits
  void write_bit(uint32_t col, uint32_t reads_after, uint32_t from_plane,
                 bool from_acc) {
    // value is in plane `from_plane` (measurement) or in acc (XOR)
    if (bit_slot[col] != kNone) {  // previous value no longer needed
      free_slots.push_back(bit_slot[col]);
      bit_slot[col] = kNone;
    }
    uint32_t keep = kNone;
    if (reads_after > 0) {
      keep = alloc_slot();
      bit_slot[col] = keep;
      reads_left[col] = reads_after;
    }
    if (from_acc)
      emit(I_ACC_STORE, col, keep);
    else
      emit(I_MEAS, from_plane, col, keep);
    written[col] = 1;
  }

  void measure(uint32_t q, uint32_t bit, uint32_t reads_after) {
    if (bit == 0) return;
    write_bit(bit - 1, reads_after, xs[q], false);
    P->bitops += 1;
  }

  void reset(uint32_t q) {
    emit(I_RESETZ, xs[q], zs[q], rs++);
    P->bitops += 2;
  }

  // ---- noise
  int noise1(uint32_t q, double c1, double c2, double c3, uint32_t bitops) {
    auto clamp = [](double v) { return std::isnan(v) ? 0.0 : std::min(1.0, std::max(0.0, v)); };
    c1 = clamp(c1);
    c2 = std::max(c1, clamp(c2));
    c3 = std::max(c2, clamp(c3));
    Site s{};
    s.kind = 1;
    s.rslot = rs;
    rs += 2;
    if (c3 > 0) {  // 2^15-scaled: X if u < t1, Z if u < t2, else Y
      s.t1 = c1 >= c3 ? (1u << 15) : (uint32_t)std::floor(c1 / c3 * 32768.0);
      s.t2 = c2 >= c3 ? (1u << 15) : (uint32_t)std::floor(c2 / c3 * 32768.0);
    }
    s.bitops = bitops;
    s.c1 = c1;
    s.c2 = c2;
    s.c3 = c3;
    const uint32_t id = (uint32_t)P->sites.size();
    P->sites.push_back(s);
    site_pe.push_back(c3);
    emit(I_NOISE1, xs[q], zs[q], id);
    P->bitops += bitops;
    return QCF_OK;
  }

  int noise2(uint32_t q1, uint32_t q2, double e) {
    e = std::isnan(e) ? 0.0 : std::min(1.0, std::max(0.0, e));
    Site s{};
    s.kind = 2;
    s.rslot = rs;
    rs += 4;
    s.bitops = 4;
    s.c3 = e;
    const uint32_t id = (uint32_t)P->sites.size();
    P->sites.push_back(s);
    site_pe.push_back(e);
    emit(I_NOISE2, xs[q1] | (zs[q1] << 16), xs[q2] | (zs[q2] << 16), id);
    P->bitops += 4;
    return QCF_OK;
  }

  int assign_classes() {
    std::vector<double> distinct;
    for (double p : site_pe)
      if (p > 0) distinct.push_back(p);
    std::sort(distinct.begin(), distinct.end());
    distinct.erase(std::unique(distinct.begin(), distinct.end()), distinct.end());
    std::vector<int> bucket(distinct.size());
    std::vector<double> cls_p;
    if (distinct.size() <= kMaxClasses) {
      for (size_t i = 0; i < distinct.size(); ++i) {
        bucket[i] = (int)i;
        cls_p.push_back(distinct[i]);
      }
    } else {
      // equal-width buckets in log(p); each class runs at its largest rate
      const double lo = std::log(distinct.front()), hi = std::log(distinct.back());
      const double w = (hi - lo) / kMaxClasses;
      std::vector<int> raw(distinct.size());
      for (size_t i = 0; i < distinct.size(); ++i) {
        int b = w > 0 ? (int)((std::log(distinct[i]) - lo) / w) : 0;
        raw[i] = std::min<int>(b, kMaxClasses - 1);
      }
      int cur = -1, prev = -1;
      for (size_t i = 0; i < distinct.size(); ++i) {
        if (raw[i] != prev) {
          ++cur;
          prev = raw[i];
          cls_p.push_back(0);
        }
        bucket[i] = cur;
        cls_p[cur] = std::max(cls_p[cur], distinct[i]);
      }
    }
    P->classes.clear();
    for (double p : cls_p) {
      ClassInfo ci;
      ci.p = p;
      const double l2 = std::log1p(-p) / std::log(2.0);  // log2(1-p) <= 0
      ci.c = (float)(1.0 / l2);
      P->classes.push_back(ci);
    }
    std::vector<uint32_t> ord(cls_p.size(), 0);
    for (size_t i = 0; i < P->sites.size(); ++i) {
      Site &s = P->sites[i];
      const double pe = site_pe[i];
      if (cls_p.empty()) {
        s.cls = kNoFire;
        continue;
      }
      int c = 0;
      if (pe > 0) {
        const size_t k = std::lower_bound(distinct.begin(), distinct.end(), pe) - distinct.begin();
        c = bucket[k];
      }
      s.cls = (uint32_t)c;
      s.ordinal = ord[c]++;
      const double r = pe > 0 ? pe / cls_p[c] : 0.0;
      s.accept = r >= 1.0 ? (1u << 31) : (uint32_t)std::llround(r * 2147483648.0);
    }
    for (uint32_t n : ord)
      if (n > kMaxOrdinals)
        return set_error(QCF_ERR_UNSUPPORTED,
                         "pftrajectories: %u noise sites share one rate class (limit %u)",
                         n, kMaxOrdinals);
    return QCF_OK;
  }

  // Move each noise site back to the run of noise sites before it when no
  // operation in between reads or writes the planes of its qubits: a Pauli
  // flip commutes with every operation that leaves those planes alone, so for
  // any drawn flips the record and the frame come out the same, and fewer,
  // larger runs mean fewer event passes (x_diag_circuit_noisy_measurement:
  // each PauliError sits between gates on other qubits, and its 1000 sites
  // become one run instead of 1000).  A site only ever joins the run of the
  // noise sites just before it in program order, so the ordinals inside a
  // run stay consecutive.
  void hoist_noise() {
    const std::vector<Instr> &B = P->body;
    std::vector<Instr> out, run, tail;
    out.reserve(B.size());
    std::vector<uint32_t> stamp;  // plane -> id of the tail that touched it
    uint32_t tail_id = 1;
    std::vector<uint32_t> pl;
    auto touch = [&](uint32_t p) {
      if (stamp.size() <= p) stamp.resize(p + 1, 0);
      stamp[p] = tail_id;
    };
    auto touched = [&](uint32_t p) { return p < stamp.size() && stamp[p] == tail_id; };
    for (size_t i = 0; i < B.size(); ++i) {
      const Instr &in = B[i];
      if (in.op == I_NOISE1 || in.op == I_NOISE2) {
        const uint32_t q[4] = {in.op == I_NOISE2 ? in.a & 0xFFFFu : in.a,
                               in.op == I_NOISE2 ? in.a >> 16 : in.b,
                               in.op == I_NOISE2 ? in.b & 0xFFFFu : kNone,
                               in.op == I_NOISE2 ? in.b >> 16 : kNone};
        bool free = !run.empty();
        for (uint32_t p : q)
          if (p != kNone && touched(p)) free = false;
        if (!free) {
          out.insert(out.end(), run.begin(), run.end());
          out.insert(out.end(), tail.begin(), tail.end());
          run.clear();
          tail.clear();
          ++tail_id;
        }
        run.push_back(in);
        continue;
      }
      if (run.empty()) {
        out.push_back(in);
        continue;
      }
      tail.push_back(in);
      pl.clear();
      instr_planes(B, i, &pl);
      for (uint32_t p : pl) touch(p);
    }
    out.insert(out.end(), run.begin(), run.end());
    out.insert(out.end(), tail.begin(), tail.end());
    P->body.swap(out);
  }

  // Turn every run of consecutive noise instructions into groups: a header
  // I_NOISEG followed by one I_SITE data entry per site, never straddling a
  // 64-instruction chunk (NOPs pad to the next chunk when needed).  Pauli
  // noise commutes with Pauli noise, so a run is one unit of work: the native
  // sampler checks the events of all its sites at once.  The body is assumed
  // to start on a chunk boundary (the prologues are padded).
  void group_sites() {
    std::vector<Instr> out;
    out.reserve(P->body.size() * 2);
    const auto &B = P->body;
    size_t i = 0;
    while (i < B.size()) {
      if (B[i].op != I_NOISE1 && B[i].op != I_NOISE2) {
        out.push_back(B[i++]);
        continue;
      }
      size_t j = i;
      while (j < B.size() && (B[j].op == I_NOISE1 || B[j].op == I_NOISE2) &&
             j - i < kChunk - 1) {
        // ordinals must stay consecutive inside a group
        if (j > i && P->sites[B[j].c].ordinal != P->sites[B[j - 1].c].ordinal + 1) break;
        ++j;
      }
      const uint32_t G = (uint32_t)(j - i);
      const size_t pos = out.size() % kChunk;
      if (pos + 1 + G > kChunk)
        for (size_t k = pos; k < kChunk; ++k) out.push_back(Instr{I_END, 0, 0, 0, 0, 0, 0, 0});
      out.push_back(Instr{I_NOISEG | (G << 8), P->sites[B[i].c].ordinal, 0, 0, 0, 0, 0, 0});
      for (size_t k = i; k < j; ++k) {
        const Site &s = P->sites[B[k].c];
        const uint32_t pb = B[k].op == I_NOISE2 ? B[k].b : 0u;
        const uint32_t pa = B[k].op == I_NOISE2 ? B[k].a : (B[k].a | (B[k].b << 16));
        out.push_back(Instr{I_SITE, pa, pb, s.rslot, s.t1 | (s.t2 << 16), s.accept, 0, 0});
      }
      i = j;
    }
    P->body.swap(out);
  }

  // Peepholes: bundle runs of plane XORs into XORN (up to 7 pairs, a pair is
  // added only if its source is not a target already in the bundle, so
  // reading all sources first equals the sequential order), and fuse the
  // ClassicalXOR sequences ACC,ACCX,STORE -> XSTORE2 and ACC,STORE -> MEAS.
  void peephole() {
    std::vector<Instr> out;
    out.reserve(P->body.size());
    std::vector<uint32_t> pairs, dsts;
    auto flush = [&]() {
      if (pairs.empty()) return;
      Instr in{I_XORN | ((uint32_t)pairs.size() << 8), 0, 0, 0, 0, 0, 0, 0};
      uint32_t *w = &in.a;
      for (size_t i = 0; i < pairs.size(); ++i) w[i] = pairs[i];
      out.push_back(in);
      pairs.clear();
      dsts.clear();
    };
    const auto &B = P->body;
    for (size_t i = 0; i < B.size(); ++i) {
      const Instr &in = B[i];
      if (in.op == (I_XORN | (1u << 8))) {
        const uint32_t dst = in.a & 0xFFFFu, src = in.a >> 16;
        if (pairs.size() == 7 || std::find(dsts.begin(), dsts.end(), src) != dsts.end()) flush();
        pairs.push_back(in.a);
        dsts.push_back(dst);
        continue;
      }
      flush();
      if (in.op == I_ACC_LDS && i + 2 < B.size() && B[i + 1].op == I_ACCX_LDS &&
          B[i + 2].op == I_ACC_STORE) {
        out.push_back(Instr{I_XSTORE2, in.a, B[i + 1].a, B[i + 2].a, B[i + 2].b, 0, 0, 0});
        i += 2;
        continue;
      }
      if (in.op == I_ACC_LDS && i + 1 < B.size() && B[i + 1].op == I_ACC_STORE) {
        out.push_back(Instr{I_MEAS, in.a, B[i + 1].a, B[i + 1].b, 0, 0, 0, 0});
        i += 1;
        continue;
      }
      out.push_back(in);
    }
    flush();
    P->body.swap(out);
    // measure-and-reset: MEAS(x) + RESETZ(x, z) -> MR; a lone RESETZ -> MR
    // without a record; MR followed by the XOR of its value with another
    // kept bit (a detector against the previous round) -> MRX
    out.clear();
    const auto &C2 = P->body;
    for (size_t i = 0; i < C2.size(); ++i) {
      const Instr &in = C2[i];
      Instr mr{};
      bool have = false;
      if (in.op == I_MEAS && i + 1 < C2.size() && C2[i + 1].op == I_RESETZ && C2[i + 1].a == in.a) {
        mr = Instr{I_MR, in.a, C2[i + 1].b, C2[i + 1].c, in.b, in.c, kNone, kNone};
        have = true;
        i += 1;
      } else if (in.op == I_RESETZ) {
        mr = Instr{I_MR, in.a, in.b, in.c, kNone, kNone, kNone, kNone};
        have = true;
      }
      if (!have) {
        out.push_back(in);
        continue;
      }
      if (mr.e != kNone && i + 1 < C2.size() && C2[i + 1].op == I_XSTORE2 && C2[i + 1].d == kNone &&
          (C2[i + 1].a == mr.e || C2[i + 1].b == mr.e) && C2[i + 1].a != C2[i + 1].b) {
        const Instr &x = C2[i + 1];
        mr.op = I_MRX;
        mr.f = x.a == mr.e ? x.b : x.a;
        mr.g = x.c;
        i += 1;
      }
      out.push_back(mr);
    }
    P->body.swap(out);
  }
};

// Classical slots (planes >= base) read and written by one instruction.
// Every instruction reads all of its slots before it writes any.
void slot_fields(Instr &in, uint32_t base, std::vector<uint32_t *> *reads,
                 std::vector<uint32_t *> *defs) {
  auto rd = [&](uint32_t &f) {
    if (f != kNone && f >= base) reads->push_back(&f);
  };
  auto df = [&](uint32_t &f) {
    if (f != kNone && f >= base) defs->push_back(&f);
  };
  switch (in.op & 0xFFu) {
    case I_MEAS: rd(in.a); df(in.c); break;
    case I_XSTORE2: rd(in.a); rd(in.b); df(in.d); break;
    case I_ACC_LDS:
    case I_ACCX_LDS: rd(in.a); break;
    case I_ACC_STORE: df(in.b); break;
    case I_MR: df(in.e); break;
    case I_MRX: rd(in.f); df(in.e); break;
    default: break;
  }
}

// Reassign the classical slots of the body from the live ranges of their
// values: at each instruction the slots whose values are read for the last
// time are freed before the values it defines take slots (most recently freed
// first).  A measurement that replaces its own previous-round value in a
// detector then lands in the same slot every round, so the rounds of a memory
// experiment compile to identical instruction blocks.  Returns the slot count.
uint32_t renumber_slots(std::vector<Instr> &body, uint32_t base) {
  std::vector<int64_t> cur;      // old slot -> current value
  std::vector<size_t> last;      // value -> index of its last read (or def)
  std::vector<uint32_t *> reads, defs;
  auto val_of = [&](uint32_t slot) -> int64_t & {
    if (cur.size() <= slot - base) cur.resize(slot - base + 1, -1);
    return cur[slot - base];
  };
  for (size_t i = 0; i < body.size(); ++i) {
    reads.clear();
    defs.clear();
    slot_fields(body[i], base, &reads, &defs);
    for (uint32_t *f : reads) {
      const int64_t v = val_of(*f);
      if (v >= 0) last[v] = i;
    }
    for (uint32_t *f : defs) {
      val_of(*f) = (int64_t)last.size();
      last.push_back(i);
    }
  }
  std::fill(cur.begin(), cur.end(), -1);
  std::vector<uint32_t> slot_of(last.size(), kNone), free_list;
  uint32_t next = base;
  int64_t nval = 0;
  std::vector<int64_t> freeing;
  for (size_t i = 0; i < body.size(); ++i) {
    reads.clear();
    defs.clear();
    slot_fields(body[i], base, &reads, &defs);
    freeing.clear();
    for (uint32_t *f : reads) {
      const int64_t v = val_of(*f);
      if (v < 0) continue;  // read of a slot never written: leave as is
      *f = slot_of[v];
      if (last[v] == i && std::find(freeing.begin(), freeing.end(), v) == freeing.end())
        freeing.push_back(v);
    }
    for (int64_t v : freeing) free_list.push_back(slot_of[v]);
    for (uint32_t *f : defs) {
      const int64_t v = nval++;
      val_of(*f) = v;
      uint32_t s;
      if (!free_list.empty()) {
        s = free_list.back();
        free_list.pop_back();
      } else {
        s = next++;
      }
      slot_of[v] = s;
      *f = s;
      if (last[v] == i) free_list.push_back(s);  // never read
    }
  }
  return next;
}

int op_qubit_check(const qcf_op &op, int nqargs, uint32_t nq) {
  if (op.q1 == 0 || (nqargs == 2 && op.q2 == 0))
    return set_error(QCF_ERR_ARGUMENT,
                     "Qubit indices have to be larger than zero (op kind %u)", op.kind);
  if (nqargs == 2 && op.q1 == op.q2)
    return set_error(QCF_ERR_ARGUMENT,
                     "Failed to create a two qubit gate because the two qubits it "
                     "acts on have the same index (%u).", op.q1);
  const uint32_t qmax = nqargs == 2 ? std::max(op.q1, op.q2) : op.q1;
  if (qmax > nq)
    return set_error(QCF_ERR_BOUNDS,
                     "qubit %u is outside the frame of %u qubits (incl. ancilla)", qmax, nq);
  return QCF_OK;
}

bool is_gate(uint32_t kind) { return find_gate(kind) != nullptr; }

}  // namespace
}  // namespace qcf

using namespace qcf;

qcf_program::~qcf_program() {
  for (auto &d : dev) {
    if (d.device < 0) continue;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(d.device);
    for (void *p : d.prog)
      if (p) (void)hipFree(p);
    (void)hipSetDevice(cur);
  }
  for (auto &j : jit) {
    if (!j.module) continue;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(j.device);
    (void)hipModuleUnload((hipModule_t)j.module);
    (void)hipSetDevice(cur);
  }
}

std::vector<qcf::Instr> qcf::variant_code(const qcf_program *P, int variant) {
  std::vector<Instr> code;
  const auto &pro = (variant & 1) ? P->prologue_fresh : P->prologue_inplace;
  // a fresh run that returns only the record leaves out the gauge draws that
  // cannot change its distribution (mark_dead_gauges)
  const bool thin = (variant & 3) == 1;
  const size_t p0 = code.size();
  code.insert(code.end(), pro.begin(), pro.end());
  if (thin && P->gauge_dead_pro.size() == pro.size())
    for (size_t i = 0; i < pro.size(); ++i)
      if (P->gauge_dead_pro[i]) code[p0 + i].op |= kNoGauge;
  while (code.size() % kChunk) code.push_back(Instr{I_END, 0, 0, 0, 0, 0, 0, 0});
  const size_t b0 = code.size();
  code.insert(code.end(), P->body.begin(), P->body.end());
  if (thin && P->gauge_dead_body.size() == P->body.size())
    for (size_t i = 0; i < P->body.size(); ++i)
      if (P->gauge_dead_body[i]) code[b0 + i].op |= kNoGauge;
  if (variant & 2) code.insert(code.end(), P->epilogue_store.begin(), P->epilogue_store.end());
  if (code.empty()) code.push_back(Instr{I_END, 0, 0, 0, 0, 0, 0, 0});
  return code;
}

// Planes instruction code[i] reads or writes (a noise group: its sites').
void qcf::instr_planes(const std::vector<Instr> &code, size_t i, std::vector<uint32_t> *out) {
  const Instr &in = code[i];
  auto use = [&](uint32_t p) {
    if (p != kNone) out->push_back(p);
  };
  switch (in.op & 0xFFu) {
    case I_XORN: {
      const uint32_t *w = &in.a;
      for (uint32_t k = 0; k < (in.op >> 8); ++k) use(w[k] >> 16), use(w[k] & 0xFFFFu);
      break;
    }
    case I_NOISEG:
      for (uint32_t k = 1; k <= (in.op >> 8) && i + k < code.size(); ++k) {
        const Instr &s = code[i + k];
        use(s.a & 0xFFFFu), use(s.a >> 16);
        if (s.b) use(s.b & 0xFFFFu), use(s.b >> 16);
      }
      break;
    case I_MR: use(in.a), use(in.b), use(in.e); break;
    case I_MRX: use(in.a), use(in.b), use(in.e), use(in.f); break;
    case I_XSTORE2: use(in.a), use(in.b), use(in.d); break;
    case I_MEAS: use(in.a), use(in.c); break;
    case I_RESETZ:
    case I_INIT:
    case I_LIN1: use(in.a), use(in.b); break;
    case I_ACC_LDS:
    case I_ACCX_LDS:
    case I_LOAD:
    case I_STORE:
    case I_ZERO: use(in.a); break;
    case I_ACC_STORE: use(in.b); break;
    default: break;
  }
}

// Which gauge draws a fresh, record-only run needs.
//
// Over GF(2) the frame is linear in its random inputs: gates, measurements,
// resets and classical XORs are linear maps of the planes, a noise site
// XORs flips drawn independently of the frame, and each gauge draw (initZ!,
// the sMRZ/sMRX z-flip) XORs an independent uniform word into a z plane.  So
// the record of a shot is R = G g + N, with g the vector of gauge bits
// (uniform on GF(2)^k), G the record bits each gauge bit reaches, and N the
// noise part, independent of g.  G g is uniform on the column span of G, and
// so is G' g' for any set of columns of G that spans it.  Keeping the draws
// whose columns are independent of the columns kept before them (in program
// order) and dropping the rest therefore leaves the distribution of every
// record, jointly over all its bits, exactly as it was: only which uniform
// word lands where changes.  The final frame is not covered (its gauge part
// spans more), so launches that return the frame draw everything.
//
// G comes from running the program symbolically: each plane and record
// column holds the set of gauge draws XORed into it (a bitset), following the
// instruction semantics of the interpreter (qcf_frames.hip run_tile).  The
// column basis is a greedy XOR basis over record bits.  Programs whose
// bitsets or elimination would be very large keep every draw.
void qcf::mark_dead_gauges(qcf_program *P) {
  P->gauge_dead_pro.clear();
  P->gauge_dead_body.clear();
  auto is_gauge = [](const Instr &in) {
    const uint32_t op = in.op & 0xFFu;
    return op == I_INIT || op == I_RESETZ || op == I_MR || op == I_MRX;
  };
  size_t k = 0;
  for (const Instr &in : P->prologue_fresh) k += is_gauge(in);
  for (const Instr &in : P->body) k += is_gauge(in);
  P->ngauge = (uint32_t)k;
  P->ngauge_live = (uint32_t)k;
  P->linear_ok = false;
  P->rec_gauge.clear();
  if (k == 0) return;
  const size_t W = (k + 63) / 64;           // words per gauge bitset
  const size_t NB = std::max<uint32_t>(P->nbits, 1u);
  const size_t WB = (NB + 63) / 64;         // words per record-bit bitset
  const double cells = (double)(P->nplanes + NB + 1) * W + (double)k * WB;
  const double elim = (double)k * (double)std::min(k, NB) * (double)WB;
  if (cells > 32e6 || elim > 4e9) return;   // keep every draw
  std::vector<uint64_t> pl((size_t)P->nplanes * W, 0), rec(NB * W, 0), acc(W, 0);
  auto PLp = [&](uint32_t p) { return &pl[(size_t)p * W]; };
  auto RCp = [&](uint32_t b) { return &rec[(size_t)b * W]; };
  auto cpy = [&](uint64_t *d, const uint64_t *s) { std::copy(s, s + W, d); };
  auto xr = [&](uint64_t *d, const uint64_t *s) {
    for (size_t w = 0; w < W; ++w) d[w] ^= s[w];
  };
  auto zero = [&](uint64_t *d) { std::fill(d, d + W, 0ull); };
  std::vector<uint64_t> v(W), t(W), src(7 * W);
  size_t g = 0;  // next gauge index
  auto run = [&](const std::vector<Instr> &code) {
    for (size_t i = 0; i < code.size(); ++i) {
      const Instr &in = code[i];
      const uint32_t op = in.op & 0xFFu;
      switch (op) {
        case I_XORN: {
          const uint32_t n = in.op >> 8;
          const uint32_t *w = &in.a;
          for (uint32_t j = 0; j < n; ++j) cpy(&src[j * W], PLp(w[j] >> 16));
          for (uint32_t j = 0; j < n; ++j) xr(PLp(w[j] & 0xFFFFu), &src[j * W]);
          break;
        }
        case I_NOISEG: i += in.op >> 8; break;  // noise: independent of g
        case I_MEAS:
          cpy(v.data(), PLp(in.a));
          cpy(RCp(in.b), v.data());
          if (in.c != kNone) cpy(PLp(in.c), v.data());
          break;
        case I_XSTORE2:
          cpy(v.data(), PLp(in.a));
          xr(v.data(), PLp(in.b));
          cpy(RCp(in.c), v.data());
          if (in.d != kNone) cpy(PLp(in.d), v.data());
          break;
        case I_RESETZ:
          zero(PLp(in.a));
          PLp(in.b)[g / 64] ^= 1ull << (g % 64);
          ++g;
          break;
        case I_ACC_LDS: cpy(acc.data(), PLp(in.a)); break;
        case I_ACCX_LDS: xr(acc.data(), PLp(in.a)); break;
        case I_ACC_REC: cpy(acc.data(), RCp(in.a)); break;
        case I_ACCX_REC: xr(acc.data(), RCp(in.a)); break;
        case I_ACC_STORE:
          cpy(RCp(in.a), acc.data());
          if (in.b != kNone) cpy(PLp(in.b), acc.data());
          break;
        case I_REC_ZERO: zero(RCp(in.a)); break;
        case I_INIT:
          zero(PLp(in.a));
          zero(PLp(in.b));
          PLp(in.b)[g / 64] ^= 1ull << (g % 64);
          ++g;
          break;
        case I_LOAD:
        case I_ZERO: zero(PLp(in.a)); break;
        case I_LIN1: {
          cpy(v.data(), PLp(in.a));
          cpy(t.data(), PLp(in.b));
          uint64_t *x = PLp(in.a), *z = PLp(in.b);
          zero(x);
          zero(z);
          if (in.c & 1u) xr(x, v.data());
          if (in.c & 2u) xr(x, t.data());
          if (in.c & 4u) xr(z, v.data());
          if (in.c & 8u) xr(z, t.data());
          break;
        }
        case I_MR:
        case I_MRX: {
          cpy(v.data(), PLp(in.a));
          if (in.d != kNone) cpy(RCp(in.d), v.data());
          if (op == I_MRX) {
            cpy(t.data(), v.data());
            xr(t.data(), PLp(in.f));
            cpy(RCp(in.g), t.data());
          }
          if (in.e != kNone) cpy(PLp(in.e), v.data());
          zero(PLp(in.a));
          PLp(in.b)[g / 64] ^= 1ull << (g % 64);
          ++g;
          break;
        }
        default: break;  // I_END, I_STORE
      }
    }
  };
  run(P->prologue_fresh);
  run(P->body);
  // columns of G: gauge draw j -> the record bits it reaches
  std::vector<uint64_t> col((size_t)k * WB, 0);
  for (size_t b = 0; b < P->nbits; ++b) {
    const uint64_t *r = RCp((uint32_t)b);
    for (size_t w = 0; w < W; ++w)
      for (uint64_t m = r[w]; m; m &= m - 1) {
        const size_t j = w * 64 + __builtin_ctzll(m);
        col[j * WB + b / 64] |= 1ull << (b % 64);
      }
  }
  // greedy XOR basis in program order, pivots on the lowest set bit
  std::vector<int64_t> piv(NB, -1);  // record bit -> basis vector
  std::vector<uint64_t> basis;
  std::vector<uint8_t> live(k, 0);
  size_t nb = 0;
  // (a basis vector's lowest set bit is its pivot, so reducing by it clears
  // that bit and touches no lower one)
  for (size_t j = 0; j < k; ++j) {
    uint64_t *c = &col[j * WB];
    for (size_t w = 0; w < WB && !live[j];) {
      if (!c[w]) {
        ++w;
        continue;
      }
      const size_t b = w * 64 + __builtin_ctzll(c[w]);
      if (piv[b] < 0) {  // independent of the draws kept so far: kept
        piv[b] = (int64_t)nb++;
        basis.insert(basis.end(), c, c + WB);
        live[j] = 1;
      } else {
        const uint64_t *e = &basis[(size_t)piv[b] * WB];
        for (size_t u = w; u < WB; ++u) c[u] ^= e[u];
      }
    }
  }
  size_t j = 0, nlive = 0;
  P->gauge_dead_pro.assign(P->prologue_fresh.size(), 0);
  for (size_t i = 0; i < P->prologue_fresh.size(); ++i)
    if (is_gauge(P->prologue_fresh[i])) P->gauge_dead_pro[i] = !live[j++];
  P->gauge_dead_body.assign(P->body.size(), 0);
  for (size_t i = 0; i < P->body.size(); ++i)
    if (is_gauge(P->body[i])) P->gauge_dead_body[i] = !live[j++];
  for (uint8_t x : live) nlive += x;
  P->ngauge_live = (uint32_t)nlive;
  // no site with a nonzero rate (none at all, or every one at zero, which
  // native draws never fire): record column b = XOR of the kept draws among
  // those reaching it
  P->linear_ok = false;
  P->rec_gauge.clear();
  if (P->classes.empty()) {
    std::vector<uint32_t> pos(k, 0);
    for (size_t i = 0, n = 0; i < k; ++i) pos[i] = live[i] ? (uint32_t)n++ : 0u;
    P->rec_gauge.assign(P->nbits, {});
    for (uint32_t b = 0; b < P->nbits; ++b) {
      const uint64_t *r = RCp(b);
      for (size_t w = 0; w < W; ++w)
        for (uint64_t m = r[w]; m; m &= m - 1) {
          const size_t i = w * 64 + __builtin_ctzll(m);
          if (live[i]) P->rec_gauge[b].push_back(pos[i]);
        }
    }
    P->linear_ok = true;
  }
}

extern "C" int qcf_program_build(const qcf_op *ops, size_t nops,
                                 const uint32_t *args, size_t nargs,
                                 uint32_t nqubits, uint32_t nbits,
                                 qcf_program **out) {
  clear_error();
  if (!out) return set_error(QCF_ERR_ARGUMENT, "null output pointer");
  *out = nullptr;
  if (nops && !ops) return set_error(QCF_ERR_ARGUMENT, "null op array");
  // ---- size the frame like _create_pauliframe (src/pauli_frames.jl:197-201)
  uint32_t qmax = 1, bmax = 1;
  for (size_t i = 0; i < nops; ++i) {
    const qcf_op &op = ops[i];
    const GateMap *g = find_gate(op.kind);
    if (g) {
      qmax = std::max(qmax, op.q1);
      if (g->nq == 2) qmax = std::max(qmax, op.q2);
    } else {
      switch (op.kind) {
        case QCF_OP_SINGLE_QUBIT_OPERATOR:
        case QCF_OP_PAULI_ERROR:
        case QCF_OP_PAULI_NOISE:
          qmax = std::max(qmax, op.q1);
          break;
        case QCF_OP_DEPOLARIZE2:
          qmax = std::max(qmax, std::max(op.q1, op.q2));
          break;
        case QCF_OP_MZ: case QCF_OP_MRZ: case QCF_OP_MX: case QCF_OP_MRX:
          qmax = std::max(qmax, op.q1);
          bmax = std::max(bmax, op.bit);
          break;
        case QCF_OP_PAULI_MEASUREMENT:
          qmax = std::max(qmax, op.arg_len ? op.arg_len - 1 : 0);
          bmax = std::max(bmax, op.bit);
          break;
        case QCF_OP_CLASSICAL_XOR:
          bmax = std::max(bmax, op.bit);
          if (op.arg_off + (size_t)op.arg_len > nargs)
            return set_error(QCF_ERR_ARGUMENT, "ClassicalXOR arguments out of range");
          for (uint32_t k = 0; k < op.arg_len; ++k) bmax = std::max(bmax, args[op.arg_off + k]);
          break;
        case QCF_OP_NOP:
          break;
        default:
          return set_error(QCF_ERR_UNSUPPORTED,
                           "no PauliFrame method for operation kind %u", op.kind);
      }
    }
  }
  const uint32_t n = nqubits ? nqubits : qmax;
  const uint32_t B = nbits ? nbits : bmax;
  const uint32_t nq = n + 1;  // hidden ancilla (src/pauli_frames.jl:30-37)
  if (nq > 0x7FFF)
    return set_error(QCF_ERR_UNSUPPORTED, "frame of %u qubits exceeds the plane index range", nq);

  qcf_program *P = new (std::nothrow) qcf_program();
  if (!P) return set_error(QCF_ERR_ALLOC, "out of host memory");
  P->nq = nq;
  P->nbits = B;
  Compiler C(P, nq, B);

  // ---- classical liveness: reads of each written value before it is
  //      overwritten (ClassicalXOR reads precede its own store)
  std::vector<uint32_t> reads_after_write(nops, 0), pending(B, 0);
  auto write_bit_of = [&](const qcf_op &op) -> uint32_t {
    switch (op.kind) {
      case QCF_OP_MZ: case QCF_OP_MX: case QCF_OP_MRZ: case QCF_OP_MRX:
      case QCF_OP_PAULI_MEASUREMENT: case QCF_OP_CLASSICAL_XOR:
        return op.bit;
      default:
        return 0;
    }
  };
  for (size_t i = 0; i < nops; ++i) {
    const qcf_op &op = ops[i];
    const uint32_t w = write_bit_of(op);
    if (w > B) {
      delete P;
      return set_error(QCF_ERR_BOUNDS, "classical bit %u outside the record of %u bits", w, B);
    }
    if (op.kind == QCF_OP_CLASSICAL_XOR) {
      if (op.arg_len == 0) {
        delete P;
        return set_error(QCF_ERR_BOUNDS, "ClassicalXOR needs at least one input bit");
      }
      if (op.bit == 0) {
        delete P;
        return set_error(QCF_ERR_BOUNDS, "ClassicalXOR store bit must be >= 1");
      }
      for (uint32_t k = 0; k < op.arg_len; ++k) {
        const uint32_t b = args[op.arg_off + k];
        if (b == 0 || b > B) {
          delete P;
          return set_error(QCF_ERR_BOUNDS, "ClassicalXOR input bit %u outside 1..%u", b, B);
        }
      }
    }
  }
  for (size_t i = nops; i-- > 0;) {
    const qcf_op &op = ops[i];
    const uint32_t w = write_bit_of(op);
    if (w) {
      reads_after_write[i] = pending[w - 1];
      pending[w - 1] = 0;
    }
    if (op.kind == QCF_OP_CLASSICAL_XOR)
      for (uint32_t k = 0; k < op.arg_len; ++k) pending[args[op.arg_off + k] - 1]++;
  }

  // ---- lower every operation
  for (size_t i = 0; i < nops; ++i) {
    const qcf_op &op = ops[i];
    int st = QCF_OK;
    if (const GateMap *g = find_gate(op.kind)) {
      st = op_qubit_check(op, g->nq, nq);
      if (st == QCF_OK) C.gate(*g, op.q1 - 1, g->nq == 2 ? op.q2 - 1 : 0);
    } else {
      switch (op.kind) {
        case QCF_OP_NOP:
          break;
        case QCF_OP_SINGLE_QUBIT_OPERATOR:
          st = op_qubit_check(op, 1, nq);
          if (st == QCF_OK && (op.arg_len < 4 || op.arg_off + (size_t)4 > nargs))
            st = set_error(QCF_ERR_ARGUMENT, "SingleQubitOperator needs 4 matrix entries");
          if (st == QCF_OK) st = C.single_qubit_operator(op.q1 - 1, args + op.arg_off);
          break;
        case QCF_OP_MZ:
          st = op_qubit_check(op, 1, nq);
          if (st == QCF_OK) C.measure(op.q1 - 1, op.bit, reads_after_write[i]);
          break;
        case QCF_OP_MX:
          st = op_qubit_check(op, 1, nq);
          if (st == QCF_OK && op.bit != 0) {
            C.gate(*find_gate(QCF_OP_HADAMARD), op.q1 - 1, 0);
            C.measure(op.q1 - 1, op.bit, reads_after_write[i]);
          }
          break;
        case QCF_OP_MRZ:
        case QCF_OP_MRX:
          st = op_qubit_check(op, 1, nq);
          if (st == QCF_OK) {
            if (op.kind == QCF_OP_MRX) C.gate(*find_gate(QCF_OP_HADAMARD), op.q1 - 1, 0);
            C.measure(op.q1 - 1, op.bit, reads_after_write[i]);
            C.reset(op.q1 - 1);
          }
          break;
        case QCF_OP_PAULI_MEASUREMENT: {
          // args: one code per qubit 1..len (0=I 1=X 2=Z 3=Y), then the phase
          if (op.arg_len < 1 || op.arg_off + (size_t)op.arg_len > nargs) {
            st = set_error(QCF_ERR_ARGUMENT, "PauliMeasurement arguments out of range");
            break;
          }
          const uint32_t len = op.arg_len - 1;
          const uint32_t anc = len + 1;  // nqubits(op.pauli) + 1
          if (anc > nq) {
            st = set_error(QCF_ERR_BOUNDS, "PauliMeasurement ancilla %u outside the frame of %u qubits", anc, nq);
            break;
          }
          for (uint32_t q = 1; q <= len; ++q) {
            const uint32_t code = args[op.arg_off + q - 1] & 3;
            if (code == 1) C.gate(*find_gate(QCF_OP_XCZ), q - 1, anc - 1);
            else if (code == 2) C.gate(*find_gate(QCF_OP_CNOT), q - 1, anc - 1);
            else if (code == 3) C.gate(*find_gate(QCF_OP_YCX), q - 1, anc - 1);
          }
          // a nonzero phase applies sX to the ancilla: no effect on frame bits
          C.measure(anc - 1, op.bit, reads_after_write[i]);
          C.reset(anc - 1);
          break;
        }
        case QCF_OP_PAULI_ERROR:
          st = op_qubit_check(op, 1, nq);
          if (st == QCF_OK) {
            const double t = op.p[0] / 3;
            st = C.noise1(op.q1 - 1, t, 2 * t, 3 * t, 2);
          }
          break;
        case QCF_OP_PAULI_NOISE:
          st = op_qubit_check(op, 1, nq);
          if (st == QCF_OK) {
            const double px = op.p[0], py = op.p[1], pz = op.p[2];
            const uint32_t w = (px > 0 || py > 0 ? 1 : 0) + (pz > 0 || py > 0 ? 1 : 0);
            st = C.noise1(op.q1 - 1, px, px + pz, px + pz + py, w);
          }
          break;
        case QCF_OP_DEPOLARIZE2:
          st = op_qubit_check(op, 2, nq);
          if (st == QCF_OK) st = C.noise2(op.q1 - 1, op.q2 - 1, op.p[0]);
          break;
        case QCF_OP_CLASSICAL_XOR: {
          for (uint32_t k = 0; k < op.arg_len; ++k) {
            const uint32_t col = args[op.arg_off + k] - 1;
            const uint32_t slot = C.bit_slot[col];
            if (slot != kNone) {
              C.emit(k == 0 ? I_ACC_LDS : I_ACCX_LDS, slot);
              if (--C.reads_left[col] == 0) {
                C.free_slots.push_back(slot);
                C.bit_slot[col] = kNone;
              }
            } else {
              C.emit(k == 0 ? I_ACC_REC : I_ACCX_REC, col);
              if (!C.written[col]) C.zero_first[col] = 1;
            }
          }
          C.write_bit(op.bit - 1, reads_after_write[i], 0, true);
          C.P->bitops += op.arg_len - 1;
          break;
        }
        default:
          st = set_error(QCF_ERR_UNSUPPORTED, "no PauliFrame method for operation kind %u", op.kind);
      }
    }
    if (st != QCF_OK) {
      delete P;
      return st;
    }
  }
  if (int st = C.assign_classes(); st != QCF_OK) {
    delete P;
    return st;
  }
  for (const Site &s : P->sites)
    if (s.accept != 0u && s.accept != (1u << 31)) P->thin = true;
  C.hoist_noise();
  C.peephole();
  P->nplanes = renumber_slots(P->body, 2 * nq);
  C.group_sites();
  if (P->nplanes > 0xFFFF) {
    delete P;
    return set_error(QCF_ERR_UNSUPPORTED, "circuit needs %u on-chip planes", P->nplanes);
  }
  P->nbody_rs = C.rs;
  P->ninit_rs = nq;
  // ---- prologues / epilogue
  for (uint32_t q = 0; q < nq; ++q)
    P->prologue_fresh.push_back(Instr{I_INIT, 2 * q, 2 * q + 1, q, 0, 0, 0, 0});
  for (uint32_t b = 0; b < B; ++b)
    if (!C.written[b] || C.zero_first[b])
      P->prologue_fresh.push_back(Instr{I_REC_ZERO, b, 0, 0, 0, 0, 0, 0});
  for (uint32_t q = 0; q < nq; ++q) {
    P->prologue_inplace.push_back(Instr{I_LOAD, 2 * q, q, 0, 0, 0, 0, 0});
    P->prologue_inplace.push_back(Instr{I_LOAD, 2 * q + 1, nq + q, 0, 0, 0, 0, 0});
    P->epilogue_store.push_back(Instr{I_STORE, C.xs[q], q, 0, 0, 0, 0, 0});
    P->epilogue_store.push_back(Instr{I_STORE, C.zs[q], nq + q, 0, 0, 0, 0, 0});
  }
  mark_dead_gauges(P);
  *out = P;
  return QCF_OK;
}

extern "C" int qcf_program_free(qcf_program *prog) {
  delete prog;
  return QCF_OK;
}

extern "C" int qcf_program_get_info(const qcf_program *P, qcf_program_info *info) {
  clear_error();
  if (!P || !info) return set_error(QCF_ERR_ARGUMENT, "null pointer");
  info->nqubits = P->nq;
  info->nbits = P->nbits;
  info->nplanes = P->nplanes;
  info->ninstr = (uint32_t)P->body.size();
  info->nsites = (uint32_t)P->sites.size();
  info->nclasses = (uint32_t)P->classes.size();
  info->nreplay = P->ninit_rs + P->nbody_rs;
  info->lds_bytes = P->nplanes * 64u * 4u;
  info->bitops_per_shot = P->bitops;
  return QCF_OK;
}

extern "C" int qcf_program_gauge_planes(const qcf_program *P, uint8_t *out, size_t n) {
  clear_error();
  if (!P || (!out && n)) return set_error(QCF_ERR_ARGUMENT, "null pointer");
  const size_t nrep = (size_t)P->ninit_rs + P->nbody_rs;
  if (n < nrep) return set_error(QCF_ERR_DIMENSION, "gauge plane array %zu < %zu", n, nrep);
  std::fill(out, out + n, (uint8_t)0);
  auto mark = [&](const std::vector<Instr> &code, const std::vector<uint8_t> &dead,
                  uint32_t base) {
    for (size_t i = 0; i < code.size(); ++i) {
      const uint32_t op = code[i].op & 0xFFu;
      if (op != I_INIT && op != I_RESETZ && op != I_MR && op != I_MRX) continue;
      const size_t r = (size_t)base + code[i].c;
      if (r < n) out[r] = (i < dead.size() && dead[i]) ? 2 : 1;
    }
  };
  mark(P->prologue_fresh, P->gauge_dead_pro, 0u);
  mark(P->body, P->gauge_dead_body, P->ninit_rs);
  return QCF_OK;
}

extern "C" double qcf_program_bitops(const qcf_program *P) { return P ? P->bitops : 0.0; }
