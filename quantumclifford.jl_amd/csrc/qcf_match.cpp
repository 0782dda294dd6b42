// Detector error model of a compiled frame program and a minimum-weight
// perfect-matching decoder over it (host code).
//
// The reference decodes through `AbstractSyndromeDecoder`
// (src/ecc/decoder_pipeline.jl:1-24, batchdecode :12-24, evaluate_guesses
// :172-189) with matching delegated to PyMatching
// (ext/QuantumCliffordPyQDecodersExt/QuantumCliffordPyQDecodersExt.jl:83-130).
// This file is the repository's own matching decoder for circuit-level
// noise:
//
//   * the error model is read off the compiled program itself: every noise
//     site and Pauli choice is one fault; 64 faults at a time (one per bit of
//     a 64-bit word) are pushed through the instruction stream with all
//     gauge draws zero, and the detector / observable columns they flip are
//     the fault's signature.  Faults with the same signature merge.
//   * faults that flip one or two detectors are edges (one detector: an edge
//     to the boundary); larger ones are split into existing edges when the
//     parts' detectors and observables add up (the X and Z halves of a Y).
//   * edge weight log((1-p)/p) in fixed point; all-pairs shortest paths by
//     Dijkstra with the observable parity of each path.
//   * per shot: the defects plus one boundary copy each form a complete
//     graph; an O(n^3) primal-dual blossom algorithm (Edmonds; dense form)
//     finds the minimum-weight perfect matching, whose paths give the
//     predicted observable flips.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <queue>
#include <thread>
#include <vector>

#include "../../include/qcf.h"
#include "qcf_common.hpp"
#include "qcf_program.hpp"

namespace qcf {
namespace {

constexpr int64_t kInf = std::numeric_limits<int64_t>::max() / 4;
constexpr double kWeightScale = 65536.0;  // fixed-point edge weights

// ---------------------------------------------------------------- blossom
// Maximum-weight matching on a dense graph with positive integer weights
// (0 = no edge).  Vertices 1..n; ids n+1..2n name blossoms.  Duals are kept
// doubled so every update stays integral: an edge (u,v) is tight when
// dual[u] + dual[v] == 2 w(u,v) (blossom duals included through the edge's
// endpoints being inside the blossoms).
class Blossom {
 public:
  explicit Blossom(int n) { reset(n); }

  void reset(int n) {
    n_ = n;
    const int m = 2 * n + 1;
    if ((int)W_.size() < m * m) W_.assign((size_t)m * m, Edge{});
    for (int u = 0; u < m; ++u)
      for (int v = 0; v < m; ++v) E(u, v) = Edge{u, v, 0};
    dual_.assign(m, 0);
    mate_.assign(m, 0);
    slack_.assign(m, 0);
    top_.assign(m, 0);
    parent_.assign(m, 0);
    label_.assign(m, -1);
    seen_.assign(m, 0);
    kids_.assign(m, {});
    if ((int)entry_.size() < m * (n + 1)) entry_.assign((size_t)m * (n + 1), 0);
    std::fill(entry_.begin(), entry_.begin() + (size_t)m * (n + 1), 0);
    stamp_ = 0;
  }

  void set_weight(int u, int v, int64_t w) {
    E(u, v).w = w;
    E(v, u).w = w;
  }

  // Returns mate[1..n] (0 = unmatched).
  const std::vector<int> &solve() {
    nb_ = n_;
    int64_t wmax = 0;
    for (int u = 1; u <= n_; ++u) {
      top_[u] = u;
      kids_[u].clear();
      for (int v = 1; v <= n_; ++v) {
        from(u, v) = u == v ? u : 0;
        wmax = std::max(wmax, E(u, v).w);
      }
    }
    for (int u = 1; u <= n_; ++u) dual_[u] = wmax;
    while (phase()) {
    }
    return mate_;
  }

 private:
  struct Edge {
    int u = 0, v = 0;
    int64_t w = 0;
  };
  int n_ = 0, nb_ = 0, stamp_ = 0;
  std::vector<Edge> W_;
  std::vector<int64_t> dual_;
  std::vector<int> mate_, slack_, top_, parent_, label_, seen_;
  std::vector<std::vector<int>> kids_;  // blossom -> sub-blossoms in cycle order
  std::vector<int> entry_;              // (blossom, vertex) -> sub-blossom holding it
  std::vector<int> queue_;

  Edge &E(int u, int v) { return W_[(size_t)u * (2 * n_ + 1) + v]; }
  int &from(int b, int x) { return entry_[(size_t)b * (n_ + 1) + x]; }
  int64_t slack_of(const Edge &e) { return dual_[e.u] + dual_[e.v] - 2 * E(e.u, e.v).w; }

  void note_slack(int u, int x) {
    if (!slack_[x] || slack_of(E(u, x)) < slack_of(E(slack_[x], x))) slack_[x] = u;
  }
  void recompute_slack(int x) {
    slack_[x] = 0;
    for (int u = 1; u <= n_; ++u)
      if (E(u, x).w > 0 && top_[u] != x && label_[top_[u]] == 0) note_slack(u, x);
  }
  void enqueue(int x) {
    if (x <= n_) {
      queue_.push_back(x);
    } else {
      for (int t : kids_[x]) enqueue(t);
    }
  }
  void set_top(int x, int b) {
    top_[x] = b;
    if (x > n_)
      for (int t : kids_[x]) set_top(t, b);
  }
  // Position of sub-blossom xr in b's cycle, reorienting the cycle so that
  // the path from the base to xr has even length.
  int even_position(int b, int xr) {
    auto &k = kids_[b];
    const int pr = (int)(std::find(k.begin(), k.end(), xr) - k.begin());
    if (pr % 2 == 1) {
      std::reverse(k.begin() + 1, k.end());
      return (int)k.size() - pr;
    }
    return pr;
  }
  void match_edge(int u, int v) {
    mate_[u] = E(u, v).v;
    if (u <= n_) return;
    const Edge e = E(u, v);
    const int xr = from(u, e.u);
    const int pr = even_position(u, xr);
    for (int i = 0; i < pr; ++i) match_edge(kids_[u][i], kids_[u][i ^ 1]);
    match_edge(xr, v);
    std::rotate(kids_[u].begin(), kids_[u].begin() + pr, kids_[u].end());
  }
  void augment(int u, int v) {
    for (;;) {
      const int xnv = top_[mate_[u]];
      match_edge(u, v);
      if (!xnv) return;
      match_edge(xnv, top_[parent_[xnv]]);
      u = top_[parent_[xnv]];
      v = xnv;
    }
  }
  int common_ancestor(int u, int v) {
    ++stamp_;
    while (u || v) {
      if (u) {
        if (seen_[u] == stamp_) return u;
        seen_[u] = stamp_;
        u = top_[mate_[u]];
        if (u) u = top_[parent_[u]];
      }
      std::swap(u, v);
    }
    return 0;
  }
  void make_blossom(int u, int base, int v) {
    int b = n_ + 1;
    while (b <= nb_ && top_[b]) ++b;
    if (b > nb_) ++nb_;
    dual_[b] = 0;
    label_[b] = 0;
    mate_[b] = mate_[base];
    auto &k = kids_[b];
    k.clear();
    k.push_back(base);
    for (int x = u, y; x != base; x = top_[parent_[y]]) {
      k.push_back(x);
      k.push_back(y = top_[mate_[x]]);
      enqueue(y);
    }
    std::reverse(k.begin() + 1, k.end());
    for (int x = v, y; x != base; x = top_[parent_[y]]) {
      k.push_back(x);
      k.push_back(y = top_[mate_[x]]);
      enqueue(y);
    }
    set_top(b, b);
    for (int x = 1; x <= nb_; ++x) {
      E(b, x).w = 0;
      E(x, b).w = 0;
    }
    for (int x = 1; x <= n_; ++x) from(b, x) = 0;
    for (int xs : k) {
      for (int x = 1; x <= nb_; ++x)
        if (E(b, x).w == 0 || slack_of(E(xs, x)) < slack_of(E(b, x))) {
          E(b, x) = E(xs, x);
          E(x, b) = E(x, xs);
        }
      for (int x = 1; x <= n_; ++x)
        if (from(xs, x)) from(b, x) = xs;
    }
    recompute_slack(b);
  }
  void expand_blossom(int b) {
    for (int t : kids_[b]) set_top(t, t);
    const int xr = from(b, E(b, parent_[b]).u);
    const int pr = even_position(b, xr);
    for (int i = 0; i < pr; i += 2) {
      const int xs = kids_[b][i], xns = kids_[b][i + 1];
      parent_[xs] = E(xns, xs).u;
      label_[xs] = 1;
      label_[xns] = 0;
      slack_[xs] = 0;
      recompute_slack(xns);
      enqueue(xns);
    }
    label_[xr] = 1;
    parent_[xr] = parent_[b];
    for (size_t i = pr + 1; i < kids_[b].size(); ++i) {
      const int xs = kids_[b][i];
      label_[xs] = -1;
      recompute_slack(xs);
    }
    top_[b] = 0;
  }
  // A tight edge from an outer vertex: grow the tree, shrink a blossom, or
  // augment (returns true).
  bool tight_edge(const Edge &e) {
    const int u = top_[e.u], v = top_[e.v];
    if (label_[v] == -1) {
      parent_[v] = e.u;
      label_[v] = 1;
      const int nu = top_[mate_[v]];
      slack_[v] = slack_[nu] = 0;
      label_[nu] = 0;
      enqueue(nu);
    } else if (label_[v] == 0) {
      const int base = common_ancestor(u, v);
      if (!base) {
        augment(u, v);
        augment(v, u);
        return true;
      }
      make_blossom(u, base, v);
    }
    return false;
  }
  // One search for an augmenting path with dual adjustments; false when the
  // matching is maximum.
  bool phase() {
    std::fill(label_.begin(), label_.begin() + nb_ + 1, -1);
    std::fill(slack_.begin(), slack_.begin() + nb_ + 1, 0);
    queue_.clear();
    for (int x = 1; x <= nb_; ++x)
      if (top_[x] == x && !mate_[x]) {
        parent_[x] = 0;
        label_[x] = 0;
        enqueue(x);
      }
    if (queue_.empty()) return false;
    for (;;) {
      for (size_t qi = 0; qi < queue_.size(); ++qi) {
        const int u = queue_[qi];
        if (label_[top_[u]] == 1) continue;
        for (int v = 1; v <= n_; ++v)
          if (E(u, v).w > 0 && top_[u] != top_[v]) {
            if (slack_of(E(u, v)) == 0) {
              if (tight_edge(E(u, v))) return true;
            } else {
              note_slack(u, top_[v]);
            }
          }
      }
      queue_.clear();
      int64_t d = kInf;
      for (int b = n_ + 1; b <= nb_; ++b)
        if (top_[b] == b && label_[b] == 1) d = std::min(d, dual_[b] / 2);
      for (int x = 1; x <= nb_; ++x)
        if (top_[x] == x && slack_[x]) {
          if (label_[x] == -1)
            d = std::min(d, slack_of(E(slack_[x], x)));
          else if (label_[x] == 0)
            d = std::min(d, slack_of(E(slack_[x], x)) / 2);
        }
      for (int u = 1; u <= n_; ++u) {
        if (label_[top_[u]] == 0) {
          if (dual_[u] <= d) return false;
          dual_[u] -= d;
        } else if (label_[top_[u]] == 1) {
          dual_[u] += d;
        }
      }
      for (int b = n_ + 1; b <= nb_; ++b)
        if (top_[b] == b) {
          if (label_[b] == 0)
            dual_[b] += 2 * d;
          else if (label_[b] == 1)
            dual_[b] -= 2 * d;
        }
      for (int x = 1; x <= nb_; ++x)
        if (top_[x] == x && slack_[x] && top_[slack_[x]] != x &&
            slack_of(E(slack_[x], x)) == 0)
          if (tight_edge(E(slack_[x], x))) return true;
      for (int b = n_ + 1; b <= nb_; ++b)
        if (top_[b] == b && label_[b] == 1 && dual_[b] == 0) expand_blossom(b);
    }
  }
};

}  // namespace
}  // namespace qcf

// ------------------------------------------------------------ error model
struct qcf_dem {
  struct Mech {
    std::vector<uint32_t> dets;  // sorted detector indices
    uint64_t obs = 0;            // observable mask
    double p = 0;
  };
  uint32_t ndet = 0, nobs = 0;
  std::vector<Mech> mechs;
};

struct qcf_matcher {
  uint32_t ndet = 0, nobs = 0;
  // graph edges (boundary = node ndet)
  struct GEdge {
    uint32_t a, b;
    int64_t w;
    uint64_t obs;
    double p;
    std::vector<uint32_t> mechs;  // error-model mechanisms this edge stands for
    uint32_t rep = 0;             // graph-like mechanism with the edge's observables
  };
  std::vector<GEdge> edges;
  std::vector<std::vector<std::pair<uint32_t, uint32_t>>> adj;  // node -> (nbr, edge)
  // all-pairs shortest paths over nodes 0..ndet (ndet = boundary)
  std::vector<int64_t> dist;     // (ndet+1)^2
  std::vector<uint64_t> par;     // observable parity of the path
  std::vector<int32_t> pred;     // predecessor edge on the path from the row node
  uint32_t dropped = 0;          // mechanisms that did not decompose into edges
  size_t nmech = 0;
};

namespace qcf {
namespace {

// Push faults through the program, 64 at a time.
int build_dem(const qcf_program *P, const uint32_t *det_bits, size_t ndet,
              const uint32_t *obs_bits, size_t nobs, qcf_dem *D) {
  if (nobs > 64) return set_error(QCF_ERR_UNSUPPORTED, "at most 64 observables");
  std::vector<int32_t> det_of(P->nbits, -1), obs_of(P->nbits, -1);
  for (size_t i = 0; i < ndet; ++i) {
    if (det_bits[i] == 0 || det_bits[i] > P->nbits)
      return set_error(QCF_ERR_BOUNDS, "detector bit %u outside the record (%u bits)",
                       det_bits[i], P->nbits);
    det_of[det_bits[i] - 1] = (int32_t)i;
  }
  for (size_t i = 0; i < nobs; ++i) {
    if (obs_bits[i] == 0 || obs_bits[i] > P->nbits)
      return set_error(QCF_ERR_BOUNDS, "observable bit %u outside the record (%u bits)",
                       obs_bits[i], P->nbits);
    obs_of[obs_bits[i] - 1] = (int32_t)i;
  }
  D->ndet = (uint32_t)ndet;
  D->nobs = (uint32_t)nobs;
  // faults: (site, pauli, p); 1q pauli 1=X 2=Z 3=Y; 2q base-4 digits
  struct Fault {
    uint32_t site, pauli;
    double p;
  };
  std::vector<Fault> faults;
  std::vector<uint32_t> site_of_ordinal;
  for (uint32_t i = 0; i < P->sites.size(); ++i) {
    const Site &s = P->sites[i];
    if (s.cls != 0) continue;
    if (site_of_ordinal.size() <= s.ordinal) site_of_ordinal.resize(s.ordinal + 1, kNone);
    site_of_ordinal[s.ordinal] = i;
    if (s.kind == 1) {
      const double px = s.c1, pz = s.c2 - s.c1, py = s.c3 - s.c2;
      if (px > 0) faults.push_back({i, 1, px});
      if (pz > 0) faults.push_back({i, 2, pz});
      if (py > 0) faults.push_back({i, 3, py});
    } else if (s.c3 > 0) {
      for (uint32_t k = 1; k < 16; ++k) faults.push_back({i, k, s.c3 / 15.0});
    }
  }
  const std::vector<Instr> code = variant_code(P, 1);
  std::vector<uint64_t> pl(std::max<uint32_t>(P->nplanes, 1)), rec(std::max<uint32_t>(P->nbits, 1));
  std::map<std::pair<std::vector<uint32_t>, uint64_t>, size_t> index;
  std::vector<std::vector<std::pair<uint32_t, uint32_t>>> at_site(P->sites.size());
  for (size_t f0 = 0; f0 < faults.size(); f0 += 64) {
    const size_t nf = std::min<size_t>(64, faults.size() - f0);
    for (auto &v : at_site) v.clear();
    for (size_t k = 0; k < nf; ++k)
      at_site[faults[f0 + k].site].push_back({(uint32_t)k, faults[f0 + k].pauli});
    std::fill(pl.begin(), pl.end(), 0);
    std::fill(rec.begin(), rec.end(), 0);
    uint64_t acc = 0;
    for (size_t i = 0; i < code.size(); ++i) {
      const Instr &in = code[i];
      const uint32_t op = in.op & 0xFFu;
      switch (op) {
        case I_END:
          break;
        case I_XORN: {
          const uint32_t n = in.op >> 8;
          const uint32_t *w = &in.a;
          uint64_t v[7];
          for (uint32_t k = 0; k < n; ++k) v[k] = pl[w[k] >> 16];
          for (uint32_t k = 0; k < n; ++k) pl[w[k] & 0xFFFFu] ^= v[k];
          break;
        }
        case I_NOISEG: {
          const uint32_t G = in.op >> 8;
          for (uint32_t k = 0; k < G; ++k) {
            const Instr &s = code[i + 1 + k];
            const uint32_t ord = in.a + k;
            if (ord >= site_of_ordinal.size() || site_of_ordinal[ord] == kNone) continue;
            const uint32_t sid = site_of_ordinal[ord];
            const uint32_t pls[4] = {s.a & 0xFFFFu, s.a >> 16, s.b & 0xFFFFu, s.b >> 16};
            for (const auto &fb : at_site[sid]) {
              const uint64_t bit = 1ull << fb.first;
              const uint32_t d1 = fb.second & 3u, d2 = fb.second >> 2;
              if (d1 & 1u) pl[pls[0]] ^= bit;  // X or Y
              if (d1 >= 2u) pl[pls[1]] ^= bit; // Z or Y
              if (s.b) {
                if (d2 & 1u) pl[pls[2]] ^= bit;
                if (d2 >= 2u) pl[pls[3]] ^= bit;
              }
            }
          }
          i += G;
          break;
        }
        case I_SITE:
          break;
        case I_MEAS: {
          const uint64_t v = pl[in.a];
          rec[in.b] = v;
          if (in.c != kNone) pl[in.c] = v;
          break;
        }
        case I_XSTORE2: {
          const uint64_t v = pl[in.a] ^ pl[in.b];
          rec[in.c] = v;
          if (in.d != kNone) pl[in.d] = v;
          break;
        }
        case I_RESETZ:
          pl[in.a] = 0;
          break;
        case I_ACC_LDS: acc = pl[in.a]; break;
        case I_ACCX_LDS: acc ^= pl[in.a]; break;
        case I_ACC_REC: acc = rec[in.a]; break;
        case I_ACCX_REC: acc ^= rec[in.a]; break;
        case I_ACC_STORE:
          rec[in.a] = acc;
          if (in.b != kNone) pl[in.b] = acc;
          break;
        case I_REC_ZERO: rec[in.a] = 0; break;
        case I_INIT:
          pl[in.a] = 0;
          pl[in.b] = 0;
          break;
        case I_ZERO: pl[in.a] = 0; break;
        case I_LIN1: {
          const uint64_t x = pl[in.a], z = pl[in.b];
          pl[in.a] = ((in.c & 1u) ? x : 0) ^ ((in.c & 2u) ? z : 0);
          pl[in.b] = ((in.c & 4u) ? x : 0) ^ ((in.c & 8u) ? z : 0);
          break;
        }
        case I_MR:
        case I_MRX: {
          const uint64_t v = pl[in.a];
          if (in.d != kNone) rec[in.d] = v;
          if (op == I_MRX) rec[in.g] = v ^ pl[in.f];
          if (in.e != kNone) pl[in.e] = v;
          pl[in.a] = 0;
          break;
        }
        default:
          return set_error(QCF_ERR_UNSUPPORTED, "error model: instruction %u", op);
      }
    }
    for (size_t k = 0; k < nf; ++k) {
      std::vector<uint32_t> dets;
      uint64_t obs = 0;
      for (uint32_t b = 0; b < P->nbits; ++b) {
        if (!((rec[b] >> k) & 1u)) continue;
        if (det_of[b] >= 0) dets.push_back((uint32_t)det_of[b]);
        if (obs_of[b] >= 0) obs |= 1ull << obs_of[b];
      }
      if (dets.empty() && !obs) continue;  // undetectable and harmless
      std::sort(dets.begin(), dets.end());
      const double p = faults[f0 + k].p;
      auto key = std::make_pair(dets, obs);
      auto it = index.find(key);
      if (it == index.end()) {
        index.emplace(key, D->mechs.size());
        qcf_dem::Mech m;
        m.dets = std::move(dets);
        m.obs = obs;
        m.p = p;
        D->mechs.push_back(std::move(m));
      } else {
        double &q = D->mechs[it->second].p;
        q = q * (1 - p) + p * (1 - q);
      }
    }
  }
  return QCF_OK;
}

int64_t weight_of(double p) {
  p = std::min(std::max(p, 1e-300), 0.5 - 1e-12);
  return std::max<int64_t>(1, (int64_t)std::llround(std::log((1 - p) / p) * kWeightScale));
}

int build_matcher(const qcf_dem *D, qcf_matcher *M) {
  M->ndet = D->ndet;
  M->nobs = D->nobs;
  M->nmech = D->mechs.size();
  const uint32_t B = D->ndet;  // boundary node
  std::map<std::pair<uint32_t, uint32_t>, size_t> eidx;
  auto add_edge = [&](uint32_t a, uint32_t b, uint64_t obs, double p, uint32_t mech) {
    if (a > b) std::swap(a, b);
    auto it = eidx.find({a, b});
    if (it == eidx.end()) {
      eidx.emplace(std::make_pair(a, b), M->edges.size());
      M->edges.push_back({a, b, 0, obs, p, {mech}});
      return;
    }
    auto &e = M->edges[it->second];
    if (e.obs != obs && p > e.p) e.obs = obs;  // disagreeing parallels: the likelier wins
    e.p = e.p * (1 - p) + p * (1 - e.p);
    e.mechs.push_back(mech);
  };
  // graph-like mechanisms first
  for (uint32_t i = 0; i < D->mechs.size(); ++i) {
    const auto &m = D->mechs[i];
    if (m.dets.size() == 1) add_edge(m.dets[0], B, m.obs, m.p, i);
    if (m.dets.size() == 2) add_edge(m.dets[0], m.dets[1], m.obs, m.p, i);
  }
  // larger ones: split into existing edges whose observables add up
  std::map<std::pair<uint32_t, uint32_t>, uint64_t> known;
  for (const auto &e : M->edges) known[{e.a, e.b}] = e.obs;
  for (uint32_t i = 0; i < D->mechs.size(); ++i) {
    const auto &m = D->mechs[i];
    if (m.dets.size() <= 2) continue;
    if (m.dets.empty()) continue;
    // search pairings of the detectors into parts of size 1 (boundary) or 2
    std::vector<std::pair<uint32_t, uint32_t>> parts, best;
    std::vector<char> used(m.dets.size(), 0);
    bool found = false;
    std::function<void(uint64_t)> rec = [&](uint64_t obs) {
      if (found) return;
      size_t first = 0;
      while (first < used.size() && used[first]) ++first;
      if (first == used.size()) {
        if (obs == m.obs) {
          found = true;
          best = parts;
        }
        return;
      }
      used[first] = 1;
      for (size_t j = first + 1; j < used.size() && !found; ++j) {
        if (used[j]) continue;
        auto it = known.find({m.dets[first], m.dets[j]});
        if (it == known.end()) continue;
        used[j] = 1;
        parts.push_back({m.dets[first], m.dets[j]});
        rec(obs ^ it->second);
        parts.pop_back();
        used[j] = 0;
      }
      if (!found) {
        auto it = known.find({m.dets[first], B});
        if (it != known.end()) {
          parts.push_back({m.dets[first], B});
          rec(obs ^ it->second);
          parts.pop_back();
        }
      }
      used[first] = 0;
    };
    if (m.dets.size() <= 8) rec(0);
    if (!found) {
      ++M->dropped;
      continue;
    }
    for (const auto &pt : best) add_edge(pt.first, pt.second, known[pt], m.p, i);
  }
  for (auto &e : M->edges) {
    e.w = weight_of(e.p);
    e.rep = e.mechs[0];
    for (uint32_t mi : e.mechs)
      if (D->mechs[mi].dets.size() <= 2 && D->mechs[mi].obs == e.obs) {
        e.rep = mi;
        break;
      }
  }
  const uint32_t N = D->ndet + 1;
  M->adj.assign(N, {});
  for (uint32_t k = 0; k < M->edges.size(); ++k) {
    M->adj[M->edges[k].a].push_back({M->edges[k].b, k});
    M->adj[M->edges[k].b].push_back({M->edges[k].a, k});
  }
  M->dist.assign((size_t)N * N, kInf);
  M->par.assign((size_t)N * N, 0);
  M->pred.assign((size_t)N * N, -1);
  // Dijkstra from every node (the boundary included); rows are independent
  const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::atomic<uint32_t> next{0};
  auto worker = [&]() {
    std::vector<std::pair<int64_t, uint32_t>> heap;
    for (uint32_t s; (s = next.fetch_add(1)) < N;) {
      int64_t *d = &M->dist[(size_t)s * N];
      uint64_t *pa = &M->par[(size_t)s * N];
      int32_t *pr = &M->pred[(size_t)s * N];
      d[s] = 0;
      heap.clear();
      heap.push_back({0, s});
      while (!heap.empty()) {
        std::pop_heap(heap.begin(), heap.end(), std::greater<>());
        const auto [du, u] = heap.back();
        heap.pop_back();
        if (du != d[u]) continue;
        for (const auto &[v, ei] : M->adj[u]) {
          // paths do not pass through the boundary
          const int64_t nd = du + M->edges[ei].w;
          if (nd < d[v]) {
            d[v] = nd;
            pa[v] = pa[u] ^ M->edges[ei].obs;
            pr[v] = (int32_t)ei;
            if (v != D->ndet || s == D->ndet) {
              heap.push_back({nd, v});
              std::push_heap(heap.begin(), heap.end(), std::greater<>());
            }
          }
        }
      }
    }
  };
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nth; ++t) th.emplace_back(worker);
  for (auto &t : th) t.join();
  return QCF_OK;
}

// Decode one syndrome given as a list of defect detectors.  Writes the
// predicted observable mask and, if `edges_out` is given, the graph edges
// along the matched paths.
uint64_t decode_defects(const qcf_matcher *M, const std::vector<uint32_t> &defects, Blossom &bl,
                        std::vector<uint32_t> *edges_out) {
  const uint32_t N = M->ndet + 1, B = M->ndet;
  const int k = (int)defects.size();
  if (k == 0) return 0;
  auto dist = [&](uint32_t a, uint32_t b) { return M->dist[(size_t)a * N + b]; };
  auto par = [&](uint32_t a, uint32_t b) { return M->par[(size_t)a * N + b]; };
  std::vector<std::pair<uint32_t, uint32_t>> pairs;
  if (k == 1) {
    pairs.push_back({defects[0], B});
  } else if (k == 2 && dist(defects[0], defects[1]) <=
                           dist(defects[0], B) + dist(defects[1], B)) {
    pairs.push_back({defects[0], defects[1]});
  } else if (k == 2) {
    pairs.push_back({defects[0], B});
    pairs.push_back({defects[1], B});
  } else {
    // defects 1..k and their boundary copies k+1..2k; M - w turns the
    // minimum-weight perfect matching into a maximum-weight one
    int64_t wmax = 0;
    for (int i = 0; i < k; ++i) {
      wmax = std::max(wmax, std::min(dist(defects[i], B), kInf / 8));
      for (int j = i + 1; j < k; ++j)
        wmax = std::max(wmax, std::min(dist(defects[i], defects[j]), kInf / 8));
    }
    const int64_t big = (int64_t)(k + 1) * wmax + 1;
    bl.reset(2 * k);
    for (int i = 0; i < k; ++i) {
      const int64_t db = dist(defects[i], B);
      if (db < kInf) bl.set_weight(1 + i, 1 + k + i, big - db);
      for (int j = i + 1; j < k; ++j) {
        const int64_t dij = dist(defects[i], defects[j]);
        if (dij < kInf) bl.set_weight(1 + i, 1 + j, big - dij);
        bl.set_weight(1 + k + i, 1 + k + j, big);
      }
    }
    const std::vector<int> &mate = bl.solve();
    for (int i = 0; i < k; ++i) {
      const int m = mate[1 + i];
      if (m == 0) continue;  // no perfect matching: leave unmatched
      if (m > k)
        pairs.push_back({defects[i], B});
      else if (m - 1 > i)
        pairs.push_back({defects[i], defects[m - 1]});
    }
  }
  uint64_t obs = 0;
  for (const auto &pr : pairs) {
    obs ^= par(pr.first, pr.second);
    if (edges_out) {
      // walk back from pr.second along predecessor edges of row pr.first
      uint32_t v = pr.second;
      while (v != pr.first) {
        const int32_t ei = M->pred[(size_t)pr.first * N + v];
        if (ei < 0) break;
        edges_out->push_back((uint32_t)ei);
        const auto &e = M->edges[ei];
        v = e.a == v ? e.b : e.a;
      }
    }
  }
  return obs;
}

template <class GetDefects>
int decode_many(const qcf_matcher *M, uint64_t shots, int nthreads, GetDefects get,
                uint8_t *obs_out, uint8_t *mech_out) {
  const unsigned nth = nthreads > 0 ? (unsigned)nthreads
                                    : std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::atomic<uint64_t> next{0};
  constexpr uint64_t kChunkShots = 256;
  auto worker = [&]() {
    Blossom bl(1);
    std::vector<uint32_t> defects, edges;
    for (uint64_t s0; (s0 = next.fetch_add(kChunkShots)) < shots;) {
      const uint64_t s1 = std::min(shots, s0 + kChunkShots);
      for (uint64_t s = s0; s < s1; ++s) {
        defects.clear();
        get(s, &defects);
        edges.clear();
        const uint64_t o = decode_defects(M, defects, bl, mech_out ? &edges : nullptr);
        for (uint32_t j = 0; j < M->nobs; ++j) obs_out[s * M->nobs + j] = (uint8_t)((o >> j) & 1u);
        if (mech_out) {
          uint8_t *row = mech_out + s * M->nmech;
          std::memset(row, 0, M->nmech);
          for (uint32_t ei : edges) row[M->edges[ei].rep] ^= 1u;
        }
      }
    }
  };
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nth; ++t) th.emplace_back(worker);
  for (auto &t : th) t.join();
  return QCF_OK;
}

}  // namespace
}  // namespace qcf

using namespace qcf;

extern "C" int qcf_dem_build(const qcf_program *prog, const uint32_t *det_bits, size_t ndet,
                             const uint32_t *obs_bits, size_t nobs, qcf_dem **out) {
  clear_error();
  if (!prog || !out || (ndet && !det_bits) || (nobs && !obs_bits))
    return set_error(QCF_ERR_ARGUMENT, "null argument");
  *out = nullptr;
  qcf_dem *D = new qcf_dem();
  const int st = build_dem(prog, det_bits, ndet, obs_bits, nobs, D);
  if (st != QCF_OK) {
    delete D;
    return st;
  }
  *out = D;
  return QCF_OK;
}

extern "C" int qcf_dem_free(qcf_dem *dem) {
  delete dem;
  return QCF_OK;
}

extern "C" int qcf_dem_size(const qcf_dem *dem, size_t *nmech, uint32_t *ndet, uint32_t *nobs) {
  clear_error();
  if (!dem) return set_error(QCF_ERR_ARGUMENT, "null error model");
  if (nmech) *nmech = dem->mechs.size();
  if (ndet) *ndet = dem->ndet;
  if (nobs) *nobs = dem->nobs;
  return QCF_OK;
}

extern "C" int qcf_dem_mechanism(const qcf_dem *dem, size_t i, uint32_t *dets, size_t cap,
                                 size_t *ndets, uint64_t *obs, double *p) {
  clear_error();
  if (!dem) return set_error(QCF_ERR_ARGUMENT, "null error model");
  if (i >= dem->mechs.size()) return set_error(QCF_ERR_BOUNDS, "mechanism %zu of %zu", i, dem->mechs.size());
  const auto &m = dem->mechs[i];
  if (ndets) *ndets = m.dets.size();
  if (dets)
    for (size_t k = 0; k < std::min(cap, m.dets.size()); ++k) dets[k] = m.dets[k];
  if (obs) *obs = m.obs;
  if (p) *p = m.p;
  return QCF_OK;
}

extern "C" int qcf_matcher_create(const qcf_dem *dem, qcf_matcher **out) {
  clear_error();
  if (!dem || !out) return set_error(QCF_ERR_ARGUMENT, "null argument");
  *out = nullptr;
  qcf_matcher *M = new qcf_matcher();
  const int st = build_matcher(dem, M);
  if (st != QCF_OK) {
    delete M;
    return st;
  }
  *out = M;
  return QCF_OK;
}

extern "C" int qcf_matcher_free(qcf_matcher *m) {
  delete m;
  return QCF_OK;
}

extern "C" int qcf_matcher_info(const qcf_matcher *m, size_t *nedges, uint32_t *dropped) {
  clear_error();
  if (!m) return set_error(QCF_ERR_ARGUMENT, "null matcher");
  if (nedges) *nedges = m->edges.size();
  if (dropped) *dropped = m->dropped;
  return QCF_OK;
}

extern "C" int qcf_matcher_decode(const qcf_matcher *m, const uint8_t *syndromes, uint64_t shots,
                                  uint8_t *obs_out, uint8_t *mech_out, int nthreads) {
  clear_error();
  if (!m || (shots && (!syndromes || !obs_out))) return set_error(QCF_ERR_ARGUMENT, "null argument");
  const uint32_t nd = m->ndet;
  return decode_many(
      m, shots, nthreads,
      [&](uint64_t s, std::vector<uint32_t> *d) {
        const uint8_t *row = syndromes + s * nd;
        for (uint32_t j = 0; j < nd; ++j)
          if (row[j]) d->push_back(j);
      },
      obs_out, mech_out);
}

extern "C" int qcf_matcher_decode_packed(const qcf_matcher *m, const uint64_t *dets, size_t ld,
                                         uint64_t shots, uint8_t *obs_out, int nthreads) {
  clear_error();
  if (!m || (shots && (!dets || !obs_out))) return set_error(QCF_ERR_ARGUMENT, "null argument");
  if (ld * 64 < shots) return set_error(QCF_ERR_DIMENSION, "ld %zu words < %llu shots", ld,
                                         (unsigned long long)shots);
  // one 64-shot word at a time: each detector word is read once and its set
  // bits distributed to the shots' defect lists
  const uint32_t nd = m->ndet;
  const uint64_t nwords = (shots + 63) / 64;
  const unsigned nth = nthreads > 0 ? (unsigned)nthreads
                                    : std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::atomic<uint64_t> next{0};
  auto worker = [&]() {
    Blossom bl(1);
    std::vector<std::vector<uint32_t>> lists(64);
    for (uint64_t wi; (wi = next.fetch_add(1)) < nwords;) {
      for (auto &l : lists) l.clear();
      for (uint32_t j = 0; j < nd; ++j) {
        uint64_t w = dets[(size_t)j * ld + wi];
        while (w) {
          lists[__builtin_ctzll(w)].push_back(j);
          w &= w - 1;
        }
      }
      const uint64_t s0 = wi * 64, s1 = std::min(shots, s0 + 64);
      for (uint64_t s = s0; s < s1; ++s) {
        const uint64_t o = decode_defects(m, lists[s - s0], bl, nullptr);
        for (uint32_t j = 0; j < m->nobs; ++j) obs_out[s * m->nobs + j] = (uint8_t)((o >> j) & 1u);
      }
    }
  };
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nth; ++t) th.emplace_back(worker);
  for (auto &t : th) t.join();
  return QCF_OK;
}

extern "C" int qcf_debug_max_weight_matching(int n, const int64_t *w, int32_t *mate) {
  clear_error();
  if (n < 0 || (n && (!w || !mate))) return set_error(QCF_ERR_ARGUMENT, "bad argument");
  if (n == 0) return QCF_OK;
  Blossom bl(n);
  for (int u = 0; u < n; ++u)
    for (int v = u + 1; v < n; ++v)
      if (w[(size_t)u * n + v] > 0) bl.set_weight(u + 1, v + 1, w[(size_t)u * n + v]);
  const std::vector<int> &m = bl.solve();
  for (int u = 0; u < n; ++u) mate[u] = m[u + 1] - 1;
  return QCF_OK;
}
