// Device program of the Pauli-frame sampler.
//
// A circuit is lowered on the host into a flat stream of 32-byte
// instructions over "planes".  A plane is one 32-bit word per lane of a
// 64-lane wave, i.e. the x bit or the z bit of one qubit for 2048 shots, or a
// classical bit kept on chip.  Gates whose bit map is a permutation of planes
// (H swaps x and z, SWAP exchanges qubits, ...) cost nothing on the device:
// the compiler renames planes instead.  Every other symbolic Clifford becomes
// a few in-place plane XORs.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/qcf.h"

namespace qcf {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kBlockShots = 2048;  // shots per wave tile (64 lanes x 32)
constexpr uint32_t kMaxClasses = 1;     // noise-rate classes per program (others thinned)
constexpr uint32_t kNoFire = 0xFFu;     // site class of a zero-rate site
constexpr uint32_t kMaxOrdinals = 1u << 20;  // sites per class
constexpr uint32_t kChunk = 64;          // instructions fetched per wave step

enum InstrOp : uint32_t {
  I_END = 0,
  I_XORN,       // n = op>>8 (1..7) pairs in words a..g: P[lo16] ^= P[hi16]; all
                // sources are read before any target is written (the host only
                // bundles pairs where that equals sequential order)
  I_NOISE1,     // (compiler internal) single-qubit site c on planes x=a, z=b
  I_NOISE2,     // (compiler internal) two-qubit site c on x1|z1<<16 = a, x2|z2<<16 = b
  I_NOISEG,     // G = op>>8 noise sites with consecutive ordinals starting at a;
                // their data are the next G instructions (I_SITE), always in
                // the same 64-instruction chunk
  I_SITE,       // site data: a = x1|z1<<16, b = x2|z2<<16 (0: one qubit),
                // c = replay plane, d = t1|t2<<16, e = accept (see qcf_device.h)
  I_MEAS,       // v = P[a]; rec[b] = v; if c != none: P[c] = v
  I_XSTORE2,    // v = P[a] ^ P[b]; rec[c] = v; if d != none: P[d] = v
  I_RESETZ,     // P[a] = 0; P[b] ^= draw(c)           (sMRZ reset part)
  I_ACC_LDS,    // acc = P[a]
  I_ACCX_LDS,   // acc ^= P[a]
  I_ACC_REC,    // acc = rec[a]
  I_ACCX_REC,   // acc ^= rec[a]
  I_ACC_STORE,  // rec[a] = acc; if b != none: P[b] = acc
  I_REC_ZERO,   // rec[a] = 0
  I_INIT,       // P[a] = 0; P[b] = draw(c)            (PauliFrame + initZ!)
  I_LOAD,       // P[a] = frame[b]
  I_STORE,      // frame[b] = P[a]
  I_ZERO,       // P[a] = 0
  I_LIN1,       // (P[a],P[b]) := linear map c (bit0 x<-x, bit1 x<-z, bit2 z<-x, bit3 z<-z)
  I_MR,         // sMRZ: v = P[a]; if d != none: rec[d] = v; if e != none: P[e] = v;
                //       P[a] = 0; P[b] ^= draw(c)
  I_MRX,        // I_MR, then the detector rec[g] = v ^ P[f]
};

// Flag in the op word of I_INIT, I_RESETZ, I_MR and I_MRX (fresh record-only
// code only, see variant_code): the instruction's gauge draw reaches no record
// bit beyond what the draws kept before it already span, so a native launch
// makes no draw there (the z plane is left alone, or zero for I_INIT) and a
// mask export leaves that mask plane as the caller zeroed it.  Replay still
// applies the mask it is given.
constexpr uint32_t kNoGauge = 1u << 8;

// 32-byte device instruction.  Thresholds of I_SITE are 2^31-scaled and
// compared with the top 31 bits of a draw.
struct Instr {
  uint32_t op, a, b, c, d, e, f, g;
};

// One noise site.  Native sampling draws events at rate p_class along the
// class's (site ordinal, shot) sequence and keeps each with probability
// accept / 2^31 (thinning to this site's own rate).  Given an event, the
// Pauli is X if u < t1, Z if u < t2, else Y (single qubit), or one of the 15
// non-identity two-qubit Paulis uniformly (kind 2).
struct Site {
  uint32_t cls;      // class index, or kNoFire
  uint32_t ordinal;  // index of this site within its class
  uint32_t accept;   // thinning threshold in [0, 2^31]
  uint32_t rslot;    // first replay plane (body relative)
  uint32_t t1, t2;   // 2^15-scaled cumulative thresholds
  uint32_t kind;     // 1 or 2 qubits
  uint32_t bitops;   // accounting weight
  // exact rates (host only): one qubit X = c1, Z = c2 - c1, Y = c3 - c2;
  // two qubits: any non-identity Pauli with total rate c3
  double c1, c2, c3;
};

struct ClassInfo {
  double p;      // event rate of the class
  float c;       // 1 / log2(1 - p)  (<= 0)
};

}  // namespace qcf

struct qcf_program;
namespace qcf {
// Device code of one launch variant: bit 0 = fresh prologue (else in-place
// loads), bit 1 = frame store epilogue; the body starts on a chunk boundary.
std::vector<Instr> variant_code(const qcf_program *P, int variant);
// Marks the gauge draws of the fresh prologue and the body that a fresh,
// record-only run can leave out (gauge_dead_pro / gauge_dead_body).
void mark_dead_gauges(qcf_program *P);
// Appends the planes instruction code[i] reads or writes (a noise group: its
// sites'; a compiler-internal noise site: none -- callers handle those).
void instr_planes(const std::vector<Instr> &code, size_t i, std::vector<uint32_t> *out);
}  // namespace qcf

// Host-side compiled program (opaque to C callers).
struct qcf_program {
  uint32_t nq = 0;      // frame qubits incl. hidden ancilla
  uint32_t nbits = 0;   // record columns
  uint32_t nplanes = 0; // LDS planes (2*nq + classical slots)
  uint32_t ninit_rs = 0;   // replay planes used by the fresh prologue (= nq)
  uint32_t nbody_rs = 0;   // replay planes drawn by the body
  double bitops = 0;
  std::vector<qcf::Instr> prologue_fresh;   // INIT + REC_ZERO
  std::vector<qcf::Instr> prologue_inplace; // LOAD
  std::vector<qcf::Instr> body;
  std::vector<qcf::Instr> epilogue_store;   // STORE
  // per instruction of prologue_fresh / body: 1 = gauge draw that a fresh,
  // record-only run skips (empty when the analysis was not run)
  std::vector<uint8_t> gauge_dead_pro, gauge_dead_body;
  uint32_t ngauge = 0, ngauge_live = 0;  // gauge draws of a fresh run; those kept
  // noiseless programs (no noise site): the record of a fresh record-only run
  // is linear in the kept gauge words -- rec_gauge[b] lists, for record
  // column b, the positions (in draw order) of the kept draws it XORs
  bool linear_ok = false;
  std::vector<std::vector<uint32_t>> rec_gauge;
  std::vector<qcf::Site> sites;
  std::vector<qcf::ClassInfo> classes;
  // some site keeps only part of its class's events (0 < accept < 2^31): the
  // native event stream then draws a thinning uniform per event
  bool thin = false;
  // device copies, created lazily per device and variant
  struct DeviceCopy {
    int device = -1;
    void *prog[4] = {nullptr, nullptr, nullptr, nullptr};  // variant -> Instr*
    uint32_t ninstr[4] = {0, 0, 0, 0};
  };
  std::vector<DeviceCopy> dev;
  // kernels specialised to this program by hiprtc, per device and variant
  struct JitKernel {
    int device = -1;
    int variant = -1;
    void *module = nullptr;   // hipModule_t
    void *function = nullptr; // hipFunction_t
    uint32_t lds_bytes = 0;
    uint32_t blocks_per_cu = 0;
    uint64_t key = 0;  // hash of the kernel source
    std::string knobs;  // generator knob values it was built under
    bool failed = false;
  };
  std::vector<JitKernel> jit;
  ~qcf_program();
};
