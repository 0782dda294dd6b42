// Device program of the Pauli-frame sampler.
//
// A circuit is lowered on the host into a flat stream of 16-byte
// instructions over "planes".  A plane is one 32-bit word per lane of a
// 64-lane wave, i.e. the x bit or the z bit of one qubit for 2048 shots, or a
// classical bit kept on chip.  Gates whose bit map is a permutation of planes
// (H swaps x and z, SWAP exchanges qubits, ...) cost nothing on the device:
// the compiler renames planes instead.  Every other symbolic Clifford becomes
// a few in-place plane XORs.
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/qcf.h"

namespace qcf {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kBlockShots = 2048;  // shots per wave tile (64 lanes x 32)
constexpr uint32_t kMaxClasses = 4;     // noise-rate classes per program
constexpr uint32_t kNoFire = 0xFFu;     // site class of a zero-rate site
constexpr uint32_t kMaxOrdinals = 1u << 20;  // sites per class

enum InstrOp : uint32_t {
  I_END = 0,
  I_XOR,        // P[a] ^= P[b]
  I_XOR2,       // P[a] ^= P[b]; P[c & 0xffff] ^= P[c >> 16]   (independent pairs)
  I_ZERO,       // P[a] = 0
  I_INIT,       // P[a] = 0; P[b] = draw(c)            (PauliFrame + initZ!)
  I_LOAD,       // P[a] = frame[b]
  I_STORE,      // frame[b] = P[a]
  I_MEAS,       // v = P[a]; rec[b] = v; if c != none: P[c] = v
  I_RESETZ,     // P[a] = 0; P[b] ^= draw(c)           (sMRZ reset part)
  I_NOISE1,     // site c on planes x=a, z=b
  I_NOISE2,     // site c on planes x1=a&0xffff z1=a>>16 x2=b&0xffff z2=b>>16
  I_ACC_LDS,    // acc = P[a]
  I_ACCX_LDS,   // acc ^= P[a]
  I_ACC_REC,    // acc = rec[a]
  I_ACCX_REC,   // acc ^= rec[a]
  I_ACC_STORE,  // rec[a] = acc; if b != none: P[b] = acc
  I_REC_ZERO,   // rec[a] = 0
  I_LIN1,       // (P[a],P[b]) := linear map c (bit0 x<-x, bit1 x<-z, bit2 z<-x, bit3 z<-z)
};

struct Instr {
  uint32_t op, a, b, c;
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
  uint64_t t1, t2;   // 2^32-scaled cumulative thresholds
  uint32_t kind;     // 1 or 2 qubits
  uint32_t bitops;   // accounting weight
};

struct ClassInfo {
  double p;      // event rate of the class
  float c;       // 1 / log2(1 - p)  (<= 0)
};

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
  std::vector<qcf::Site> sites;
  std::vector<qcf::ClassInfo> classes;
  // device copies, created lazily per device and variant
  struct DeviceCopy {
    int device = -1;
    void *prog[4] = {nullptr, nullptr, nullptr, nullptr};  // variant -> Instr*
    uint32_t ninstr[4] = {0, 0, 0, 0};
    void *sites = nullptr;
  };
  std::vector<DeviceCopy> dev;
  ~qcf_program();
};
