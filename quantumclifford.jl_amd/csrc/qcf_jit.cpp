// Per-circuit frame-sampling kernels compiled at run time with hiprtc.
//
// The interpreter in qcf_frames.hip decodes every device instruction on the
// wave (v_readlane, branches, address arithmetic), which the instruction
// counters show is where its time goes.  For the native-RNG launches this
// file instead writes the program out as straight-line HIP source -- plane
// XORs become ds_read/ds_xor with constant LDS offsets, noise groups call the
// shared event pass with their site tables in constant memory -- compiles it
// once per program and variant, and launches that kernel.  The RNG, event
// pass and gauge streams come from qcf_device.h, the same code the
// interpreter runs, so both produce the same bits for the same seed.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "qcf_common.hpp"
#include "qcf_jit.hpp"
#include "qcf_program.hpp"

namespace qcf {
namespace {
const char kDeviceSrc[] =
#include "qcf_device_src.inc"
    ;

// Kernel parameters; the generated source declares the same struct.
const char kKernelPrelude[] = R"QCFJIT(
struct KJ {
  uint32_t *rec;
  uint32_t *frame;
  uint64_t seed;
  uint64_t ld32;
  uint32_t nblocks, nw32, nwbuf, lastmask;
  float c;
  uint32_t has_class, pad0, pad1;
};
struct FlipLDS {
  uint32_t *L;
  __device__ __forceinline__ void operator()(uint32_t p, uint32_t tl, uint32_t bit, uint32_t,
                                             uint32_t) const {
    atomicXor(&L[p * 64u + tl], bit);
  }
};
)QCFJIT";

namespace {
std::string plane(uint32_t p) { return "L[" + std::to_string(p * 64u) + "u + l]"; }
std::string rec(uint32_t col) { return "P.rec[(uint64_t)" + std::to_string(col) + "u * P.ld32 + gw]"; }
}  // namespace

}  // namespace

// Write the kernel source of one variant; returns false for programs the
// generator does not handle (the interpreter runs those).
bool generate(const qcf_program *P, int variant, std::string *out) {
  const std::vector<Instr> code = variant_code(P, variant);
  std::ostringstream tab, body;
  std::vector<std::pair<uint32_t, uint32_t>> groups;  // (first site, count)
  uint32_t nsite = 0;
  for (size_t i = 0; i < code.size(); ++i) {
    const Instr &in = code[i];
    if ((in.op & 0xFFu) != I_NOISEG) continue;
    const uint32_t G = in.op >> 8;
    groups.push_back({nsite, G});
    for (uint32_t k = 1; k <= G; ++k) {
      const Instr &s = code[i + k];
      tab << s.a << "u," << s.b << "u," << s.d << "u," << s.e << "u,\n";
    }
    nsite += G;
    i += G;
  }
  auto load_group = [&](size_t g, const char *pfx) {
    std::ostringstream o;
    const uint32_t off = groups[g].first, G = groups[g].second;
    o << "    { const uint32_t *q = QS + 4u * (" << off << "u + (l < " << G << "u ? l : " << G - 1
      << "u));\n      " << pfx << "0 = q[0]; " << pfx << "1 = q[1]; " << pfx << "2 = q[2]; "
      << pfx << "3 = q[3]; }\n";
    return o.str();
  };
  size_t gi = 0;
  if (!groups.empty()) body << "    uint32_t t0, t1, t2, t3;\n" << load_group(0, "t");
  for (size_t i = 0; i < code.size(); ++i) {
    const Instr &in = code[i];
    const uint32_t op = in.op & 0xFFu;
    const uint32_t a = in.a, b = in.b, c = in.c;
    switch (op) {
      case I_END:
        break;
      case I_XORN: {
        const uint32_t n = in.op >> 8;
        const uint32_t *w = &in.a;
        body << "    {";
        for (uint32_t k = 0; k < n; ++k)
          body << " const uint32_t v" << k << " = " << plane(w[k] >> 16) << ";";
        for (uint32_t k = 0; k < n; ++k)
          body << " atomicXor(&" << plane(w[k] & 0xFFFFu) << ", v" << k << ");";
        body << " }\n";
        break;
      }
      case I_NOISEG: {
        const uint32_t G = in.op >> 8;
        body << "    {\n";
        if (gi + 1 < groups.size()) body << "    uint32_t n0, n1, n2, n3;\n" << load_group(gi + 1, "n");
        body << "    if (P.has_class) qcfd::group_events(ev, " << a << "u, " << G
             << "u, 0u, t0, t1, t2, t3, 0u, P.seed, P.c, blk, l, flip);\n";
        if (gi + 1 < groups.size()) body << "    t0 = n0; t1 = n1; t2 = n2; t3 = n3;\n";
        body << "    }\n";
        ++gi;
        i += G;
        break;
      }
      case I_MR:
      case I_MRX: {
        const uint32_t col = in.d, keep = in.e;
        body << "    { const uint32_t v = " << plane(a) << ";";
        if (col != kNone) body << " " << rec(col) << " = v;";
        if (op == I_MRX)
          body << " " << rec(in.g) << " = v ^ " << plane(in.f) << ";";
        if (keep != kNone) body << " " << plane(keep) << " = v;";
        body << " " << plane(a) << " = 0u; atomicXor(&" << plane(b) << ", g.next()); }\n";
        break;
      }
      case I_XSTORE2:
        body << "    { const uint32_t v = " << plane(a) << " ^ " << plane(b) << "; " << rec(c)
             << " = v;";
        if (in.d != kNone) body << " " << plane(in.d) << " = v;";
        body << " }\n";
        break;
      case I_MEAS:
        body << "    { const uint32_t v = " << plane(a) << "; " << rec(b) << " = v;";
        if (c != kNone) body << " " << plane(c) << " = v;";
        body << " }\n";
        break;
      case I_RESETZ:
        body << "    " << plane(a) << " = 0u; atomicXor(&" << plane(b) << ", g.next());\n";
        break;
      case I_ACC_LDS:
        body << "    acc = " << plane(a) << ";\n";
        break;
      case I_ACCX_LDS:
        body << "    acc ^= " << plane(a) << ";\n";
        break;
      case I_ACC_REC:
        body << "    acc = " << rec(a) << ";\n";
        break;
      case I_ACCX_REC:
        body << "    acc ^= " << rec(a) << ";\n";
        break;
      case I_ACC_STORE:
        body << "    " << rec(a) << " = acc;";
        if (b != kNone) body << " " << plane(b) << " = acc;";
        body << "\n";
        break;
      case I_REC_ZERO:
        body << "    " << rec(a) << " = 0u;\n";
        break;
      case I_INIT:
        body << "    " << plane(a) << " = 0u; " << plane(b) << " = g.next();\n";
        break;
      case I_LOAD:
        body << "    " << plane(a) << " = P.frame[(uint64_t)" << b << "u * P.ld32 + gw];\n";
        break;
      case I_STORE:
        body << "    P.frame[(uint64_t)" << b << "u * P.ld32 + gw] = " << plane(a) << ";\n";
        break;
      case I_ZERO:
        body << "    " << plane(a) << " = 0u;\n";
        break;
      case I_LIN1:
        body << "    { const uint32_t x = " << plane(a) << ", z = " << plane(b) << "; " << plane(a)
             << " = " << ((c & 1u) ? "x" : "0u") << " ^ " << ((c & 2u) ? "z" : "0u") << "; "
             << plane(b) << " = " << ((c & 4u) ? "x" : "0u") << " ^ " << ((c & 8u) ? "z" : "0u")
             << "; }\n";
        break;
      default:
        return false;
    }
  }
  std::ostringstream src;
  src << "#define QCF_RTC 1\n" << kDeviceSrc << kKernelPrelude;
  src << "__constant__ uint32_t QS[" << std::max<uint32_t>(4, 4 * nsite) << "] = {\n"
      << (nsite ? tab.str() : std::string("0u,0u,0u,0u")) << "};\n";
  src << R"QCFJIT(
extern "C" __global__ void __launch_bounds__(64) qcf_jit_kernel(KJ P) {
  extern __shared__ uint32_t L[];
  const uint32_t l = threadIdx.x;
  const FlipLDS flip{L};
  // complete tiles only: every lane's 32 shots exist
  for (uint32_t blk = blockIdx.x; blk < P.nblocks; blk += gridDim.x) {
    const uint32_t gw = blk * 64u + l;
    qcfd::Xoshiro g = qcfd::gauge_stream(P.seed, gw);
    qcfd::EvState ev = qcfd::ev_init();
    if (P.has_class) ev = qcfd::gen_batch(ev, P.seed, P.c, blk, l);
    uint32_t acc = 0u;
    (void)acc;
)QCFJIT";
  src << body.str();
  src << "  }\n}\n";
  *out = src.str();
  return true;
}

int compile(const std::string &src, std::vector<char> *code, std::string *log) {
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "qcf_jit.hip", 0, nullptr, nullptr) !=
      HIPRTC_SUCCESS)
    return set_error(QCF_ERR_HIP, "hiprtcCreateProgram failed");
  const char *opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang",
                        "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof opts / sizeof *opts), opts);
  size_t n = 0;
  hiprtcGetProgramLogSize(prog, &n);
  if (n > 1) {
    log->resize(n);
    hiprtcGetProgramLog(prog, &(*log)[0]);
  }
  if (r != HIPRTC_SUCCESS) {
    hiprtcDestroyProgram(&prog);
    return set_error(QCF_ERR_HIP, "hiprtc compile failed: %.900s", log->c_str());
  }
  hiprtcGetCodeSize(prog, &n);
  code->resize(n);
  hiprtcGetCode(prog, code->data());
  hiprtcDestroyProgram(&prog);
  return QCF_OK;
}

namespace {
std::mutex g_jit_mu;
constexpr size_t kCodeCacheMax = 32;
std::unordered_map<std::string, std::vector<char>> g_code_cache;  // source -> code object
}  // namespace

// Programs longer than this many device instructions stay on the interpreter:
// compile time grows faster than linearly with the straight-line code.
size_t jit_max_instr() {
  const char *e = getenv("QCF_JIT_MAX_INSTR");
  return e ? (size_t)strtoull(e, nullptr, 10) : (size_t)12000;
}

bool jit_disabled() {
  const char *e = getenv("QCF_NO_JIT");
  return e && e[0] == '1';
}

int jit_get(qcf_program *P, int variant, void **function) {
  *function = nullptr;
  int dev = 0;
  QCF_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_jit_mu);
  for (auto &j : P->jit)
    if (j.device == dev && j.variant == variant) {
      if (j.failed) return set_error(QCF_ERR_UNSUPPORTED, "specialised kernel unavailable");
      *function = j.function;
      return QCF_OK;
    }
  qcf_program::JitKernel jk;
  jk.device = dev;
  jk.variant = variant;
  std::string src, log;
  std::vector<char> bin;
  int st = QCF_OK;
  if (variant_code(P, variant).size() > jit_max_instr())
    st = set_error(QCF_ERR_UNSUPPORTED, "program longer than QCF_JIT_MAX_INSTR instructions");
  else if (!generate(P, variant, &src))
    st = set_error(QCF_ERR_UNSUPPORTED, "op not specialisable");
  if (st == QCF_OK) {
    // programs with the same source (the same circuit built twice) share one
    // compilation
    auto it = g_code_cache.find(src);
    if (it != g_code_cache.end()) {
      bin = it->second;
    } else {
      st = compile(src, &bin, &log);
      if (st == QCF_OK) {
        if (g_code_cache.size() >= kCodeCacheMax) g_code_cache.clear();
        g_code_cache.emplace(src, bin);
      }
    }
  }
  if (st == QCF_OK) {
    hipModule_t m;
    hipFunction_t f;
    if (hipModuleLoadData(&m, bin.data()) != hipSuccess) {
      st = set_error(QCF_ERR_HIP, "hipModuleLoadData failed");
    } else if (hipModuleGetFunction(&f, m, "qcf_jit_kernel") != hipSuccess) {
      (void)hipModuleUnload(m);
      st = set_error(QCF_ERR_HIP, "hipModuleGetFunction failed");
    } else {
      jk.module = m;
      jk.function = f;
    }
  }
  jk.failed = st != QCF_OK;
  P->jit.push_back(jk);
  *function = jk.function;
  return st;
}

int jit_launch(void *function, const JitParams &kp, uint32_t grid, uint32_t lds_bytes,
               hipStream_t stream) {
  JitParams p = kp;
  void *args[] = {&p};
  const hipError_t e = hipModuleLaunchKernel((hipFunction_t)function, grid, 1, 1, 64, 1, 1,
                                             lds_bytes, stream, args, nullptr);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return set_error(QCF_ERR_HIP, "specialised kernel launch: %s", hipGetErrorString(e));
  }
  return QCF_OK;
}

}  // namespace qcf

extern "C" int qcf_debug_jit_source(const qcf_program *prog, int variant, char *buf, size_t cap,
                                    size_t *needed) {
  qcf::clear_error();
  if (!prog) return qcf::set_error(QCF_ERR_ARGUMENT, "null program");
  std::string src;
  if (!qcf::generate(prog, variant & 3, &src))
    return qcf::set_error(QCF_ERR_UNSUPPORTED, "op not specialisable");
  if (needed) *needed = src.size() + 1;
  if (buf && cap) {
    const size_t n = std::min(cap - 1, src.size());
    memcpy(buf, src.data(), n);
    buf[n] = 0;
  }
  return QCF_OK;
}

extern "C" int qcf_debug_jit_compile(const qcf_program *prog, int variant, size_t *code_bytes) {
  qcf::clear_error();
  if (!prog) return qcf::set_error(QCF_ERR_ARGUMENT, "null program");
  std::string src, log;
  std::vector<char> bin;
  if (!qcf::generate(prog, variant & 3, &src))
    return qcf::set_error(QCF_ERR_UNSUPPORTED, "op not specialisable");
  QCF_TRY(qcf::compile(src, &bin, &log));
  if (code_bytes) *code_bytes = bin.size();
  return QCF_OK;
}
