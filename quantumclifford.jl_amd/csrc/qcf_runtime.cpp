// Error text, per-device scratch and small C-ABI utilities of libqcf.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "qcf_common.hpp"

namespace qcf {

static thread_local std::string g_last_error;

int set_error(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

void clear_error() { g_last_error.clear(); }

static std::mutex g_dev_mu;
static std::vector<DeviceState *> g_dev;

int device_state(DeviceState **out) {
  int dev = 0;
  QCF_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_dev_mu);
  if ((int)g_dev.size() <= dev) g_dev.resize(dev + 1, nullptr);
  if (!g_dev[dev]) {
    auto *st = new DeviceState();
    QCF_HIP_CHECK(hipMalloc(&st->counters, kCounterSlots * 2 * sizeof(uint32_t)));
    QCF_HIP_CHECK(hipMemset(st->counters, 0, kCounterSlots * 2 * sizeof(uint32_t)));
    hipDeviceProp_t prop;
    QCF_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
    st->cu_count = prop.multiProcessorCount;
    g_dev[dev] = st;
  }
  *out = g_dev[dev];
  return QCF_OK;
}

}  // namespace qcf

extern "C" const char *qcf_last_error(void) {
  return qcf::g_last_error.c_str();
}

extern "C" const char *qcf_version(void) {
  return "qcf 0.1.0 (gfx950)";
}
