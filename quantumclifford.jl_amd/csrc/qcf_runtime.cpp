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

namespace {
std::mutex g_bounds_mu;
std::vector<std::pair<const void *, int>> g_bounds;  // kernel -> maxThreadsPerBlock

int bounds_check(const void *key, int maxt, unsigned threads, const char *name) {
  if (maxt > 0 && threads > (unsigned)maxt)
    return set_error(QCF_ERR_ARGUMENT, "%s: %u threads per block exceed the kernel's launch bound of %d",
                     name, threads, maxt);
  (void)key;
  return QCF_OK;
}

bool cached_bound(const void *key, int *maxt) {
  std::lock_guard<std::mutex> lk(g_bounds_mu);
  for (const auto &e : g_bounds)
    if (e.first == key) {
      *maxt = e.second;
      return true;
    }
  return false;
}

void cache_bound(const void *key, int maxt) {
  std::lock_guard<std::mutex> lk(g_bounds_mu);
  g_bounds.emplace_back(key, maxt);
}
}  // namespace

int check_block(const void *kernel, unsigned threads, const char *name) {
  int maxt = 0;
  if (!cached_bound(kernel, &maxt)) {
    hipFuncAttributes a{};
    QCF_HIP_CHECK(hipFuncGetAttributes(&a, kernel));
    maxt = a.maxThreadsPerBlock;
    cache_bound(kernel, maxt);
  }
  return bounds_check(kernel, maxt, threads, name);
}

int check_block_module(hipFunction_t kernel, unsigned threads, const char *name) {
  int maxt = 0;
  if (!cached_bound((const void *)kernel, &maxt)) {
    QCF_HIP_CHECK(hipFuncGetAttribute(&maxt, HIP_FUNC_ATTRIBUTE_MAX_THREADS_PER_BLOCK, kernel));
    cache_bound((const void *)kernel, maxt);
  }
  return bounds_check((const void *)kernel, maxt, threads, name);
}

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

struct ScratchKey {
  int dev, kind;
  hipStream_t stream;
  bool operator==(const ScratchKey &o) const {
    return dev == o.dev && kind == o.kind && stream == o.stream;
  }
};
struct ScratchBuf {
  ScratchKey key;
  void *ptr;
  size_t bytes;
};
static std::mutex g_scratch_mu;
static std::vector<ScratchBuf> g_scratch;

int stream_scratch(int kind, hipStream_t stream, size_t bytes, void **out) {
  int dev = 0;
  QCF_HIP_CHECK(hipGetDevice(&dev));
  const ScratchKey key{dev, kind, stream};
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  ScratchBuf *b = nullptr;
  for (auto &s : g_scratch)
    if (s.key == key) b = &s;
  if (!b) {
    g_scratch.push_back(ScratchBuf{key, nullptr, 0});
    b = &g_scratch.back();
  }
  if (b->bytes < bytes) {
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    QCF_HIP_CHECK(hipStreamIsCapturing(stream, &cap));
    if (cap != hipStreamCaptureStatusNone)
      return set_error(QCF_ERR_ARGUMENT,
                       "workspace of %zu bytes must grow to %zu while the stream is being "
                       "captured; make one uncaptured call of this size first",
                       b->bytes, bytes);
    if (b->ptr) {
      QCF_HIP_CHECK(hipStreamSynchronize(stream));  // the last user of it
      QCF_HIP_CHECK(hipFree(b->ptr));
      b->ptr = nullptr;
      b->bytes = 0;
    }
    QCF_HIP_CHECK(hipMalloc(&b->ptr, bytes));
    b->bytes = bytes;
  }
  *out = b->ptr;
  return QCF_OK;
}

}  // namespace qcf

extern "C" const char *qcf_last_error(void) {
  return qcf::g_last_error.c_str();
}

extern "C" const char *qcf_version(void) {
  return "qcf 0.1.0 (gfx950)";
}
