// Error plumbing, device buffers and small C-ABI utilities.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "host_math.hpp"

namespace cyr {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

cyr_status fail(cyr_status s, const std::string &msg) {
  g_last_error = msg;
  return s;
}

cyr_status hip_fail(hipError_t e, const char *what, const char *file, int line) {
  char buf[512];
  std::snprintf(buf, sizeof(buf), "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
  g_last_error = buf;
  if (e == hipErrorOutOfMemory) return CYR_E_NOMEM;
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return CYR_E_NODEVICE;
  return CYR_E_DEVICE;
}

cyr_status ensure_buffer(DeviceBuffer &b, size_t bytes) {
  if (bytes == 0) bytes = 256;
  if (b.ptr && b.bytes >= bytes) return CYR_OK;
  if (b.ptr) {
    hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
  }
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    char buf[160];
    std::snprintf(buf, sizeof(buf), "hipMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
    return fail(e == hipErrorOutOfMemory ? CYR_E_NOMEM : CYR_E_DEVICE, buf);
  }
  b.ptr = p;
  b.bytes = bytes;
  return CYR_OK;
}

void free_buffer(DeviceBuffer &b) {
  if (b.ptr) hipFree(b.ptr);
  b.ptr = nullptr;
  b.bytes = 0;
}

namespace {
std::atomic<bool> g_kt_on{false};
std::mutex g_kt_mu;
struct KtPending {
  std::string name;
  hipEvent_t a, b;
  double bytes;
};
struct KtStat {
  double ms = 0, bytes = 0;
  uint64_t calls = 0;
};
thread_local std::vector<KtPending> g_kt_pending;
thread_local hipEvent_t g_kt_open = nullptr;
std::vector<hipEvent_t> g_kt_pool;
std::map<std::string, KtStat> g_kt_stats;

hipEvent_t kt_event() {
  {
    std::lock_guard<std::mutex> lk(g_kt_mu);
    if (!g_kt_pool.empty()) {
      hipEvent_t e = g_kt_pool.back();
      g_kt_pool.pop_back();
      return e;
    }
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
}  // namespace

bool kt_enabled() { return g_kt_on.load(std::memory_order_relaxed); }

void kt_begin(hipStream_t st) {
  if (!kt_enabled()) return;
  g_kt_open = kt_event();
  if (g_kt_open) hipEventRecord(g_kt_open, st);
}

void kt_end(const char *name, hipStream_t st, double bytes) {
  if (!kt_enabled() || !g_kt_open) return;
  hipEvent_t e = kt_event();
  if (!e) return;
  hipEventRecord(e, st);
  g_kt_pending.push_back(KtPending{name, g_kt_open, e, bytes});
  g_kt_open = nullptr;
}

void kt_collect() {
  if (g_kt_pending.empty()) return;
  std::lock_guard<std::mutex> lk(g_kt_mu);
  for (auto &p : g_kt_pending) {
    float ms = 0;
    if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      KtStat &s = g_kt_stats[p.name];
      s.ms += ms;
      s.bytes += p.bytes;
      s.calls += 1;
    }
    g_kt_pool.push_back(p.a);
    g_kt_pool.push_back(p.b);
  }
  g_kt_pending.clear();
}

}  // namespace cyr

extern "C" {

cyr_status cyr_kernel_timing(int enable) {
  std::lock_guard<std::mutex> lk(cyr::g_kt_mu);
  if (enable) cyr::g_kt_stats.clear();
  cyr::g_kt_on.store(enable != 0);
  return CYR_OK;
}

int cyr_kernel_timing_report(char *buf, size_t len) {
  std::lock_guard<std::mutex> lk(cyr::g_kt_mu);
  std::string out;
  char line[256];
  for (const auto &kv : cyr::g_kt_stats) {
    const cyr::KtStat &s = kv.second;
    const double n = s.calls ? (double)s.calls : 1.0;
    std::snprintf(line, sizeof(line), "%s %llu %.3f %.3f %.0f\n", kv.first.c_str(), (unsigned long long)s.calls,
                  s.ms * 1000.0, s.ms * 1000.0 / n, s.bytes / n);
    out += line;
  }
  if (buf && len) {
    std::strncpy(buf, out.c_str(), len - 1);
    buf[len - 1] = 0;
  }
  return (int)cyr::g_kt_stats.size();
}

const char *cyr_status_string(cyr_status s) {
  switch (s) {
    case CYR_OK: return "ok";
    case CYR_E_INVALID: return "invalid argument";
    case CYR_E_NOMEM: return "out of memory";
    case CYR_E_DEVICE: return "device error";
    case CYR_E_ADMISSION: return "request exceeds the HBM budget";
    case CYR_E_CANCELLED: return "cancelled";
    case CYR_E_UNSUPPORTED: return "unsupported";
    case CYR_E_NODEVICE: return "no HIP device";
    case CYR_E_TIMEOUT: return "timed out";
    case CYR_E_NOTFOUND: return "unknown request";
  }
  return "unknown status";
}

const char *cyr_last_error(void) { return cyr::g_last_error.c_str(); }

int cyr_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

uint32_t cyr_num_angles(double start, double increment, double end) {
  if (increment == 0) increment = 1.0;
  return cyr::num_angles(start, increment, end);
}

const char *cyr_version(void) { return "cyrsoxs-mi355x 0.1.0 (gfx950)"; }

}  // extern "C"
