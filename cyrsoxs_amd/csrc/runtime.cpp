// Error plumbing, device buffers and small C-ABI utilities.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <thread>
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

static std::shared_mutex g_workspace_mu;
static std::atomic<int> g_live_handles{0};

std::shared_lock<std::shared_mutex> workspace_use() { return std::shared_lock<std::shared_mutex>(g_workspace_mu); }
std::unique_lock<std::shared_mutex> workspace_exclusive() { return std::unique_lock<std::shared_mutex>(g_workspace_mu); }
void live_handles_add(int delta) { g_live_handles.fetch_add(delta); }
int live_handles() { return g_live_handles.load(); }

static thread_local ScratchArena *t_arena = nullptr;

void set_scratch_arena(ScratchArena *a) { t_arena = a; }
ScratchArena *scratch_arena() { return t_arena; }

void *ScratchArena::carve(size_t bytes) {
  const size_t off = (top + 255) & ~(size_t)255;
  if (off > cap || bytes > cap - off) return nullptr;
  top = off + bytes;
  last_off = off;
  last_ptr = base + off;
  return last_ptr;
}

// ensure_buffer under an installed arena: slices only, never an allocation.
static cyr_status ensure_in_arena(ScratchArena &a, DeviceBuffer &b, size_t bytes) {
  if (b.ptr && b.bytes >= bytes) return CYR_OK;
  if (b.ptr && b.ptr == a.last_ptr && a.last_off + bytes <= a.cap) {  // topmost slice: grow in place
    a.top = a.last_off + bytes;
    b.bytes = bytes;
    return CYR_OK;
  }
  void *p = a.carve(bytes);
  if (!p) {
    char buf[200];
    std::snprintf(buf, sizeof(buf),
                  "allocation-free call: its work buffers need more than the %zu-byte scratch array "
                  "(%zu bytes in use, %zu more requested)",
                  a.cap, a.top, bytes);
    return fail(CYR_E_NOMEM, buf);
  }
  b.ptr = p;
  b.bytes = bytes;
  return CYR_OK;
}

// Test hook for the error path of a failed device allocation: with
// CYR_TEST_FAIL_ALLOC=k in the environment, the device allocation after the next
// k ones fails with CYR_E_NOMEM (and every later one, until the value changes)
// without calling hipMalloc.
static bool injected_alloc_failure() {
  static std::mutex mu;
  static std::string seen;
  static long left = -1;
  const char *v = std::getenv("CYR_TEST_FAIL_ALLOC");
  std::lock_guard<std::mutex> lk(mu);
  if (!v || !*v) {
    seen.clear();
    left = -1;
    return false;
  }
  if (seen != v) {
    seen = v;
    left = std::atol(v);
  }
  if (left < 0) return false;
  if (left == 0) return true;
  left--;
  return false;
}

cyr_status ensure_buffer(DeviceBuffer &b, size_t bytes) {
  if (bytes == 0) bytes = 256;
  if (t_arena) return ensure_in_arena(*t_arena, b, bytes);
  if (b.ptr && b.bytes >= bytes) return CYR_OK;
  if (b.ptr) {
    hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
  }
  if (injected_alloc_failure()) return fail(CYR_E_NOMEM, "injected device allocation failure (CYR_TEST_FAIL_ALLOC)");
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
  if (b.ptr && !(t_arena && t_arena->owns(b.ptr))) hipFree(b.ptr);
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

// ---- host -> device upload of large pageable buffers ------------------------
// kUpThreads host threads each take every kUpThreads-th kUpChunk piece: copy it
// into one of their two pinned buffers, queue its DMA on their own stream, and
// copy the next piece while that DMA runs.  Staging (pinned buffers, streams,
// events) is created once per device and kept until cyr_release_workspace.
namespace {
constexpr int kUpThreads = 4;
constexpr size_t kUpChunk = 32ull << 20;

struct UploadStage {
  void *pin[kUpThreads][2] = {};
  hipStream_t st[kUpThreads] = {};
  hipEvent_t ev[kUpThreads][2] = {};
};
std::mutex g_up_mu;
std::map<int, UploadStage *> g_up;

hipError_t make_stage(UploadStage &u) {
  for (int t = 0; t < kUpThreads; t++) {
    hipError_t e = hipStreamCreateWithFlags(&u.st[t], hipStreamNonBlocking);  // default priority
    for (int b = 0; b < 2 && e == hipSuccess; b++) {
      e = hipHostMalloc(&u.pin[t][b], kUpChunk, hipHostMallocDefault);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&u.ev[t][b], hipEventDisableTiming);
    }
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

void free_stage(UploadStage &u) {
  for (int t = 0; t < kUpThreads; t++) {
    if (u.st[t]) (void)hipStreamSynchronize(u.st[t]);
    for (int b = 0; b < 2; b++) {
      if (u.pin[t][b]) (void)hipHostFree(u.pin[t][b]);
      if (u.ev[t][b]) (void)hipEventDestroy(u.ev[t][b]);
      u.pin[t][b] = nullptr;
      u.ev[t][b] = nullptr;
    }
    if (u.st[t]) (void)hipStreamDestroy(u.st[t]);
    u.st[t] = nullptr;
  }
}
}  // namespace

cyr_status upload_host(void *dst, const void *src, size_t n, hipStream_t st) {
  if (n < 2 * kUpChunk) {  // small: one pageable copy, ordered on st
    CYR_HIP(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, st));
    return CYR_OK;
  }
  int dev = 0;
  CYR_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_up_mu);  // one staged upload per process at a time
  UploadStage *&u = g_up[dev];
  if (!u) {
    u = new UploadStage();
    const hipError_t e = make_stage(*u);
    if (e != hipSuccess) {
      free_stage(*u);
      delete u;
      u = nullptr;
      return hip_fail(e, "pinned upload staging", __FILE__, __LINE__);
    }
  }
  // work queued earlier on st (e.g. a kernel still reading dst) finishes first
  CYR_HIP(hipStreamSynchronize(st));
  const size_t pieces = (n + kUpChunk - 1) / kUpChunk;
  hipError_t err[kUpThreads];
  auto work = [&](int t) {
    hipError_t e = hipSetDevice(dev);
    size_t k = 0;
    for (size_t c = (size_t)t; c < pieces && e == hipSuccess; c += kUpThreads, k++) {
      const int b = (int)(k & 1);
      if (k >= 2) e = hipEventSynchronize(u->ev[t][b]);  // that buffer's previous DMA is done
      if (e != hipSuccess) break;
      const size_t off = c * kUpChunk, len = std::min(kUpChunk, n - off);
      std::memcpy(u->pin[t][b], static_cast<const unsigned char *>(src) + off, len);
      e = hipMemcpyAsync(static_cast<unsigned char *>(dst) + off, u->pin[t][b], len, hipMemcpyHostToDevice, u->st[t]);
      if (e == hipSuccess) e = hipEventRecord(u->ev[t][b], u->st[t]);
    }
    const hipError_t e2 = hipStreamSynchronize(u->st[t]);
    err[t] = e != hipSuccess ? e : e2;
  };
  std::vector<std::thread> th;
  for (int t = 1; t < kUpThreads; t++) th.emplace_back(work, t);
  work(0);
  for (auto &x : th) x.join();
  for (int t = 0; t < kUpThreads; t++)
    if (err[t] != hipSuccess) return hip_fail(err[t], "staged upload", __FILE__, __LINE__);
  return CYR_OK;
}

void release_upload_staging() {
  std::lock_guard<std::mutex> lk(g_up_mu);
  int cur = 0;
  (void)hipGetDevice(&cur);
  for (auto &kv : g_up) {
    if (!kv.second) continue;
    (void)hipSetDevice(kv.first);
    free_stage(*kv.second);
    delete kv.second;
  }
  g_up.clear();
  (void)hipSetDevice(cur);
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
