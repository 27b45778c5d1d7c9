// Error plumbing, device buffers and small C-ABI utilities.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

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

}  // namespace cyr

extern "C" {

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
