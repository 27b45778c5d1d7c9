// Allocation-free whole simulation: cyr_simulate_device_noalloc.
//
// The same executors as cyr_simulate_device, with every device buffer the call
// needs beyond the caller's inputs and outputs (constant tables, twiddle tables,
// work volumes, band values, partial sums, averages) carved from one file-scope
// device array instead of allocated.  Its device address is looked up once per
// device with hipGetSymbolAddress; host code uses that address only.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <map>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "fft.hpp"
#include "host_math.hpp"
#include "pipeline.hpp"

namespace cyr {
namespace {

constexpr size_t kScratchBytes = CYR_NOALLOC_SCRATCH_BYTES;
static_assert(kScratchBytes % 256 == 0, "scratch size must keep 256-byte alignment");

// One copy per device, held from module load until the process ends.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct ScratchSlot {
  std::mutex mu;  // serialises the calls on one device (they share the array)
  unsigned char *ptr = nullptr;
};
std::mutex g_slots_mu;
std::map<int, ScratchSlot *> g_slots;

// The current device's scratch address (looked up on its first call).
cyr_status scratch_slot(ScratchSlot **out) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
    (void)hipGetLastError();
    return fail(CYR_E_NODEVICE, "no HIP device found");
  }
  int dev = 0;
  CYR_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_slots_mu);
  auto it = g_slots.find(dev);
  if (it == g_slots.end()) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
    if (!p) return fail(CYR_E_DEVICE, "hipGetSymbolAddress(scratch) returned no address");
    ScratchSlot *s = new ScratchSlot();
    s->ptr = reinterpret_cast<unsigned char *>(p);
    it = g_slots.emplace(dev, s).first;
  }
  *out = it->second;
  return CYR_OK;
}

size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

// Scratch bytes a simulation needs: tables, the twiddle tables of its axes and
// the executor's work buffers (work_bytes bounds them; a material-basis energy
// group whose Ewald band outgrows the model fails later with CYR_E_NOMEM).
uint64_t scratch_need(const cyr_sim_desc *d, const SimGeometry &g) {
  const size_t cs = d->precision == CYR_FP64 ? 16 : 8, blocks = d->precision == CYR_FP64 ? 1 : 2;
  uint64_t b = al256(table_bytes(d, g));
  const uint32_t ax[3] = {g.nx, g.ny, g.nz};
  for (int i = 0; i < 3; i++) {
    bool seen = ax[i] <= 1;
    for (int j = 0; j < i; j++) seen = seen || ax[j] == ax[i];
    if (!seen) b += al256(cs * blocks * ax[i]);
  }
  return b + work_bytes(d, g);
}

}  // namespace
}  // namespace cyr

using namespace cyr;

extern "C" cyr_status cyr_simulate_device_noalloc(const cyr_sim_desc *desc, const void *d_voxels, const double *mats,
                                                  void *d_out, void *stream) {
  SimGeometry g;
  CYR_TRY(build_geometry(desc, g));
  if (!d_voxels || !mats || !d_out) return fail(CYR_E_INVALID, "null buffer");
  const uint64_t need = scratch_need(desc, g);
  if (need > kScratchBytes) {
    char buf[200];
    std::snprintf(buf, sizeof(buf),
                  "allocation-free call: this simulation needs %llu bytes of device scratch; the limit is %zu "
                  "(use cyr_simulate_device)",
                  (unsigned long long)need, kScratchBytes);
    return fail(CYR_E_INVALID, buf);
  }
  ScratchSlot *slot;
  CYR_TRY(scratch_slot(&slot));
  std::lock_guard<std::mutex> lk(slot->mu);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  // host images of the tables and twiddles live until the stream has drained
  std::vector<unsigned char> host_tables(table_bytes(desc, g));
  ScratchArena arena;
  arena.base = slot->ptr;
  arena.cap = kScratchBytes;
  arena.stream = st;
  // constant tables first (the executors read them as resident tables)
  fill_tables(desc, g, mats, host_tables.data());
  void *tables = arena.carve(host_tables.size());
  if (!tables) return fail(CYR_E_NOMEM, "allocation-free call: tables exceed the scratch array");
  CYR_HIP(hipMemcpyAsync(tables, host_tables.data(), host_tables.size(), hipMemcpyHostToDevice, st));
  WorkBuffers wb;      // every buffer a slice of the arena
  wb.max_lanes = 1;    // no second stream or events
  set_scratch_arena(&arena);
  const cyr_status s = run_simulation(desc, g, d_voxels, mats, d_out, wb, st, kRunResident, tables);
  set_scratch_arena(nullptr);
  // drain the stream on errors too: no queued launch may outlive the host images
  const hipError_t se = hipStreamSynchronize(st);
  kt_collect();
  if (s != CYR_OK) return s;
  if (se != hipSuccess) return hip_fail(se, "hipStreamSynchronize", __FILE__, __LINE__);
  return CYR_OK;
}
