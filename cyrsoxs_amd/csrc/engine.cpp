// Batched execution and the serving engine.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "host_math.hpp"
#include "pipeline.hpp"

using namespace cyr;

extern "C" {

cyr_status cyr_simulate_batch(uint32_t count, const cyr_sim_desc *descs, const void *const *d_voxels,
                              const double *const *mats, void *const *d_outs, void *stream) {
  if (count == 0) return CYR_OK;
  if (!descs || !d_voxels || !mats || !d_outs) return fail(CYR_E_INVALID, "null batch arrays");
  std::vector<SimGeometry> geo(count);
  for (uint32_t i = 0; i < count; i++) CYR_TRY(build_geometry(&descs[i], geo[i]));
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  WorkBuffers wb;
  cyr_status s = CYR_OK;
  for (uint32_t i = 0; i < count && s == CYR_OK; i++)
    s = run_simulation(&descs[i], geo[i], d_voxels[i], mats[i], d_outs[i], wb, st);
  if (s == CYR_OK && hipStreamSynchronize(st) != hipSuccess) s = fail(CYR_E_DEVICE, "sync failed");
  wb.release();
  return s;
}

cyr_status cyr_engine_create(int, uint64_t, uint32_t, cyr_engine *) { return fail(CYR_E_UNSUPPORTED, "engine"); }
cyr_status cyr_engine_submit(cyr_engine, const cyr_sim_desc *, const void *, const double *, void *, uint64_t *) {
  return fail(CYR_E_UNSUPPORTED, "engine");
}
cyr_status cyr_engine_wait(cyr_engine, uint64_t, double, cyr_status *) { return fail(CYR_E_UNSUPPORTED, "engine"); }
cyr_status cyr_engine_poll(cyr_engine, uint64_t, int *, cyr_status *) { return fail(CYR_E_UNSUPPORTED, "engine"); }
cyr_status cyr_engine_cancel(cyr_engine, uint64_t) { return fail(CYR_E_UNSUPPORTED, "engine"); }
cyr_status cyr_engine_get_stats(cyr_engine, cyr_engine_stats *) { return fail(CYR_E_UNSUPPORTED, "engine"); }
cyr_status cyr_engine_destroy(cyr_engine) { return fail(CYR_E_UNSUPPORTED, "engine"); }

}  // extern "C"
