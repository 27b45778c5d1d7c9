// Executors of the scattering pipeline (plain per-angle and E-basis) over
// device-resident inputs.
#pragma once

#include "common.hpp"
#include "host_math.hpp"

namespace cyr {

struct WorkBuffers {
  DeviceBuffer pol, nt, scatter, proj, acc, count, avg, tables, dc, bins, part;
  void release();
  size_t bytes() const;
};

// Device bytes a simulation needs beyond its morphology and output.
uint64_t work_bytes(const cyr_sim_desc *d, const SimGeometry &g);
// Bytes of the device morphology in its storage format.
uint64_t morphology_bytes(const cyr_sim_desc *d);
uint64_t output_bytes(const cyr_sim_desc *d, const SimGeometry &g);

// Runs one simulation whose morphology (storage format) is resident at d_vox;
// writes [nE][nK][Ny][Nx] of the compute precision to d_out.  Asynchronous on
// `st` except for small table uploads; the caller synchronises.
cyr_status run_simulation(const cyr_sim_desc *d, const SimGeometry &g, const void *d_vox, const double *mats,
                          void *d_out, WorkBuffers &wb, hipStream_t st);

// Converts host morphology of the compute precision into the storage format.
void convert_storage(const cyr_sim_desc *d, const void *host_vox, std::vector<unsigned char> &out);

}  // namespace cyr
