// Executors of the scattering pipeline (plain per-angle and E-basis) over
// device-resident inputs.
#pragma once

#include "common.hpp"
#include "host_math.hpp"

namespace cyr {

// Work buffers of one in-flight (energy, k) unit: polarization volumes (plain)
// or tensor fields (E-basis), projections, average, and the E-basis bin table,
// partial sums and DC sums.
struct LaneBuffers {
  DeviceBuffer pol, proj, avg, bins, part, dc;
  void release();
  size_t bytes() const;
};

struct WorkBuffers {
  LaneBuffers lane[2];
  DeviceBuffer nt, scatter, acc, count, tables;
  // second stream (default priority) and events of the two-lane E-basis executor
  hipStream_t aux = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_polx[2] = {nullptr, nullptr};
  cyr_status ensure_aux();
  void release();
  size_t bytes() const;
};

// (energy, k) units of the E-basis executor in flight (1 or 2, default 1).
uint32_t energy_lanes();
void set_energy_lanes(uint32_t n);

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
