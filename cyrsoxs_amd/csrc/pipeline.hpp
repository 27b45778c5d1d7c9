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
  // pinned staging of the per-simulation table upload; stage_done marks when the
  // last upload from it has been consumed
  void *stage = nullptr;
  size_t stage_bytes = 0;
  hipEvent_t stage_done = nullptr;
  bool stage_pending = false;
  // cap on the E-basis energy lanes for work queued through this object (the
  // batch executor runs members concurrently instead)
  uint32_t max_lanes = 2;
  cyr_status ensure_aux();
  // Asynchronous host -> device copy of n bytes through the pinned staging.
  cyr_status upload(void *dst, const void *src, size_t n, hipStream_t st);
  void release();
  size_t bytes() const;
};

// How run_simulation treats the stream and the constant tables.
enum RunMode {
  kRunSync = 0,      // upload tables, queue the work, wait for the stream
  kRunAsync = 1,     // upload tables, queue the work, return (the caller synchronises)
  kRunResident = 2   // tables already resident in wb.tables (graph capture): queue kernels only
};

// (energy, k) units of the E-basis executor in flight (1 or 2, default 1).
uint32_t energy_lanes();
void set_energy_lanes(uint32_t n);

// Frees the per-device streams and work buffers of the batch executor.
void release_batch_buffers();

// Device bytes a simulation needs beyond its morphology and output.
uint64_t work_bytes(const cyr_sim_desc *d, const SimGeometry &g);
// Bytes of the device morphology in its storage format.
uint64_t morphology_bytes(const cyr_sim_desc *d);
uint64_t output_bytes(const cyr_sim_desc *d, const SimGeometry &g);

// Runs one simulation whose morphology (storage format) is resident at d_vox;
// writes [nE][nK][Ny][Nx] of the compute precision to d_out.  Asynchronous on
// `st` except for small table uploads; the caller synchronises.
// kRunResident reads the constant tables at resident_tables (laid out as
// fill_tables writes them); the other modes build and upload them.
cyr_status run_simulation(const cyr_sim_desc *d, const SimGeometry &g, const void *d_vox, const double *mats,
                          void *d_out, WorkBuffers &wb, hipStream_t st, RunMode mode = kRunSync,
                          const void *resident_tables = nullptr);

// The simulation's constant tables (material constants, Hanning, E-basis
// coefficients, rotation inverses): their size, and their contents on the host.
size_t table_bytes(const cyr_sim_desc *d, const SimGeometry &g);
void fill_tables(const cyr_sim_desc *d, const SimGeometry &g, const double *mats, void *host);

// True when the simulation runs the fused E-basis kernels (else the plain executor).
bool ebasis_usable(const cyr_sim_desc *d, const SimGeometry &g);
// The executor a simulation runs: CYR_PIPE_PLAIN, CYR_PIPE_EBASIS or CYR_PIPE_MBASIS
// (CYR_PIPE_AUTO resolved by HBM bytes; fused executors need ebasis_usable).
int resolved_pipeline(const cyr_sim_desc *d, const SimGeometry &g);

// Autotunes the E-basis Y/z launch shape of the simulation's grid (see cyr_autotune).
struct YzTuning;
cyr_status autotune_shape(const cyr_sim_desc *d, const SimGeometry &g, WorkBuffers &wb, hipStream_t st,
                          YzTuning *chosen);

// Converts host morphology of the compute precision into the storage format.
void convert_storage(const cyr_sim_desc *d, const void *host_vox, std::vector<unsigned char> &out);

}  // namespace cyr
