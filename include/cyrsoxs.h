/*
 * cyrsoxs.h — C ABI of the MI355X-native CyRSoXS Born-approximation pipeline.
 *
 * Plain pointers, sizes and POD descriptors only; every call returns a
 * cyr_status and never exits the process (the reference calls exit() on every
 * device error: include/cudaHeaders.h:43-63).  Citations name the reference
 * interface each entry point replaces (paths relative to the usnistgov/cyrsoxs
 * tree).
 *
 * Layouts (shared with the reference):
 *   morphology   [M][Nz][Ny][Nx] x 4 reals per voxel: Euler (S, theta, psi, Vfrac) or
 *                vector (sx, sy, sz, phi_unaligned)             (include/Input/Input.h:59-94)
 *   constants    double[nE][M][4] = (npara.re, npara.im, nperp.re, nperp.im),
 *                n = (1 - delta) + i beta                       (include/Input/InputData.h:371-375)
 *   patterns     [nE][nK][Ny][Nx] of the compute precision      (src/cudaMain.cu:968-971)
 *   polarization [Nz][Ny][Nx] complex (re, im) of the compute precision
 */
#ifndef CYRSOXS_H
#define CYRSOXS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  CYR_OK = 0,
  CYR_E_INVALID = 1,     /* bad descriptor or argument */
  CYR_E_NOMEM = 2,       /* device or host allocation failed */
  CYR_E_DEVICE = 3,      /* HIP runtime / kernel launch error */
  CYR_E_ADMISSION = 4,   /* request larger than the engine's HBM budget */
  CYR_E_CANCELLED = 5,   /* request cancelled before it ran */
  CYR_E_UNSUPPORTED = 6, /* option combination the reference does not support either */
  CYR_E_NODEVICE = 7,    /* no HIP device (reference: src/cudaMain.cu:402-405) */
  CYR_E_TIMEOUT = 8,     /* wait timed out; the request is still in flight */
  CYR_E_NOTFOUND = 9     /* unknown request id */
} cyr_status;

/* Compute precision (the reference's compile-time DOUBLE_PRECISION switch,
 * include/Datatypes.h:33-46, made a runtime parameter). */
enum { CYR_FP32 = 0, CYR_FP64 = 1 };
/* Device storage of the morphology: native = the compute precision. */
enum { CYR_STORE_NATIVE = 0, CYR_STORE_FP16 = 1, CYR_STORE_BF16 = 2 };
/* Execution strategy.  PLAIN keeps the reference's per-angle structure
 * (polarization -> 3 full 3D FFTs -> DC/shift -> Ewald -> rotate/accumulate for
 * every E angle).  EBASIS uses the linearity of the pipeline in E: two
 * polarizations per energy, an XY FFT, and only the Ewald-sphere z bins of the
 * z transform; every angle is then a 2D operation.  MBASIS also uses the
 * linearity in the optical constants: per k vector it transforms seven
 * energy-independent real fields per material once (packed two per complex
 * volume), keeps their values in the band of Ewald bins all energies touch, and
 * forms each energy's fields there as combinations with that energy's
 * constants; the per-energy work is then 2D.  All three produce the same
 * patterns within floating-point rounding. */
enum { CYR_PIPE_PLAIN = 0, CYR_PIPE_EBASIS = 1, CYR_PIPE_MBASIS = 2, CYR_PIPE_AUTO = 3 };
/* CYR_PIPE_AUTO picks EBASIS or MBASIS per simulation from the HBM bytes each would
 * move (MBASIS once a simulation has more than a few energies per k vector). */

/* Mirrors InputData (include/Input/InputData.h:186-232) plus the compile-time
 * switches of the reference build.  Enumerations use the reference's values
 * (include/Datatypes.h:56-182). */
typedef struct {
  uint32_t dims[3];          /* Nx, Ny, Nz (X fastest); Nz == 1 selects the 2D path */
  uint32_t num_materials;
  double phys_size;          /* nm */
  uint32_t num_energies;
  const double *energies;    /* eV, length num_energies */
  double angle_start;        /* E-angle rotation, degrees */
  double angle_increment;
  double angle_end;
  uint32_t num_kvectors;
  const double *kvectors;    /* [nK][3], normalised; (0,0,1) for case type 0 */
  double detector[3];        /* normalised detector direction (case type 2) */
  int32_t case_type;         /* 0 default, 1 beam divergence, 2 grazing incidence */
  int32_t morphology_type;   /* 0 Euler angles, 1 vector morphology */
  int32_t reference_frame;   /* 0 MATERIAL, 1 LAB */
  int32_t interpolation;     /* 0 nearest neighbour, 1 linear */
  int32_t windowing;         /* 0 none, 1 Hanning */
  int32_t rot_mask;          /* 0/1 */
  int32_t scatter_approach;  /* 0 partial, 1 full */
  int32_t algorithm;         /* 0 communication minimizing, 1 memory minimizing (Nt) */
  int32_t precision;         /* CYR_FP32 / CYR_FP64 */
  int32_t storage;           /* CYR_STORE_* (device morphology element type) */
  int32_t pipeline;          /* CYR_PIPE_* */
  int32_t reserved;
} cyr_sim_desc;

/* ---- whole simulation ---------------------------------------------------- */

/* Replaces int cudaMain(...) / int cudaMainStreams(...) (include/cudaMain.h:53-68,
 * src/cudaMain.cu:377-1667).  voxels: host [M][V][4] of the compute precision;
 * out: host [nE][nK][Ny][Nx] of the compute precision.  Host buffers are
 * caller-owned and not retained after return. */
cyr_status cyr_simulate(const cyr_sim_desc *desc, const void *voxels, const double *mats, void *out);

/* Same, with inputs and outputs already resident in HBM.  d_voxels is
 * [M][V][4] of the storage type (float/double, or 16-bit for FP16/BF16
 * storage); d_out is [nE][nK][Ny][Nx] of the compute precision.  Work is
 * queued on `stream` (a hipStream_t, or NULL for the default stream) and the
 * call returns after the stream has finished. */
cyr_status cyr_simulate_device(const cyr_sim_desc *desc, const void *d_voxels, const double *mats,
                               void *d_out, void *stream);

/* Allocation-free form of cyr_simulate_device, with the same arguments and
 * results.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  It calls no hipMalloc/hipFree and no
 * allocating library routine: the constant tables, twiddle tables and work
 * buffers live in one static device array of CYR_NOALLOC_SCRATCH_BYTES per GPU,
 * held from library load until the process ends.  Simulations that need more
 * (e.g. 64^3 fits, 512^2 x 128 does not) are rejected with CYR_E_INVALID before
 * any device work.  Every call on one GPU shares that GPU's array, which keeps
 * the last call's contents (each call initialises what it reads); calls on one
 * GPU from several threads are serialised by a lock, and each returns after
 * `stream` has drained.  Not for use under graph capture of `stream` (the call
 * synchronises it). */
#define CYR_NOALLOC_SCRATCH_BYTES (128u * 1024u * 1024u)
cyr_status cyr_simulate_device_noalloc(const cyr_sim_desc *desc, const void *d_voxels, const double *mats,
                                       void *d_out, void *stream);

/* Replaces int computePolarization(...) (include/cudaMain.h:84-87,
 * src/cudaMain.cu:1669-1738): one polarization for (energy_id, eangle_rad),
 * k vector 0, copied to three host complex arrays [V] of the compute precision. */
cyr_status cyr_polarization(const cyr_sim_desc *desc, const void *voxels, const double *mats,
                            uint32_t energy_id, double eangle_rad, void *px, void *py, void *pz);

/* Device form of the polarization launcher computePolarization(d_materialConstants,
 * d_voxelInput, ...) (src/cudaMain.cu:299-332): d_p* are [V] complex. */
cyr_status cyr_polarization_device(const cyr_sim_desc *desc, const void *d_voxels,
                                   const double *mats_e, double eangle_rad, void *d_px,
                                   void *d_py, void *d_pz, void *stream);

/* ---- stage-level device entry points (the reference's host launchers) ----- */

/* Forward, unnormalised, in-place 3D C2C FFT of `count` volumes; replaces
 * performFFT / cufftExecC2C|Z2Z (include/cudaMain.h:184-190). */
cyr_status cyr_fft3d_device(uint32_t nx, uint32_t ny, uint32_t nz, int32_t precision,
                            void *const *d_volumes, uint32_t count, void *stream);

/* One axis (0 = X, 1 = Y, 2 = Z) of the same transform, queued on `stream`
 * without synchronising (the pass cufftPlan3d performs internally). */
cyr_status cyr_fft_axis_device(uint32_t nx, uint32_t ny, uint32_t nz, int32_t precision,
                               void *const *d_volumes, uint32_t count, int32_t axis, void *stream);

/* DC replacement then Igor shift, in place; replaces replaceDCComponent +
 * performFFTShift (src/cudaMain.cu:57-60,178-181). */
cyr_status cyr_dc_shift_device(uint32_t nx, uint32_t ny, uint32_t nz, int32_t precision,
                               void *d_volume, int32_t do_dc, int32_t do_shift, void *stream);

/* Ewald projection for (energy_id, k_id) from UNshifted FFT volumes (after DC
 * replacement); the Igor shift is applied in the gather addressing.  Replaces
 * peformEwaldProjectionGPU (src/cudaMain.cu:203-243); d_proj is [Ny][Nx]. */
cyr_status cyr_ewald_device(const cyr_sim_desc *desc, uint32_t energy_id, uint32_t k_id,
                            const void *d_px, const void *d_py, const void *d_pz, void *d_proj,
                            void *stream);

/* Bilinear affine warp of an [Ny][Nx] image with forward coefficients
 * c = [[c0,c1,c2],[c3,c4,c5]] (source -> destination); destination pixels whose
 * back-mapped point falls outside the source stay as they were.  Replaces
 * nppiWarpAffine_{32f,64f}_C1R_Ctx with NPPI_INTER_LINEAR (src/cudaMain.cu:91-121). */
cyr_status cyr_warp_affine_device(uint32_t nx, uint32_t ny, int32_t precision, const void *d_src,
                                  void *d_dst, const double coeffs[6], void *stream);

/* ---- batched execution ---------------------------------------------------- */

/* Runs `count` independent simulations (uneven grids allowed) resident in HBM as
 * one schedule: members are spread over several streams (longest first onto the
 * least-loaded stream, each stream with its own work buffers), their constant
 * tables are uploaded together, and every E-basis grid shape without a tuning is
 * autotuned first.  Returns after `stream` has finished.  No counterpart in the
 * reference (it runs one simulation per call). */
cyr_status cyr_simulate_batch(uint32_t count, const cyr_sim_desc *descs,
                              const void *const *d_voxels, const double *const *mats,
                              void *const *d_outs, void *stream);

/* A batch plan: the same schedule, set up once (tables resident, shapes tuned,
 * buffers sized by one eager run whose outputs land in d_outs) and captured in a
 * hipGraph over `lanes` streams (0 = automatic).  The device pointers are bound
 * at creation; cyr_batch_launch replays the graph on `stream` (for instance
 * after the caller has refreshed the morphologies in place) and returns
 * without waiting.  Replays of one plan must not overlap. */
typedef struct cyr_batch_s *cyr_batch;
cyr_status cyr_batch_create(uint32_t count, const cyr_sim_desc *descs, const void *const *d_voxels,
                            const double *const *mats, void *const *d_outs, uint32_t lanes,
                            cyr_batch *out);
cyr_status cyr_batch_launch(cyr_batch b, void *stream);
cyr_status cyr_batch_destroy(cyr_batch b);

/* ---- serving engine ------------------------------------------------------- */

typedef struct cyr_engine_s *cyr_engine;

typedef struct {
  uint64_t submitted, completed, rejected, cancelled;
  uint64_t resident_bytes;   /* HBM counted against the budget: what the slots' buffers hold
                                between requests plus the footprints of requests now running */
  uint64_t budget_bytes;
  uint32_t in_flight;
  uint32_t queued;
} cyr_engine_stats;

/* Continuous-batching engine on one device: requests are admitted in arrival
 * order once their HBM footprint (cyr_footprint_bytes) fits in budget_bytes
 * (0 = 90% of the device memory) next to the requests already running; each of
 * `slots` (1..16) host threads drives its own stream, so requests that arrive
 * while others run start as soon as a slot and memory free up and their work
 * overlaps on the device.  Each request's morphology stays resident in HBM
 * across its energies.  No counterpart in the reference. */
cyr_status cyr_engine_create(int device, uint64_t budget_bytes, uint32_t slots, cyr_engine *out);
/* Queues a simulation.  desc and mats are copied; voxels (host, [M][V][4] of the
 * compute precision) and out (host, [nE][nK][Ny][Nx]) are used in place and
 * must stay valid until the request completes (pinned host memory makes the
 * transfers asynchronous).  CYR_E_ADMISSION if the footprint exceeds the budget. */
cyr_status cyr_engine_submit(cyr_engine eng, const cyr_sim_desc *desc, const void *voxels,
                             const double *mats, void *out, uint64_t *request_id);
/* Waits up to timeout_s (< 0: forever; CYR_E_TIMEOUT if still in flight) and
 * reports the request's own status (CYR_E_CANCELLED if it was cancelled). */
cyr_status cyr_engine_wait(cyr_engine eng, uint64_t request_id, double timeout_s,
                           cyr_status *request_status);
cyr_status cyr_engine_poll(cyr_engine eng, uint64_t request_id, int *done,
                           cyr_status *request_status);
/* Time in the queue, time running (upload, simulation, download) and the whole
 * submit-to-result latency of a completed request, in seconds. */
cyr_status cyr_engine_request_times(cyr_engine eng, uint64_t request_id, double *queued_s,
                                    double *run_s, double *latency_s);
/* Forgets a finished request (its id becomes unknown). */
cyr_status cyr_engine_release(cyr_engine eng, uint64_t request_id);
/* Cancels a request that has not started (CYR_E_INVALID once it runs). */
cyr_status cyr_engine_cancel(cyr_engine eng, uint64_t request_id);
cyr_status cyr_engine_get_stats(cyr_engine eng, cyr_engine_stats *stats);
/* Cancels queued requests, lets running ones finish, frees everything. */
cyr_status cyr_engine_destroy(cyr_engine eng);

/* ---- autotuning ------------------------------------------------------------ */

/* Launch choices of the E-basis executor's Y-FFT / z-bin kernel for one grid
 * shape (Nx, Ny, Nz) and precision: columns per work group (0 or a power of two
 * <= 16) and the grid size its split of z aims for (0 = built-in defaults).
 * No counterpart in the reference (cuFFT plans its own launches). */
typedef struct {
  uint32_t columns;
  uint32_t target_blocks;
} cyr_yz_tuning;

/* Records / reads the process-wide choice for a shape ({0, 0} clears it). */
cyr_status cyr_set_tuning(const uint32_t dims[3], int32_t precision, const cyr_yz_tuning *t);
cyr_status cyr_get_tuning(const uint32_t dims[3], int32_t precision, cyr_yz_tuning *t);
/* Times the candidate launch shapes for desc's grid on the current device (HIP
 * events, zero-filled fields) and records the fastest; *chosen (may be NULL)
 * receives it ({0, 0} when the grid has nothing to tune, e.g. 2D). */
cyr_status cyr_autotune(const cyr_sim_desc *desc, cyr_yz_tuning *chosen);

/* (energy, k) units the E-basis executor keeps in flight on separate streams,
 * each with its own work buffers: 1 (default, serial) or 2 (one unit's
 * polarization/X-FFT kernel may overlap the previous unit's Y/z chain; pays off
 * only when single kernels do not fill the device).  Results are identical
 * either way.  No counterpart in the reference. */
cyr_status cyr_set_energy_lanes(uint32_t lanes);
uint32_t cyr_energy_lanes(void);

/* ---- utilities ------------------------------------------------------------- */

const char *cyr_status_string(cyr_status s);
/* Detail of the last failure on the calling thread ("" if none). */
const char *cyr_last_error(void);
int cyr_device_count(void);
/* numAnglesRotation = round((end - start)/increment + 1) (src/cudaMain.cu:393-394). */
uint32_t cyr_num_angles(double start, double increment, double end);
/* The executor a descriptor runs: CYR_PIPE_PLAIN, CYR_PIPE_EBASIS or
 * CYR_PIPE_MBASIS (CYR_PIPE_AUTO resolved; grids without fused-kernel plans run
 * PLAIN); -1 for an invalid descriptor.  No counterpart in the reference. */
int32_t cyr_resolved_pipeline(const cyr_sim_desc *desc);
/* Device bytes a simulation holds while it runs (morphology + work volumes). */
uint64_t cyr_footprint_bytes(const cyr_sim_desc *desc);
/* Frees the library's per-device caches (work buffers, twiddle tables, batch
 * streams, upload staging).  It waits for calls that are using them and returns
 * CYR_E_INVALID while a batch plan or serving engine exists (their graphs and
 * threads keep pointers into the caches): destroy those first.  Work queued
 * through the library must have finished (cyr_fft_axis_device returns before
 * its kernels do). */
cyr_status cyr_release_workspace(void);
/* Library version string. */
const char *cyr_version(void);

/* Opt-in per-kernel timing: while enabled, the executors bracket every launch
 * with HIP events on the launch stream and accumulate the durations per kernel.
 * cyr_kernel_timing(1) also clears the accumulated figures.  No counterpart in
 * the reference (its PROFILING build times stages with host clocks,
 * src/cudaMain.cu:407-434). */
cyr_status cyr_kernel_timing(int enable);
/* Writes "name calls total_us avg_us avg_bytes" lines into buf (NUL-terminated,
 * truncated to len), avg_bytes being the launch's algorithmic HBM traffic (each
 * input read once, each output written once); returns the number of kernels. */
int cyr_kernel_timing_report(char *buf, size_t len);

#ifdef __cplusplus
}
#endif
#endif /* CYRSOXS_H */
