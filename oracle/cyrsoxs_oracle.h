/*
 * FP64 CPU oracle of the CyRSoXS Born-approximation pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (cyrsoxs_amd/) never links, loads or calls it.
 *
 * It restates, function by function and in IEEE double precision, what the
 * reference computes on the GPU.  Every function cites the reference file:line
 * it follows (paths relative to the usnistgov/cyrsoxs source tree).  Where the
 * reference's behaviour is defined by a closed NVIDIA library (cuFFT, NPP
 * warpAffine, cuBLAS) the oracle implements the documented mathematical
 * contract instead: the unnormalised forward DFT for cuFFT, and the bilinear
 * inverse-map convention documented at orc_warp_affine() for NPP.
 *
 * Data conventions (all arrays are caller-owned):
 *   - voxel volumes are X-fastest: idx = X + Nx*(Y + Ny*Z)   (include/uniaxial.h:531-533)
 *   - complex arrays are interleaved doubles [re, im]
 *   - morphology: double[M][V][4] = (s1.x, s1.y, s1.z, s1.w) per voxel per material
 *       Euler:  (S, theta, psi, Vfrac)          (include/Input/Input.h:61-76)
 *       Vector: (sx, sy, sz, phi_unaligned)
 *   - optical constants: double[nE][M][4] = (npara.re, npara.im, nperp.re, nperp.im)
 *     with n = (1 - delta) + i*beta            (include/Input/InputData.h:371-375)
 *   - patterns: double[nE][nK][Ny][Nx]          (src/cudaMain.cu:968-971)
 *
 * Parity pins (tests/test_oracle_*.py):
 *   rotation matrices   -> closed-form Rz(theta), orthogonality, Rodrigues identities
 *   polarization        -> vacuum, isotropic, single-aligned-voxel closed forms, linearity in E
 *   FFT                 -> numpy.fft.fftn, brute-force O(N^2) DFT, delta/plane wave, Parseval
 *   DC / Igor shift     -> integer index maps (involution, neighbour rule)
 *   Ewald projection    -> per-pixel closed form evaluated independently in numpy
 *   warp affine         -> identity at 0 deg, exact permutation at 90 deg (square, even N)
 *   whole pipeline      -> Born sphere form factor, contrast scaling |dn^2|^2, linearity checks
 * The NPP warpAffine boundary/NaN-tap rule is "parity unpinned": the reference
 * ships no outputs and NPP is closed; the convention here is the build's own.
 */
#ifndef CYRSOXS_ORACLE_H
#define CYRSOXS_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  uint32_t nx, ny, nz;           /* voxel dims, X fastest */
  uint32_t num_materials;
  double phys_size;              /* nm */
  uint32_t num_energies;
  const double *energies;        /* eV */
  double angle_start;            /* degrees */
  double angle_increment;        /* degrees */
  double angle_end;              /* degrees */
  uint32_t num_kvectors;
  const double *kvectors;        /* [nK][3], normalised */
  double detector[3];            /* normalised detector direction */
  int32_t case_type;             /* 0 default, 1 beam divergence, 2 grazing incidence */
  int32_t morphology_type;       /* 0 Euler, 1 vector */
  int32_t reference_frame;       /* 0 MATERIAL, 1 LAB */
  int32_t interpolation;         /* 0 nearest neighbour, 1 linear */
  int32_t windowing;             /* 0 none, 1 Hanning */
  int32_t rot_mask;              /* 0/1 */
  int32_t scatter_approach;      /* 0 partial, 1 full */
  int32_t algorithm;             /* 0 communication minimizing, 1 memory minimizing (Nt) */
} orc_desc;

/* ---- host rotation math (include/Rotation.h) ---- */
void orc_rotation_vec_to_vec(const double a[3], const double b[3], double R[9]);
void orc_rodrigues(double out[3], const double v[3], const double axis[3], double angle);
void orc_rotation_matrix_k(const double k[3], double RK[9]);
void orc_base_configuration(const double k[3], const double RK[9], double R[9], double *base_angle);
void orc_rotation_matrix_e(const double k[3], const double RK[9], double angle, double R[9]);
void orc_warp_affine_matrix(const double src[6], const double dst[6], double coeffs[6]);
uint32_t orc_num_angles(double start, double increment, double end);

/* ---- per-stage kernels ---- */
void orc_polarization(const orc_desc *d, const double *voxels, const double *mats_e,
                      const double R[9], double *px, double *py, double *pz);
void orc_compute_nt(const orc_desc *d, const double *voxels, const double *mats_e, double *nt);
void orc_polarization_nt(const orc_desc *d, const double *nt, const double R[9],
                         double *px, double *py, double *pz);
void orc_fft3d(double *data, uint32_t nx, uint32_t ny, uint32_t nz);
void orc_fft1d(double *data, uint32_t n);
void orc_dft1d_direct(const double *in, double *out, uint32_t n);
void orc_dc_replace(double *data, uint32_t nx, uint32_t ny, uint32_t nz);
void orc_igor_shift(double *data, uint32_t nx, uint32_t ny, uint32_t nz);
void orc_scatter3d(const orc_desc *d, const double *px, const double *py, const double *pz,
                   double k, const double kvec[3], double *scatter);
void orc_ewald_partial(const orc_desc *d, const double *px, const double *py, const double *pz,
                       double k, const double kvec[3], double *proj);
void orc_ewald_full(const orc_desc *d, const double *scatter, double k, const double kvec[3],
                    double *proj);
void orc_warp_affine(const double *src, double *dst, uint32_t nx, uint32_t ny,
                     const double coeffs[6]);
void orc_rotation_coeffs(double angle_rad, uint32_t nx, uint32_t ny, double coeffs[6]);

/* ---- whole pipeline (src/cudaMain.cu:377-1035 and :1037-1667) ---- */
int orc_simulate(const orc_desc *d, const double *voxels, const double *mats, double *out);
/* Same, restricted to energies [e0, e1) and angles [a0, a1): the bounded
 * CPU-baseline sample.  out holds (e1-e0)*nK patterns. */
int orc_simulate_subset(const orc_desc *d, const double *voxels, const double *mats, double *out,
                        uint32_t e0, uint32_t e1, uint32_t a0, uint32_t a1);
int orc_max_threads(void);
void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
