// FP64 host-side geometry of the pipeline: E-field rotation matrices, k-vector
// base configurations, per-angle image-rotation coefficients and the detector
// warp.  Everything here runs once per simulation (or per angle) on the host, is
// compiled with -ffp-contract=off, and is passed to the kernels as tables so
// that every geometric decision on the GPU is the IEEE-double value of the
// expression as written.
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/cyrsoxs.h"

namespace cyr {

struct M3 { double m[9]; };
struct V3 { double x, y, z; };

// include/Rotation.h:272-315 (F G F^-1 rotation taking a onto b, FEQUALS shortcut)
M3 rotation_vec_to_vec(const V3 &a, const V3 &b);
// include/Rotation.h:326-338
V3 rodrigues(const V3 &v, const V3 &axis, double angle);
// include/Rotation.h:402-413
M3 rotation_matrix_k(const V3 &k);
// include/Rotation.h:422-465 (returns the base angle; R written to *R if non-null)
double base_configuration(const V3 &k, const M3 &RK, M3 *R);
// include/Rotation.h:474-484
M3 rotation_matrix_e(const V3 &k, const M3 &RK, double angle);
// include/Rotation.h:499-516
void warp_affine_matrix(const double src[6], const double dst[6], double c[6]);
M3 matmul(const M3 &A, const M3 &B);
V3 matvec(const M3 &A, const V3 &v);

// Inverse of a forward affine warp (source -> destination):
// (i00, i01, b0, i10, i11, b1) such that src = (i00 X + i01 Y + b0, i10 X + i11 Y + b1).
void affine_inverse(const double c[6], double inv[6]);
// OpenCV-form rotation about (Nx/2, Ny/2) used for the E-angle image rotation
// (src/cudaMain.cu:848-855).
void rotation_coeffs(double angle, uint32_t nx, uint32_t ny, double c[6]);

uint32_t num_angles(double start, double inc, double end);

// Everything the executors need that is fixed for one simulation.
struct SimGeometry {
  uint32_t nx = 0, ny = 0, nz = 0, M = 0, nE = 0, nK = 0, nA = 0;
  // angle steps per 90 degrees when the E-angle list steps an exact divisor of
  // 90 degrees and spans whole quarter turns (0 otherwise)
  uint32_t quarter = 0;
  uint64_t V = 0, P = 0;
  std::vector<M3> RK;          // per k
  std::vector<double> base;    // per k base angle (radians)
  M3 detector;                 // detector rotation
  std::vector<double> det_inv; // per k: 6 inverse coefficients of the detector warp
  // per (k, angle): E-rotation, E vector, image-rotation inverse coefficients,
  // and the E-basis coefficients (E = c e1 + s e2).
  std::vector<M3> R;           // [nK][nA]
  std::vector<double> rot_inv; // [nK][nA][6]
  std::vector<double> ebasis;  // [nK][nA][2]
  std::vector<V3> e1, e2;      // per k
  std::vector<V3> kvec;        // per k, normalised
  std::vector<double> kmag;    // per energy
};

cyr_status validate_desc(const cyr_sim_desc *d);
cyr_status build_geometry(const cyr_sim_desc *d, SimGeometry &g);

}  // namespace cyr
