/*
 * FP64 CPU oracle of the CyRSoXS Born-approximation pipeline.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  See
 * cyrsoxs_oracle.h for the data conventions and the list of parity pins.
 * Built with -ffp-contract=off so every geometric decision (Ewald validity,
 * nearest-neighbour z index, rotation validity, bilinear tap selection) is the
 * IEEE-double result of the expression as written.
 */
#include "cyrsoxs_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef double complex cplx;

/* ------------------------------------------------------------------------- */
/* Host rotation math                                                         */
/* ------------------------------------------------------------------------- */

/* FEQUALS(x,y): |x-y| < 1e-6  (include/Datatypes.h:143) */
static int fequals(double x, double y) { return fabs(x - y) < 1E-6; }

static double dot3(const double a[3], const double b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* computeCrossProduct (include/Rotation.h:188-194) */
static void cross3(const double a[3], const double b[3], double c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = -(a[0] * b[2] - a[2] * b[0]);
  c[2] = a[0] * b[1] - a[1] * b[0];
}

/* Matrix::performMatrixMultiplication<false,false> (include/Rotation.h:104-114) */
static void matmul3(const double A[9], const double B[9], double C[9]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += A[i * 3 + k] * B[k * 3 + j];
      C[i * 3 + j] = s;
    }
}

/* doMatVec<false> (include/Rotation.h:353-356) */
static void matvec3(const double A[9], const double v[3], double o[3]) {
  o[0] = A[0] * v[0] + A[1] * v[1] + A[2] * v[2];
  o[1] = A[3] * v[0] + A[4] * v[1] + A[5] * v[2];
  o[2] = A[6] * v[0] + A[7] * v[1] + A[8] * v[2];
}

/* computeInverseMatrix (include/Rotation.h:233-252) */
static void inverse3(const double A[9], double Inv[9]) {
  double det = A[0] * (A[4] * A[8] - A[7] * A[5]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
               A[2] * (A[3] * A[7] - A[4] * A[6]);
  double invdet = 1. / det;
  Inv[0] = (A[4] * A[8] - A[7] * A[5]) * invdet;
  Inv[1] = (A[2] * A[7] - A[1] * A[8]) * invdet;
  Inv[2] = (A[1] * A[5] - A[2] * A[4]) * invdet;
  Inv[3] = (A[5] * A[6] - A[3] * A[8]) * invdet;
  Inv[4] = (A[0] * A[8] - A[2] * A[6]) * invdet;
  Inv[5] = (A[3] * A[2] - A[0] * A[5]) * invdet;
  Inv[6] = (A[3] * A[7] - A[6] * A[4]) * invdet;
  Inv[7] = (A[6] * A[1] - A[0] * A[7]) * invdet;
  Inv[8] = (A[0] * A[4] - A[3] * A[1]) * invdet;
}

/* computeRotationMatrix(originalVec, transformedVec, R): the F*G*F^-1 rotation
 * taking a onto b, with the FEQUALS identity shortcut (include/Rotation.h:272-315). */
void orc_rotation_vec_to_vec(const double a[3], const double b[3], double R[9]) {
  if (fequals(a[0], b[0]) && fequals(a[1], b[1]) && fequals(a[2], b[2])) {
    memset(R, 0, 9 * sizeof(double));
    R[0] = R[4] = R[8] = 1;
    return;
  }
  double dp = dot3(a, b);
  double cr[3];
  cross3(a, b, cr);
  double nc = sqrt(dot3(cr, cr));
  double G[9] = {dp, -nc, 0, nc, dp, 0, 0, 0, 1};
  double sub[3] = {b[0] - dp * a[0], b[1] - dp * a[1], b[2] - dp * a[2]};
  double ns = sqrt(dot3(sub, sub));
  double F[9] = {a[0], sub[0] / ns, -cr[0], a[1], sub[1] / ns, -cr[1], a[2], sub[2] / ns, -cr[2]};
  double iF[9], t[9];
  inverse3(F, iF);
  matmul3(G, iF, t);
  matmul3(F, t, R);
}

/* performRodriguesRotation (include/Rotation.h:326-338) */
void orc_rodrigues(double out[3], const double v[3], const double axis[3], double angle) {
  double c = cos(angle), s = sin(angle);
  double cr[3];
  cross3(axis, v, cr);
  double dp = dot3(axis, v);
  double f = (1 - c) * dp;
  for (int i = 0; i < 3; ++i) out[i] = c * v[i] + s * cr[i] + axis[i] * f;
}

/* computeRotationMatrixK: rotation taking z onto k (include/Rotation.h:402-413) */
void orc_rotation_matrix_k(const double k[3], double RK[9]) {
  static const double z[3] = {0, 0, 1};
  orc_rotation_vec_to_vec(z, k, RK);
}

/* computeRotationMatrixBaseConfiguration: 100000-step search for the E angle
 * that makes the rotated x axis have |y| minimal and x > 0
 * (include/Rotation.h:422-465). */
void orc_base_configuration(const double k[3], const double RK[9], double R[9], double *base_angle) {
  static const double X[3] = {1, 0, 0};
  const unsigned num_interval = 100000;
  const double dtheta = M_PI / (num_interval * 1.0);
  double sx[3], rx[3];
  matvec3(RK, X, sx);
  double maxdiff = INFINITY, ang = 0;
  for (unsigned i = 0; i < num_interval; i++) {
    orc_rodrigues(rx, sx, k, i * dtheta);
    double diff = fabs(rx[1]);
    if (maxdiff > diff) {
      maxdiff = diff;
      ang = i * dtheta;
    }
  }
  orc_rodrigues(rx, sx, k, ang);
  ang = rx[0] > 0 ? ang : ang + M_PI;
  orc_rodrigues(rx, sx, k, ang);
  double RX[9];
  orc_rotation_vec_to_vec(sx, rx, RX);
  matmul3(RX, RK, R);
  *base_angle = ang;
}

/* computeRotationMatrix(k, RK, R, angle): E-field rotation (include/Rotation.h:474-484) */
void orc_rotation_matrix_e(const double k[3], const double RK[9], double angle, double R[9]) {
  static const double X[3] = {1, 0, 0};
  double sx[3], rx[3], RX[9];
  matvec3(RK, X, sx);
  orc_rodrigues(rx, sx, k, angle);
  orc_rotation_vec_to_vec(sx, rx, RX);
  matmul3(RX, RK, R);
}

/* computeWarpAffineMatrix: affine map through 3 point pairs (include/Rotation.h:499-516).
 * src/dst are [3][2]; coeffs is [2][3] row-major. */
void orc_warp_affine_matrix(const double src[6], const double dst[6], double c[6]) {
  double x0 = src[0], y0 = src[1], x1 = src[2], y1 = src[3], x2 = src[4], y2 = src[5];
  double u0 = dst[0], v0 = dst[1], u1 = dst[2], v1 = dst[3], u2 = dst[4], v2 = dst[5];
  double denom = x0 * y1 - x1 * y0 - x0 * y2 + x2 * y0 + x1 * y2 - x2 * y1;
  c[0] = (u0 * y1 - u1 * y0 - u0 * y2 + u2 * y0 + u1 * y2 - u2 * y1) / (denom);
  c[1] = -(u0 * x1 - u1 * x0 - u0 * x2 + u2 * x0 + u1 * x2 - u2 * x1) / (denom);
  c[2] = (u0 * x1 * y2 - u0 * x2 * y1 - u1 * x0 * y2 + u1 * x2 * y0 + u2 * x0 * y1 - u2 * x1 * y0) / (denom);
  c[3] = (v0 * y1 - v1 * y0 - v0 * y2 + v2 * y0 + v1 * y2 - v2 * y1) / (denom);
  c[4] = -(v0 * x1 - v1 * x0 - v0 * x2 + v2 * x0 + v1 * x2 - v2 * x1) / (denom);
  c[5] = (v0 * x1 * y2 - v0 * x2 * y1 - v1 * x0 * y2 + v1 * x2 * y0 + v2 * x0 * y1 - v2 * x1 * y0) / (denom);
}

/* numAnglesRotation = round((end-start)/inc + 1)  (src/cudaMain.cu:393-394) */
uint32_t orc_num_angles(double start, double increment, double end) {
  return (uint32_t)round((end - start) / increment + 1);
}

/* Per-angle warp coefficients of the E-angle image rotation, OpenCV
 * getRotationMatrix2D form about (Nx/2, Ny/2) (src/cudaMain.cu:848-855). */
void orc_rotation_coeffs(double angle, uint32_t nx, uint32_t ny, double c[6]) {
  double alpha = cos(angle), beta = sin(angle);
  c[0] = alpha;
  c[1] = beta;
  c[2] = ((1 - alpha) * nx / 2 - beta * ny / 2.);
  c[3] = -beta;
  c[4] = alpha;
  c[5] = (beta * nx / 2. + (1 - alpha) * ny / 2.);
}

/* ------------------------------------------------------------------------- */
/* Polarization                                                               */
/* ------------------------------------------------------------------------- */

static cplx csq(cplx v) {  /* computeComplexSquare (include/cudaUtils.h:243-249) */
  double re = creal(v), im = cimag(v);
  return (re * re - im * im) + I * (2 * re * im);
}

/* Hanning weight h(i,N) = 0.5(1-cos(2 pi i/N))  (src/cudaMain.cu:276-295) */
static double hanning(uint32_t i, uint32_t n) { return 0.5 * (1 - cos(2 * M_PI * i / (n))); }

/* computePolarizationEulerAngles / computePolarizationVectorMorphologyOptimized
 * (include/uniaxial.h:58-169, :184-292), wrapped by the computePolarization
 * kernel that applies the Hanning window (src/cudaMain.cu:245-297). */
void orc_polarization(const orc_desc *d, const double *voxels, const double *mats_e,
                      const double R[9], double *px, double *py, double *pz) {
  const uint64_t V = (uint64_t)d->nx * d->ny * d->nz;
  const uint32_t M = d->num_materials;
  const double one_by_4pi = 1.0 / (4.0 * M_PI);
  const double mvx = R[0], mvy = R[3], mvz = R[6]; /* R * (1,0,0) */
  const int enable2d = d->nz == 1;
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < (int64_t)V; t++) {
    cplx pX = 0, pY = 0, pZ = 0;
    for (uint32_t m = 0; m < M; m++) {
      cplx npar = mats_e[4 * m + 0] + I * mats_e[4 * m + 1];
      cplx nper = mats_e[4 * m + 2] + I * mats_e[4 * m + 3];
      const double *s1 = voxels + 4 * ((uint64_t)m * V + t);
      double sx, sy, sz, phi_a, phi_ui, phi;
      if (d->morphology_type == 1) {
        sx = s1[0];
        sy = s1[1];
        sz = s1[2];
        phi_ui = s1[3];
        phi = phi_ui + sx * sx + sy * sy + sz * sz;
        phi_a = 1.0;
      } else {
        double psi = s1[2], theta = s1[1], vfrac = s1[3];
        phi_a = vfrac * s1[0];
        sx = cos(psi) * sin(theta);
        sy = sin(psi) * sin(theta);
        sz = cos(theta);
        phi_ui = vfrac - phi_a;
        phi = vfrac;
      }
      cplx nsum = npar + 2 * nper;
      nsum = csq(nsum);
      npar = csq(npar);
      nper = csq(nper);
      double npr = creal(npar), npi = cimag(npar), nqr = creal(nper), nqi = cimag(nper);
      double nsr = creal(nsum), nsi = cimag(nsum);
      cplx r;
      r = (phi_a * (npr * sx * sx + nqr * (sy * sy + sz * sz)) + ((phi_ui * nsr) / 9.0) - phi) +
          I * (phi_a * (npi * sx * sx + nqi * (sy * sy + sz * sz)) + ((phi_ui * nsi) / 9.0));
      pX += r * mvx;
      r = (phi_a * (npr - nqr) * sx * sy) + I * (phi_a * (npi - nqi) * sx * sy);
      pX += r * mvy;
      pY += r * mvx;
      r = (phi_a * (npr - nqr) * sx * sz) + I * (phi_a * (npi - nqi) * sx * sz);
      pX += r * mvz;
      pZ += r * mvx;
      r = (phi_a * (npr * sy * sy + nqr * (sx * sx + sz * sz)) + ((phi_ui * nsr) / 9.0) - phi) +
          I * (phi_a * (npi * sy * sy + nqi * (sx * sx + sz * sz)) + ((phi_ui * nsi) / 9.0));
      pY += r * mvy;
      r = (phi_a * (npr - nqr) * sy * sz) + I * (phi_a * (npi - nqi) * sy * sz);
      pY += r * mvz;
      pZ += r * mvy;
      r = (phi_a * (npr * sz * sz + nqr * (sx * sx + sy * sy)) + ((phi_ui * nsr) / 9.0) - phi) +
          I * (phi_a * (npi * sz * sz + nqi * (sx * sx + sy * sy)) + ((phi_ui * nsi) / 9.0));
      pZ += r * mvz;
    }
    pX *= one_by_4pi;
    pY *= one_by_4pi;
    pZ *= one_by_4pi;
    if (d->reference_frame == 0) { /* MATERIAL: p <- R^T p  (rotate<true>, include/Rotation.h:376-384) */
      cplx tX = R[0] * pX + R[3] * pY + R[6] * pZ;
      cplx tY = R[1] * pX + R[4] * pY + R[7] * pZ;
      cplx tZ = R[2] * pX + R[5] * pY + R[8] * pZ;
      pX = tX;
      pY = tY;
      pZ = tZ;
    }
    if (d->windowing == 1) {
      uint32_t Z = (uint32_t)(t / ((uint64_t)d->nx * d->ny));
      uint32_t Y = (uint32_t)((t - (uint64_t)Z * d->nx * d->ny) / d->nx);
      uint32_t X = (uint32_t)(t - (uint64_t)Y * d->nx - (uint64_t)Z * d->nx * d->ny);
      double w = hanning(X, d->nx) * hanning(Y, d->ny) * (enable2d ? 1.0 : hanning(Z, d->nz));
      pX *= w;
      pY *= w;
      pZ *= w;
    }
    px[2 * t] = creal(pX);
    px[2 * t + 1] = cimag(pX);
    py[2 * t] = creal(pY);
    py[2 * t + 1] = cimag(pY);
    pz[2 * t] = creal(pZ);
    pz[2 * t + 1] = cimag(pZ);
  }
}

/* computeNtEulerAngles / computeNtVectorMorphology accumulated over materials
 * (include/uniaxial.h:305-430, driver src/cudaMain.cu:1241-1256).
 * nt layout here: double[6][V][2], entries in order 00,01,02,11,12,22. */
void orc_compute_nt(const orc_desc *d, const double *voxels, const double *mats_e, double *nt) {
  const uint64_t V = (uint64_t)d->nx * d->ny * d->nz;
  const uint32_t M = d->num_materials;
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < (int64_t)V; t++) {
    cplx acc[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t m = 0; m < M; m++) {
      cplx npar = mats_e[4 * m + 0] + I * mats_e[4 * m + 1];
      cplx nper = mats_e[4 * m + 2] + I * mats_e[4 * m + 3];
      const double *s1 = voxels + 4 * ((uint64_t)m * V + t);
      double sx, sy, sz, phi_a, phi_ui, phi;
      if (d->morphology_type == 1) {
        sx = s1[0];
        sy = s1[1];
        sz = s1[2];
        phi_ui = s1[3];
        phi = phi_ui + sx * sx + sy * sy + sz * sz;
        phi_a = 1.0;
      } else {
        double psi = s1[2], theta = s1[1], vfrac = s1[3];
        phi_a = vfrac * s1[0];
        sx = cos(psi) * sin(theta);
        sy = sin(psi) * sin(theta);
        sz = cos(theta);
        phi_ui = vfrac - phi_a;
        phi = vfrac;
      }
      cplx nsum = csq(npar + 2 * nper);
      npar = csq(npar);
      nper = csq(nper);
      double npr = creal(npar), npi = cimag(npar), nqr = creal(nper), nqi = cimag(nper);
      double nsr = creal(nsum), nsi = cimag(nsum);
      acc[0] += (phi_a * (npr * sx * sx + nqr * (sy * sy + sz * sz)) + ((phi_ui * nsr) / 9.0) - phi) +
                I * (phi_a * (npi * sx * sx + nqi * (sy * sy + sz * sz)) + ((phi_ui * nsi) / 9.0));
      acc[1] += (phi_a * (npr - nqr) * sx * sy) + I * (phi_a * (npi - nqi) * sx * sy);
      acc[2] += (phi_a * (npr - nqr) * sx * sz) + I * (phi_a * (npi - nqi) * sx * sz);
      acc[3] += (phi_a * (npr * sy * sy + nqr * (sx * sx + sz * sz)) + ((phi_ui * nsr) / 9.0) - phi) +
                I * (phi_a * (npi * sy * sy + nqi * (sx * sx + sz * sz)) + ((phi_ui * nsi) / 9.0));
      acc[4] += (phi_a * (npr - nqr) * sy * sz) + I * (phi_a * (npi - nqi) * sy * sz);
      acc[5] += (phi_a * (npr * sz * sz + nqr * (sx * sx + sy * sy)) + ((phi_ui * nsr) / 9.0) - phi) +
                I * (phi_a * (npi * sz * sz + nqi * (sx * sx + sy * sy)) + ((phi_ui * nsi) / 9.0));
    }
    for (int j = 0; j < 6; j++) {
      nt[2 * ((uint64_t)j * V + t)] = creal(acc[j]);
      nt[2 * ((uint64_t)j * V + t) + 1] = cimag(acc[j]);
    }
  }
}

/* computePolarizationVectorMorphologyLowMemory: p = Nt E / 4pi, no Hanning
 * (include/uniaxial.h:441-521). */
void orc_polarization_nt(const orc_desc *d, const double *nt, const double R[9], double *px,
                         double *py, double *pz) {
  const uint64_t V = (uint64_t)d->nx * d->ny * d->nz;
  const double one_by_4pi = 1.0 / (4.0 * M_PI);
  const double mvx = R[0], mvy = R[3], mvz = R[6];
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < (int64_t)V; t++) {
    cplx A[6];
    for (int j = 0; j < 6; j++)
      A[j] = nt[2 * ((uint64_t)j * V + t)] + I * nt[2 * ((uint64_t)j * V + t) + 1];
    cplx pX = 0, pY = 0, pZ = 0;
    pX += A[0] * mvx;
    pX += A[1] * mvy;
    pY += A[1] * mvx;
    pX += A[2] * mvz;
    pZ += A[2] * mvx;
    pY += A[3] * mvy;
    pY += A[4] * mvz;
    pZ += A[4] * mvy;
    pZ += A[5] * mvz;
    pX *= one_by_4pi;
    pY *= one_by_4pi;
    pZ *= one_by_4pi;
    if (d->reference_frame == 0) {
      cplx tX = R[0] * pX + R[3] * pY + R[6] * pZ;
      cplx tY = R[1] * pX + R[4] * pY + R[7] * pZ;
      cplx tZ = R[2] * pX + R[5] * pY + R[8] * pZ;
      pX = tX;
      pY = tY;
      pZ = tZ;
    }
    px[2 * t] = creal(pX);
    px[2 * t + 1] = cimag(pX);
    py[2 * t] = creal(pY);
    py[2 * t + 1] = cimag(pY);
    pz[2 * t] = creal(pZ);
    pz[2 * t + 1] = cimag(pZ);
  }
}

/* ------------------------------------------------------------------------- */
/* FFT: forward, unnormalised, complex-to-complex.  The reference calls cuFFT
 * (cufftPlan3d(Nz,Ny,Nx) + cufftExecC2C/Z2Z FORWARD, include/cudaMain.h:184-190,
 * src/cudaMain.cu:480); the contract is the DFT
 *   F[k] = sum_n x[n] exp(-2 pi i n k / N)   along each axis.
 * Implemented here as a mixed-radix Stockham autosort FFT in double precision
 * with exact-angle twiddles (cos/sin of 2 pi m / N).                          */
/* ------------------------------------------------------------------------- */

typedef struct {
  uint32_t n;
  int nf;
  uint32_t fac[64];
  cplx *tw; /* tw[m] = exp(-2 pi i m / n) */
} fft_plan;

static void plan_init(fft_plan *p, uint32_t n) {
  p->n = n;
  p->nf = 0;
  uint32_t r = n;
  while (r % 4 == 0) { p->fac[p->nf++] = 4; r /= 4; }
  while (r % 2 == 0) { p->fac[p->nf++] = 2; r /= 2; }
  for (uint32_t f = 3; r > 1; f += 2) {
    while (r % f == 0) { p->fac[p->nf++] = f; r /= f; }
  }
  p->tw = (cplx *)malloc(sizeof(cplx) * (n ? n : 1));
  for (uint32_t m = 0; m < n; m++) {
    double a = 2 * M_PI * (double)m / (double)n;
    p->tw[m] = cos(a) - I * sin(a);
  }
}

static void plan_free(fft_plan *p) { free(p->tw); }

static void fft_exec(const fft_plan *p, cplx *x, cplx *work) {
  const uint32_t n = p->n;
  if (n <= 1) return;
  cplx *src = x, *dst = work;
  uint32_t Ls = 1;
  cplx a[64], b[64];
  for (int s = 0; s < p->nf; s++) {
    const uint32_t R = p->fac[s];
    const uint32_t Ns = Ls * R, m = n / R, tstep = n / Ns, rstep = n / R;
    cplx *big_a = NULL, *big_b = NULL;
    cplx *pa = a, *pb = b;
    if (R > 64) {
      big_a = (cplx *)malloc(sizeof(cplx) * R);
      big_b = (cplx *)malloc(sizeof(cplx) * R);
      pa = big_a;
      pb = big_b;
    }
    for (uint32_t j = 0; j < m; j++) {
      const uint32_t k = j % Ls;
      for (uint32_t r = 0; r < R; r++) {
        cplx v = src[j + r * m];
        if (k) v *= p->tw[(uint64_t)r * k * tstep];
        pa[r] = v;
      }
      if (R == 2) {
        pb[0] = pa[0] + pa[1];
        pb[1] = pa[0] - pa[1];
      } else if (R == 4) {
        cplx t0 = pa[0] + pa[2], t1 = pa[0] - pa[2];
        cplx t2 = pa[1] + pa[3], t3 = (pa[1] - pa[3]) * (-I);
        pb[0] = t0 + t2;
        pb[1] = t1 + t3;
        pb[2] = t0 - t2;
        pb[3] = t1 - t3;
      } else {
        for (uint32_t q = 0; q < R; q++) {
          cplx acc = 0;
          for (uint32_t r = 0; r < R; r++) acc += pa[r] * p->tw[(uint64_t)((r * q) % R) * rstep];
          pb[q] = acc;
        }
      }
      const uint32_t base = (j / Ls) * Ns + k;
      for (uint32_t q = 0; q < R; q++) dst[base + q * Ls] = pb[q];
    }
    free(big_a);
    free(big_b);
    cplx *t = src;
    src = dst;
    dst = t;
    Ls = Ns;
  }
  if (src != x) memcpy(x, src, sizeof(cplx) * n);
}

void orc_fft1d(double *data, uint32_t n) {
  fft_plan p;
  plan_init(&p, n);
  cplx *w = (cplx *)malloc(sizeof(cplx) * (n ? n : 1));
  fft_exec(&p, (cplx *)data, w);
  free(w);
  plan_free(&p);
}

void orc_dft1d_direct(const double *in, double *out, uint32_t n) {
  const cplx *x = (const cplx *)in;
  cplx *y = (cplx *)out;
  for (uint32_t k = 0; k < n; k++) {
    cplx acc = 0;
    for (uint32_t t = 0; t < n; t++) {
      double a = 2 * M_PI * (double)(((uint64_t)t * k) % n) / (double)n;
      acc += x[t] * (cos(a) - I * sin(a));
    }
    y[k] = acc;
  }
}

/* One axis of the 3D transform: pencils of length n, element stride es,
 * `count` pencils whose bases are produced by base(i). */
static void fft_axis(cplx *data, uint32_t nx, uint32_t ny, uint32_t nz, int axis) {
  const uint32_t n = axis == 0 ? nx : axis == 1 ? ny : nz;
  if (n <= 1) return;
  fft_plan p;
  plan_init(&p, n);
  const uint64_t es = axis == 0 ? 1 : axis == 1 ? nx : (uint64_t)nx * ny;
  const uint64_t count = axis == 0 ? (uint64_t)ny * nz : axis == 1 ? (uint64_t)nx * nz : (uint64_t)nx * ny;
#pragma omp parallel
  {
    cplx *buf = (cplx *)malloc(sizeof(cplx) * n);
    cplx *work = (cplx *)malloc(sizeof(cplx) * n);
#pragma omp for schedule(static)
    for (int64_t i = 0; i < (int64_t)count; i++) {
      uint64_t base;
      if (axis == 0) base = (uint64_t)i * nx;
      else if (axis == 1) base = (uint64_t)(i % nx) + (uint64_t)(i / nx) * nx * ny;
      else base = (uint64_t)i;
      for (uint32_t t = 0; t < n; t++) buf[t] = data[base + t * es];
      fft_exec(&p, buf, work);
      for (uint32_t t = 0; t < n; t++) data[base + t * es] = buf[t];
    }
    free(buf);
    free(work);
  }
  plan_free(&p);
}

void orc_fft3d(double *data, uint32_t nx, uint32_t ny, uint32_t nz) {
  fft_axis((cplx *)data, nx, ny, nz, 0);
  fft_axis((cplx *)data, nx, ny, nz, 1);
  fft_axis((cplx *)data, nx, ny, nz, 2);
}

/* replaceDCComponentWithAverage (src/cudaMain.cu:123-176): F[0] <- mean of the
 * 4 in-plane periodic neighbours, plus the 2 z neighbours when Nz > 1. */
void orc_dc_replace(double *data, uint32_t nx, uint32_t ny, uint32_t nz) {
  cplx *F = (cplx *)data;
  cplx sum = 0;
  int count = 0;
  sum += F[1]; count++;
  sum += F[nx]; count++;
  sum += F[nx - 1]; count++;
  sum += F[(uint64_t)(ny - 1) * nx]; count++;
  if (nz > 1) {
    sum += F[(uint64_t)nx * ny]; count++;
    sum += F[(uint64_t)(nz - 1) * nx * ny]; count++;
  }
  F[0] = creal(sum) / (double)count + I * (cimag(sum) / (double)count);
}

/* FFTIgor: in-place involutive shift out[X,Y,Z] = F[(N/2 - X) mod N, ...]
 * performed by swapping each pair once (include/uniaxial.h:570-613). */
void orc_igor_shift(double *data, uint32_t nx, uint32_t ny, uint32_t nz) {
  cplx *F = (cplx *)data;
  const uint64_t V = (uint64_t)nx * ny * nz;
  const uint32_t mx = nx / 2, my = ny / 2, mz = nz / 2;
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < (int64_t)V; t++) {
    uint32_t Z = (uint32_t)(t / ((uint64_t)nx * ny));
    uint32_t Y = (uint32_t)((t - (uint64_t)Z * nx * ny) / nx);
    uint32_t X = (uint32_t)(t - (uint64_t)Y * nx - (uint64_t)Z * nx * ny);
    uint32_t ix = X <= mx ? mx - X : nx + mx - X;
    uint32_t iy = Y <= my ? my - Y : ny + my - Y;
    uint32_t iz = Z <= mz ? mz - Z : nz + mz - Z;
    uint64_t c = (uint64_t)ix + (uint64_t)iy * nx + (uint64_t)iz * nx * ny;
    if (c > (uint64_t)t) {
      cplx tmp = F[c];
      F[c] = F[t];
      F[t] = tmp;
    }
  }
}

/* ------------------------------------------------------------------------- */
/* Scatter magnitude and Ewald projection                                     */
/* ------------------------------------------------------------------------- */

/* computeMagVec1TimesVec1TTimesVec2: ||(k^2 I - v v^T) p||^2  (include/cudaUtils.h:268-296) */
static double mag_vvp(const double v[3], const cplx p[3], double k) {
  const double d = k * k;
  const double a = v[0], b = v[1], c = v[2];
  cplx t0 = -a * a * p[0] + d * p[0] - a * (b * p[1] + c * p[2]);
  cplx t1 = -a * b * p[0] - b * b * p[1] + d * p[1] - b * c * p[2];
  cplx t2 = -a * c * p[0] - b * c * p[1] + (-c * c + d) * p[2];
  return creal(t0) * creal(t0) + cimag(t0) * cimag(t0) + creal(t1) * creal(t1) +
         cimag(t1) * cimag(t1) + creal(t2) * creal(t2) + cimag(t2) * cimag(t2);
}

typedef struct {
  double start, dx, dy, dz;
  int enable2d;
} qgrid;

static qgrid make_qgrid(const orc_desc *d) {
  qgrid g;
  g.start = -(M_PI / d->phys_size);
  g.dx = (2.0 * M_PI / d->phys_size) / ((d->nx - 1) * 1.0);
  g.dy = (2.0 * M_PI / d->phys_size) / ((d->ny - 1) * 1.0);
  g.enable2d = d->nz == 1;
  g.dz = g.enable2d ? 0 : (2.0 * M_PI / d->phys_size) / ((d->nz - 1) * 1.0);
  return g;
}

/* computeScatter3D (device function, include/uniaxial.h:630-659): I at voxel id
 * of the shifted volume, q from the voxel's grid coordinates. */
static double scatter_at(const orc_desc *d, const qgrid *g, const cplx *px, const cplx *py,
                         const cplx *pz, double k, const double kv[3], uint64_t id) {
  uint64_t nxy = (uint64_t)d->nx * d->ny;
  uint32_t Z = (uint32_t)(id / nxy);
  uint32_t Y = (uint32_t)((id - Z * nxy) / d->nx);
  uint32_t X = (uint32_t)(id - (uint64_t)Y * d->nx - Z * nxy);
  double q[3];
  q[0] = g->start + X * g->dx;
  q[1] = g->start + Y * g->dy;
  q[2] = g->enable2d ? 0 : g->start + Z * g->dz;
  double v[3] = {k * kv[0] + q[0], k * kv[1] + q[1], k * kv[2] + q[2]};
  cplx p[3] = {px[id], py[id], pz[id]};
  return mag_vvp(v, p, k);
}

/* Ewald-sphere geometry shared by the partial and full projections
 * (include/uniaxial.h:821-947).  Returns 0 for a NaN pixel, 1 for a single
 * gather at z0 (nearest neighbour / 2D), 2 for linear interpolation between
 * z0 and z0+1 with weight f on z0+1.
 * Out-of-range z indices (which the reference would read out of bounds; never
 * reached for k along z) are NaN here. */
static int ewald_geometry(const orc_desc *d, const qgrid *g, double k, const double kv[3],
                          uint32_t X, uint32_t Y, uint32_t *z0, double *f) {
  double posx = (g->start + X * g->dx);
  double posy = (g->start + Y * g->dy);
  double kx = k * kv[0], ky = k * kv[1], kz = k * kv[2];
  double val = k * k - (kx + posx) * (kx + posx) - (ky + posy) * (ky + posy);
  if ((val < 0) || (X == (d->nx - 1)) || (Y == (d->ny - 1))) return 0;
  double posz = -kz + sqrt(val);
  if (d->interpolation == 0) {
    if (g->enable2d) { *z0 = 0; return 1; }
    double r = round((posz - g->start) / (g->dz));
    if (r < 0 || r > (double)(d->nz - 1)) return 0;
    *z0 = (uint32_t)r;
    return 1;
  }
  if (g->enable2d) { *z0 = 0; return 1; }
  double t = ((posz - g->start) / (g->dz));
  if (!(t > -1.0)) return 0;
  uint32_t Z = (uint32_t)t;
  if (Z + 1 > d->nz - 1) return 0;
  *z0 = Z;
  *f = (posz - (g->start + Z * g->dz)) / g->dz;
  return 2;
}

/* computeEwaldProjectionGPU, PARTIAL approach from the shifted p~ (include/uniaxial.h:876-947) */
void orc_ewald_partial(const orc_desc *d, const double *px, const double *py, const double *pz,
                       double k, const double kvec[3], double *proj) {
  const qgrid g = make_qgrid(d);
  const cplx *cx = (const cplx *)px, *cy = (const cplx *)py, *cz = (const cplx *)pz;
  const uint64_t nxy = (uint64_t)d->nx * d->ny;
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < (int64_t)nxy; t++) {
    uint32_t Y = (uint32_t)(t / d->nx), X = (uint32_t)(t - (uint64_t)Y * d->nx);
    uint32_t z0 = 0;
    double f = 0;
    int kind = ewald_geometry(d, &g, k, kvec, X, Y, &z0, &f);
    if (kind == 0) {
      proj[t] = NAN;
    } else if (kind == 1) {
      proj[t] = scatter_at(d, &g, cx, cy, cz, k, kvec, (uint64_t)z0 * nxy + t);
    } else {
      double d1 = scatter_at(d, &g, cx, cy, cz, k, kvec, (uint64_t)z0 * nxy + t);
      double d2 = scatter_at(d, &g, cx, cy, cz, k, kvec, (uint64_t)(z0 + 1) * nxy + t);
      proj[t] = ((1 - f) * d1 + (f)*d2);
    }
  }
}

/* computeScatter3D kernel, FULL approach (include/uniaxial.h:676-722) */
void orc_scatter3d(const orc_desc *d, const double *px, const double *py, const double *pz,
                   double k, const double kvec[3], double *scatter) {
  const qgrid g = make_qgrid(d);
  const uint64_t V = (uint64_t)d->nx * d->ny * d->nz;
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < (int64_t)V; t++)
    scatter[t] = scatter_at(d, &g, (const cplx *)px, (const cplx *)py, (const cplx *)pz, k, kvec, t);
}

/* computeEwaldProjectionGPU from the stored scatter volume (include/uniaxial.h:821-874) */
void orc_ewald_full(const orc_desc *d, const double *scatter, double k, const double kvec[3],
                    double *proj) {
  const qgrid g = make_qgrid(d);
  const uint64_t nxy = (uint64_t)d->nx * d->ny;
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < (int64_t)nxy; t++) {
    uint32_t Y = (uint32_t)(t / d->nx), X = (uint32_t)(t - (uint64_t)Y * d->nx);
    uint32_t z0 = 0;
    double f = 0;
    int kind = ewald_geometry(d, &g, k, kvec, X, Y, &z0, &f);
    if (kind == 0) proj[t] = NAN;
    else if (kind == 1) proj[t] = scatter[(uint64_t)z0 * nxy + t];
    else proj[t] = ((1 - f) * scatter[(uint64_t)z0 * nxy + t] + (f)*scatter[(uint64_t)(z0 + 1) * nxy + t]);
  }
}

/* ------------------------------------------------------------------------- */
/* Bilinear affine warp (stands in for nppiWarpAffine_*_C1R, NPPI_INTER_LINEAR,
 * src/cudaMain.cu:91-121,858-875).  NPP is closed; the convention is the
 * build's own (parity unpinned against NPP):
 *   coeffs c map source -> destination: x' = c0 x + c1 y + c2, y' = c3 x + c4 y + c5
 *   inverse: det = c0 c4 - c1 c3, (i00,i01,i10,i11) = (c4,-c1,-c3,c0)/det,
 *            b0 = -(i00 c2 + i01 c5), b1 = -(i10 c2 + i11 c5)
 *   destination pixel (X,Y) samples sx = i00 X + i01 Y + b0, sy = i10 X + i11 Y + b1
 *   it is written iff 0 <= sx <= nx-1 and 0 <= sy <= ny-1 (otherwise untouched)
 *   value = sum over the 4 taps of w * src with w the bilinear weights;
 *   taps whose weight is exactly 0 are skipped (so 0 deg is an exact copy and
 *   no tap outside the image is ever read).                                   */
/* ------------------------------------------------------------------------- */

void orc_warp_affine(const double *src, double *dst, uint32_t nx, uint32_t ny, const double c[6]) {
  const double det = c[0] * c[4] - c[1] * c[3];
  const double i00 = c[4] / det, i01 = -c[1] / det, i10 = -c[3] / det, i11 = c[0] / det;
  const double b0 = -(i00 * c[2] + i01 * c[5]);
  const double b1 = -(i10 * c[2] + i11 * c[5]);
  const uint64_t nxy = (uint64_t)nx * ny;
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < (int64_t)nxy; t++) {
    uint32_t Y = (uint32_t)(t / nx), X = (uint32_t)(t - (uint64_t)Y * nx);
    double sx = i00 * X + i01 * Y + b0;
    double sy = i10 * X + i11 * Y + b1;
    if (!(sx >= 0 && sx <= (double)(nx - 1) && sy >= 0 && sy <= (double)(ny - 1))) continue;
    double fx0 = floor(sx), fy0 = floor(sy);
    uint32_t x0 = (uint32_t)fx0, y0 = (uint32_t)fy0;
    double fx = sx - fx0, fy = sy - fy0;
    double w00 = (1 - fx) * (1 - fy), w10 = fx * (1 - fy), w01 = (1 - fx) * fy, w11 = fx * fy;
    double v = 0;
    if (w00 != 0) v += w00 * src[(uint64_t)y0 * nx + x0];
    if (w10 != 0) v += w10 * src[(uint64_t)y0 * nx + x0 + 1];
    if (w01 != 0) v += w01 * src[(uint64_t)(y0 + 1) * nx + x0];
    if (w11 != 0) v += w11 * src[(uint64_t)(y0 + 1) * nx + x0 + 1];
    dst[t] = v;
  }
}

/* ------------------------------------------------------------------------- */
/* Whole pipeline (src/cudaMain.cu:377-1035 CommunicationMinimizing and
 * :1037-1667 MemoryMinimizing): energy -> kVector -> E-angle loop.            */
/* ------------------------------------------------------------------------- */

int orc_simulate_subset(const orc_desc *d, const double *voxels, const double *mats, double *out,
                        uint32_t e0, uint32_t e1, uint32_t a0, uint32_t a1) {
  const uint64_t V = (uint64_t)d->nx * d->ny * d->nz;
  const uint64_t P = (uint64_t)d->nx * d->ny;
  const uint32_t M = d->num_materials, nK = d->num_kvectors;
  const uint32_t nA_total = orc_num_angles(d->angle_start, d->angle_increment, d->angle_end);
  if (a1 > nA_total) a1 = nA_total;
  if (e1 > d->num_energies) e1 = d->num_energies;
  const uint32_t nA = a1 > a0 ? a1 - a0 : 0;

  double *px = (double *)malloc(sizeof(double) * 2 * V);
  double *py = (double *)malloc(sizeof(double) * 2 * V);
  double *pz = (double *)malloc(sizeof(double) * 2 * V);
  double *nt = d->algorithm == 1 ? (double *)malloc(sizeof(double) * 12 * V) : NULL;
  double *scatter = d->scatter_approach == 1 ? (double *)malloc(sizeof(double) * V) : NULL;
  double *proj = (double *)malloc(sizeof(double) * P);
  double *rot = (double *)malloc(sizeof(double) * P);
  double *avg = (double *)malloc(sizeof(double) * P);
  uint32_t *mask = (uint32_t *)malloc(sizeof(uint32_t) * P);
  double *RK = (double *)malloc(sizeof(double) * 9 * (nK ? nK : 1));
  double *base = (double *)malloc(sizeof(double) * (nK ? nK : 1));
  if (!px || !py || !pz || (d->algorithm == 1 && !nt) || (d->scatter_approach == 1 && !scatter) ||
      !proj || !rot || !avg || !mask || !RK || !base) {
    free(px); free(py); free(pz); free(nt); free(scatter); free(proj); free(rot); free(avg);
    free(mask); free(RK); free(base);
    return 1;
  }
  /* RotationMatrix::initComputation (src/RotationMatrix.cpp:30-51) */
  for (uint32_t kk = 0; kk < nK; kk++) {
    double tmp[9];
    orc_rotation_matrix_k(d->kvectors + 3 * kk, RK + 9 * kk);
    orc_base_configuration(d->kvectors + 3 * kk, RK + 9 * kk, tmp, base + kk);
  }
  double Dm[9];
  orc_rotation_matrix_k(d->detector, Dm);

  for (uint32_t e = e0; e < e1; e++) {
    const double *mats_e = mats + 4 * (uint64_t)e * M;
    const double energy = d->energies[e];
    if (d->algorithm == 1) orc_compute_nt(d, voxels, mats_e, nt);
    for (uint32_t kk = 0; kk < nK; kk++) {
      const double *kv = d->kvectors + 3 * kk;
      const double *rk = RK + 9 * kk;
      for (uint64_t t = 0; t < P; t++) { avg[t] = 0; mask[t] = 0; }
      const double wavelength = 1239.84197 / energy;
      const double kmag = 2 * M_PI / wavelength;
      for (uint32_t i = a0; i < a1; i++) {
        const double eangle = (base[kk] + d->angle_start + i * d->angle_increment) * M_PI / 180.0;
        double R[9];
        orc_rotation_matrix_e(kv, rk, eangle, R);
        if (d->algorithm == 1) orc_polarization_nt(d, nt, R, px, py, pz);
        else orc_polarization(d, voxels, mats_e, R, px, py, pz);
        orc_fft3d(px, d->nx, d->ny, d->nz);
        orc_fft3d(py, d->nx, d->ny, d->nz);
        orc_fft3d(pz, d->nx, d->ny, d->nz);
        orc_dc_replace(px, d->nx, d->ny, d->nz);
        orc_dc_replace(py, d->nx, d->ny, d->nz);
        orc_dc_replace(pz, d->nx, d->ny, d->nz);
        orc_igor_shift(px, d->nx, d->ny, d->nz);
        orc_igor_shift(py, d->nx, d->ny, d->nz);
        orc_igor_shift(pz, d->nx, d->ny, d->nz);
        if (d->scatter_approach == 1) {
          orc_scatter3d(d, px, py, pz, kmag, kv, scatter);
          orc_ewald_full(d, scatter, kmag, kv, proj);
        } else {
          orc_ewald_partial(d, px, py, pz, kmag, kv, proj);
        }
        for (uint64_t t = 0; t < P; t++) rot[t] = NAN; /* cublasScale(NaN) (src/cudaMain.cu:830-839) */
        double c[6];
        orc_rotation_coeffs(eangle, d->nx, d->ny, c);
        orc_warp_affine(proj, rot, d->nx, d->ny, c);
        if (d->rot_mask) { /* computeRotationMask (include/uniaxial.h:958-970) */
          for (uint64_t t = 0; t < P; t++) {
            if (isnan(rot[t])) rot[t] = 0.0;
            else mask[t]++;
          }
        }
        for (uint64_t t = 0; t < P; t++) avg[t] += rot[t]; /* cublasSaxpy (src/cudaMain.cu:882-887) */
      }
      if (d->rot_mask) { /* averageRotation (include/uniaxial.h:979-992) */
        for (uint64_t t = 0; t < P; t++)
          avg[t] = mask[t] == 0 ? 0 : avg[t] * (1.0 / (mask[t] * 1.0));
      } else {
        const double fac = 1.0 / (nA ? nA : 1);
        for (uint64_t t = 0; t < P; t++) avg[t] *= fac;
      }
      /* detector warp (src/cudaMain.cu:916-960) */
      {
        const double srcp[6] = {d->nx / 2., d->ny / 2., d->nx * 0.5, d->ny * 1.0, d->nx * 1.0, d->ny * 0.5};
        const double center[2] = {d->nx / 2., d->ny / 2.};
        double rotm[9], dstp[6];
        matmul3(Dm, rk, rotm);
        for (int q = 0; q < 3; q++) {
          double s[3] = {srcp[2 * q] - center[0], srcp[2 * q + 1] - center[1], 0}, o[3];
          matvec3(rotm, s, o);
          dstp[2 * q] = o[0] + center[0];
          dstp[2 * q + 1] = o[1] + center[1];
        }
        double c[6];
        orc_warp_affine_matrix(srcp, dstp, c);
        memcpy(proj, avg, sizeof(double) * P);
        for (uint64_t t = 0; t < P; t++) avg[t] = d->rot_mask ? 0.0 * avg[t] : NAN;
        orc_warp_affine(proj, avg, d->nx, d->ny, c);
      }
      memcpy(out + ((uint64_t)(e - e0) * nK + kk) * P, avg, sizeof(double) * P);
    }
  }
  free(px); free(py); free(pz); free(nt); free(scatter); free(proj); free(rot); free(avg);
  free(mask); free(RK); free(base);
  return 0;
}

int orc_simulate(const orc_desc *d, const double *voxels, const double *mats, double *out) {
  return orc_simulate_subset(d, voxels, mats, out, 0, d->num_energies, 0,
                             orc_num_angles(d->angle_start, d->angle_increment, d->angle_end));
}

int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}
