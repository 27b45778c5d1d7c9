// FP64 host geometry (see host_math.hpp).  Compiled with -ffp-contract=off.
#include "host_math.hpp"

#include <array>
#include <cmath>
#include <map>
#include <mutex>
#include <cstring>
#include <limits>
#include <string>

#include "common.hpp"

namespace cyr {

static bool fequals(double x, double y) { return std::fabs(x - y) < 1E-6; }  // include/Datatypes.h:143
static double dot(const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static V3 cross(const V3 &a, const V3 &b) {  // include/Rotation.h:188-194
  return V3{a.y * b.z - a.z * b.y, -(a.x * b.z - a.z * b.x), a.x * b.y - a.y * b.x};
}

M3 matmul(const M3 &A, const M3 &B) {
  M3 C;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += A.m[i * 3 + k] * B.m[k * 3 + j];
      C.m[i * 3 + j] = s;
    }
  return C;
}

V3 matvec(const M3 &A, const V3 &v) {
  return V3{A.m[0] * v.x + A.m[1] * v.y + A.m[2] * v.z, A.m[3] * v.x + A.m[4] * v.y + A.m[5] * v.z,
            A.m[6] * v.x + A.m[7] * v.y + A.m[8] * v.z};
}

// include/Rotation.h:233-252
static M3 inverse(const M3 &A) {
  const double *a = A.m;
  double det = a[0] * (a[4] * a[8] - a[7] * a[5]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
               a[2] * (a[3] * a[7] - a[4] * a[6]);
  double inv = 1. / det;
  M3 r;
  r.m[0] = (a[4] * a[8] - a[7] * a[5]) * inv;
  r.m[1] = (a[2] * a[7] - a[1] * a[8]) * inv;
  r.m[2] = (a[1] * a[5] - a[2] * a[4]) * inv;
  r.m[3] = (a[5] * a[6] - a[3] * a[8]) * inv;
  r.m[4] = (a[0] * a[8] - a[2] * a[6]) * inv;
  r.m[5] = (a[3] * a[2] - a[0] * a[5]) * inv;
  r.m[6] = (a[3] * a[7] - a[6] * a[4]) * inv;
  r.m[7] = (a[6] * a[1] - a[0] * a[7]) * inv;
  r.m[8] = (a[0] * a[4] - a[3] * a[1]) * inv;
  return r;
}

M3 rotation_vec_to_vec(const V3 &a, const V3 &b) {
  M3 R{};
  if (fequals(a.x, b.x) && fequals(a.y, b.y) && fequals(a.z, b.z)) {
    R.m[0] = R.m[4] = R.m[8] = 1;
    return R;
  }
  const double dp = dot(a, b);
  const V3 cr = cross(a, b);
  const double nc = std::sqrt(dot(cr, cr));
  M3 G{{dp, -nc, 0, nc, dp, 0, 0, 0, 1}};
  const V3 sub{b.x - dp * a.x, b.y - dp * a.y, b.z - dp * a.z};
  const double ns = std::sqrt(dot(sub, sub));
  M3 F{{a.x, sub.x / ns, -cr.x, a.y, sub.y / ns, -cr.y, a.z, sub.z / ns, -cr.z}};
  return matmul(F, matmul(G, inverse(F)));
}

V3 rodrigues(const V3 &v, const V3 &axis, double angle) {
  const double c = std::cos(angle), s = std::sin(angle);
  const V3 cr = cross(axis, v);
  const double f = (1 - c) * dot(axis, v);
  return V3{c * v.x + s * cr.x + axis.x * f, c * v.y + s * cr.y + axis.y * f,
            c * v.z + s * cr.z + axis.z * f};
}

M3 rotation_matrix_k(const V3 &k) { return rotation_vec_to_vec(V3{0, 0, 1}, k); }

namespace {
// The search below is a pure function of (k, RK) and costs ~1 ms of host time;
// every simulation with the same k vector reuses its result.
std::mutex g_base_mu;
std::map<std::array<double, 12>, double> g_base_search;
}  // namespace

double base_configuration(const V3 &k, const M3 &RK, M3 *R) {
  std::array<double, 12> key{k.x, k.y, k.z};
  for (int i = 0; i < 9; i++) key[3 + i] = RK.m[i];
  const V3 sx = matvec(RK, V3{1, 0, 0});
  double ang = 0;
  bool cached = false;
  {
    std::lock_guard<std::mutex> lk(g_base_mu);
    auto it = g_base_search.find(key);
    if (it != g_base_search.end()) {
      ang = it->second;
      cached = true;
    }
  }
  if (!cached) {
    const unsigned num_interval = 100000;
    const double dtheta = M_PI / (num_interval * 1.0);
    double maxdiff = std::numeric_limits<double>::infinity();
    for (unsigned i = 0; i < num_interval; i++) {
      const V3 rx = rodrigues(sx, k, i * dtheta);
      const double diff = std::fabs(rx.y);
      if (maxdiff > diff) {
        maxdiff = diff;
        ang = i * dtheta;
      }
    }
    std::lock_guard<std::mutex> lk(g_base_mu);
    g_base_search[key] = ang;
  }
  V3 rx = rodrigues(sx, k, ang);
  ang = rx.x > 0 ? ang : ang + M_PI;
  rx = rodrigues(sx, k, ang);
  if (R) *R = matmul(rotation_vec_to_vec(sx, rx), RK);
  return ang;
}

M3 rotation_matrix_e(const V3 &k, const M3 &RK, double angle) {
  const V3 sx = matvec(RK, V3{1, 0, 0});
  const V3 rx = rodrigues(sx, k, angle);
  return matmul(rotation_vec_to_vec(sx, rx), RK);
}

void warp_affine_matrix(const double src[6], const double dst[6], double c[6]) {
  const double x0 = src[0], y0 = src[1], x1 = src[2], y1 = src[3], x2 = src[4], y2 = src[5];
  const double u0 = dst[0], v0 = dst[1], u1 = dst[2], v1 = dst[3], u2 = dst[4], v2 = dst[5];
  const double denom = x0 * y1 - x1 * y0 - x0 * y2 + x2 * y0 + x1 * y2 - x2 * y1;
  c[0] = (u0 * y1 - u1 * y0 - u0 * y2 + u2 * y0 + u1 * y2 - u2 * y1) / (denom);
  c[1] = -(u0 * x1 - u1 * x0 - u0 * x2 + u2 * x0 + u1 * x2 - u2 * x1) / (denom);
  c[2] = (u0 * x1 * y2 - u0 * x2 * y1 - u1 * x0 * y2 + u1 * x2 * y0 + u2 * x0 * y1 - u2 * x1 * y0) / (denom);
  c[3] = (v0 * y1 - v1 * y0 - v0 * y2 + v2 * y0 + v1 * y2 - v2 * y1) / (denom);
  c[4] = -(v0 * x1 - v1 * x0 - v0 * x2 + v2 * x0 + v1 * x2 - v2 * x1) / (denom);
  c[5] = (v0 * x1 * y2 - v0 * x2 * y1 - v1 * x0 * y2 + v1 * x2 * y0 + v2 * x0 * y1 - v2 * x1 * y0) / (denom);
}

void affine_inverse(const double c[6], double inv[6]) {
  const double det = c[0] * c[4] - c[1] * c[3];
  const double i00 = c[4] / det, i01 = -c[1] / det, i10 = -c[3] / det, i11 = c[0] / det;
  inv[0] = i00;
  inv[1] = i01;
  inv[2] = -(i00 * c[2] + i01 * c[5]);
  inv[3] = i10;
  inv[4] = i11;
  inv[5] = -(i10 * c[2] + i11 * c[5]);
}

void rotation_coeffs(double angle, uint32_t nx, uint32_t ny, double c[6]) {
  const double alpha = std::cos(angle), beta = std::sin(angle);
  c[0] = alpha;
  c[1] = beta;
  c[2] = ((1 - alpha) * nx / 2 - beta * ny / 2.);
  c[3] = -beta;
  c[4] = alpha;
  c[5] = (beta * nx / 2. + (1 - alpha) * ny / 2.);
}

uint32_t num_angles(double start, double inc, double end) {
  return (uint32_t)std::round((end - start) / inc + 1);
}

cyr_status validate_desc(const cyr_sim_desc *d) {
  if (!d) return fail(CYR_E_INVALID, "null descriptor");
  if (d->dims[0] < 2 || d->dims[1] < 2 || d->dims[2] < 1)
    return fail(CYR_E_INVALID, "dims must satisfy Nx >= 2, Ny >= 2, Nz >= 1");
  if ((uint64_t)d->dims[0] * d->dims[1] * d->dims[2] >= (1ull << 40))
    return fail(CYR_E_INVALID, "grid too large");
  if (d->num_materials < 1) return fail(CYR_E_INVALID, "num_materials must be >= 1");
  if (!(d->phys_size > 0)) return fail(CYR_E_INVALID, "phys_size must be > 0");
  if (d->num_energies < 1 || !d->energies) return fail(CYR_E_INVALID, "no energies");
  for (uint32_t i = 0; i < d->num_energies; i++)
    if (!(d->energies[i] > 0) || !std::isfinite(d->energies[i]))
      return fail(CYR_E_INVALID, "energies must be finite and > 0");
  if (!std::isfinite(d->angle_start) || !std::isfinite(d->angle_end) || !std::isfinite(d->angle_increment))
    return fail(CYR_E_INVALID, "E-angle triple must be finite");
  if (d->angle_start > d->angle_end)  // include/Input/InputData.h:277-279
    return fail(CYR_E_INVALID, "Start EAngle can not be greater than end EAngle");
  if (d->angle_increment == 0 && !fequals(d->angle_start, d->angle_end))
    return fail(CYR_E_INVALID, "Increment angle = 0");
  if (d->case_type < 0 || d->case_type > 2) return fail(CYR_E_INVALID, "case_type out of range");
  if (d->morphology_type < 0 || d->morphology_type > 1) return fail(CYR_E_INVALID, "morphology_type out of range");
  if (d->reference_frame < 0 || d->reference_frame > 1) return fail(CYR_E_INVALID, "reference_frame out of range");
  if (d->interpolation < 0 || d->interpolation > 1) return fail(CYR_E_INVALID, "interpolation out of range");
  if (d->windowing < 0 || d->windowing > 1) return fail(CYR_E_INVALID, "windowing out of range");
  if (d->rot_mask < 0 || d->rot_mask > 1) return fail(CYR_E_INVALID, "rot_mask out of range");
  if (d->scatter_approach < 0 || d->scatter_approach > 1) return fail(CYR_E_INVALID, "scatter_approach out of range");
  if (d->algorithm < 0 || d->algorithm > 1) return fail(CYR_E_INVALID, "algorithm out of range");
  if (d->precision != CYR_FP32 && d->precision != CYR_FP64) return fail(CYR_E_INVALID, "precision out of range");
  if (d->storage < 0 || d->storage > 2) return fail(CYR_E_INVALID, "storage out of range");
  if (d->pipeline < 0 || d->pipeline > 3) return fail(CYR_E_INVALID, "pipeline out of range");
  if (d->case_type != 0 && (d->num_kvectors < 1 || !d->kvectors))
    return fail(CYR_E_INVALID, "case types 1 and 2 need at least one k vector");
  if (d->num_kvectors > 0 && !d->kvectors) return fail(CYR_E_INVALID, "kvectors is null");
  for (uint32_t i = 0; i < d->num_kvectors; i++) {
    const double *k = d->kvectors + 3 * i;
    const double n = std::sqrt(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
    if (!(std::fabs(n - 1.0) < 1e-6)) return fail(CYR_E_INVALID, "k vectors must be normalised");
  }
  double inc = d->angle_increment;
  if (inc == 0) inc = 1.0;  // start == end: include/Input/InputData.h:280-286
  const uint32_t nA = num_angles(d->angle_start, inc, d->angle_end);
  if (nA < 1 || nA > 100000) return fail(CYR_E_INVALID, "number of E angles out of range");
  return CYR_OK;
}

cyr_status build_geometry(const cyr_sim_desc *d, SimGeometry &g) {
  CYR_TRY(validate_desc(d));
  g.nx = d->dims[0];
  g.ny = d->dims[1];
  g.nz = d->dims[2];
  g.M = d->num_materials;
  g.nE = d->num_energies;
  g.V = (uint64_t)g.nx * g.ny * g.nz;
  g.P = (uint64_t)g.nx * g.ny;
  const double inc = d->angle_increment == 0 ? 1.0 : d->angle_increment;
  g.nA = num_angles(d->angle_start, inc, d->angle_end);
  g.quarter = 0;
  if (inc > 0 && inc <= 90.0) {
    const double steps = std::round(90.0 / inc);
    // (a tight bound: the orbit rotation kernel assumes the tables' quarter turns
    // agree to far below its pixel margin of 1e-9 + N 1e-12)
    if (std::fabs(steps * inc - 90.0) < 1e-12 && g.nA % (uint32_t)steps == 0 && g.nA / (uint32_t)steps >= 2)
      g.quarter = (uint32_t)steps;
  }
  std::vector<V3> ks;
  if (d->case_type == 0 && d->num_kvectors == 0) ks.push_back(V3{0, 0, 1});
  for (uint32_t i = 0; i < d->num_kvectors; i++)
    ks.push_back(V3{d->kvectors[3 * i], d->kvectors[3 * i + 1], d->kvectors[3 * i + 2]});
  g.nK = (uint32_t)ks.size();
  g.kvec = ks;
  g.RK.resize(g.nK);
  g.base.resize(g.nK);
  g.e1.resize(g.nK);
  g.e2.resize(g.nK);
  g.R.resize((size_t)g.nK * g.nA);
  g.rot_inv.resize((size_t)g.nK * g.nA * 6);
  g.ebasis.resize((size_t)g.nK * g.nA * 2);
  g.det_inv.resize((size_t)g.nK * 6);
  // RotationMatrix::initComputation (src/RotationMatrix.cpp:30-51)
  const V3 det{d->detector[0], d->detector[1], d->detector[2]};
  g.detector = rotation_matrix_k(d->case_type == 2 ? det : V3{0, 0, 1});
  for (uint32_t k = 0; k < g.nK; k++) {
    g.RK[k] = rotation_matrix_k(ks[k]);
    g.base[k] = base_configuration(ks[k], g.RK[k], nullptr);
    g.e1[k] = matvec(g.RK[k], V3{1, 0, 0});
    g.e2[k] = cross(ks[k], g.e1[k]);
    for (uint32_t a = 0; a < g.nA; a++) {
      // Eangle = (baseRotAngle + start + i*increment) * pi / 180  (src/cudaMain.cu:612)
      const double eangle = (g.base[k] + d->angle_start + a * inc) * M_PI / 180.0;
      const M3 R = rotation_matrix_e(ks[k], g.RK[k], eangle);
      g.R[(size_t)k * g.nA + a] = R;
      double c[6];
      rotation_coeffs(eangle, g.nx, g.ny, c);
      affine_inverse(c, &g.rot_inv[((size_t)k * g.nA + a) * 6]);
      const V3 E{R.m[0], R.m[3], R.m[6]};
      g.ebasis[((size_t)k * g.nA + a) * 2 + 0] = dot(E, g.e1[k]);
      g.ebasis[((size_t)k * g.nA + a) * 2 + 1] = dot(E, g.e2[k]);
    }
    // detector warp through 3 points (src/cudaMain.cu:916-938)
    const double src[6] = {g.nx / 2., g.ny / 2., g.nx * 0.5, g.ny * 1.0, g.nx * 1.0, g.ny * 0.5};
    const double center[2] = {g.nx / 2., g.ny / 2.};
    const M3 rotm = matmul(g.detector, g.RK[k]);
    double dst[6];
    for (int q = 0; q < 3; q++) {
      const V3 o = matvec(rotm, V3{src[2 * q] - center[0], src[2 * q + 1] - center[1], 0});
      dst[2 * q] = o.x + center[0];
      dst[2 * q + 1] = o.y + center[1];
    }
    double c[6];
    warp_affine_matrix(src, dst, c);
    affine_inverse(c, &g.det_inv[(size_t)k * 6]);
  }
  g.kmag.resize(g.nE);
  for (uint32_t e = 0; e < g.nE; e++) {
    const double wavelength = 1239.84197 / d->energies[e];  // src/cudaMain.cu:607-608
    g.kmag[e] = 2 * M_PI / wavelength;
  }
  return CYR_OK;
}

}  // namespace cyr
