// FP64 device geometry shared by the Ewald and rotation kernels.  Contraction
// is disabled so every expression is evaluated exactly as written, matching the
// host-side definition bit for bit (masks and nearest-neighbour z indices are
// decided here).
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace cyr {

// Correctly rounded double square root: the hardware/library result is checked
// against its neighbours with exact FMA residuals.
__device__ inline double sqrt_rn(double x) {
#pragma clang fp contract(off)
  double r = sqrt(x);
  if (!(x > 0) || !isfinite(x)) return r;
  double best = r, berr = fabs(fma(-r, r, x));
  const double cand[2] = {nextafter(r, 0.0), nextafter(r, INFINITY)};
  for (int i = 0; i < 2; i++) {
    const double c = cand[i];
    const double e = fabs(fma(-c, c, x));
    if (e < berr) {
      berr = e;
      best = c;
    }
  }
  return best;
}

// Ewald-sphere geometry of pixel (X, Y) (include/uniaxial.h:821-947).
// Returns 0 (NaN pixel), 1 (single z bin z0) or 2 (linear between z0 and
// z0+1 with weight f on z0+1).  Out-of-range z, which the reference reads out
// of bounds (unreachable for k along z), is a NaN pixel.
__device__ inline int ewald_geometry(const EwaldGeom &g, uint32_t X, uint32_t Y, uint32_t *z0, double *f) {
#pragma clang fp contract(off)
  const double posx = (g.start + X * g.dx);
  const double posy = (g.start + Y * g.dy);
  const double kx = g.k * g.kx, ky = g.k * g.ky, kz = g.k * g.kz;
  const double val = g.k * g.k - (kx + posx) * (kx + posx) - (ky + posy) * (ky + posy);
  if ((val < 0) || (X == (g.nx - 1)) || (Y == (g.ny - 1))) return 0;
  const double posz = -kz + sqrt_rn(val);
  if (g.enable2d) {
    *z0 = 0;
    return 1;
  }
  if (g.interpolation == 0) {
    const double r = round((posz - g.start) / (g.dz));
    if (r < 0 || r > (double)(g.nz - 1)) return 0;
    *z0 = (uint32_t)r;
    return 1;
  }
  const double t = ((posz - g.start) / (g.dz));
  if (!(t > -1.0)) return 0;
  const uint32_t Z = (uint32_t)t;
  if (Z + 1 > g.nz - 1) return 0;
  *z0 = Z;
  *f = (posz - (g.start + Z * g.dz)) / g.dz;
  return 2;
}

// q-vector of shifted grid point (X, Y, Z) plus k k^ (include/uniaxial.h:642-653),
// returned as the symmetric matrix k^2 I - v v^T (6 entries: xx xy xz yy yz zz).
__device__ inline void scatter_matrix(const EwaldGeom &g, uint32_t X, uint32_t Y, uint32_t Z, double m[6]) {
#pragma clang fp contract(off)
  const double qx = g.start + X * g.dx;
  const double qy = g.start + Y * g.dy;
  const double qz = g.enable2d ? 0.0 : g.start + Z * g.dz;
  const double vx = g.k * g.kx + qx, vy = g.k * g.ky + qy, vz = g.k * g.kz + qz;
  const double d = g.k * g.k;
  m[0] = d - vx * vx;
  m[1] = -vx * vy;
  m[2] = -vx * vz;
  m[3] = d - vy * vy;
  m[4] = -vy * vz;
  m[5] = d - vz * vz;
}

// || (k^2 I - v v^T) p ||^2 (include/cudaUtils.h:268-296)
template <typename T> __device__ inline T scatter_intensity(const double md[6], const cx<T> p[3]) {
  const T m0 = (T)md[0], m1 = (T)md[1], m2 = (T)md[2], m3 = (T)md[3], m4 = (T)md[4], m5 = (T)md[5];
  const T t0r = m0 * p[0].x + m1 * p[1].x + m2 * p[2].x, t0i = m0 * p[0].y + m1 * p[1].y + m2 * p[2].y;
  const T t1r = m1 * p[0].x + m3 * p[1].x + m4 * p[2].x, t1i = m1 * p[0].y + m3 * p[1].y + m4 * p[2].y;
  const T t2r = m2 * p[0].x + m4 * p[1].x + m5 * p[2].x, t2i = m2 * p[0].y + m4 * p[1].y + m5 * p[2].y;
  return t0r * t0r + t0i * t0i + t1r * t1r + t1i * t1i + t2r * t2r + t2i * t2i;
}

// Igor shift index map (involution): shifted coordinate -> FFT bin.
__device__ inline uint32_t unshift(uint32_t X, uint32_t n) {
  const uint32_t m = n / 2;
  return X <= m ? m - X : n + m - X;
}

// Bilinear sample at the back-mapped point of destination pixel (X, Y) with
// inverse coefficients inv = (i00, i01, b0, i10, i11, b1).  Returns false if the
// point is outside [0, nx-1] x [0, ny-1]; taps with zero weight are skipped.
// Back-mapped source point, evaluated exactly as written: it decides which
// destination pixels are written (bit-identical to the host definition).
__device__ inline void back_map_point(const double inv[6], uint32_t X, uint32_t Y, double &sx, double &sy) {
#pragma clang fp contract(off)
  sx = inv[0] * X + inv[1] * Y + inv[2];
  sy = inv[3] * X + inv[4] * Y + inv[5];
}

template <typename T>
__device__ inline bool bilinear_sample(const T *src, uint32_t nx, uint32_t ny, const double inv[6], uint32_t X,
                                       uint32_t Y, double *out) {
  double sx, sy;
  back_map_point(inv, X, Y, sx, sy);
  const bool inside = sx >= 0 && sx <= (double)(nx - 1) && sy >= 0 && sy <= (double)(ny - 1);
  // outside points sample (0, 0) and are reported invalid: no branch around the loads
  const double fx0 = inside ? floor(sx) : 0.0, fy0 = inside ? floor(sy) : 0.0;
  const uint32_t x0 = (uint32_t)fx0, y0 = (uint32_t)fy0;
  const double fx = inside ? sx - fx0 : 0.0, fy = inside ? sy - fy0 : 0.0;
  const double w00 = (1 - fx) * (1 - fy), w10 = fx * (1 - fy), w01 = (1 - fx) * fy, w11 = fx * fy;
  // the four taps are loaded unconditionally (clamped into the image, so several
  // gathers can be in flight); a tap of weight 0 is then skipped, not multiplied
  const uint32_t x1 = x0 + 1 < nx ? x0 + 1 : x0, y1 = y0 + 1 < ny ? y0 + 1 : y0;
  const double a00 = (double)src[(uint64_t)y0 * nx + x0], a10 = (double)src[(uint64_t)y0 * nx + x1];
  const double a01 = (double)src[(uint64_t)y1 * nx + x0], a11 = (double)src[(uint64_t)y1 * nx + x1];
  double v = 0;
  v = w00 != 0 ? v + w00 * a00 : v;
  v = w10 != 0 ? v + w10 * a10 : v;
  v = w01 != 0 ? v + w01 * a01 : v;
  v = w11 != 0 ? v + w11 * a11 : v;
  *out = v;
  return inside;
}

}  // namespace cyr
