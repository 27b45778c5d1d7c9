// E-angle image rotation, rotation mask, angle averaging and detector warp
// (replaces cublasSscal(NaN) + nppiWarpAffine + computeRotationMask +
// cublasSaxpy + averageRotation/cublasSscal + the detector warp:
// src/cudaMain.cu:830-971, include/uniaxial.h:958-992).
//
// Sums over angles are FP64 and run in angle order (deterministic, no atomics).
// Bilinear convention (NPP is closed; the build's own, shared with the oracle):
// destination pixel (X, Y) samples the source at the back-mapped point
// (inverse affine, FP64); it is written iff that point lies in
// [0, Nx-1] x [0, Ny-1]; taps of exactly zero weight are skipped.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <utility>
#include <vector>

#include "geom.hpp"
#include "kernels.hpp"

namespace cyr {

namespace {

struct Inv6 { double c[6]; };

// Pixel of this thread in the rotation kernels: 256-thread blocks cover 16 x 16
// tiles, each wave an 8 x 8 quadrant, so a wave's rotated source footprint is
// compact (a 64-pixel row segment maps to a 64-pixel line of source rows at
// most angles, touching many more cache lines).  Tiles come from `tiles`
// (ty << 16 | tx, sorted by distance from the rotation centre) in XCD order:
// XCD x gets the x-th annulus, and a rotation about the centre maps an annulus
// onto itself, so each XCD's L2 holds only its annulus of the source planes.
__device__ inline bool tile_pixel(const uint32_t *__restrict__ tiles, uint32_t ntiles, uint32_t nx, uint32_t ny,
                                  uint32_t &X, uint32_t &Y) {
  const uint32_t q = xcd_tile(blockIdx.x, gridDim.x);
  if (q >= ntiles) return false;
  const uint32_t tt = tiles[q];
  const uint32_t w = threadIdx.x >> 6, l = threadIdx.x & 63;
  X = (tt & 0xffffu) * 16 + (w & 1) * 8 + (l & 7);
  Y = (tt >> 16) * 16 + (w >> 1) * 8 + (l >> 3);
  return X < nx && Y < ny;
}

template <typename T>
__global__ void __launch_bounds__(256) rotate_accumulate_kernel(const T *__restrict__ proj, double *__restrict__ acc,
                                                                uint32_t *__restrict__ count, uint32_t nx, uint32_t ny,
                                                                Inv6 inv, int rot_mask) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)nx * ny) return;
  const uint32_t Y = (uint32_t)(t / nx), X = (uint32_t)(t - (uint64_t)Y * nx);
  double v;
  if (!bilinear_sample<T>(proj, nx, ny, inv.c, X, Y, &v)) v = __longlong_as_double(0x7ff8000000000000ll);
  if (rot_mask) {
    if (!isnan(v)) {
      acc[t] += v;
      count[t] += 1;
    }
  } else {
    acc[t] += v;
  }
}

__global__ void __launch_bounds__(256) average_kernel(const double *__restrict__ acc, const uint32_t *__restrict__ count,
                                                      double *__restrict__ avg, uint64_t P, uint32_t nA, int rot_mask) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P) return;
  if (rot_mask) {
    const uint32_t c = count[t];
    avg[t] = c == 0 ? 0.0 : acc[t] * (1.0 / (c * 1.0));
  } else {
    avg[t] = acc[t] * (1.0 / (double)nA);
  }
}

template <typename T>
__global__ void __launch_bounds__(256) rotate_all_kernel(const T *__restrict__ proj, const double *__restrict__ inv_tab,
                                                         uint32_t nA, uint32_t nx, uint32_t ny, int rot_mask,
                                                         double *__restrict__ avg, const uint32_t *__restrict__ tiles,
                                                         uint32_t ntiles) {
  const uint64_t P = (uint64_t)nx * ny;
  uint32_t X, Y;
  if (!tile_pixel(tiles, ntiles, nx, ny, X, Y)) return;
  const uint64_t t = (uint64_t)Y * nx + X;
  double acc = 0;
  uint32_t cnt = 0;
  // independent gathers per angle; unrolling lets several be in flight (the sum
  // itself stays in angle order)
#pragma unroll 4
  for (uint32_t a = 0; a < nA; a++) {
    double inv[6];
#pragma unroll
    for (int i = 0; i < 6; i++) inv[i] = inv_tab[6 * (uint64_t)a + i];
    double v;
    if (!bilinear_sample<T>(proj + (uint64_t)a * P, nx, ny, inv, X, Y, &v)) v = __longlong_as_double(0x7ff8000000000000ll);
    if (rot_mask) {
      if (!isnan(v)) {
        acc += v;
        cnt++;
      }
    } else {
      acc += v;
    }
  }
  if (rot_mask) avg[t] = cnt == 0 ? 0.0 : acc * (1.0 / (cnt * 1.0));
  else avg[t] = acc * (1.0 / (double)nA);
}

// Laboratory frame: every angle's intensity from the per-pixel quadratic form
// q = (q0, q1, q2): I(theta) = c^2 q0 + s^2 q1 + 2 c s q2.  Bilinear sampling is
// linear, so the three components are sampled (same taps, weights and validity
// as bilinear_sample; a tap of zero weight contributes nothing, NaN included)
// and combined once per angle, in the storage precision (FP64 runs: FP64).  Each
// angle's back-mapped point, taps and weights serve NE energies.

// Back-mapped point of one (pixel, angle): validity, top-left tap, fractions.
struct QPoint {
  uint32_t x0, y0;
  double sx, sy, fx, fy;
  bool inside;
};

__device__ inline QPoint quad_point(const double *inv, uint32_t X, uint32_t Y, uint32_t nx, uint32_t ny) {
  QPoint p;
  back_map_point(inv, X, Y, p.sx, p.sy);
  const double sx = p.sx, sy = p.sy;
  p.inside = sx >= 0 && sx <= (double)(nx - 1) && sy >= 0 && sy <= (double)(ny - 1);
  const double fx0 = p.inside ? floor(sx) : 0.0, fy0 = p.inside ? floor(sy) : 0.0;
  p.x0 = (uint32_t)fx0;
  p.y0 = (uint32_t)fy0;
  p.fx = p.inside ? sx - fx0 : 0.0;
  p.fy = p.inside ? sy - fy0 : 0.0;
  return p;
}

// One record (q0, q1, q2) of a plane of quadratic forms.
template <typename T> struct Tap3;
template <> struct Tap3<float> { typedef float v __attribute__((ext_vector_type(3))); };
template <> struct Tap3<double> { typedef double v __attribute__((ext_vector_type(3))); };
template <typename T> using Rec = typename Tap3<T>::v;

// The four taps of p (clamped into the image) for NE energy planes.
template <typename T, int NE>
__device__ inline void quad_gather(const T *const (&qp)[NE], const QPoint &p, uint32_t nx, uint32_t ny,
                                   Rec<T> (&tp)[NE][4]) {
  const uint32_t x1 = p.x0 + 1 < nx ? p.x0 + 1 : p.x0, y1 = p.y0 + 1 < ny ? p.y0 + 1 : p.y0;
  const uint32_t off[4] = {quad_index(p.x0, p.y0, nx, ny), quad_index(x1, p.y0, nx, ny), quad_index(p.x0, y1, nx, ny),
                           quad_index(x1, y1, nx, ny)};  // P < 2^32
#pragma unroll
  for (int e = 0; e < NE; e++)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const T *r = qp[e] + (uint64_t)off[i] * kQuadRecord;
      tp[e][i] = Rec<T>{r[0], r[1], r[2]};
    }
}

// Bilinear sums (q0, q1, q2) of NE energies at point p from its taps tp.
template <typename T, int NE>
__device__ inline void quad_sums(const QPoint &p, const Rec<T> (&tp)[NE][4], T (&sm)[NE][3]) {
  const double gx = 1 - p.fx, gy = 1 - p.fy;
  const double wd[4] = {gx * gy, p.fx * gy, gx * p.fy, p.fx * p.fy};
  // taps are interpolated in the storage precision (the weights rounded once per
  // angle); which taps count (weight exactly 0 or not) is decided on the FP64
  // weights, and the sum over angles stays FP64
  T w[4];
  bool z[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    w[i] = (T)wd[i];
    z[i] = wd[i] == 0;
  }
#pragma unroll
  for (int e = 0; e < NE; e++) {
    T s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const Rec<T> t = tp[e][i];
      s0 += w[i] * (z[i] ? (T)0 : t.x);
      s1 += w[i] * (z[i] ? (T)0 : t.y);
      s2 += w[i] * (z[i] ? (T)0 : t.z);
    }
    sm[e][0] = s0;
    sm[e][1] = s1;
    sm[e][2] = s2;
  }
}

// Adds one angle's intensity (coefficients c = (c^2, s^2, 2cs)) from the sums sm
// (NaN where the point is outside) to the FP64 sums of NE energies.
template <typename T, int NE>
__device__ inline void quad_combine(bool inside, const T (&sm)[NE][3], const double *c, int rot_mask,
                                    double (&acc)[NE], uint32_t (&cnt)[NE]) {
  const T c0 = (T)c[0], c1 = (T)c[1], c2 = (T)c[2];
#pragma unroll
  for (int e = 0; e < NE; e++) {
    const double v =
        inside ? (double)(c0 * sm[e][0] + c1 * sm[e][1] + c2 * sm[e][2]) : __longlong_as_double(0x7ff8000000000000ll);
    if (rot_mask) {
      const bool ok = !isnan(v);
      acc[e] += ok ? v : 0.0;
      cnt[e] += ok ? 1u : 0u;
    } else {
      acc[e] += v;
    }
  }
}

template <typename T, int NE>
__device__ inline void quad_add(const QPoint &p, const Rec<T> (&tp)[NE][4], const double *c, int rot_mask,
                                double (&acc)[NE], uint32_t (&cnt)[NE]) {
  T sm[NE][3];
  quad_sums<T, NE>(p, tp, sm);
  quad_combine<T, NE>(p.inside, sm, c, rot_mask, acc, cnt);
}

template <int NE>
__device__ inline void quad_store(double *avg, uint64_t P, uint64_t t, uint32_t e0, uint32_t ne, uint32_t nA,
                                  int rot_mask, const double (&acc)[NE], const uint32_t (&cnt)[NE]) {
#pragma unroll
  for (int e = 0; e < NE; e++) {
    if (e0 + e >= ne) break;
    double *o = avg + (uint64_t)(e0 + e) * P;
    if (rot_mask) o[t] = cnt[e] == 0 ? 0.0 : acc[e] * (1.0 / (cnt[e] * 1.0));
    else o[t] = acc[e] * (1.0 / (double)nA);
  }
}

// Every angle of one pixel, in angle order.
template <typename T, int NE>
__device__ inline void quad_pixel(const T *const (&qp)[NE], const double *__restrict__ inv_tab,
                                  const double *__restrict__ coef, uint32_t nA, uint32_t nx, uint32_t ny, uint32_t X,
                                  uint32_t Y, int rot_mask, double (&acc)[NE], uint32_t (&cnt)[NE]) {
#pragma unroll 2
  for (uint32_t a = 0; a < nA; a++) {
    const QPoint p = quad_point(inv_tab + 6 * (uint64_t)a, X, Y, nx, ny);
    Rec<T> tp[NE][4];
    quad_gather<T, NE>(qp, p, nx, ny, tp);
    quad_add<T, NE>(p, tp, coef + 3 * (uint64_t)a, rot_mask, acc, cnt);
  }
}

// grid y: groups of NE energies (ne in total)
template <typename T, int NE>
__global__ void __launch_bounds__(256) rotate_quad_kernel(const T *__restrict__ quad, const double *__restrict__ inv_tab,
                                                          const double *__restrict__ coef, uint32_t nA, uint32_t nx,
                                                          uint32_t ny, int rot_mask, double *__restrict__ avg,
                                                          uint32_t ne, const uint32_t *__restrict__ tiles,
                                                          uint32_t ntiles) {
  const uint64_t P = (uint64_t)nx * ny;
  uint32_t X, Y;
  if (!tile_pixel(tiles, ntiles, nx, ny, X, Y)) return;
  const uint32_t e0 = blockIdx.z * NE;
  const T *qp[NE];
#pragma unroll
  for (int e = 0; e < NE; e++) qp[e] = quad + (uint64_t)(e0 + e < ne ? e0 + e : e0) * P * kQuadRecord;
  double acc[NE] = {};
  uint32_t cnt[NE] = {};
  quad_pixel<T, NE>(qp, inv_tab, coef, nA, nx, ny, X, Y, rot_mask, acc, cnt);
  quad_store<NE>(avg, P, (uint64_t)Y * nx + X, e0, ne, nA, rot_mask, acc, cnt);
}

// Tiles of the general order beyond every finite plane value: the averages
// rotate_quad_kernel would form there (NaN, or 0 with the rotation mask), as in
// orbit_fill_kernel.  Same tile and pixel layout as tile_pixel.
template <int NE>
__global__ void __launch_bounds__(256) quad_fill_kernel(double *__restrict__ avg, uint32_t nx, uint32_t ny,
                                                        uint32_t ne, int rot_mask, const uint32_t *__restrict__ tiles,
                                                        uint32_t first) {
  const uint64_t P = (uint64_t)nx * ny;
  const uint32_t tt = tiles[first + blockIdx.x];
  const uint32_t w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const uint32_t X = (tt & 0xffffu) * 16 + (w & 1) * 8 + (l & 7);
  const uint32_t Y = (tt >> 16) * 16 + (w >> 1) * 8 + (l >> 3);
  if (X >= nx || Y >= ny) return;
  const uint32_t e0 = blockIdx.z * NE;
  const double fill = rot_mask ? 0.0 : __longlong_as_double(0x7ff8000000000000ll);
#pragma unroll
  for (int e = 0; e < NE; e++)
    if (e0 + e < ne) avg[(uint64_t)(e0 + e) * P + (uint64_t)Y * nx + X] = fill;
}

// Quarter-turn orbits (square N x N planes whose angle list holds Q quarter turns
// of A4 steps: angle a + q A4 is angle a plus 90q degrees).  The image rotation
// turns about (N/2, N/2), so with s(X, Y) = (N - Y, X) (a 90-degree turn of the
// pixel grid) the point pixel x samples at angle a + q A4 is the point pixel
// s^q(x) samples at angle a.  One thread owns the orbit x, s x, s^2 x, s^3 x of a
// pixel with X - N/2 in [1, N/2 - 1] and Y - N/2 in [0, N/2 - 1] (these orbits tile
// [1, N-1]^2 minus the centre): it back-maps and gathers each (s^i x, a < A4) once
// and serves the Q (pixel, angle) pairs that sample there.
//
// The pairs' own FP64 back-maps (contraction off) differ from the shared point,
// evaluated with fused multiply-adds, by the rounding of the angle tables and
// of the evaluation (~N 1e-15 pixels: at most 4e-12 over every pair of a 4096^2
// plane with 720 angles).  When the shared point lies farther than the margin
// 1e-9 + N 1e-12 from every cell edge and from the image border, all Q pairs have
// its tap cell, validity and nonzero weights, so they reuse its bilinear sums
// (only the weights' last bits can differ from a pair's own).  Such points take
// the lean path: no zero-weight tests, taps through 32-bit buffer offsets (the
// plane is below 2 GiB), the intensity in the compute precision, and one value
// for the pairs q and q + 2 when their coefficients are equal (`share`).  The
// rest (a point on or near a pixel edge: angles at exact multiples of 90 degrees,
// or special angles such as 60 degrees on the axes) is recorded per angle in a
// 64-bit mask and replayed after each run of 64 angles: there each pair back-maps
// its own point and gathers its own taps, as quad_pixel does.
// Validity and NaN taps are therefore decided exactly as per pair; only the order
// of the FP64 sums over angles differs.  Without the rotation mask a point
// outside the image marks its pairs' pixels NaN (the value the sum would reach).
// Row 0, column 0 and the centre (the tail blocks after the orbit tiles) take
// every angle directly.

// (q0, q1, q2) of the record at byte offset `off` of a plane; offsets past the
// buffer's size read zeros (only taps of points that take the exact path do).
template <typename T> __device__ inline typename Tap3<T>::v tap_load(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_amdgcn_raw_buffer_load_b96(r, off, 0, 0);
  } else {
    // 24-byte records sit at 8-byte alignment: three 8-byte loads
    typedef int i2 __attribute__((ext_vector_type(2)));
    const i2 a = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
    const i2 b = __builtin_amdgcn_raw_buffer_load_b64(r, off + 8, 0, 0);
    const i2 c = __builtin_amdgcn_raw_buffer_load_b64(r, off + 16, 0, 0);
    typename Tap3<double>::v v;
    v.x = __builtin_bit_cast(double, ((unsigned long long)(unsigned)a.y << 32) | (unsigned)a.x);
    v.y = __builtin_bit_cast(double, ((unsigned long long)(unsigned)b.y << 32) | (unsigned)b.x);
    v.z = __builtin_bit_cast(double, ((unsigned long long)(unsigned)c.y << 32) | (unsigned)c.x);
    return v;
  }
}

// Buffer resource word 3 of a raw (untyped) buffer on gfx9-family devices.
constexpr int kRawBufferFlags = 0x00020000;

template <typename T, int NE, int Q, bool RM>
__global__ void __launch_bounds__(256) rotate_orbit_kernel(const T *__restrict__ quad, const double *__restrict__ inv_tab,
                                                           const double *__restrict__ coef,
                                                           const T *__restrict__ coef_t, uint32_t nA, uint32_t A4,
                                                           uint32_t n, double *__restrict__ avg, uint32_t ne,
                                                           const uint32_t *__restrict__ otiles, uint32_t n_otiles,
                                                           int share) {
  const uint64_t P = (uint64_t)n * n;
  const uint32_t h = n / 2;
  const uint32_t qb = xcd_tile(blockIdx.x, gridDim.x);
  const uint32_t e0 = blockIdx.z * NE;
  const T *qp[NE];
#pragma unroll
  for (int e = 0; e < NE; e++) qp[e] = quad + (uint64_t)(e0 + e < ne ? e0 + e : e0) * P * kQuadRecord;
  if (qb >= n_otiles) {
    // tail: row 0 (n pixels), column 0 below it (n - 1) and the centre
    const uint32_t j = (qb - n_otiles) * 256 + threadIdx.x;
    if (j >= 2 * n) return;
    const uint32_t X = j < n ? j : (j < 2 * n - 1 ? 0 : h);
    const uint32_t Y = j < n ? 0 : (j < 2 * n - 1 ? j - n + 1 : h);
    double acc[NE] = {};
    uint32_t cnt[NE] = {};
    quad_pixel<T, NE>(qp, inv_tab, coef, nA, n, n, X, Y, RM, acc, cnt);
    quad_store<NE>(avg, P, (uint64_t)Y * n + X, e0, ne, nA, RM, acc, cnt);
    return;
  }
  const uint32_t tt = otiles[qb];
  const uint32_t w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const uint32_t u = (tt & 0xffffu) * 16 + (w & 1) * 8 + (l & 7) + 1;
  const uint32_t v = (tt >> 16) * 16 + (w >> 1) * 8 + (l >> 3);
  if (u >= h || v >= h) return;
  // orbit pixels s^i(x), i = 0..3
  const uint32_t X0 = h + u, Y0 = h + v;
  const uint32_t PX[4] = {X0, n - Y0, n - X0, Y0};
  const uint32_t PY[4] = {Y0, X0, n - Y0, n - X0};
  const double hi = (double)(n - 1);
  const double m = 1e-9 + 1e-12 * (double)n;
  constexpr uint32_t rb = kQuadRecord * (uint32_t)sizeof(T);
  __amdgpu_buffer_rsrc_t rs[NE];
#pragma unroll
  for (int e = 0; e < NE; e++)
    rs[e] = __builtin_amdgcn_make_buffer_rsrc((void *)qp[e], (short)0, (int)(P * rb), kRawBufferFlags);
  double acc[4][NE] = {};
  uint32_t cnt[4][NE] = {};
  bool dead[4] = {false, false, false, false};
  static_for<0, 4>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    // angles in chunks of 64: bit (a - a0) of `defer` marks angle a of the chunk
    // whose point takes the exact path
    for (uint32_t a0 = 0; a0 < A4; a0 += 64) {
      const uint32_t a1 = A4 - a0 > 64 ? a0 + 64 : A4;
      uint64_t defer = 0;
      for (uint32_t a = a0; a < a1; a++) {
        const double *inv = inv_tab + 6 * (uint64_t)a;
        const double sx = __builtin_fma(inv[0], (double)PX[i], __builtin_fma(inv[1], (double)PY[i], inv[2]));
        const double sy = __builtin_fma(inv[3], (double)PX[i], __builtin_fma(inv[4], (double)PY[i], inv[5]));
        const bool inside = sx >= 0 && sx <= hi && sy >= 0 && sy <= hi;
        const double fx0 = floor(sx), fy0 = floor(sy);
        const double fx = sx - fx0, fy = sy - fy0;
        const bool safe = inside ? (fx >= m && fx <= 1 - m && fy >= m && fy <= 1 - m)
                                 : (sx < -m || sx > hi + m || sy < -m || sy > hi + m);
        const uint32_t x0 = inside ? (uint32_t)fx0 : 0u, y0 = inside ? (uint32_t)fy0 : 0u;
        const uint32_t o00 = quad_index(x0, y0, n, n) * rb, o10 = quad_index(x0 + 1, y0, n, n) * rb;
        const uint32_t o01 = quad_index(x0, y0 + 1, n, n) * rb, o11 = quad_index(x0 + 1, y0 + 1, n, n) * rb;
        typename Tap3<T>::v t[NE][4];
#pragma unroll
        for (int e = 0; e < NE; e++) {
          t[e][0] = tap_load<T>(rs[e], o00);
          t[e][1] = tap_load<T>(rs[e], o10);
          t[e][2] = tap_load<T>(rs[e], o01);
          t[e][3] = tap_load<T>(rs[e], o11);
        }
        if (!safe) {
          defer |= 1ull << (a - a0);
          continue;
        }
        // every pair shares this point's cell, validity and nonzero weights
        const T fxT = (T)fx, fyT = (T)fy, gx = (T)1 - fxT, gy = (T)1 - fyT;
        const T w00 = gx * gy, w10 = fxT * gy, w01 = gx * fyT, w11 = fxT * fyT;
        T sm[NE][3];
#pragma unroll
        for (int e = 0; e < NE; e++) {
          sm[e][0] = w00 * t[e][0].x + w10 * t[e][1].x + w01 * t[e][2].x + w11 * t[e][3].x;
          sm[e][1] = w00 * t[e][0].y + w10 * t[e][1].y + w01 * t[e][2].y + w11 * t[e][3].y;
          sm[e][2] = w00 * t[e][0].z + w10 * t[e][1].z + w01 * t[e][2].z + w11 * t[e][3].z;
        }
        auto add = [&](auto qc, const T(&vq)[NE]) {
          constexpr int q = decltype(qc)::value;
          constexpr int j = (i - q + 4) & 3;  // s^q(pixel j) = pixel i
#pragma unroll
          for (int e = 0; e < NE; e++) {
            if constexpr (RM) {
              const double vv = (double)vq[e];
              const bool ok = inside && !isnan(vv);
              acc[j][e] += ok ? vv : 0.0;
              cnt[j][e] += ok ? 1u : 0u;
            } else {
              acc[j][e] += (double)vq[e];
            }
          }
          if constexpr (!RM) dead[j] = dead[j] || !inside;
        };
        auto value = [&](int q, T(&vq)[NE]) {
          const T *c = coef_t + 3 * (uint64_t)(a + q * A4);
#pragma unroll
          for (int e = 0; e < NE; e++) vq[e] = c[0] * sm[e][0] + c[1] * sm[e][1] + c[2] * sm[e][2];
        };
        bool shared = false;
        if constexpr (Q == 4) {
          if (share) {  // pairs q and q + 2 have equal coefficients
            T v0[NE], v1[NE];
            value(0, v0);
            value(1, v1);
            add(std::integral_constant<int, 0>{}, v0);
            add(std::integral_constant<int, 1>{}, v1);
            add(std::integral_constant<int, 2>{}, v0);
            add(std::integral_constant<int, 3>{}, v1);
            shared = true;
          }
        }
        if (!shared) {
          static_for<0, Q>([&](auto qc) {
            T vq[NE];
            value(decltype(qc)::value, vq);
            add(qc, vq);
          });
        }
      }
      // points near a cell edge or the image border: every pair back-maps its own
      // point exactly and gathers its own taps (one pair at a time, so the rare path
      // holds few registers)
      while (defer) {
        const uint32_t a = a0 + (uint32_t)__builtin_ctzll(defer);
        defer &= defer - 1;
#pragma unroll 1
        for (uint32_t q = 0; q < (uint32_t)Q; q++) {
          const uint32_t j = (i - q + 4) & 3;
          const uint32_t aa = a + q * A4;
          const QPoint d = quad_point(inv_tab + 6 * (uint64_t)aa, PX[j], PY[j], n, n);
          Rec<T> td[NE][4];
          quad_gather<T, NE>(qp, d, n, n, td);
          T sm[NE][3];
          quad_sums<T, NE>(d, td, sm);
          const double *c = coef + 3 * (uint64_t)aa;
          switch (j) {  // constant indices keep the accumulators in registers
            case 0: quad_combine<T, NE>(d.inside, sm, c, RM, acc[0], cnt[0]); break;
            case 1: quad_combine<T, NE>(d.inside, sm, c, RM, acc[1], cnt[1]); break;
            case 2: quad_combine<T, NE>(d.inside, sm, c, RM, acc[2], cnt[2]); break;
            default: quad_combine<T, NE>(d.inside, sm, c, RM, acc[3], cnt[3]); break;
          }
        }
      }
    }
  });
#pragma unroll
  for (int i = 0; i < 4; i++) {
    if constexpr (!RM) {
      if (dead[i]) {
#pragma unroll
        for (int e = 0; e < NE; e++) acc[i][e] = __longlong_as_double(0x7ff8000000000000ll);
      }
    }
    quad_store<NE>(avg, P, (uint64_t)PY[i] * n + PX[i], e0, ne, nA, RM, acc[i], cnt[i]);
  }
}

// Orbit tiles beyond every finite plane value (their pixels' taps, within 1.5
// pixels of points as far from the centre as the pixel, are all NaN): the
// averages rotate_orbit_kernel would form there, NaN, or 0 with the rotation
// mask (no valid sample).  Same tile and pixel layout as rotate_orbit_kernel.
template <int NE, bool RM>
__global__ void __launch_bounds__(256) orbit_fill_kernel(double *__restrict__ avg, uint32_t n, uint32_t ne,
                                                         const uint32_t *__restrict__ otiles, uint32_t first) {
  const uint64_t P = (uint64_t)n * n;
  const uint32_t h = n / 2;
  const uint32_t e0 = blockIdx.z * NE;
  const uint32_t tt = otiles[first + blockIdx.x];
  const uint32_t w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const uint32_t u = (tt & 0xffffu) * 16 + (w & 1) * 8 + (l & 7) + 1;
  const uint32_t v = (tt >> 16) * 16 + (w >> 1) * 8 + (l >> 3);
  if (u >= h || v >= h) return;
  const uint32_t X0 = h + u, Y0 = h + v;
  const uint32_t PX[4] = {X0, n - Y0, n - X0, Y0};
  const uint32_t PY[4] = {Y0, X0, n - Y0, n - X0};
  const double fill = RM ? 0.0 : __longlong_as_double(0x7ff8000000000000ll);
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int e = 0; e < NE; e++)
      if (e0 + e < ne) avg[(uint64_t)(e0 + e) * P + (uint64_t)PY[i] * n + PX[i]] = fill;
}

// grid y = pattern: avg + y P -> out + y out_stride
template <typename T>
__global__ void __launch_bounds__(256) detector_warp_kernel(const double *__restrict__ avg, T *__restrict__ out,
                                                            uint32_t nx, uint32_t ny, Inv6 inv, int rot_mask,
                                                            uint64_t out_stride) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)nx * ny) return;
  avg += (uint64_t)blockIdx.y * nx * ny;
  out += (uint64_t)blockIdx.y * out_stride;
  const uint32_t Y = (uint32_t)(t / nx), X = (uint32_t)(t - (uint64_t)Y * nx);
  double v;
  if (!bilinear_sample<double>(avg, nx, ny, inv.c, X, Y, &v))
    v = rot_mask ? 0.0 : __longlong_as_double(0x7ff8000000000000ll);
  out[t] = (T)v;
}

// The same for an identity warp (the detector normal along the beam): every
// destination pixel samples its own source pixel with weight 1, so the pattern
// is the average converted (0 + a, as the weighted sum forms it).
template <typename T>
__global__ void __launch_bounds__(256) detector_copy_kernel(const double *__restrict__ avg, T *__restrict__ out,
                                                            uint64_t P, uint64_t out_stride) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P) return;
  out[(uint64_t)blockIdx.y * out_stride + t] = (T)(0.0 + avg[(uint64_t)blockIdx.y * P + t]);
}

bool is_identity(const double inv[6]) {
  return inv[0] == 1.0 && inv[1] == 0.0 && inv[2] == 0.0 && inv[3] == 0.0 && inv[4] == 1.0 && inv[5] == 0.0;
}

template <typename T>
__global__ void __launch_bounds__(256) warp_kernel(const T *__restrict__ src, T *__restrict__ dst, uint32_t nx,
                                                   uint32_t ny, Inv6 inv) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)nx * ny) return;
  const uint32_t Y = (uint32_t)(t / nx), X = (uint32_t)(t - (uint64_t)Y * nx);
  double v;
  if (bilinear_sample<T>(src, nx, ny, inv.c, X, Y, &v)) dst[t] = (T)v;
}

Inv6 mk_inv(const double inv[6]) {
  Inv6 r;
  for (int i = 0; i < 6; i++) r.c[i] = inv[i];
  return r;
}

}  // namespace

uint32_t rotation_tile_count(uint32_t nx, uint32_t ny) { return ceil_div(nx, 16) * ceil_div(ny, 16); }

// Tiles of rotation_tile_order whose centre lies within r pixels of the plane's
// centre: the first ones of that order (same key).
uint32_t rotation_tiles_within(uint32_t nx, uint32_t ny, double r) {
  const uint32_t tx = ceil_div(nx, 16), ty = ceil_div(ny, 16);
  uint32_t c = 0;
  for (uint32_t y = 0; y < ty; y++)
    for (uint32_t x = 0; x < tx; x++) {
      const double dx = (x * 16.0 + 8.0) - nx / 2.0, dy = (y * 16.0 + 8.0) - ny / 2.0;
      c += dx * dx + dy * dy <= r * r ? 1u : 0u;
    }
  return c;
}

void rotation_tile_order(uint32_t nx, uint32_t ny, uint32_t *out) {
  const uint32_t tx = ceil_div(nx, 16), ty = ceil_div(ny, 16);
  std::vector<std::pair<double, uint32_t>> r;
  r.reserve((size_t)tx * ty);
  for (uint32_t y = 0; y < ty; y++)
    for (uint32_t x = 0; x < tx; x++) {
      const double dx = (x * 16.0 + 8.0) - nx / 2.0, dy = (y * 16.0 + 8.0) - ny / 2.0;
      r.emplace_back(dx * dx + dy * dy, (y << 16) | x);
    }
  std::stable_sort(r.begin(), r.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
  for (size_t i = 0; i < r.size(); i++) out[i] = r[i].second;
}

uint32_t rotation_orbit_tile_count(uint32_t n) { return n >= 4 && n % 2 == 0 ? ceil_div(n / 2 - 1, 16) * ceil_div(n / 2, 16) : 0; }

void rotation_orbit_tile_order(uint32_t n, uint32_t *out) {
  // orbit representatives: u = X - n/2 in [1, n/2 - 1] (tile x covers u - 1),
  // v = Y - n/2 in [0, n/2 - 1]; ordered by distance from the centre (0, 0)
  const uint32_t h = n / 2, tx = ceil_div(h - 1, 16), ty = ceil_div(h, 16);
  std::vector<std::pair<double, uint32_t>> r;
  r.reserve((size_t)tx * ty);
  for (uint32_t y = 0; y < ty; y++)
    for (uint32_t x = 0; x < tx; x++) {
      const double du = x * 16.0 + 9.0, dv = y * 16.0 + 8.0;
      r.emplace_back(du * du + dv * dv, (y << 16) | x);
    }
  std::stable_sort(r.begin(), r.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
  for (size_t i = 0; i < r.size(); i++) out[i] = r[i].second;
}

// Orbit tiles whose centre (du, dv of rotation_orbit_tile_order) lies within r
// pixels of the plane's centre: the first ones of that order.
uint32_t rotation_orbit_tiles_within(uint32_t n, double r) {
  const uint32_t h = n / 2, tx = ceil_div(h - 1, 16), ty = ceil_div(h, 16);
  uint32_t c = 0;
  for (uint32_t y = 0; y < ty; y++)
    for (uint32_t x = 0; x < tx; x++) {
      const double du = x * 16.0 + 9.0, dv = y * 16.0 + 8.0;
      c += du * du + dv * dv <= r * r ? 1u : 0u;
    }
  return c;
}

uint32_t rotation_quarter_turns(uint32_t nx, uint32_t ny, uint32_t nA, uint32_t quarter) {
  if (quarter == 0 || nx != ny || nx < 4 || nx % 2 != 0 || nA % quarter != 0) return 0;
  const uint32_t q = nA / quarter;
  return q >= 2 && q <= 4 ? q : 0;
}

template <typename T>
cyr_status launch_rotate_accumulate(const T *proj, double *acc, uint32_t *count, uint32_t nx, uint32_t ny,
                                    const double inv[6], int rot_mask, hipStream_t st) {
  const uint64_t P = (uint64_t)nx * ny;
  hipLaunchKernelGGL((rotate_accumulate_kernel<T>), dim3(ceil_div(P, 256)), dim3(256), 0, st, proj, acc, count, nx,
                     ny, mk_inv(inv), rot_mask);
  CYR_LAUNCHED();
  return CYR_OK;
}

cyr_status launch_average(const double *acc, const uint32_t *count, double *avg, uint64_t P, uint32_t nA, int rot_mask,
                          hipStream_t st) {
  hipLaunchKernelGGL(average_kernel, dim3(ceil_div(P, 256)), dim3(256), 0, st, acc, count, avg, P, nA, rot_mask);
  CYR_LAUNCHED();
  return CYR_OK;
}

template <typename T>
cyr_status launch_rotate_all(const T *proj, const double *inv_tab, uint32_t nA, uint32_t nx, uint32_t ny, int rot_mask,
                             double *avg, const uint32_t *tiles, double rfin, hipStream_t st) {
  const uint32_t ntiles = rotation_tile_count(nx, ny);
  if (!tiles) return fail(CYR_E_INVALID, "rotation: tile order missing");
  // the finite-radius cut of launch_rotate_quad (the projections are NaN where the
  // forms are)
  const uint32_t n_act = rfin >= 0 ? std::min(ntiles, rotation_tiles_within(nx, ny, rfin + 14.0)) : ntiles;
  hipLaunchKernelGGL((rotate_all_kernel<T>), dim3(xcd_grid(n_act)), dim3(256), 0, st, proj, inv_tab, nA, nx, ny,
                     rot_mask, avg, tiles, n_act);
  CYR_LAUNCHED();
  if (n_act < ntiles) {
    hipLaunchKernelGGL((quad_fill_kernel<1>), dim3(ntiles - n_act), dim3(256), 0, st, avg, nx, ny, 1u, rot_mask, tiles,
                       n_act);
    CYR_LAUNCHED();
  }
  return CYR_OK;
}

template <typename T, int NE, bool RM>
hipError_t launch_orbit_rm(uint32_t Q, dim3 grid, hipStream_t st, const T *quad, const double *inv_tab,
                           const double *coef, const T *coef_t, uint32_t nA, uint32_t A4, uint32_t n, double *avg,
                           uint32_t ne, const uint32_t *otiles, uint32_t n_otiles, int share) {
  switch (Q) {
    case 2:
      hipLaunchKernelGGL((rotate_orbit_kernel<T, NE, 2, RM>), grid, dim3(256), 0, st, quad, inv_tab, coef, coef_t, nA,
                         A4, n, avg, ne, otiles, n_otiles, share);
      break;
    case 3:
      hipLaunchKernelGGL((rotate_orbit_kernel<T, NE, 3, RM>), grid, dim3(256), 0, st, quad, inv_tab, coef, coef_t, nA,
                         A4, n, avg, ne, otiles, n_otiles, share);
      break;
    default:
      hipLaunchKernelGGL((rotate_orbit_kernel<T, NE, 4, RM>), grid, dim3(256), 0, st, quad, inv_tab, coef, coef_t, nA,
                         A4, n, avg, ne, otiles, n_otiles, share);
      break;
  }
  return hipGetLastError();
}

template <typename T>
cyr_status launch_rotate_quad(const T *quad, const double *inv_tab, const double *coef, const T *coef_t, uint32_t nA,
                              uint32_t nx, uint32_t ny, int rot_mask, double *avg, uint32_t ne, const uint32_t *tiles,
                              uint32_t quarter, const uint32_t *otiles, bool share, double rfin, hipStream_t st) {
  if (ne == 0) return CYR_OK;
  if ((uint64_t)nx * ny > 0xffffffffull) return fail(CYR_E_UNSUPPORTED, "rotation: more than 2^32 pixels");
  if (!tiles) return fail(CYR_E_INVALID, "rotation: tile order missing");
  // the orbit kernel addresses a plane through 32-bit buffer offsets
  const bool small_plane = (uint64_t)nx * ny * kQuadRecord * sizeof(T) < (1ull << 31);
  const uint32_t Q = otiles && small_plane ? rotation_quarter_turns(nx, ny, nA, quarter) : 0;
  if (Q) {
    // quarter-turn orbits: each gathered tap cell serves Q (pixel, angle) pairs
    const uint32_t n_otiles = rotation_orbit_tile_count(nx);
    // tiles (in order of distance from the centre) whose pixels all lie more than
    // rfin + 2 pixels out are filled, not sampled: a tile's pixels are within
    // 8 sqrt(2) < 11.4 pixels of its centre, hence the cut at rfin + 14
    const uint32_t n_act = rfin >= 0 ? std::min(n_otiles, rotation_orbit_tiles_within(nx, rfin + 14.0)) : n_otiles;
    // 2 energies per pass (measured against 1 and 4: 512^2 x 20 energies 415 / 384 us
    // for 4 / 2; 2048^2 film 7.30 / 6.75 / 8.92 ms per 10 energies for 1 / 2 / 4)
    const uint64_t blocks = (uint64_t)n_act + ceil_div(2ull * nx, 256);
    const dim3 grid(xcd_grid(blocks), 1, ceil_div(ne, 2));
    hipError_t e =
        rot_mask ? launch_orbit_rm<T, 2, true>(Q, grid, st, quad, inv_tab, coef, coef_t, nA, quarter, nx, avg, ne,
                                                otiles, n_act, share ? 1 : 0)
                 : launch_orbit_rm<T, 2, false>(Q, grid, st, quad, inv_tab, coef, coef_t, nA, quarter, nx, avg, ne,
                                                 otiles, n_act, share ? 1 : 0);
    if (e != hipSuccess) return hip_fail(e, "rotate_orbit_kernel", __FILE__, __LINE__);
    if (n_act < n_otiles) {
      const dim3 gf(n_otiles - n_act, 1, ceil_div(ne, 2));
      if (rot_mask) hipLaunchKernelGGL((orbit_fill_kernel<2, true>), gf, dim3(256), 0, st, avg, nx, ne, otiles, n_act);
      else hipLaunchKernelGGL((orbit_fill_kernel<2, false>), gf, dim3(256), 0, st, avg, nx, ne, otiles, n_act);
      e = hipGetLastError();
      if (e != hipSuccess) return hip_fail(e, "orbit_fill_kernel", __FILE__, __LINE__);
    }
    return CYR_OK;
  }
  const uint32_t ntiles = rotation_tile_count(nx, ny);
  // tiles (in order of distance from the centre (Nx/2, Ny/2), about which the
  // image turns) whose pixels all lie more than rfin + 2 pixels out are filled:
  // a 16 x 16 tile's pixels are within 8 sqrt(2) of its centre
  const uint32_t n_act = rfin >= 0 ? std::min(ntiles, rotation_tiles_within(nx, ny, rfin + 14.0)) : ntiles;
  // Energies per pass: 4 share each angle's taps; once a plane of quadratic forms
  // exceeds 16 MB (2048^2: 64 MB), one energy per pass keeps the gathers' working
  // set in cache (measured 19% faster there, equal at 512^2).
  const bool one = (uint64_t)nx * ny * kQuadRecord * sizeof(T) > (16ull << 20);
  if (one)
    hipLaunchKernelGGL((rotate_quad_kernel<T, 1>), dim3(xcd_grid(n_act), 1, ne), dim3(256), 0, st, quad, inv_tab,
                       coef, nA, nx, ny, rot_mask, avg, ne, tiles, n_act);
  else
    hipLaunchKernelGGL((rotate_quad_kernel<T, 4>), dim3(xcd_grid(n_act), 1, ceil_div(ne, 4)), dim3(256), 0, st, quad,
                       inv_tab, coef, nA, nx, ny, rot_mask, avg, ne, tiles, n_act);
  CYR_LAUNCHED();
  if (n_act < ntiles) {
    if (one)
      hipLaunchKernelGGL((quad_fill_kernel<1>), dim3(ntiles - n_act, 1, ne), dim3(256), 0, st, avg, nx, ny, ne, rot_mask,
                         tiles, n_act);
    else
      hipLaunchKernelGGL((quad_fill_kernel<4>), dim3(ntiles - n_act, 1, ceil_div(ne, 4)), dim3(256), 0, st, avg, nx, ny,
                         ne, rot_mask, tiles, n_act);
    CYR_LAUNCHED();
  }
  return CYR_OK;
}

template <typename T>
cyr_status launch_detector_warp(const double *avg, T *out, uint32_t nx, uint32_t ny, const double inv[6], int rot_mask,
                                hipStream_t st) {
  const uint64_t P = (uint64_t)nx * ny;
  if (is_identity(inv))
    hipLaunchKernelGGL((detector_copy_kernel<T>), dim3(ceil_div(P, 256)), dim3(256), 0, st, avg, out, P, (uint64_t)0);
  else
    hipLaunchKernelGGL((detector_warp_kernel<T>), dim3(ceil_div(P, 256)), dim3(256), 0, st, avg, out, nx, ny,
                       mk_inv(inv), rot_mask, (uint64_t)0);
  CYR_LAUNCHED();
  return CYR_OK;
}

template <typename T>
cyr_status launch_detector_warp_n(const double *avg, T *out, uint64_t out_stride, uint32_t ne, uint32_t nx,
                                  uint32_t ny, const double inv[6], int rot_mask, hipStream_t st) {
  const uint64_t P = (uint64_t)nx * ny;
  if (ne == 0) return CYR_OK;
  if (is_identity(inv))
    hipLaunchKernelGGL((detector_copy_kernel<T>), dim3(ceil_div(P, 256), ne), dim3(256), 0, st, avg, out, P,
                       out_stride);
  else
    hipLaunchKernelGGL((detector_warp_kernel<T>), dim3(ceil_div(P, 256), ne), dim3(256), 0, st, avg, out, nx, ny,
                       mk_inv(inv), rot_mask, out_stride);
  CYR_LAUNCHED();
  return CYR_OK;
}

template <typename T>
cyr_status launch_warp(const T *src, T *dst, uint32_t nx, uint32_t ny, const double inv[6], hipStream_t st) {
  const uint64_t P = (uint64_t)nx * ny;
  hipLaunchKernelGGL((warp_kernel<T>), dim3(ceil_div(P, 256)), dim3(256), 0, st, src, dst, nx, ny, mk_inv(inv));
  CYR_LAUNCHED();
  return CYR_OK;
}

#define CYR_INST(T)                                                                                                \
  template cyr_status launch_rotate_accumulate<T>(const T *, double *, uint32_t *, uint32_t, uint32_t,             \
                                                  const double *, int, hipStream_t);                               \
  template cyr_status launch_rotate_all<T>(const T *, const double *, uint32_t, uint32_t, uint32_t, int, double *, \
                                           const uint32_t *, double, hipStream_t);                                 \
  template cyr_status launch_rotate_quad<T>(const T *, const double *, const double *, const T *, uint32_t,      \
                                            uint32_t, uint32_t, int, double *, uint32_t, const uint32_t *,        \
                                            uint32_t, const uint32_t *, bool, double, hipStream_t);                \
  template cyr_status launch_detector_warp_n<T>(const double *, T *, uint64_t, uint32_t, uint32_t, uint32_t,       \
                                                const double *, int, hipStream_t);                                 \
  template cyr_status launch_detector_warp<T>(const double *, T *, uint32_t, uint32_t, const double *, int,        \
                                              hipStream_t);                                                        \
  template cyr_status launch_warp<T>(const T *, T *, uint32_t, uint32_t, const double *, hipStream_t);
CYR_INST(float)
CYR_INST(double)
#undef CYR_INST

}  // namespace cyr
