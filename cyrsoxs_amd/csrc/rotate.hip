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

#include <type_traits>

#include "geom.hpp"
#include "kernels.hpp"

namespace cyr {

namespace {

struct Inv6 { double c[6]; };

// Pixel of this thread in the rotation kernels: 256-thread blocks cover 16 x 16
// tiles, each wave an 8 x 8 quadrant, so a wave's rotated source footprint is
// compact (a 64-pixel row segment maps to a 64-pixel line of source rows at
// most angles, touching many more cache lines).
__device__ inline bool tile_pixel(uint32_t nx, uint32_t ny, uint32_t &X, uint32_t &Y) {
  const uint32_t w = threadIdx.x >> 6, l = threadIdx.x & 63;
  X = blockIdx.x * 16 + (w & 1) * 8 + (l & 7);
  Y = blockIdx.y * 16 + (w >> 1) * 8 + (l >> 3);
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
                                                         double *__restrict__ avg) {
  const uint64_t P = (uint64_t)nx * ny;
  uint32_t X, Y;
  if (!tile_pixel(nx, ny, X, Y)) return;
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
// and combined once per angle.
template <typename T>
__global__ void __launch_bounds__(256) rotate_quad_kernel(const T *__restrict__ quad, const double *__restrict__ inv_tab,
                                                          const double *__restrict__ coef, uint32_t nA, uint32_t nx,
                                                          uint32_t ny, int rot_mask, double *__restrict__ avg) {
  uint32_t X, Y;
  if (!tile_pixel(nx, ny, X, Y)) return;
  const uint64_t t = (uint64_t)Y * nx + X;
  using V4 = typename std::conditional<sizeof(T) == 4, float4, double4>::type;
  const V4 *q4 = reinterpret_cast<const V4 *>(quad);
  double acc = 0;
  uint32_t cnt = 0;
#pragma unroll 4
  for (uint32_t a = 0; a < nA; a++) {
    double sx, sy;
    back_map_point(inv_tab + 6 * (uint64_t)a, X, Y, sx, sy);
    const bool inside = sx >= 0 && sx <= (double)(nx - 1) && sy >= 0 && sy <= (double)(ny - 1);
    const double fx0 = inside ? floor(sx) : 0.0, fy0 = inside ? floor(sy) : 0.0;
    const uint32_t x0 = (uint32_t)fx0, y0 = (uint32_t)fy0;
    const double fx = inside ? sx - fx0 : 0.0, fy = inside ? sy - fy0 : 0.0;
    const uint32_t x1 = x0 + 1 < nx ? x0 + 1 : x0, y1 = y0 + 1 < ny ? y0 + 1 : y0;
    const uint32_t r0 = y0 * nx, r1 = y1 * nx;  // P < 2^32
    const V4 a00 = q4[r0 + x0], a10 = q4[r0 + x1], a01 = q4[r1 + x0], a11 = q4[r1 + x1];
    const double gx = 1 - fx, gy = 1 - fy;
    const double w[4] = {gx * gy, fx * gy, gx * fy, fx * fy};
    const V4 tap[4] = {a00, a10, a01, a11};
    double s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const bool z = w[i] == 0;  // selected in the storage type: one move per component
      s0 += w[i] * (double)(z ? (T)0 : tap[i].x);
      s1 += w[i] * (double)(z ? (T)0 : tap[i].y);
      s2 += w[i] * (double)(z ? (T)0 : tap[i].z);
    }
    const double *c = coef + 3 * (uint64_t)a;
    const double v = inside ? c[0] * s0 + c[1] * s1 + c[2] * s2 : __longlong_as_double(0x7ff8000000000000ll);
    if (rot_mask) {
      const bool ok = !isnan(v);
      acc += ok ? v : 0.0;
      cnt += ok ? 1u : 0u;
    } else {
      acc += v;
    }
  }
  if (rot_mask) avg[t] = cnt == 0 ? 0.0 : acc * (1.0 / (cnt * 1.0));
  else avg[t] = acc * (1.0 / (double)nA);
}

template <typename T>
__global__ void __launch_bounds__(256) detector_warp_kernel(const double *__restrict__ avg, T *__restrict__ out,
                                                            uint32_t nx, uint32_t ny, Inv6 inv, int rot_mask) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)nx * ny) return;
  const uint32_t Y = (uint32_t)(t / nx), X = (uint32_t)(t - (uint64_t)Y * nx);
  double v;
  if (!bilinear_sample<double>(avg, nx, ny, inv.c, X, Y, &v))
    v = rot_mask ? 0.0 : __longlong_as_double(0x7ff8000000000000ll);
  out[t] = (T)v;
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
                             double *avg, hipStream_t st) {
  const uint64_t P = (uint64_t)nx * ny;
  (void)P;
  hipLaunchKernelGGL((rotate_all_kernel<T>), dim3(ceil_div(nx, 16), ceil_div(ny, 16)), dim3(256), 0, st, proj,
                     inv_tab, nA, nx, ny, rot_mask, avg);
  CYR_LAUNCHED();
  return CYR_OK;
}

template <typename T>
cyr_status launch_rotate_quad(const T *quad, const double *inv_tab, const double *coef, uint32_t nA, uint32_t nx,
                              uint32_t ny, int rot_mask, double *avg, hipStream_t st) {
  const uint64_t P = (uint64_t)nx * ny;
  (void)P;
  hipLaunchKernelGGL((rotate_quad_kernel<T>), dim3(ceil_div(nx, 16), ceil_div(ny, 16)), dim3(256), 0, st, quad,
                     inv_tab, coef, nA, nx, ny, rot_mask, avg);
  CYR_LAUNCHED();
  return CYR_OK;
}

template <typename T>
cyr_status launch_detector_warp(const double *avg, T *out, uint32_t nx, uint32_t ny, const double inv[6], int rot_mask,
                                hipStream_t st) {
  const uint64_t P = (uint64_t)nx * ny;
  hipLaunchKernelGGL((detector_warp_kernel<T>), dim3(ceil_div(P, 256)), dim3(256), 0, st, avg, out, nx, ny,
                     mk_inv(inv), rot_mask);
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
                                           hipStream_t);                                                           \
  template cyr_status launch_rotate_quad<T>(const T *, const double *, const double *, uint32_t, uint32_t,        \
                                            uint32_t, int, double *, hipStream_t);                                 \
  template cyr_status launch_detector_warp<T>(const double *, T *, uint32_t, uint32_t, const double *, int,        \
                                              hipStream_t);                                                        \
  template cyr_status launch_warp<T>(const T *, T *, uint32_t, uint32_t, const double *, hipStream_t);
CYR_INST(float)
CYR_INST(double)
#undef CYR_INST

}  // namespace cyr
