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

#include "geom.hpp"
#include "kernels.hpp"

namespace cyr {

namespace {

struct Inv6 { double c[6]; };

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
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P) return;
  const uint32_t Y = (uint32_t)(t / nx), X = (uint32_t)(t - (uint64_t)Y * nx);
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
  hipLaunchKernelGGL((rotate_all_kernel<T>), dim3(ceil_div(P, 256)), dim3(256), 0, st, proj, inv_tab, nA, nx, ny,
                     rot_mask, avg);
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
  template cyr_status launch_detector_warp<T>(const double *, T *, uint32_t, uint32_t, const double *, int,        \
                                              hipStream_t);                                                        \
  template cyr_status launch_warp<T>(const T *, T *, uint32_t, uint32_t, const double *, hipStream_t);
CYR_INST(float)
CYR_INST(double)
#undef CYR_INST

}  // namespace cyr
