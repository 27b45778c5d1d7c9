// Ewald-sphere projection, DC replacement and FFT shift (replaces
// computeEwaldProjectionGPU / computeScatter3D / replaceDCComponentWithAverage /
// FFTIgor: include/uniaxial.h:570-947, src/cudaMain.cu:123-243).
//
// The Igor shift is never materialised on the hot path: a shifted coordinate
// (X, Y, Z) reads FFT bin (unshift(X), unshift(Y), unshift(Z)).
#include <hip/hip_runtime.h>

#include <cmath>

#include "geom.hpp"
#include "kernels.hpp"

namespace cyr {

EwaldGeom make_ewald_geom(uint32_t nx, uint32_t ny, uint32_t nz, double phys_size, double k, const double kvec[3],
                          int interpolation) {
  EwaldGeom g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.start = -(M_PI / phys_size);
  g.dx = (2.0 * M_PI / phys_size) / ((nx - 1) * 1.0);
  g.dy = (2.0 * M_PI / phys_size) / ((ny - 1) * 1.0);
  g.enable2d = nz == 1;
  g.dz = g.enable2d ? 0.0 : (2.0 * M_PI / phys_size) / ((nz - 1) * 1.0);
  g.k = k;
  g.kx = kvec[0];
  g.ky = kvec[1];
  g.kz = kvec[2];
  g.interpolation = interpolation;
  return g;
}

namespace {

template <typename T> __device__ inline T nan_t();
template <> __device__ inline float nan_t<float>() { return __int_as_float(0x7fc00000); }
template <> __device__ inline double nan_t<double>() { return __longlong_as_double(0x7ff8000000000000ll); }

template <typename T>
__device__ inline T intensity_at(const EwaldGeom &g, const cx<T> *px, const cx<T> *py, const cx<T> *pz, uint32_t X,
                                 uint32_t Y, uint32_t Z) {
  const uint64_t b = (uint64_t)unshift(X, g.nx) + (uint64_t)unshift(Y, g.ny) * g.nx +
                     (uint64_t)(g.enable2d ? 0 : unshift(Z, g.nz)) * g.nx * g.ny;
  const cx<T> p[3] = {px[b], py[b], pz[b]};
  double m[6];
  scatter_matrix(g, X, Y, Z, m);
  return scatter_intensity<T>(m, p);
}

template <typename T>
__global__ void __launch_bounds__(256) ewald_partial_kernel(EwaldGeom g, const cx<T> *__restrict__ px,
                                                            const cx<T> *__restrict__ py,
                                                            const cx<T> *__restrict__ pz, T *__restrict__ proj) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)g.nx * g.ny) return;
  const uint32_t Y = (uint32_t)(t / g.nx), X = (uint32_t)(t - (uint64_t)Y * g.nx);
  uint32_t z0 = 0;
  double f = 0;
  const int kind = ewald_geometry(g, X, Y, &z0, &f);
  T out;
  if (kind == 0) {
    out = nan_t<T>();
  } else if (kind == 1) {
    out = intensity_at<T>(g, px, py, pz, X, Y, z0);
  } else {
    const T d1 = intensity_at<T>(g, px, py, pz, X, Y, z0);
    const T d2 = intensity_at<T>(g, px, py, pz, X, Y, z0 + 1);
    const T w = (T)f;
    out = (1 - w) * d1 + w * d2;
  }
  proj[t] = out;
}

template <typename T>
__global__ void __launch_bounds__(256) scatter3d_kernel(EwaldGeom g, const cx<T> *__restrict__ px,
                                                        const cx<T> *__restrict__ py, const cx<T> *__restrict__ pz,
                                                        T *__restrict__ scatter) {
  const uint64_t P = (uint64_t)g.nx * g.ny;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P * g.nz) return;
  const uint32_t Z = (uint32_t)(t / P);
  const uint64_t r = t - (uint64_t)Z * P;
  const uint32_t Y = (uint32_t)(r / g.nx), X = (uint32_t)(r - (uint64_t)Y * g.nx);
  scatter[t] = intensity_at<T>(g, px, py, pz, X, Y, Z);
}

template <typename T>
__global__ void __launch_bounds__(256) ewald_full_kernel(EwaldGeom g, const T *__restrict__ scatter,
                                                         T *__restrict__ proj) {
  const uint64_t P = (uint64_t)g.nx * g.ny;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P) return;
  const uint32_t Y = (uint32_t)(t / g.nx), X = (uint32_t)(t - (uint64_t)Y * g.nx);
  uint32_t z0 = 0;
  double f = 0;
  const int kind = ewald_geometry(g, X, Y, &z0, &f);
  T out;
  if (kind == 0) out = nan_t<T>();
  else if (kind == 1) out = scatter[(uint64_t)z0 * P + t];
  else {
    const T w = (T)f;
    out = (1 - w) * scatter[(uint64_t)z0 * P + t] + w * scatter[(uint64_t)(z0 + 1) * P + t];
  }
  proj[t] = out;
}

// F[0] <- mean of the periodic neighbours (4 in-plane, +2 along z when Nz > 1);
// one thread per volume.
template <typename T> struct DcVols { cx<T> *v[8]; };

template <typename T>
__global__ void dc_replace_kernel(DcVols<T> vols, uint32_t count, uint32_t nx, uint32_t ny, uint32_t nz) {
  const uint32_t c = threadIdx.x;
  if (c >= count) return;
  cx<T> *F = vols.v[c];
  T sr = 0, si = 0;
  int n = 0;
  const uint64_t idx[6] = {1, nx, (uint64_t)nx - 1, (uint64_t)(ny - 1) * nx, (uint64_t)nx * ny,
                           (uint64_t)(nz - 1) * nx * ny};
  const int cnt = nz > 1 ? 6 : 4;
  for (int i = 0; i < cnt; i++) {
    sr += F[idx[i]].x;
    si += F[idx[i]].y;
    n++;
  }
  F[0] = mk<T>(sr / (T)n, si / (T)n);
}

template <typename T>
__global__ void __launch_bounds__(256) igor_shift_kernel(cx<T> *vol, uint32_t nx, uint32_t ny, uint32_t nz) {
  const uint64_t P = (uint64_t)nx * ny;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P * nz) return;
  const uint32_t Z = (uint32_t)(t / P);
  const uint64_t r = t - (uint64_t)Z * P;
  const uint32_t Y = (uint32_t)(r / nx), X = (uint32_t)(r - (uint64_t)Y * nx);
  const uint64_t c = (uint64_t)unshift(X, nx) + (uint64_t)unshift(Y, ny) * nx + (uint64_t)unshift(Z, nz) * P;
  if (c > t) {
    const cx<T> a = vol[c];
    vol[c] = vol[t];
    vol[t] = a;
  }
}

}  // namespace

template <typename T>
cyr_status launch_ewald_partial(const EwaldGeom &g, const cx<T> *px, const cx<T> *py, const cx<T> *pz, T *proj,
                                hipStream_t st) {
  const uint64_t P = (uint64_t)g.nx * g.ny;
  hipLaunchKernelGGL((ewald_partial_kernel<T>), dim3(ceil_div(P, 256)), dim3(256), 0, st, g, px, py, pz, proj);
  CYR_LAUNCHED();
  return CYR_OK;
}

template <typename T>
cyr_status launch_scatter3d(const EwaldGeom &g, const cx<T> *px, const cx<T> *py, const cx<T> *pz, T *scatter,
                            hipStream_t st) {
  const uint64_t V = (uint64_t)g.nx * g.ny * g.nz;
  hipLaunchKernelGGL((scatter3d_kernel<T>), dim3(ceil_div(V, 256)), dim3(256), 0, st, g, px, py, pz, scatter);
  CYR_LAUNCHED();
  return CYR_OK;
}

template <typename T> cyr_status launch_ewald_full(const EwaldGeom &g, const T *scatter, T *proj, hipStream_t st) {
  const uint64_t P = (uint64_t)g.nx * g.ny;
  hipLaunchKernelGGL((ewald_full_kernel<T>), dim3(ceil_div(P, 256)), dim3(256), 0, st, g, scatter, proj);
  CYR_LAUNCHED();
  return CYR_OK;
}

template <typename T>
cyr_status launch_dc_replace(cx<T> *const *vols, uint32_t count, uint32_t nx, uint32_t ny, uint32_t nz,
                             hipStream_t st) {
  for (uint32_t c0 = 0; c0 < count; c0 += 8) {
    DcVols<T> dv{};
    const uint32_t nc = count - c0 < 8 ? count - c0 : 8;
    for (uint32_t i = 0; i < nc; i++) dv.v[i] = vols[c0 + i];
    hipLaunchKernelGGL((dc_replace_kernel<T>), dim3(1), dim3(64), 0, st, dv, nc, nx, ny, nz);
    CYR_LAUNCHED();
  }
  return CYR_OK;
}

template <typename T> cyr_status launch_igor_shift(cx<T> *vol, uint32_t nx, uint32_t ny, uint32_t nz, hipStream_t st) {
  const uint64_t V = (uint64_t)nx * ny * nz;
  hipLaunchKernelGGL((igor_shift_kernel<T>), dim3(ceil_div(V, 256)), dim3(256), 0, st, vol, nx, ny, nz);
  CYR_LAUNCHED();
  return CYR_OK;
}

#define CYR_INST(T)                                                                                                  \
  template cyr_status launch_ewald_partial<T>(const EwaldGeom &, const cx<T> *, const cx<T> *, const cx<T> *, T *,   \
                                              hipStream_t);                                                          \
  template cyr_status launch_scatter3d<T>(const EwaldGeom &, const cx<T> *, const cx<T> *, const cx<T> *, T *,       \
                                          hipStream_t);                                                              \
  template cyr_status launch_ewald_full<T>(const EwaldGeom &, const T *, T *, hipStream_t);                          \
  template cyr_status launch_dc_replace<T>(cx<T> *const *, uint32_t, uint32_t, uint32_t, uint32_t, hipStream_t);     \
  template cyr_status launch_igor_shift<T>(cx<T> *, uint32_t, uint32_t, uint32_t, hipStream_t);
CYR_INST(float)
CYR_INST(double)
#undef CYR_INST

}  // namespace cyr
