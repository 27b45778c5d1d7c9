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

// DC replacement values of the six XY-transformed volumes G(kx, ky, z): the
// neighbours of bin (0,0,0) are sums over z of G at kz = 0 (in-plane
// neighbours) or kz = 1, Nz-1 (z neighbours).  One block per volume.
template <typename T> struct Six { cx<T> *v[6]; };

template <typename T>
__global__ void __launch_bounds__(256) ebasis_dc_kernel(Six<T> g6, uint32_t nx, uint32_t ny, uint32_t nz,
                                                        const cx<T> *__restrict__ twz, cx<T> *dcval) {
  __shared__ T red[256][12];
  const cx<T> *G = g6.v[blockIdx.x];
  const uint64_t P = (uint64_t)nx * ny;
  const uint64_t cols[4] = {1, nx, (uint64_t)nx - 1, (uint64_t)(ny - 1) * nx};
  T acc[12];
#pragma unroll
  for (int i = 0; i < 12; i++) acc[i] = 0;
  for (uint32_t z = threadIdx.x; z < nz; z += blockDim.x) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const cx<T> v = G[(uint64_t)z * P + cols[i]];
      acc[2 * i] += v.x;
      acc[2 * i + 1] += v.y;
    }
    if (nz > 1) {
      const cx<T> v = G[(uint64_t)z * P];
      const cx<T> w1 = twz[z % nz];                                    // kz = 1
      const cx<T> w2 = twz[(uint32_t)(((uint64_t)z * (nz - 1)) % nz)];  // kz = nz - 1
      const cx<T> a = cmul<T>(v, w1), b = cmul<T>(v, w2);
      acc[8] += a.x;
      acc[9] += a.y;
      acc[10] += b.x;
      acc[11] += b.y;
    }
  }
#pragma unroll
  for (int i = 0; i < 12; i++) red[threadIdx.x][i] = acc[i];
  __syncthreads();
  for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
#pragma unroll
      for (int i = 0; i < 12; i++) red[threadIdx.x][i] += red[threadIdx.x + s][i];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const int cnt = nz > 1 ? 6 : 4;
    T sr = 0, si = 0;
    for (int i = 0; i < cnt; i++) {
      sr += red[0][2 * i];
      si += red[0][2 * i + 1];
    }
    dcval[blockIdx.x] = mk<T>(sr / (T)cnt, si / (T)cnt);
  }
}

// E-basis Ewald: per detector pixel, the z DFT of the six XY-transformed
// volumes is evaluated only at the pixel's Ewald bin(s), then every angle's
// intensity follows from P(theta) = c P1 + s P2 (x R^T in the MATERIAL frame).
template <typename T, bool LINEAR3D, bool THREED, bool MATFRAME>
__global__ void __launch_bounds__(256) ebasis_ewald_kernel(EbasisArgs<T> a) {
  const EwaldGeom &g = a.g;
  const uint64_t P = (uint64_t)g.nx * g.ny;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P) return;
  const uint32_t Y = (uint32_t)(t / g.nx), X = (uint32_t)(t - (uint64_t)Y * g.nx);
  uint32_t z0 = 0;
  double f = 0;
  const int kind = ewald_geometry(g, X, Y, &z0, &f);
  if (kind == 0) {
    for (uint32_t an = 0; an < a.nA; an++) a.proj[(uint64_t)an * P + t] = nan_t<T>();
    return;
  }
  const uint32_t ix = unshift(X, g.nx), iy = unshift(Y, g.ny);
  const uint64_t col = (uint64_t)iy * g.nx + ix;
  constexpr int NG = LINEAR3D ? 2 : 1;
  uint32_t kz[NG];
  kz[0] = THREED ? unshift(z0, g.nz) : 0;
  if constexpr (LINEAR3D) kz[1] = unshift(z0 + 1, g.nz);
  cx<T> acc[NG][6];
#pragma unroll
  for (int q = 0; q < NG; q++)
#pragma unroll
    for (int c = 0; c < 6; c++) acc[q][c] = mk<T>(0, 0);
  if constexpr (THREED) {
    uint32_t e[NG];
#pragma unroll
    for (int q = 0; q < NG; q++) e[q] = 0;
    for (uint32_t z = 0; z < g.nz; z++) {
      cx<T> w[NG];
#pragma unroll
      for (int q = 0; q < NG; q++) w[q] = a.twz[e[q]];
#pragma unroll
      for (int c = 0; c < 6; c++) {
        const cx<T> v = a.vol[c][(uint64_t)z * P + col];
#pragma unroll
        for (int q = 0; q < NG; q++) {
          acc[q][c].x += v.x * w[q].x - v.y * w[q].y;
          acc[q][c].y += v.x * w[q].y + v.y * w[q].x;
        }
      }
#pragma unroll
      for (int q = 0; q < NG; q++) {
        e[q] += kz[q];
        if (e[q] >= g.nz) e[q] -= g.nz;
      }
    }
  } else {
#pragma unroll
    for (int c = 0; c < 6; c++) acc[0][c] = a.vol[c][col];
  }
  if (ix == 0 && iy == 0) {
#pragma unroll
    for (int q = 0; q < NG; q++)
      if (kz[q] == 0)
#pragma unroll
        for (int c = 0; c < 6; c++) acc[q][c] = a.dcval[c];
  }
  double md[NG][6];
#pragma unroll
  for (int q = 0; q < NG; q++) scatter_matrix(g, X, Y, THREED ? z0 + q : 0, md[q]);
  const T w1 = (T)f;
  for (uint32_t an = 0; an < a.nA; an++) {
    const T cc = a.coef[2 * an], ss = a.coef[2 * an + 1];
    T I[NG];
#pragma unroll
    for (int q = 0; q < NG; q++) {
      cx<T> p[3];
#pragma unroll
      for (int c = 0; c < 3; c++)
        p[c] = mk<T>(cc * acc[q][c].x + ss * acc[q][c + 3].x, cc * acc[q][c].y + ss * acc[q][c + 3].y);
      if constexpr (MATFRAME) {
        const T *R = a.rmat + 9 * (uint64_t)an;
        cx<T> r[3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
          r[i].x = R[0 + i] * p[0].x + R[3 + i] * p[1].x + R[6 + i] * p[2].x;
          r[i].y = R[0 + i] * p[0].y + R[3 + i] * p[1].y + R[6 + i] * p[2].y;
        }
        p[0] = r[0];
        p[1] = r[1];
        p[2] = r[2];
      }
      I[q] = scatter_intensity<T>(md[q], p);
    }
    T out;
    if constexpr (LINEAR3D) out = (1 - w1) * I[0] + w1 * I[1];
    else out = I[0];
    a.proj[(uint64_t)an * P + t] = out;
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

template <typename T>
cyr_status launch_ebasis_dc(cx<T> *const *g6, uint32_t nx, uint32_t ny, uint32_t nz, const cx<T> *twz, cx<T> *dcval,
                            hipStream_t st) {
  Six<T> s{};
  for (int i = 0; i < 6; i++) s.v[i] = g6[i];
  hipLaunchKernelGGL((ebasis_dc_kernel<T>), dim3(6), dim3(256), 0, st, s, nx, ny, nz, twz, dcval);
  CYR_LAUNCHED();
  return CYR_OK;
}

template <typename T> cyr_status launch_ebasis_ewald(const EbasisArgs<T> &a, bool matframe, hipStream_t st) {
  const uint64_t P = (uint64_t)a.g.nx * a.g.ny;
  const dim3 grid(ceil_div(P, 256)), block(256);
  const bool three = !a.g.enable2d;
  const bool lin = three && a.g.interpolation == 1;
#define CYR_EB(L, TH, MF) \
  if (lin == L && three == TH && matframe == MF) { hipLaunchKernelGGL((ebasis_ewald_kernel<T, L, TH, MF>), grid, block, 0, st, a); CYR_LAUNCHED(); return CYR_OK; }
  CYR_EB(true, true, false)
  CYR_EB(true, true, true)
  CYR_EB(false, true, false)
  CYR_EB(false, true, true)
  CYR_EB(false, false, false)
  CYR_EB(false, false, true)
#undef CYR_EB
  return fail(CYR_E_INVALID, "ebasis ewald: bad mode");
}

#define CYR_INST(T)                                                                                                  \
  template cyr_status launch_ewald_partial<T>(const EwaldGeom &, const cx<T> *, const cx<T> *, const cx<T> *, T *,   \
                                              hipStream_t);                                                          \
  template cyr_status launch_scatter3d<T>(const EwaldGeom &, const cx<T> *, const cx<T> *, const cx<T> *, T *,       \
                                          hipStream_t);                                                              \
  template cyr_status launch_ewald_full<T>(const EwaldGeom &, const T *, T *, hipStream_t);                          \
  template cyr_status launch_dc_replace<T>(cx<T> *const *, uint32_t, uint32_t, uint32_t, uint32_t, hipStream_t);     \
  template cyr_status launch_igor_shift<T>(cx<T> *, uint32_t, uint32_t, uint32_t, hipStream_t);                      \
  template cyr_status launch_ebasis_dc<T>(cx<T> *const *, uint32_t, uint32_t, uint32_t, const cx<T> *, cx<T> *,      \
                                          hipStream_t);                                                              \
  template cyr_status launch_ebasis_ewald<T>(const EbasisArgs<T> &, bool, hipStream_t);
CYR_INST(float)
CYR_INST(double)
#undef CYR_INST

}  // namespace cyr
