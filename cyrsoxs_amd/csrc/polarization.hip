// Per-voxel uniaxial polarization (replaces computePolarization and
// computeNt* / computePolarizationVectorMorphologyLowMemory:
// include/uniaxial.h:58-521, src/cudaMain.cu:245-375).
//
// For each voxel the dielectric contrast tensor N = sum_m A_m (symmetric, 6
// complex entries 00,01,02,11,12,22) is accumulated over materials:
//   Euler:  s = (cos psi sin theta, sin psi sin theta, cos theta),
//           phi_a = Vfrac S, phi_u = Vfrac - phi_a
//           A_ii = phi_a (a s_i^2 + b (s_j^2 + s_k^2)) + phi_u c
//           A_ij = phi_a d s_i s_j
//   Vector: s = (sx, sy, sz) (not unit), phi_u = w
//           A_ii = a s_i^2 + b (s_j^2 + s_k^2) + phi_u c,   A_ij = d s_i s_j
// with a = n_par^2 - 1, b = n_perp^2 - 1, c = (n_par + 2 n_perp)^2/9 - 1,
// d = n_par^2 - n_perp^2.  This is the reference's tensor with its "- Vfrac"
// vacuum term cancelled analytically (phi_a |s|^2 + phi_u - Vfrac = 0 for Euler
// angles; exact for vectors), which keeps FP32 free of 1 - n^2 cancellation.
// Then p = N E / (4 pi) for each requested E basis vector; the MATERIAL frame
// rotates p by R^T; the Hanning window (Algorithm 0 only) multiplies p.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "kernels.hpp"

namespace cyr {

namespace {

__device__ inline void sincos_t(float x, float *s, float *c) {
  *s = __sinf(x);
  *c = __cosf(x);
}
__device__ inline void sincos_t(double x, double *s, double *c) { sincos(x, s, c); }

template <typename T, typename S, int MORPH>
__device__ inline void accumulate_tensor(const S *__restrict__ vox, const MatConst<T> *__restrict__ mats,
                                         uint32_t M, uint64_t V, uint64_t idx, T nr[6], T ni[6]) {
#pragma unroll
  for (int j = 0; j < 6; j++) nr[j] = ni[j] = 0;
  for (uint32_t m = 0; m < M; m++) {
    const MatConst<T> mc = mats[m];
    T v[4];
    VoxLoad<T, S>::get(vox, (uint64_t)m * V + idx, v);
    T sx, sy, sz, pa, pu;
    if (MORPH == 1) {
      sx = v[0];
      sy = v[1];
      sz = v[2];
      pu = v[3];
      pa = (T)1;
    } else {
      T st, ct, sp, cp;
      sincos_t(v[1], &st, &ct);
      sincos_t(v[2], &sp, &cp);
      sx = cp * st;
      sy = sp * st;
      sz = ct;
      pa = v[3] * v[0];
      pu = v[3] - pa;
    }
    const T xx = sx * sx, yy = sy * sy, zz = sz * sz;
    const T xy = pa * sx * sy, xz = pa * sx * sz, yz = pa * sy * sz;
    // diagonal
    nr[0] += pa * (mc.ar * xx + mc.br * (yy + zz)) + pu * mc.cr;
    ni[0] += pa * (mc.ai * xx + mc.bi * (yy + zz)) + pu * mc.ci;
    nr[3] += pa * (mc.ar * yy + mc.br * (xx + zz)) + pu * mc.cr;
    ni[3] += pa * (mc.ai * yy + mc.bi * (xx + zz)) + pu * mc.ci;
    nr[5] += pa * (mc.ar * zz + mc.br * (xx + yy)) + pu * mc.cr;
    ni[5] += pa * (mc.ai * zz + mc.bi * (xx + yy)) + pu * mc.ci;
    // off-diagonal
    nr[1] += mc.dr * xy;
    ni[1] += mc.di * xy;
    nr[2] += mc.dr * xz;
    ni[2] += mc.di * xz;
    nr[4] += mc.dr * yz;
    ni[4] += mc.di * yz;
  }
}

// p = N e / 4pi (N symmetric: [0 1 2; 1 3 4; 2 4 5])
template <typename T>
__device__ inline void apply_tensor(const T nr[6], const T ni[6], T ex, T ey, T ez, cx<T> p[3]) {
  const T k = (T)(1.0 / (4.0 * M_PI));
  p[0] = mk<T>((nr[0] * ex + nr[1] * ey + nr[2] * ez) * k, (ni[0] * ex + ni[1] * ey + ni[2] * ez) * k);
  p[1] = mk<T>((nr[1] * ex + nr[3] * ey + nr[4] * ez) * k, (ni[1] * ex + ni[3] * ey + ni[4] * ez) * k);
  p[2] = mk<T>((nr[2] * ex + nr[4] * ey + nr[5] * ez) * k, (ni[2] * ex + ni[4] * ey + ni[5] * ez) * k);
}

// p <- R^T p  (rotate<true>, include/Rotation.h:376-384)
template <typename T> __device__ inline void rotate_t(const Mat3<T> &R, cx<T> p[3]) {
  cx<T> q[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    q[i].x = R.m[0 + i] * p[0].x + R.m[3 + i] * p[1].x + R.m[6 + i] * p[2].x;
    q[i].y = R.m[0 + i] * p[0].y + R.m[3 + i] * p[1].y + R.m[6 + i] * p[2].y;
  }
  p[0] = q[0];
  p[1] = q[1];
  p[2] = q[2];
}

// Block = a run of `blockDim.x` voxels inside one X row (scalar row decode).
template <typename T, typename S, int MORPH, int NB, bool HANN, bool MATFRAME>
__global__ void __launch_bounds__(256) polar_kernel(PolarArgs<T> a, const S *__restrict__ vox) {
  const uint32_t row = blockIdx.x / a.blocks_per_row;
  const uint32_t X = (blockIdx.x - row * a.blocks_per_row) * blockDim.x + threadIdx.x;
  if (X >= a.nx) return;
  const uint64_t idx = (uint64_t)row * a.nx + X;
  T nr[6], ni[6];
  accumulate_tensor<T, S, MORPH>(vox, a.mats, a.M, a.V, idx, nr, ni);
  T w = (T)1;
  if (HANN) {
    const uint32_t Y = row % a.ny, Z = row / a.ny;
    w = a.hx[X] * a.hy[Y] * a.hz[Z];
  }
#pragma unroll
  for (int b = 0; b < NB; b++) {
    cx<T> p[3];
    apply_tensor<T>(nr, ni, a.e[b].x, a.e[b].y, a.e[b].z, p);
    if (MATFRAME) rotate_t<T>(a.R, p);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      if (HANN) p[c] = cscale<T>(p[c], w);
      a.out[b * 3 + c][idx] = p[c];
    }
  }
}

// Nt = sum_m A_m, stored as 6 complex volumes (Algorithm 1).
template <typename T, typename S, int MORPH>
__global__ void __launch_bounds__(256) nt_kernel(PolarArgs<T> a, const S *__restrict__ vox, cx<T> *__restrict__ nt) {
  const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.V) return;
  T nr[6], ni[6];
  accumulate_tensor<T, S, MORPH>(vox, a.mats, a.M, a.V, idx, nr, ni);
#pragma unroll
  for (int j = 0; j < 6; j++) nt[(uint64_t)j * a.V + idx] = mk<T>(nr[j], ni[j]);
}

// p = Nt E / 4pi (+ R^T for MATERIAL); no Hanning (include/uniaxial.h:441-521).
template <typename T, int NB, bool MATFRAME>
__global__ void __launch_bounds__(256) polar_nt_kernel(PolarArgs<T> a, const cx<T> *__restrict__ nt) {
  const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.V) return;
  T nr[6], ni[6];
#pragma unroll
  for (int j = 0; j < 6; j++) {
    const cx<T> v = nt[(uint64_t)j * a.V + idx];
    nr[j] = v.x;
    ni[j] = v.y;
  }
#pragma unroll
  for (int b = 0; b < NB; b++) {
    cx<T> p[3];
    apply_tensor<T>(nr, ni, a.e[b].x, a.e[b].y, a.e[b].z, p);
    if (MATFRAME) rotate_t<T>(a.R, p);
#pragma unroll
    for (int c = 0; c < 3; c++) a.out[b * 3 + c][idx] = p[c];
  }
}

template <typename T, typename S, int MORPH, int NB, bool HANN>
hipError_t launch_polar_s(const PolarArgs<T> &a, const void *vox, bool matframe, hipStream_t st) {
  dim3 grid(a.blocks_per_row * a.rows), block(a.block);
  if (matframe)
    hipLaunchKernelGGL((polar_kernel<T, S, MORPH, NB, HANN, true>), grid, block, 0, st, a, (const S *)vox);
  else
    hipLaunchKernelGGL((polar_kernel<T, S, MORPH, NB, HANN, false>), grid, block, 0, st, a, (const S *)vox);
  return hipGetLastError();
}

template <typename T, typename S>
hipError_t launch_polar_dispatch(const PolarArgs<T> &a, const void *vox, int morph, int nb, bool hann,
                                 bool matframe, hipStream_t st) {
#define CYR_P(MO, NBV, H) \
  if (morph == MO && nb == NBV && hann == H) return launch_polar_s<T, S, MO, NBV, H>(a, vox, matframe, st);
  CYR_P(0, 1, false)
  CYR_P(0, 1, true)
  CYR_P(0, 2, false)
  CYR_P(0, 2, true)
  CYR_P(1, 1, false)
  CYR_P(1, 1, true)
  CYR_P(1, 2, false)
  CYR_P(1, 2, true)
#undef CYR_P
  return hipErrorInvalidValue;
}

template <typename T> struct NativeVox;
template <> struct NativeVox<float> { using type = float4; };
template <> struct NativeVox<double> { using type = double4; };

}  // namespace

template <typename T>
cyr_status launch_polarization(PolarArgs<T> a, const void *vox, int storage, int morph, int nb, bool hann,
                               bool matframe, hipStream_t st) {
  if (nb < 1 || nb > 2) return fail(CYR_E_INVALID, "polarization: 1 or 2 E bases");
  if (matframe && nb != 1) return fail(CYR_E_INVALID, "polarization: MATERIAL frame needs one basis");
  a.rows = (uint32_t)(a.V / a.nx);
  a.block = a.nx >= 256 ? 256 : ((a.nx + 63) / 64) * 64;
  a.blocks_per_row = ceil_div(a.nx, a.block);
  if ((uint64_t)a.blocks_per_row * a.rows > 0x7fffffffull) return fail(CYR_E_UNSUPPORTED, "grid too large");
  hipError_t e;
  if (storage == CYR_STORE_FP16) e = launch_polar_dispatch<T, half4>(a, vox, morph, nb, hann, matframe, st);
  else if (storage == CYR_STORE_BF16) e = launch_polar_dispatch<T, bf164>(a, vox, morph, nb, hann, matframe, st);
  else e = launch_polar_dispatch<T, typename NativeVox<T>::type>(a, vox, morph, nb, hann, matframe, st);
  if (e != hipSuccess) return hip_fail(e, "polar_kernel", __FILE__, __LINE__);
  return CYR_OK;
}

template <typename T>
cyr_status launch_nt(PolarArgs<T> a, const void *vox, int storage, int morph, cx<T> *nt, hipStream_t st) {
  dim3 grid(ceil_div(a.V, 256)), block(256);
  using NV = typename NativeVox<T>::type;
#define CYR_NT(SS)                                                                                   \
  if (morph == 1) hipLaunchKernelGGL((nt_kernel<T, SS, 1>), grid, block, 0, st, a, (const SS *)vox, nt); \
  else hipLaunchKernelGGL((nt_kernel<T, SS, 0>), grid, block, 0, st, a, (const SS *)vox, nt);
  if (storage == CYR_STORE_FP16) { CYR_NT(half4) }
  else if (storage == CYR_STORE_BF16) { CYR_NT(bf164) }
  else { CYR_NT(NV) }
#undef CYR_NT
  CYR_LAUNCHED();
  return CYR_OK;
}

template <typename T>
cyr_status launch_polarization_nt(PolarArgs<T> a, const cx<T> *nt, int nb, bool matframe, hipStream_t st) {
  dim3 grid(ceil_div(a.V, 256)), block(256);
  if (nb == 1 && matframe) hipLaunchKernelGGL((polar_nt_kernel<T, 1, true>), grid, block, 0, st, a, nt);
  else if (nb == 1) hipLaunchKernelGGL((polar_nt_kernel<T, 1, false>), grid, block, 0, st, a, nt);
  else if (nb == 2 && !matframe) hipLaunchKernelGGL((polar_nt_kernel<T, 2, false>), grid, block, 0, st, a, nt);
  else return fail(CYR_E_INVALID, "polarization_nt: unsupported basis/frame combination");
  CYR_LAUNCHED();
  return CYR_OK;
}

template cyr_status launch_polarization<float>(PolarArgs<float>, const void *, int, int, int, bool, bool, hipStream_t);
template cyr_status launch_polarization<double>(PolarArgs<double>, const void *, int, int, int, bool, bool, hipStream_t);
template cyr_status launch_nt<float>(PolarArgs<float>, const void *, int, int, cx<float> *, hipStream_t);
template cyr_status launch_nt<double>(PolarArgs<double>, const void *, int, int, cx<double> *, hipStream_t);
template cyr_status launch_polarization_nt<float>(PolarArgs<float>, const cx<float> *, int, bool, hipStream_t);
template cyr_status launch_polarization_nt<double>(PolarArgs<double>, const cx<double> *, int, bool, hipStream_t);

}  // namespace cyr
