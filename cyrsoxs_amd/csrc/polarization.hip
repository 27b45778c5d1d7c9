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
#include "tensor.hpp"

namespace cyr {

namespace {

// Block = a run of `blockDim.x` voxels inside one X row (scalar row decode).
template <typename T, typename S, int MORPH, int NB>
__global__ void __launch_bounds__(256) polar_kernel(PolarArgs<T> a, const S *__restrict__ vox, int hann,
                                                    int matframe) {
  const uint32_t row = blockIdx.x / a.blocks_per_row;
  const uint32_t X = (blockIdx.x - row * a.blocks_per_row) * blockDim.x + threadIdx.x;
  if (X >= a.nx) return;
  const uint64_t idx = (uint64_t)row * a.nx + X;
  T nr[6], ni[6];
  accumulate_tensor<T, S, MORPH>(vox, a.mats, a.M, a.V, idx, nr, ni);
  T w = (T)1;
  if (hann) {
    const uint32_t Y = row % a.ny, Z = row / a.ny;
    w = a.hx[X] * a.hy[Y] * a.hz[Z];
  }
#pragma unroll
  for (int b = 0; b < NB; b++) {
    cx<T> p[3];
    apply_tensor<T>(nr, ni, a.e[b].x, a.e[b].y, a.e[b].z, p);
    if (matframe) rotate_t<T>(a.R, p);
#pragma unroll
    for (int c = 0; c < 3; c++) a.out[b * 3 + c][idx] = cscale<T>(p[c], w);
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

template <typename T, typename S>
hipError_t launch_polar_dispatch(const PolarArgs<T> &a, const void *vox, int morph, int nb, bool hann,
                                 bool matframe, hipStream_t st) {
  const dim3 grid(a.blocks_per_row * a.rows), block(a.block);
  const int h = hann ? 1 : 0, mf = matframe ? 1 : 0;
  const S *v = (const S *)vox;
  if (morph == 0 && nb == 1) hipLaunchKernelGGL((polar_kernel<T, S, 0, 1>), grid, block, 0, st, a, v, h, mf);
  else if (morph == 0 && nb == 2) hipLaunchKernelGGL((polar_kernel<T, S, 0, 2>), grid, block, 0, st, a, v, h, mf);
  else if (morph == 1 && nb == 1) hipLaunchKernelGGL((polar_kernel<T, S, 1, 1>), grid, block, 0, st, a, v, h, mf);
  else if (morph == 1 && nb == 2) hipLaunchKernelGGL((polar_kernel<T, S, 1, 2>), grid, block, 0, st, a, v, h, mf);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

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
