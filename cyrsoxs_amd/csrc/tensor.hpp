// Per-voxel uniaxial dielectric tensor (shared by the polarization kernels and
// the fused E-basis polarization + X-FFT kernel).  See polarization.hip for
// the formulation.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace cyr {

// Device element type of a native-precision morphology.
template <typename T> struct NativeVox;
template <> struct NativeVox<float> { using type = float4; };
template <> struct NativeVox<double> { using type = double4; };

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

}  // namespace cyr
