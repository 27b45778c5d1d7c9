// Per-voxel fields + X FFT: the first kernel of both fused executors.
//
// A block holds pp rows of tpp threads (F::kE elements each); WG = 64 makes
// every row exchange wave-local (the barriers compile to LDS waits).  LDS: the
// pp x row tile, plus the X twiddle table: the 4-7 FFTs per row read their
// twiddles from LDS (the pair kernel's vector-memory instructions drop from 124 to
// 34 per wave: 731 -> 621 us at 512^2 x 128; the tensor kernel 780 -> 742 us).
// Looping a block over several rows spills registers.
//
// What a voxel contributes is selected by MODE:
//   kXTensor  the five E-basis tensor fields N'00, N'01, N'02, N'11, N'12 of one
//             energy (complex, summed over all materials);
//   kXPair    the seven energy-independent real basis fields of two materials
//             m0, m0+1, packed as (basis of m0) + i (basis of m0+1);
//   kXSingle  the seven basis fields of material m0 packed into four volumes
//             (alpha + i beta, q00 + i q11, q01 + i q02, q12).
// Basis fields of a material (all windowed and scaled by 1/4pi; s' = Q^T s):
//   alpha = phi_a (Euler) | |s|^2 (vector),  beta = phi_u (Euler) | phi (vector),
//   q00 = phi_a s'0^2, q11 = phi_a s'1^2, q01 = phi_a s'0 s'1, q02 = phi_a s'0 s'2,
//   q12 = phi_a s'1 s'2   (phi_a = 1 for vector morphologies),
// so that the tensor fields of an energy are, per material,
//   N'00 += b alpha + c beta + d q00,  N'11 += b alpha + c beta + d q11,
//   N'01 += d q01,  N'02 += d q02,  N'12 += d q12
// with b = nperp^2 - 1, c = niso^2 - 1, d = npara^2 - nperp^2 (include/uniaxial.h:106-150, 215-240).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "fft.hpp"
#include "fftpolicy.hpp"
#include "kernels.hpp"
#include "stockham.hpp"
#include "tensor.hpp"

namespace cyr {

enum XMode { kXTensor = 0, kXPair = 1, kXSingle = 2 };
// Every mode runs 4-7 FFTs per row: the X twiddles are staged in LDS once per block.
template <int MODE> constexpr bool x_lds_twiddles() { return true; }

template <int MODE> constexpr int x_outputs() { return MODE == kXTensor ? kEbFields : (MODE == kXPair ? 7 : 4); }

// ---- the five tensor fields of one voxel -------------------------------------
template <typename T, typename S, int MORPH>
__device__ inline void tensor_fields(const S *__restrict__ vox, const MatConst<T> *__restrict__ mats, uint32_t M,
                                     uint64_t V, uint64_t idx, const Mat3<T> &Q, bool rot, T fr[kEbFields],
                                     T fi[kEbFields]) {
#pragma unroll
  for (int j = 0; j < kEbFields; j++) fr[j] = fi[j] = 0;
  for (uint32_t m = 0; m < M; m++) {
    const MatConst<T> mc = mats[m];
    T v[4];
    VoxLoad<T, S>::get(vox, (uint64_t)m * V + idx, v);
    T sx, sy, sz, pa, isor, isoi;
    if (MORPH == 1) {
      sx = v[0];
      sy = v[1];
      sz = v[2];
      pa = (T)1;
      const T ss = sx * sx + sy * sy + sz * sz;
      isor = mc.br * ss + mc.cr * v[3];
      isoi = mc.bi * ss + mc.ci * v[3];
    } else {
      T st, ct, sp, cp;
      sincos_t(v[1], &st, &ct);
      sincos_t(v[2], &sp, &cp);
      sx = cp * st;
      sy = sp * st;
      sz = ct;
      pa = v[3] * v[0];
      const T pu = v[3] - pa;
      isor = mc.br * pa + mc.cr * pu;
      isoi = mc.bi * pa + mc.ci * pu;
    }
    if (rot) {  // s' = Q^T s
      const T a = Q.m[0] * sx + Q.m[3] * sy + Q.m[6] * sz;
      const T b = Q.m[1] * sx + Q.m[4] * sy + Q.m[7] * sz;
      const T c = Q.m[2] * sx + Q.m[5] * sy + Q.m[8] * sz;
      sx = a;
      sy = b;
      sz = c;
    }
    const T tx = pa * sx, ty = pa * sy;
    const T q00 = tx * sx, q01 = tx * sy, q02 = tx * sz, q11 = ty * sy, q12 = ty * sz;
    fr[0] += isor + mc.dr * q00;
    fi[0] += isoi + mc.di * q00;
    fr[1] += mc.dr * q01;
    fi[1] += mc.di * q01;
    fr[2] += mc.dr * q02;
    fi[2] += mc.di * q02;
    fr[3] += isor + mc.dr * q11;
    fi[3] += isoi + mc.di * q11;
    fr[4] += mc.dr * q12;
    fi[4] += mc.di * q12;
  }
}

// ---- the seven real basis fields of one material at one voxel ------------------
// b = (alpha, beta, q00, q11, q01, q02, q12)
template <typename T, typename S, int MORPH>
__device__ inline void basis_fields(const S *__restrict__ vox, uint64_t V, uint32_t m, uint64_t idx, const Mat3<T> &Q,
                                    bool rot, T b[7]) {
  T v[4];
  VoxLoad<T, S>::get(vox, (uint64_t)m * V + idx, v);
  T sx, sy, sz, pa;
  if (MORPH == 1) {
    sx = v[0];
    sy = v[1];
    sz = v[2];
    pa = (T)1;
    b[0] = sx * sx + sy * sy + sz * sz;
    b[1] = v[3];
  } else {
    T st, ct, sp, cp;
    sincos_t(v[1], &st, &ct);
    sincos_t(v[2], &sp, &cp);
    sx = cp * st;
    sy = sp * st;
    sz = ct;
    pa = v[3] * v[0];
    b[0] = pa;
    b[1] = v[3] - pa;
  }
  if (rot) {
    const T a = Q.m[0] * sx + Q.m[3] * sy + Q.m[6] * sz;
    const T c = Q.m[1] * sx + Q.m[4] * sy + Q.m[7] * sz;
    const T d = Q.m[2] * sx + Q.m[5] * sy + Q.m[8] * sz;
    sx = a;
    sy = c;
    sz = d;
  }
  const T tx = pa * sx, ty = pa * sy;
  b[2] = tx * sx;
  b[3] = ty * sy;
  b[4] = tx * sy;
  b[5] = tx * sz;
  b[6] = ty * sz;
}

template <typename T, typename S, int MORPH, int MODE>
__device__ inline void voxel_fields(const S *__restrict__ vox, const PolxArgs<T> &a, uint64_t idx, bool rot,
                                    T fr[x_outputs<MODE>()], T fi[x_outputs<MODE>()]) {
  if constexpr (MODE == kXTensor) {
    tensor_fields<T, S, MORPH>(vox, a.mats, a.M, a.V, idx, a.Q, rot, fr, fi);
  } else if constexpr (MODE == kXPair) {
    T b0[7], b1[7];
    basis_fields<T, S, MORPH>(vox, a.V, a.m0, idx, a.Q, rot, b0);
    basis_fields<T, S, MORPH>(vox, a.V, a.m0 + 1, idx, a.Q, rot, b1);
#pragma unroll
    for (int r = 0; r < 7; r++) {
      fr[r] = b0[r];
      fi[r] = b1[r];
    }
  } else {
    T b[7];
    basis_fields<T, S, MORPH>(vox, a.V, a.m0, idx, a.Q, rot, b);
    fr[0] = b[0];
    fi[0] = b[1];
    fr[1] = b[2];
    fi[1] = b[3];
    fr[2] = b[4];
    fi[2] = b[5];
    fr[3] = b[6];
    fi[3] = 0;
  }
}

// ---- fields + X FFT --------------------------------------------------------------
template <typename T, typename S, class F, int WG, int MODE>
__global__ void __launch_bounds__(WG, (MODE == kXTensor && F::kE <= 8) ? 4 : 2)
    xfields_kernel(StagePlan P, PolxArgs<T> a, const S *__restrict__ vox, int morph, int hann) {
  constexpr int E = F::kE, NO = x_outputs<MODE>();
  extern __shared__ unsigned char smem_raw[];
  cx<T> *lds = reinterpret_cast<cx<T> *>(smem_raw);
  const F fft(P, a.tpp, a.nx, true);
  const uint32_t tpp = a.tpp;
  const uint32_t t = threadIdx.x % tpp, p = threadIdx.x / tpp;
  const uint32_t row = blockIdx.x * a.pp + p;  // row = y + ny z
  const bool active = row < a.rows;
  const uint32_t y = active ? row % a.ny : 0, z = active ? row / a.ny : 0;
  const T k4 = (T)(1.0 / (4.0 * M_PI));
  const T wyz = hann ? a.hy[y] * a.hz[z] * k4 : k4;
  const bool rot = a.rot != 0;
  // twiddles staged in LDS.  The X transform (the first of the three) keeps
  // plain products: this kernel is register- and latency-bound, compensated
  // products here cost 12% of its time (512^2 x 128, material pair) and left the
  // FP32 pattern error unchanged to three digits; the later Y and Z passes
  // apply them.
  const cx<T> *twp = a.twx;
  if constexpr (x_lds_twiddles<MODE>()) {
    cx<T> *ltw = lds + (size_t)a.pp * a.pitch;
    // (run-time plans launch pp x tpp <= WG threads: stride by the actual block)
    for (uint32_t i = threadIdx.x; i < a.nx; i += blockDim.x) ltw[i] = a.twx[i];
    lds_sync();  // (single-wave blocks: an LDS wait only)
    twp = ltw;
  }
  if constexpr (MODE == kXPair) {
    // Material pair: a compact per-material state per element (window-weighted
    // phi_a, the second scalar field, the director s'; 5-6 values instead of the
    // 7 fields) and each output's fields formed just before its FFT, so the
    // loads are consumed early and the field products interleave with the FFTs
    // (512-point rows, 2 materials: 577 us against 625 us with all 7 x 2 fields
    // formed up front).
    auto pair_rows = [&](auto mtag) {
      constexpr int MORPH = decltype(mtag)::value;
      T wal[MORPH == 1 ? 2 : 1][E], wpa[2][E], wbe[2][E], sxa[2][E], sya[2][E], sza[2][E];
#pragma unroll
      for (int i = 0; i < E; i++) {
        const uint32_t x = fft.in_pos(t, i);
        const T w = active ? (hann ? a.hx[x] * wyz : wyz) : (T)0;
#pragma unroll
        for (int m = 0; m < 2; m++) {
          T vv[4] = {0, 0, 0, 0};
          if (active) VoxLoad<T, S>::get(vox, (uint64_t)(a.m0 + m) * a.V + (uint64_t)row * a.nx + x, vv);
          T sx, sy, sz;
          if constexpr (MORPH == 1) {  // alpha = |s|^2, beta = phi_u, phi_a = 1
            sx = vv[0];
            sy = vv[1];
            sz = vv[2];
            wal[m][i] = w * (sx * sx + sy * sy + sz * sz);
            wpa[m][i] = w;
            wbe[m][i] = w * vv[3];
          } else {  // alpha = phi_a = Vfrac S, beta = Vfrac - phi_a
            T st, ct, sp, cp;
            sincos_t(vv[1], &st, &ct);
            sincos_t(vv[2], &sp, &cp);
            sx = cp * st;
            sy = sp * st;
            sz = ct;
            const T pa = vv[3] * vv[0];
            wpa[m][i] = w * pa;
            wbe[m][i] = w * (vv[3] - pa);
          }
          if (rot) {  // s' = Q^T s
            const T q0 = a.Q.m[0] * sx + a.Q.m[3] * sy + a.Q.m[6] * sz;
            const T q1 = a.Q.m[1] * sx + a.Q.m[4] * sy + a.Q.m[7] * sz;
            const T q2 = a.Q.m[2] * sx + a.Q.m[5] * sy + a.Q.m[8] * sz;
            sx = q0;
            sy = q1;
            sz = q2;
          }
          sxa[m][i] = sx;
          sya[m][i] = sy;
          sza[m][i] = sz;
        }
      }
      // outputs c = 0..6: (alpha, beta, q00, q11, q01, q02, q12) of material m0
      // (real part) and m0 + 1 (imaginary part), q_ij = phi_a s'_i s'_j
      static_for<0, NO>([&](auto cc) {
        constexpr int c = decltype(cc)::value;
        cx<T> v[E];
#pragma unroll
        for (int i = 0; i < E; i++) {
          T f[2];
#pragma unroll
          for (int m = 0; m < 2; m++) {
            if constexpr (c == 0) f[m] = MORPH == 1 ? wal[MORPH == 1 ? m : 0][i] : wpa[m][i];
            else if constexpr (c == 1) f[m] = wbe[m][i];
            else if constexpr (c == 2) f[m] = wpa[m][i] * sxa[m][i] * sxa[m][i];
            else if constexpr (c == 3) f[m] = wpa[m][i] * sya[m][i] * sya[m][i];
            else if constexpr (c == 4) f[m] = wpa[m][i] * sxa[m][i] * sya[m][i];
            else if constexpr (c == 5) f[m] = wpa[m][i] * sxa[m][i] * sza[m][i];
            else f[m] = wpa[m][i] * sya[m][i] * sza[m][i];
          }
          v[i] = mk<T>(f[0], f[1]);
        }
        fft.template run<false>(v, t, p, lds, twp, nullptr);
        if (active) {
          cx<T> *o = a.out[c] + (uint64_t)row * a.nx;
#pragma unroll
          for (int i = 0; i < E; i++) o[fft.out_pos(t, i)] = v[i];
        }
      });
    };
    if (morph == 1) pair_rows(std::integral_constant<int, 1>{});
    else pair_rows(std::integral_constant<int, 0>{});
    return;
  }
  T fre[NO][E], fim[NO][E];
#pragma unroll
  for (int i = 0; i < E; i++) {
    const uint32_t x = fft.in_pos(t, i);
    T fr[NO], fi[NO];
    if (active) {
      const uint64_t idx = (uint64_t)row * a.nx + x;
      if (morph == 1) voxel_fields<T, S, 1, MODE>(vox, a, idx, rot, fr, fi);
      else voxel_fields<T, S, 0, MODE>(vox, a, idx, rot, fr, fi);
      const T w = hann ? a.hx[x] * wyz : wyz;
#pragma unroll
      for (int c = 0; c < NO; c++) {
        fr[c] *= w;
        fi[c] *= w;
      }
    } else {
#pragma unroll
      for (int c = 0; c < NO; c++) fr[c] = fi[c] = 0;
    }
#pragma unroll
    for (int c = 0; c < NO; c++) {
      fre[c][i] = fr[c];
      fim[c][i] = fi[c];
    }
  }
  static_for<0, NO>([&](auto cc) {
    constexpr int c = decltype(cc)::value;
    cx<T> v[E];
#pragma unroll
    for (int i = 0; i < E; i++) v[i] = mk<T>(fre[c][i], fim[c][i]);
    fft.template run<false>(v, t, p, lds, twp, nullptr);
    if (active) {
      cx<T> *o = a.out[c] + (uint64_t)row * a.nx;
#pragma unroll
      for (int i = 0; i < E; i++) o[fft.out_pos(t, i)] = v[i];
    }
    // (run() ends with a barrier after its last LDS reads, so the next field's
    // first LDS store cannot overtake them)
  });
}

// ---- plan helpers shared by the fused executors ------------------------------
// Elements-per-thread instantiations: FP32 up to 24 (radix lcm x 2^k), FP64 up to 8.
template <typename T> constexpr bool e_allowed(int E) { return sizeof(T) == 4 || E <= 8; }

// Launch shape of the compile-time X plans: rows of <= 64 threads share a wave.
constexpr int ct_polx_wg(int N, int E) { return N / E <= 64 ? 64 : N / E; }

// Elements per thread for a pencil of length n: the radix lcm, doubled while the
// pencil would need more than `tpp_cap` threads.
inline uint32_t choose_e(uint32_t n, uint32_t lcm, uint32_t tpp_cap, uint32_t e_cap) {
  uint32_t E = lcm;
  while (n / E > tpp_cap && E * 2 <= e_cap && n % (E * 2) == 0) E *= 2;
  return E;
}

// Elements-per-thread values with a compiled run-time-plan kernel.
inline bool e_instantiated(uint32_t E, int precision) {
  switch (E) {
    case 2: case 3: case 4: case 5: case 6: case 8: return true;
    case 12: case 16: case 24: return precision != CYR_FP64;
    default: return false;
  }
}

// The compile-time plan applies (its elements per thread are compiled for T).
inline bool ct_usable(uint32_t n, int precision) {
  return ct_length(n) && (precision != CYR_FP64 || ct_elements(n) <= 8);
}

template <typename T, typename S, int E, int MODE>
hipError_t launch_xfields_rt(dim3 grid, dim3 block, size_t lds, hipStream_t st, const StagePlan &P,
                             const PolxArgs<T> &a, const void *vox, int morph, int hann) {
  hipLaunchKernelGGL((xfields_kernel<T, S, RtFft<T, E>, 256, MODE>), grid, block, lds, st, P, a, (const S *)vox,
                     morph, hann);
  return hipGetLastError();
}

template <typename T, typename S, int MODE>
hipError_t launch_xfields_rt_e(uint32_t E, dim3 grid, dim3 block, size_t lds, hipStream_t st, const StagePlan &P,
                               const PolxArgs<T> &a, const void *vox, int morph, int hann, bool *found) {
  *found = true;
#define CYR_PX(EV)                                                                                             \
  case EV:                                                                                                     \
    if constexpr (e_allowed<T>(EV))                                                                            \
      return launch_xfields_rt<T, S, EV, MODE>(grid, block, lds, st, P, a, vox, morph, hann);                 \
    break;
  switch (E) {
    CYR_PX(2)
    CYR_PX(3)
    CYR_PX(4)
    CYR_PX(5)
    CYR_PX(6)
    CYR_PX(8)
    CYR_PX(12)
    CYR_PX(16)
    CYR_PX(24)
    default:
      break;
  }
#undef CYR_PX
  *found = false;
  return hipSuccess;
}

template <typename T, typename S, int N, int MODE>
hipError_t launch_xfields_ct(dim3 grid, size_t lds, hipStream_t st, const StagePlan &P, const PolxArgs<T> &a,
                             const void *vox, int morph, int hann) {
  using F = typename CtPlan<T, N, true, 1>::type;
  constexpr int WG = ct_polx_wg(N, F::kE);
  hipLaunchKernelGGL((xfields_kernel<T, S, F, WG, MODE>), grid, dim3(WG), lds, st, P, a, (const S *)vox, morph,
                     hann);
  return hipGetLastError();
}

template <typename T, typename S, int MODE>
hipError_t launch_xfields_ct_n(uint32_t n, dim3 grid, size_t lds, hipStream_t st, const StagePlan &P,
                               const PolxArgs<T> &a, const void *vox, int morph, int hann, bool *found) {
  *found = true;
  switch (n) {
    case 64: return launch_xfields_ct<T, S, 64, MODE>(grid, lds, st, P, a, vox, morph, hann);
    case 128: return launch_xfields_ct<T, S, 128, MODE>(grid, lds, st, P, a, vox, morph, hann);
    case 256: return launch_xfields_ct<T, S, 256, MODE>(grid, lds, st, P, a, vox, morph, hann);
    case 384:
      if constexpr (e_allowed<T>(12)) return launch_xfields_ct<T, S, 384, MODE>(grid, lds, st, P, a, vox, morph, hann);
      break;
    case 512: return launch_xfields_ct<T, S, 512, MODE>(grid, lds, st, P, a, vox, morph, hann);
    case 1024: return launch_xfields_ct<T, S, 1024, MODE>(grid, lds, st, P, a, vox, morph, hann);
    case 2048: return launch_xfields_ct<T, S, 2048, MODE>(grid, lds, st, P, a, vox, morph, hann);
    default: break;
  }
  *found = false;
  return hipSuccess;
}

// Plans and launches the fields + X FFT kernel of one MODE.
template <typename T, int MODE>
cyr_status launch_xfields(PolxArgs<T> a, const void *vox, int storage, int morph, bool hann, hipStream_t st) {
  StagePlan P;
  uint32_t lcm;
  if (!make_stage_plan(a.nx, P, lcm)) return fail(CYR_E_UNSUPPORTED, "fused X kernel: Nx must be 2^a 3^b 5^c");
  const int prec = sizeof(T) == 8 ? CYR_FP64 : CYR_FP32;
  const bool ct = ct_usable(a.nx, prec);
  const uint32_t E = ct ? ct_elements(a.nx) : choose_e(a.nx, lcm, 256, sizeof(T) == 8 ? 8 : 16);
  a.tpp = a.nx / E;
  if (a.tpp > 256) return fail(CYR_E_UNSUPPORTED, "fused X kernel: row needs more than 256 threads");
  // compile-time plans: rows of <= 64 threads share one wave; the run-time plan
  // packs rows into 256-thread blocks
  const uint32_t wg = ct ? (a.tpp <= 64 ? 64u : a.tpp) : 256u;
  a.pp = wg / a.tpp;
  if (a.pp < 1) a.pp = 1;
  a.pitch = ct ? ct_row_pitch(a.nx) : a.nx;
  const size_t lds = ((size_t)a.pp * a.pitch + (x_lds_twiddles<MODE>() ? a.nx : 0)) * sizeof(cx<T>);
  if (lds > device_max_lds()) return fail(CYR_E_UNSUPPORTED, "fused X kernel: row does not fit in LDS");
  cyr_status s;
  a.twx = twiddles<T>(a.nx, &s);
  if (!a.twx) return s;
  const dim3 grid(ceil_div(a.rows, a.pp)), block(a.pp * a.tpp);
  bool found = false;
  hipError_t e = hipSuccess;
  auto go = [&](auto tag) {
    using S = decltype(tag);
    if (ct) e = launch_xfields_ct_n<T, S, MODE>(a.nx, grid, lds, st, P, a, vox, morph, hann ? 1 : 0, &found);
    else e = launch_xfields_rt_e<T, S, MODE>(E, grid, block, lds, st, P, a, vox, morph, hann ? 1 : 0, &found);
  };
  if constexpr (sizeof(T) == 4) {
    if (storage == CYR_STORE_FP16) go(half4{});
    else if (storage == CYR_STORE_BF16) go(bf164{});
    else go(float4{});
  } else {
    if (storage != CYR_STORE_NATIVE) return fail(CYR_E_UNSUPPORTED, "fused X kernel FP64: native morphology storage only");
    go(double4{});
  }
  if (e != hipSuccess) return hip_fail(e, "xfields_kernel", __FILE__, __LINE__);
  if (!found) return fail(CYR_E_UNSUPPORTED, "fused X kernel: no kernel for this length");
  return CYR_OK;
}

}  // namespace cyr
