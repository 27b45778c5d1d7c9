// Fused kernels of the E-basis executor.
//
// Per (energy, k vector) the E-basis pipeline needs, for every detector pixel,
// the FFT of the two E-basis polarizations p1 = N e1 / 4pi, p2 = N e2 / 4pi
// (6 complex volumes) only at the pixel's Ewald bin(s).  Four launches:
//
//   ebasis_bins_kernel   pixel -> FFT z bin(s) of its (kx, ky) column (FP64 geometry)
//   ebasis_polx_kernel   polarization evaluated directly at the X-FFT's Stockham
//                        stage-0 positions, X FFT in LDS, X-transformed volumes Gx
//                        written once (the polarization never reaches HBM)
//   ebasis_yz_kernel     per z slice: Y FFT of a group of x columns in LDS, each
//                        output (kx, ky) accumulated into its column's z bin(s) with
//                        the z twiddle: F(kx,ky,kz) = sum_z Gy(kx,ky,z) W^(z kz).
//                        Blocks split z into chunks and write per-chunk partial sums
//                        (deterministic; no atomics).  Gx is read exactly once.
//   ebasis_angles_kernel sums the chunks, applies the DC replacement, and for every
//                        angle forms P = c P1 + s P2 (x R^T in the MATERIAL frame)
//                        and the Ewald intensity (linear in z when requested).
//
// HBM traffic per (energy, k): morphology V (16 M) + Gx write 48 V + Gx read 48 V,
// versus (16 M + 48) V + 4 x 48 V + 48 V for polarization, X FFT, Y FFT and the
// pruned z DFT as separate passes.
#include <hip/hip_runtime.h>

#include "fft.hpp"
#include "geom.hpp"
#include "kernels.hpp"
#include "stockham.hpp"
#include "tensor.hpp"

namespace cyr {

namespace {

template <typename T> __device__ inline T qnan();
template <> __device__ inline float qnan<float>() { return __int_as_float(0x7fc00000); }
template <> __device__ inline double qnan<double>() { return __longlong_as_double(0x7ff8000000000000ll); }

// ---- pixel -> FFT z bin(s) -----------------------------------------------------
__global__ void __launch_bounds__(256) ebasis_bins_kernel(EwaldGeom g, int ng, int32_t *__restrict__ kzt) {
  const uint64_t P = (uint64_t)g.nx * g.ny;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P) return;
  const uint32_t Y = (uint32_t)(t / g.nx), X = (uint32_t)(t - (uint64_t)Y * g.nx);
  uint32_t z0 = 0;
  double f = 0;
  const int kind = ewald_geometry(g, X, Y, &z0, &f);
  const uint64_t col = (uint64_t)unshift(Y, g.ny) * g.nx + unshift(X, g.nx);
  const bool three = !g.enable2d;
  kzt[col] = kind == 0 ? -1 : (int32_t)(three ? unshift(z0, g.nz) : 0);
  if (ng == 2) kzt[P + col] = kind == 2 ? (int32_t)unshift(z0 + 1, g.nz) : -1;
}

// ---- polarization + X FFT -------------------------------------------------------
template <typename T, typename S, int E>
__global__ void __launch_bounds__(256) ebasis_polx_kernel(StagePlan P, PolxArgs<T> a, const S *__restrict__ vox,
                                                          int morph, int hann) {
  extern __shared__ unsigned char smem_raw[];
  cx<T> *lds = reinterpret_cast<cx<T> *>(smem_raw);
  const uint32_t tpp = a.tpp;
  const uint32_t t = threadIdx.x % tpp, p = threadIdx.x / tpp;
  const uint32_t row = blockIdx.x * a.pp + p;  // row = y + ny z
  const bool active = row < a.rows;
  const uint32_t y = active ? row % a.ny : 0, z = active ? row / a.ny : 0;
  const uint32_t R0 = P.radix[0], m0 = P.n / R0;
  const T wyz = hann ? a.hy[y] * a.hz[z] : (T)1;
  cx<T> vall[6][E];
#pragma unroll
  for (int i = 0; i < E; i++) {
    const uint32_t x = t + ((uint32_t)i / R0) * tpp + ((uint32_t)i % R0) * m0;
    if (active) {
      T nr[6], ni[6];
      if (morph == 1) accumulate_tensor<T, S, 1>(vox, a.mats, a.M, a.V, (uint64_t)row * a.nx + x, nr, ni);
      else accumulate_tensor<T, S, 0>(vox, a.mats, a.M, a.V, (uint64_t)row * a.nx + x, nr, ni);
      const T w = hann ? a.hx[x] * wyz : (T)1;
#pragma unroll
      for (int b = 0; b < 2; b++) {
        cx<T> pv[3];
        apply_tensor<T>(nr, ni, a.e[b].x, a.e[b].y, a.e[b].z, pv);
#pragma unroll
        for (int c = 0; c < 3; c++) vall[3 * b + c][i] = cscale<T>(pv[c], w);
      }
    } else {
#pragma unroll
      for (int c = 0; c < 6; c++) vall[c][i] = mk<T>(0, 0);
    }
  }
  const LdsMap L{P.n, true};
  const uint32_t RL = P.radix[P.nstages - 1];
#pragma unroll
  for (int c = 0; c < 6; c++) {
    cx<T> v[E];
#pragma unroll
    for (int i = 0; i < E; i++) v[i] = vall[c][i];
    stockham_from_regs<T, E>(v, P, t, tpp, p, lds, L, a.twx);
    if (active) {
      cx<T> *o = a.out[c] + (uint64_t)row * a.nx;
#pragma unroll
      for (int i = 0; i < E; i++) o[final_index(t, tpp, P.n, RL, i)] = v[i];
    }
  }
}

// ---- Y FFT + z bins -----------------------------------------------------------
template <typename T, int E, int NG>
__global__ void __launch_bounds__(1024) ebasis_yz_kernel(StagePlan P, YzArgs<T> a) {
  extern __shared__ unsigned char smem_raw[];
  cx<T> *lds = reinterpret_cast<cx<T> *>(smem_raw);
  const uint32_t pp = a.pp, tpp = a.tpp;
  const uint32_t p = threadIdx.x % pp, t = threadIdx.x / pp;
  const uint32_t kx = blockIdx.x * pp + p;
  const bool active = kx < a.nx;
  const uint32_t c = blockIdx.y;
  const uint32_t zb = blockIdx.z * a.zchunk;
  const uint32_t ze = zb + a.zchunk < a.nz ? zb + a.zchunk : a.nz;
  const uint64_t P2 = (uint64_t)a.nx * a.ny;
  const uint32_t R0 = P.radix[0], m0 = P.n / R0;
  const uint32_t RL = P.radix[P.nstages - 1];
  int32_t kz[E][NG];
  uint32_t ex[E][NG];
  cx<T> acc[E][NG];
#pragma unroll
  for (int i = 0; i < E; i++) {
    const uint32_t ky = final_index(t, tpp, P.n, RL, i);
#pragma unroll
    for (int q = 0; q < NG; q++) {
      kz[i][q] = active ? a.kzt[(uint64_t)q * P2 + (uint64_t)ky * a.nx + kx] : -1;
      ex[i][q] = kz[i][q] > 0 ? (uint32_t)(((uint64_t)zb * (uint32_t)kz[i][q]) % a.nz) : 0;
      acc[i][q] = mk<T>(0, 0);
    }
  }
  const LdsMap L{pp, false};
  const cx<T> *vol = a.vol[c];
  for (uint32_t z = zb; z < ze; z++) {
    const cx<T> *base = vol + (uint64_t)z * P2 + kx;
    cx<T> v[E];
#pragma unroll
    for (int i = 0; i < E; i++) {
      const uint32_t yy = t + ((uint32_t)i / R0) * tpp + ((uint32_t)i % R0) * m0;
      v[i] = active ? base[(uint64_t)yy * a.nx] : mk<T>(0, 0);
    }
    stockham_from_regs<T, E>(v, P, t, tpp, p, lds, L, a.twy);
#pragma unroll
    for (int i = 0; i < E; i++)
#pragma unroll
      for (int q = 0; q < NG; q++) {
        if (kz[i][q] >= 0) {
          const cx<T> w = a.nz > 1 ? a.twz[ex[i][q]] : mk<T>(1, 0);  // 2D: the only bin is kz = 0
          acc[i][q].x += v[i].x * w.x - v[i].y * w.y;
          acc[i][q].y += v[i].x * w.y + v[i].y * w.x;
          uint32_t e = ex[i][q] + (uint32_t)kz[i][q];
          ex[i][q] = e >= a.nz ? e - a.nz : e;
        }
      }
  }
  if (active) {
#pragma unroll
    for (int i = 0; i < E; i++) {
      const uint32_t ky = final_index(t, tpp, P.n, RL, i);
#pragma unroll
      for (int q = 0; q < NG; q++)
        a.part[(((uint64_t)blockIdx.z * NG + q) * 6 + c) * P2 + (uint64_t)ky * a.nx + kx] = acc[i][q];
    }
  }
}

// ---- DC replacement values from the X-transformed volumes ----------------------
// F(kx,ky,kz) = sum_z sum_y Gx(kx,y,z) W_Ny^(y ky) W_Nz^(z kz) at the neighbours of
// bin (0,0,0): (1,0,0), (Nx-1,0,0), (0,1,0), (0,Ny-1,0) and, in 3D, (0,0,1), (0,0,Nz-1).
template <typename T>
__global__ void __launch_bounds__(1024) ebasis_dcx_kernel(DcxArgs<T> a) {
  __shared__ double red[1024];
  const cx<T> *G = a.vol[blockIdx.x];
  const uint64_t P2 = (uint64_t)a.nx * a.ny;
  const uint64_t n = (uint64_t)a.ny * a.nz;
  double s[12];
#pragma unroll
  for (int i = 0; i < 12; i++) s[i] = 0;
  for (uint64_t r = threadIdx.x; r < n; r += blockDim.x) {
    const uint32_t y = (uint32_t)(r % a.ny), z = (uint32_t)(r / a.ny);
    const cx<T> *row = G + (uint64_t)z * P2 + (uint64_t)y * a.nx;
    const cx<T> g1 = row[1], gm = row[a.nx - 1], g0 = row[0];
    s[0] += (double)g1.x + (double)gm.x;
    s[1] += (double)g1.y + (double)gm.y;
    const cx<T> wy1 = a.twy[y], wy2 = a.twy[(a.ny - y) % a.ny];
    const cx<T> a1 = cmul<T>(g0, wy1), a2 = cmul<T>(g0, wy2);
    s[2] += (double)a1.x + (double)a2.x;
    s[3] += (double)a1.y + (double)a2.y;
    if (a.nz > 1) {
      const cx<T> wz1 = a.twz[z], wz2 = a.twz[(a.nz - z) % a.nz];
      const cx<T> b1 = cmul<T>(g0, wz1), b2 = cmul<T>(g0, wz2);
      s[4] += (double)b1.x + (double)b2.x;
      s[5] += (double)b1.y + (double)b2.y;
    }
  }
  double tot[6];
  for (int k = 0; k < 6; k++) {
    red[threadIdx.x] = s[k];
    __syncthreads();
    for (uint32_t w = blockDim.x / 2; w > 0; w >>= 1) {
      if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    tot[k] = red[0];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double cnt = a.nz > 1 ? 6.0 : 4.0;
    a.dcval[blockIdx.x] = mk<T>((T)((tot[0] + tot[2] + tot[4]) / cnt), (T)((tot[1] + tot[3] + tot[5]) / cnt));
  }
}

// ---- chunk sums, DC, and every angle's intensity -----------------------------
template <typename T, int NG, bool MATFRAME>
__global__ void __launch_bounds__(256) ebasis_angles_kernel(AnglesArgs<T> a) {
  const EwaldGeom &g = a.g;
  const uint64_t P2 = (uint64_t)g.nx * g.ny;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P2) return;
  const uint32_t Y = (uint32_t)(t / g.nx), X = (uint32_t)(t - (uint64_t)Y * g.nx);
  uint32_t z0 = 0;
  double f = 0;
  const int kind = ewald_geometry(g, X, Y, &z0, &f);
  if (kind == 0) {
    for (uint32_t an = 0; an < a.nA; an++) a.proj[(uint64_t)an * P2 + t] = qnan<T>();
    return;
  }
  const uint32_t ix = unshift(X, g.nx), iy = unshift(Y, g.ny);
  const uint64_t col = (uint64_t)iy * g.nx + ix;
  const bool three = !g.enable2d;
  cx<T> F[NG][6];
#pragma unroll
  for (int q = 0; q < NG; q++)
#pragma unroll
    for (int c = 0; c < 6; c++) {
      cx<T> s = mk<T>(0, 0);
      for (uint32_t ch = 0; ch < a.nchunks; ch++) {
        const cx<T> v = a.part[(((uint64_t)ch * NG + q) * 6 + c) * P2 + col];
        s.x += v.x;
        s.y += v.y;
      }
      F[q][c] = s;
    }
  if (ix == 0 && iy == 0) {
#pragma unroll
    for (int q = 0; q < NG; q++) {
      const uint32_t kzq = three ? unshift(z0 + q, g.nz) : 0;
      if (kzq == 0)
#pragma unroll
        for (int c = 0; c < 6; c++) F[q][c] = a.dcval[c];
    }
  }
  double md[NG][6];
#pragma unroll
  for (int q = 0; q < NG; q++) scatter_matrix(g, X, Y, three ? z0 + q : 0, md[q]);
  const T w1 = (T)f;
  for (uint32_t an = 0; an < a.nA; an++) {
    const T cc = a.coef[2 * an], ss = a.coef[2 * an + 1];
    T I[NG];
#pragma unroll
    for (int q = 0; q < NG; q++) {
      cx<T> pv[3];
#pragma unroll
      for (int c = 0; c < 3; c++)
        pv[c] = mk<T>(cc * F[q][c].x + ss * F[q][c + 3].x, cc * F[q][c].y + ss * F[q][c + 3].y);
      if constexpr (MATFRAME) {
        const T *R = a.rmat + 9 * (uint64_t)an;
        cx<T> r[3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
          r[i].x = R[0 + i] * pv[0].x + R[3 + i] * pv[1].x + R[6 + i] * pv[2].x;
          r[i].y = R[0 + i] * pv[0].y + R[3 + i] * pv[1].y + R[6 + i] * pv[2].y;
        }
        pv[0] = r[0];
        pv[1] = r[1];
        pv[2] = r[2];
      }
      I[q] = scatter_intensity<T>(md[q], pv);
    }
    T out;
    if constexpr (NG == 2) out = (1 - w1) * I[0] + w1 * I[1];
    else out = I[0];
    a.proj[(uint64_t)an * P2 + t] = out;
  }
}

// Elements-per-thread instantiations: FP32 up to 24 (radix lcm x 2^k), FP64 up to 8.
template <typename T> constexpr bool e_allowed(int E) { return sizeof(T) == 4 || E <= 8; }

template <typename T, typename S>
hipError_t launch_polx_e(uint32_t E, dim3 grid, dim3 block, size_t lds, hipStream_t st, const StagePlan &P,
                         const PolxArgs<T> &a, const void *vox, int morph, int hann, bool *found) {
  *found = true;
#define CYR_PX(EV)                                                                                            \
  case EV:                                                                                                    \
    if constexpr (e_allowed<T>(EV)) {                                                                         \
      hipLaunchKernelGGL((ebasis_polx_kernel<T, S, EV>), grid, block, lds, st, P, a, (const S *)vox, morph, hann); \
      return hipGetLastError();                                                                               \
    }                                                                                                         \
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

template <typename T, int NG>
hipError_t launch_yz_e(uint32_t E, dim3 grid, dim3 block, size_t lds, hipStream_t st, const StagePlan &P,
                       const YzArgs<T> &a, bool *found) {
  *found = true;
#define CYR_YZ(EV)                                                                   \
  case EV:                                                                           \
    if constexpr (e_allowed<T>(EV)) {                                                \
      hipLaunchKernelGGL((ebasis_yz_kernel<T, EV, NG>), grid, block, lds, st, P, a); \
      return hipGetLastError();                                                      \
    }                                                                                \
    break;
  switch (E) {
    CYR_YZ(2)
    CYR_YZ(3)
    CYR_YZ(4)
    CYR_YZ(5)
    CYR_YZ(6)
    CYR_YZ(8)
    CYR_YZ(12)
    CYR_YZ(16)
    CYR_YZ(24)
    default:
      break;
  }
#undef CYR_YZ
  *found = false;
  return hipSuccess;
}

// z chunks of the Y/z kernel: enough blocks (x groups x 6 volumes x chunks) to
// reach target_blocks, at most one chunk per z slice.
uint32_t yz_chunks(uint32_t nx, uint32_t nz, uint32_t pp, uint32_t target_blocks) {
  const uint32_t per = ceil_div(nx, pp) * 6;
  uint32_t nch = target_blocks / per;
  if (nch < 1) nch = 1;
  return nch > nz ? nz : nch;
}

// Elements per thread for a pencil of length n: the radix lcm, doubled while the
// pencil would need more than `tpp_cap` threads.
uint32_t choose_e(uint32_t n, uint32_t lcm, uint32_t tpp_cap, uint32_t e_cap) {
  uint32_t E = lcm;
  while (n / E > tpp_cap && E * 2 <= e_cap && n % (E * 2) == 0) E *= 2;
  return E;
}

}  // namespace

// Elements-per-thread values with a compiled kernel (launch_polx_e / launch_yz_e).
static bool e_instantiated(uint32_t E, int precision) {
  switch (E) {
    case 2: case 3: case 4: case 5: case 6: case 8: return true;
    case 12: case 16: case 24: return precision != CYR_FP64;
    default: return false;
  }
}

bool ebasis_supported(uint32_t nx, uint32_t ny, int precision) {
  StagePlan P;
  uint32_t lcm;
  const size_t cs = precision == CYR_FP64 ? 16 : 8;
  const uint32_t ecap = precision == CYR_FP64 ? 8 : 16;
  if (!make_stage_plan(nx, P, lcm)) return false;
  const uint32_t ex = choose_e(nx, lcm, 256, ecap);
  if (!e_instantiated(ex, precision) || nx / ex > 256) return false;  // one X row per <= 256 threads
  if (!make_stage_plan(ny, P, lcm)) return false;
  // the Y kernel's widest case: one column per block (pp = 1)
  const uint32_t ey = choose_e(ny, lcm, 1024, ecap);
  if (!e_instantiated(ey, precision) || ny / ey > 1024) return false;
  const size_t lds = device_max_lds();
  return (size_t)nx * cs <= lds && (size_t)ny * cs <= lds && nx >= 2 && ny >= 2;
}

uint32_t ebasis_max_chunks(uint32_t nx, uint32_t nz) {
  // ebasis_yz_plan starts from 16 columns per block and only lowers it (fewer
  // blocks per slice, so fewer chunks), which makes pp = 16 the upper bound
  return yz_chunks(nx, nz, 16, kYzTargetBlocks);
}

cyr_status launch_ebasis_bins(const EwaldGeom &g, int ng, int32_t *kzt, hipStream_t st) {
  const uint64_t P = (uint64_t)g.nx * g.ny;
  hipLaunchKernelGGL(ebasis_bins_kernel, dim3(ceil_div(P, 256)), dim3(256), 0, st, g, ng, kzt);
  CYR_LAUNCHED();
  return CYR_OK;
}

template <typename T>
cyr_status launch_ebasis_polx(PolxArgs<T> a, const void *vox, int storage, int morph, bool hann, hipStream_t st) {
  StagePlan P;
  uint32_t lcm;
  if (!make_stage_plan(a.nx, P, lcm)) return fail(CYR_E_UNSUPPORTED, "E-basis: Nx must be 2^a 3^b 5^c");
  const uint32_t E = choose_e(a.nx, lcm, 256, sizeof(T) == 8 ? 8 : 16);
  a.tpp = a.nx / E;
  if (a.tpp > 256) return fail(CYR_E_UNSUPPORTED, "E-basis: X row needs more than 256 threads");
  a.pp = a.tpp >= 256 ? 1 : 256 / a.tpp;
  if (a.pp > a.rows) a.pp = a.rows;
  const size_t lds = (size_t)a.pp * a.nx * sizeof(cx<T>);
  if (lds > device_max_lds()) return fail(CYR_E_UNSUPPORTED, "E-basis: row does not fit in LDS");
  cyr_status s;
  a.twx = twiddles<T>(a.nx, &s);
  if (!a.twx) return s;
  const dim3 grid(ceil_div(a.rows, a.pp)), block(a.pp * a.tpp);
  bool found = false;
  hipError_t e = hipSuccess;
  if constexpr (sizeof(T) == 4) {
    if (storage == CYR_STORE_FP16)
      e = launch_polx_e<T, half4>(E, grid, block, lds, st, P, a, vox, morph, hann ? 1 : 0, &found);
    else if (storage == CYR_STORE_BF16)
      e = launch_polx_e<T, bf164>(E, grid, block, lds, st, P, a, vox, morph, hann ? 1 : 0, &found);
    else
      e = launch_polx_e<T, float4>(E, grid, block, lds, st, P, a, vox, morph, hann ? 1 : 0, &found);
  } else {
    if (storage != CYR_STORE_NATIVE) return fail(CYR_E_UNSUPPORTED, "E-basis FP64: native morphology storage only");
    e = launch_polx_e<T, double4>(E, grid, block, lds, st, P, a, vox, morph, hann ? 1 : 0, &found);
  }
  if (e != hipSuccess) return hip_fail(e, "ebasis_polx_kernel", __FILE__, __LINE__);
  if (!found) return fail(CYR_E_UNSUPPORTED, "E-basis: no X kernel for this length");
  return CYR_OK;
}

template <typename T> cyr_status ebasis_yz_plan(YzArgs<T> &a, uint32_t target_blocks) {
  StagePlan P;
  uint32_t lcm;
  if (!make_stage_plan(a.ny, P, lcm)) return fail(CYR_E_UNSUPPORTED, "E-basis: Ny must be 2^a 3^b 5^c");
  uint32_t pp = a.pp ? a.pp : 16;
  uint32_t E = a.e ? a.e : choose_e(a.ny, lcm, 1024 / pp, sizeof(T) == 8 ? 8 : 16);
  if (E % lcm != 0 || a.ny % E != 0) E = lcm;
  const uint32_t tpp = a.ny / E;
  while (pp > 1 && ((size_t)pp * a.ny * sizeof(cx<T>) > device_max_lds() || pp * tpp > 1024)) pp /= 2;
  if ((size_t)pp * a.ny * sizeof(cx<T>) > device_max_lds() || pp * tpp > 1024)
    return fail(CYR_E_UNSUPPORTED, "E-basis: column does not fit in LDS");
  a.e = E;
  a.tpp = tpp;
  a.pp = pp;
  if (!a.zchunk) a.zchunk = ceil_div(a.nz, yz_chunks(a.nx, a.nz, pp, target_blocks));
  a.nchunks = ceil_div(a.nz, a.zchunk);
  return CYR_OK;
}

template <typename T> cyr_status launch_ebasis_yz(YzArgs<T> a, int ng, hipStream_t st) {
  StagePlan P;
  uint32_t lcm;
  if (!make_stage_plan(a.ny, P, lcm) || !a.e || !a.pp || !a.tpp || !a.zchunk || a.ny % a.e != 0 || a.e % lcm != 0)
    return fail(CYR_E_INVALID, "E-basis: Y plan missing (call ebasis_yz_plan)");
  cyr_status s;
  a.twy = twiddles<T>(a.ny, &s);
  if (!a.twy) return s;
  const uint32_t groups = ceil_div(a.nx, a.pp);
  const dim3 grid(groups, 6, a.nchunks), block(a.pp * a.tpp);
  const size_t lds = (size_t)a.pp * a.ny * sizeof(cx<T>);
  bool found = false;
  hipError_t e = ng == 2 ? launch_yz_e<T, 2>(a.e, grid, block, lds, st, P, a, &found)
                         : launch_yz_e<T, 1>(a.e, grid, block, lds, st, P, a, &found);
  if (e != hipSuccess) return hip_fail(e, "ebasis_yz_kernel", __FILE__, __LINE__);
  if (!found) return fail(CYR_E_UNSUPPORTED, "E-basis: no Y kernel for this length");
  return CYR_OK;
}

template <typename T> cyr_status launch_ebasis_dcx(const DcxArgs<T> &a, hipStream_t st) {
  hipLaunchKernelGGL((ebasis_dcx_kernel<T>), dim3(6), dim3(1024), 0, st, a);
  CYR_LAUNCHED();
  return CYR_OK;
}

template <typename T> cyr_status launch_ebasis_angles(const AnglesArgs<T> &a, int ng, bool matframe, hipStream_t st) {
  const uint64_t P = (uint64_t)a.g.nx * a.g.ny;
  const dim3 grid(ceil_div(P, 256)), block(256);
  if (ng == 2 && matframe) hipLaunchKernelGGL((ebasis_angles_kernel<T, 2, true>), grid, block, 0, st, a);
  else if (ng == 2) hipLaunchKernelGGL((ebasis_angles_kernel<T, 2, false>), grid, block, 0, st, a);
  else if (matframe) hipLaunchKernelGGL((ebasis_angles_kernel<T, 1, true>), grid, block, 0, st, a);
  else hipLaunchKernelGGL((ebasis_angles_kernel<T, 1, false>), grid, block, 0, st, a);
  CYR_LAUNCHED();
  return CYR_OK;
}

#define CYR_INST(T)                                                                                           \
  template cyr_status launch_ebasis_polx<T>(PolxArgs<T>, const void *, int, int, bool, hipStream_t);        \
  template cyr_status ebasis_yz_plan<T>(YzArgs<T> &, uint32_t);                                              \
  template cyr_status launch_ebasis_yz<T>(YzArgs<T>, int, hipStream_t);                                      \
  template cyr_status launch_ebasis_dcx<T>(const DcxArgs<T> &, hipStream_t);                                \
  template cyr_status launch_ebasis_angles<T>(const AnglesArgs<T> &, int, bool, hipStream_t);
CYR_INST(float)
CYR_INST(double)
#undef CYR_INST

}  // namespace cyr
