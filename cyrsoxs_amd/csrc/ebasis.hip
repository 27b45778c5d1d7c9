// Fused kernels of the E-basis executor.
//
// The polarization is p(theta) = N E(theta) / 4pi with N the per-voxel tensor
// sum over materials and E(theta) = c(theta) e1 + s(theta) e2 in the plane
// normal to k.  Every later stage up to the intensity is linear (Hanning, FFT,
// DC replacement) or a permutation (Igor shift), so per (energy, k) the
// pipeline transforms the tensor field once and forms every angle's p~ from it.
// In the frame Q = [e1 e2 k^] the tensor is N' = Q^T N Q and
//     Q^T p = N' (c, s, 0)^T / 4pi,
// which needs only five fields of N': F0 = N'00, F1 = N'01, F2 = N'02,
// F3 = N'11, F4 = N'12.  For a uniaxial material (include/uniaxial.h:106-150,
// 215-240) the tensor is
//     N = sum_m [ iso_m I + d_m phi_a,m s_m s_m^T ],   d = n_par^2 - n_perp^2,
//     iso = phi_a (n_perp^2 - 1) + phi_u (n_iso^2 - 1)      (Euler, |s| = 1)
//     iso = (n_perp^2 - 1) |s|^2 + phi_u (n_iso^2 - 1)      (vector, phi_a = 1)
// (the reference's "- Vfrac" on the diagonal is folded in analytically).
//
// Four launches per (energy, k):
//   ebasis_bins_kernel   pixel -> FFT z bin of its (kx, ky) column (FP64 geometry);
//                        with linear interpolation the second bin is bin - 1 (mod Nz)
//   xfields_kernel       (xfields.hpp, tensor mode) the five fields evaluated at the
//                        X FFT's Stockham stage-0 positions, X FFT in LDS,
//                        X-transformed fields Gx written once
//   ebasis_yz_kernel     per z slice: Y FFT of a group of x columns, each output
//                        (kx, ky) accumulated into its column's z bin(s):
//                        F(kx,ky,kz) = sum_z Gy(kx,ky,z) W^(z kz); z is split into
//                        chunks with per-chunk partial sums (deterministic, no
//                        atomics).  The DC replacement's neighbour bins are
//                        accumulated on the side.  Gx is read exactly once.
//   ebasis_angles_kernel sums the chunks, applies the DC replacement, and for every
//                        angle forms p~ = A(theta) F (A = frame x E coefficients,
//                        3 x 5, FP64-built) and the Ewald intensity.
//
// HBM traffic per (energy, k): morphology V (16 M) + Gx write 40 V + Gx read 40 V.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <tuple>

#include "fft.hpp"
#include "geom.hpp"
#include "kernels.hpp"
#include "fftpolicy.hpp"
#include "stockham.hpp"
#include "tensor.hpp"
#include "xfields.hpp"

namespace cyr {

namespace {

template <typename T> __device__ inline T qnan();
template <> __device__ inline float qnan<float>() { return __int_as_float(0x7fc00000); }
template <> __device__ inline double qnan<double>() { return __longlong_as_double(0x7ff8000000000000ll); }

// ---- pixel -> FFT z bin ------------------------------------------------------
__global__ void __launch_bounds__(256) ebasis_bins_kernel(EwaldGeom g, int32_t *__restrict__ kzt) {
  const uint64_t P = (uint64_t)g.nx * g.ny;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P) return;
  const uint32_t Y = (uint32_t)(t / g.nx), X = (uint32_t)(t - (uint64_t)Y * g.nx);
  uint32_t z0 = 0;
  double f = 0;
  const int kind = ewald_geometry(g, X, Y, &z0, &f);
  const uint64_t col = (uint64_t)unshift(Y, g.ny) * g.nx + unshift(X, g.nx);
  kzt[col] = kind == 0 ? -1 : (int32_t)(g.enable2d ? 0 : unshift(z0, g.nz));
}

// ---- Y FFT + z bins -----------------------------------------------------------
// Block = pp columns (kx) x tpp threads, one field, one z chunk.  Per z slice the
// next slice's loads are issued before the current slice's FFT.
template <typename T, class F, int NG, bool THREE, int WG>
__global__ void __launch_bounds__(WG) ebasis_yz_kernel(StagePlan P, YzArgs<T> a) {
  constexpr int E = F::kE;
  extern __shared__ unsigned char smem_raw[];
  cx<T> *lds = reinterpret_cast<cx<T> *>(smem_raw);
  const F fft(P, a.tpp, a.pp, false);
  const uint32_t pp = a.pp, tpp = a.tpp;
  // tile q = (z chunk, field, column group), column group fastest, in XCD order
  const uint32_t groups = (a.nx + pp - 1) / pp;
  const uint32_t q = xcd_tile(blockIdx.x, gridDim.x);
  if (q >= groups * kEbFields * a.nchunks) return;  // grid padding: the whole block leaves
  const uint32_t grp = q % groups, c = (q / groups) % kEbFields, chunk = q / (groups * kEbFields);
  const uint32_t p = threadIdx.x % pp, t = threadIdx.x / pp;
  const uint32_t kx = grp * pp + p;
  const bool active = kx < a.nx;
  const uint64_t P2 = (uint64_t)a.nx * a.ny;
  const cx<T> *vol = a.vol[c] + (active ? kx : 0);

  if constexpr (!THREE) {
    // 2D: the only z bin is kz = 0; the Y FFT output is the projection input
    cx<T> v[E];
#pragma unroll
    for (int i = 0; i < E; i++) {
      v[i] = active ? vol[(uint64_t)fft.in_pos(t, i) * a.nx] : mk<T>(0, 0);
    }
    fft.template run<false>(v, t, p, lds, a.twy, nullptr);
    if (active) {
#pragma unroll
      for (int i = 0; i < E; i++) a.part[(uint64_t)c * P2 + (uint64_t)fft.out_pos(t, i) * a.nx + kx] = v[i];
    }
    return;
  } else {
    // LDS: [pp x ny tile][ny Y twiddles][nz z twiddles]; the Y twiddles are
    // reread every slice, so they are staged next to the tile.  This kernel is
    // VALU-bound: its Y FFT and z DFT keep plain twiddle products.
    cx<T> *lty = lds + (size_t)pp * a.ny;
    cx<T> *ltw = lty + a.ny;
    for (uint32_t i = threadIdx.x; i < a.ny; i += blockDim.x) lty[i] = a.twy[i];
    for (uint32_t i = threadIdx.x; i < a.nz; i += blockDim.x) ltw[i] = a.twz[i];
    const uint32_t zb = chunk * a.zchunk;
    const uint32_t ze = zb + a.zchunk < a.nz ? zb + a.zchunk : a.nz;
    // per element: (z-bin << 16) | running exponent z * bin mod Nz; 0xffffffff = no bin
    uint32_t kx16[E];
    cx<T> acc0[E], acc1[NG == 2 ? E : 1];
#pragma unroll
    for (int i = 0; i < E; i++) {
      const uint32_t ky = fft.out_pos(t, i);
      const int32_t b = active ? a.kzt[(uint64_t)ky * a.nx + kx] : -1;
      kx16[i] = b < 0 ? 0xffffffffu : ((uint32_t)b << 16) | (uint32_t)(((uint64_t)zb * (uint32_t)b) % a.nz);
      acc0[i] = mk<T>(0, 0);
      if constexpr (NG == 2) acc1[i] = mk<T>(0, 0);
    }
    // slice z of this column group: uniform base + 32-bit lane offset + uniform element offsets
    const uint32_t off0 = fft.in_pos(t, 0) * a.nx + (active ? kx : 0);
    auto load = [&](uint32_t z, cx<T> (&dst)[E]) {
      const cx<T> *zp = a.vol[c] + (uint64_t)z * P2;
#pragma unroll
      for (int i = 0; i < E; i++) dst[i] = active ? zp[off0 + fft.in_delta(i) * a.nx] : mk<T>(0, 0);
    };
    cx<T> vn[E];
    if (zb < ze) load(zb, vn);
    lds_sync();  // z twiddles staged (the prefetch loads stay in flight)
    for (uint32_t z = zb; z < ze; z++) {
      cx<T> v[E];
#pragma unroll
      for (int i = 0; i < E; i++) v[i] = vn[i];
      if (z + 1 < ze) load(z + 1, vn);
      fft.template run<false>(v, t, p, lds, lty, nullptr);
      cx<T> wz = mk<T>(1, 0);  // W^(-z): second bin of the linear pair
      if constexpr (NG == 2) wz = ltw[z == 0 ? 0 : a.nz - z];
#pragma unroll
      for (int i = 0; i < E; i++) {
        const uint32_t pk = kx16[i];
        if (pk != 0xffffffffu) {
          const uint32_t ex = pk & 0xffffu;
          const cx<T> w0 = ltw[ex];
          const cx<T> u = mk<T>(v[i].x * w0.x - v[i].y * w0.y, v[i].x * w0.y + v[i].y * w0.x);
          acc0[i].x += u.x;
          acc0[i].y += u.y;
          if constexpr (NG == 2) {
            acc1[i].x += u.x * wz.x - u.y * wz.y;
            acc1[i].y += u.x * wz.y + u.y * wz.x;
          }
          uint32_t e2 = ex + (pk >> 16);
          e2 = e2 >= a.nz ? e2 - a.nz : e2;
          kx16[i] = (pk & 0xffff0000u) | e2;
        }
      }
    }
    if (active) {
#pragma unroll
      for (int i = 0; i < E; i++) {
        const uint64_t o = (uint64_t)fft.out_pos(t, i) * a.nx + kx;
        a.part[(((uint64_t)chunk * NG + 0) * kEbFields + c) * P2 + o] = acc0[i];
        if constexpr (NG == 2) a.part[(((uint64_t)chunk * NG + 1) * kEbFields + c) * P2 + o] = acc1[i];
      }
    }
  }
}

// ---- DC-replacement neighbours ------------------------------------------------
// replaceDCComponentWithAverage (src/cudaMain.cu:123-176) averages F at (1,0,0),
// (Nx-1,0,0), (0,1,0), (0,Ny-1,0) and, in 3D, (0,0,1), (0,0,Nz-1).  From the
// X-transformed fields Gx their sum is
//   sum_{y,z} Gx(1,y,z) + Gx(Nx-1,y,z) + Gx(0,y,z) (2cos(2 pi y/Ny) + 2cos(2 pi z/Nz)),
// the last term only in 3D.  One block per (z, field) writes its FP64 slice sum.
template <typename T>
__global__ void __launch_bounds__(256) ebasis_dc_kernel(DcArgs<T> a) {
  __shared__ double red[2][256];
  const uint32_t z = blockIdx.x, c = blockIdx.y;
  const cx<T> *G = a.vol[c] + (uint64_t)z * a.nx * a.ny;
  // cosines from the twiddle tables, FP32 residual added back
  const double cz = a.nz > 1 ? 2.0 * tw_cos<T>(a.twz, a.nz, z) : 0.0;
  double sr = 0, si = 0;
  for (uint32_t y = threadIdx.x; y < a.ny; y += blockDim.x) {
    const cx<T> *row = G + (uint64_t)y * a.nx;
    const cx<T> g0 = row[0], g1 = row[1], gm = row[a.nx - 1];
    const double w = 2.0 * tw_cos<T>(a.twy, a.ny, y) + cz;
    sr += (double)g1.x + (double)gm.x + w * (double)g0.x;
    si += (double)g1.y + (double)gm.y + w * (double)g0.y;
  }
  red[0][threadIdx.x] = sr;
  red[1][threadIdx.x] = si;
  __syncthreads();
  for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double2 o;
    o.x = red[0][0];
    o.y = red[1][0];
    a.dcz[(uint64_t)c * a.nz + z] = o;
  }
}

// ---- chunk sums, DC, and every angle's intensity -----------------------------
template <typename T, int NG>
__global__ void __launch_bounds__(256) ebasis_angles_kernel(AnglesArgs<T> a) {
  const EwaldGeom &g = a.g;
  const uint64_t P2 = (uint64_t)g.nx * g.ny;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  // the block holding the pixel of FFT bin (0, 0) (X = Nx/2, Y = Ny/2) first sums
  // the per-slice DC-neighbour partials (fixed-order tree, deterministic)
  __shared__ double dcs[2 * kEbFields][256];
  __shared__ cx<T> dcv[kEbFields];
  const uint64_t t_dc = (uint64_t)(g.ny / 2) * g.nx + g.nx / 2;
  if (blockIdx.x == t_dc / blockDim.x) {
    for (int c = 0; c < kEbFields; c++) {
      double sr = 0, si = 0;
      for (uint32_t z = threadIdx.x; z < g.nz; z += blockDim.x) {
        const double2 v = a.dcz[(uint64_t)c * g.nz + z];
        sr += v.x;
        si += v.y;
      }
      dcs[2 * c][threadIdx.x] = sr;
      dcs[2 * c + 1][threadIdx.x] = si;
    }
    __syncthreads();
    for (uint32_t w = blockDim.x / 2; w > 0; w >>= 1) {
      if (threadIdx.x < w)
        for (int j = 0; j < 2 * kEbFields; j++) dcs[j][threadIdx.x] += dcs[j][threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      const double cnt = g.enable2d ? 4.0 : 6.0;
      for (int c = 0; c < kEbFields; c++) dcv[c] = mk<T>((T)(dcs[2 * c][0] / cnt), (T)(dcs[2 * c + 1][0] / cnt));
    }
    __syncthreads();
  }
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
  cx<T> F[NG][kEbFields];
#pragma unroll
  for (int q = 0; q < NG; q++)
#pragma unroll
    for (int c = 0; c < kEbFields; c++) {
      cx<T> s = mk<T>(0, 0);
      for (uint32_t ch = 0; ch < a.nchunks; ch++) {
        const cx<T> v = a.part[(((uint64_t)ch * NG + q) * kEbFields + c) * P2 + col];
        s.x += v.x;
        s.y += v.y;
      }
      F[q][c] = s;
    }
  if (ix == 0 && iy == 0) {
#pragma unroll
    for (int q = 0; q < NG; q++) {
      const uint32_t kzq = three ? unshift(z0 + q, g.nz) : 0;
      if (kzq != 0) continue;
#pragma unroll
      for (int c = 0; c < kEbFields; c++) F[q][c] = dcv[c];
    }
  }
  double md[NG][6];
#pragma unroll
  for (int q = 0; q < NG; q++) scatter_matrix(g, X, Y, three ? z0 + q : 0, md[q]);
  const T w1 = (T)f;
#pragma unroll 4
  for (uint32_t an = 0; an < a.nA; an++) {
    const T *A = a.amat + (uint64_t)kEbAmat * an;
    T I[NG];
#pragma unroll
    for (int q = 0; q < NG; q++) {
      cx<T> pv[3];
#pragma unroll
      for (int i = 0; i < 3; i++) {
        T re = 0, im = 0;
#pragma unroll
        for (int j = 0; j < kEbFields; j++) {
          re += A[i * kEbFields + j] * F[q][j].x;
          im += A[i * kEbFields + j] * F[q][j].y;
        }
        pv[i] = mk<T>(re, im);
      }
      I[q] = scatter_intensity<T>(md[q], pv);
    }
    T out;
    if constexpr (NG == 2) out = (1 - w1) * I[0] + w1 * I[1];
    else out = I[0];
    a.proj[(uint64_t)an * P2 + t] = out;
  }
}

// Launch shapes of the compile-time Y plans: tiles are PP columns wide (16
// columns = 128 B rows of complex64).
constexpr int ct_yz_pp(int N) { return N <= 512 ? 16 : (N == 1024 ? 8 : 4); }
uint32_t ct_yz_pp_of(uint32_t n) { return n <= 512 ? 16u : (n == 1024 ? 8u : 4u); }
// Column-tile widths compiled for the compile-time Y plans (PP x tpp <= 1024).
bool ct_yz_pp_ok(uint32_t n, uint32_t pp) {
  return (pp == 4 || pp == 8 || pp == 16) && pp * (n / ct_elements(n)) <= 1024;
}

template <int E> constexpr int rt_yz_wg() { return E <= 8 ? 1024 : 512; }

template <typename T, int NG, bool THREE>
hipError_t launch_yz_rt_e(uint32_t E, dim3 grid, dim3 block, size_t lds, hipStream_t st, const StagePlan &P,
                          const YzArgs<T> &a, bool *found) {
  *found = true;
#define CYR_YZ(EV)                                                                                              \
  case EV:                                                                                                      \
    if constexpr (e_allowed<T>(EV)) {                                                                           \
      hipLaunchKernelGGL((ebasis_yz_kernel<T, RtFft<T, EV>, NG, THREE, rt_yz_wg<EV>()>), grid, block, lds, st, P, \
                         a);                                                                                    \
      return hipGetLastError();                                                                                 \
    }                                                                                                           \
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

template <typename T, int NG, bool THREE, int N, int PP>
hipError_t launch_yz_ct_pp(dim3 grid, size_t lds, hipStream_t st, const StagePlan &P, const YzArgs<T> &a) {
  using F = typename CtPlan<T, N, false, PP>::type;
  constexpr int WG = PP * F::TPP;
  if constexpr (WG <= 1024) {
    hipLaunchKernelGGL((ebasis_yz_kernel<T, F, NG, THREE, WG>), grid, dim3(WG), lds, st, P, a);
    return hipGetLastError();
  }
  return hipErrorInvalidValue;
}

template <typename T, int NG, bool THREE, int N>
hipError_t launch_yz_ct(dim3 grid, size_t lds, hipStream_t st, const StagePlan &P, const YzArgs<T> &a) {
  switch (a.pp) {
    case 4: return launch_yz_ct_pp<T, NG, THREE, N, 4>(grid, lds, st, P, a);
    case 8: return launch_yz_ct_pp<T, NG, THREE, N, 8>(grid, lds, st, P, a);
    default: return launch_yz_ct_pp<T, NG, THREE, N, 16>(grid, lds, st, P, a);
  }
}

template <typename T, int NG, bool THREE>
hipError_t launch_yz_ct_n(uint32_t n, dim3 grid, size_t lds, hipStream_t st, const StagePlan &P, const YzArgs<T> &a,
                          bool *found) {
  *found = true;
  switch (n) {
    case 64: return launch_yz_ct<T, NG, THREE, 64>(grid, lds, st, P, a);
    case 128: return launch_yz_ct<T, NG, THREE, 128>(grid, lds, st, P, a);
    case 256: return launch_yz_ct<T, NG, THREE, 256>(grid, lds, st, P, a);
    case 384:
      if constexpr (e_allowed<T>(12)) return launch_yz_ct<T, NG, THREE, 384>(grid, lds, st, P, a);
      break;
    case 512: return launch_yz_ct<T, NG, THREE, 512>(grid, lds, st, P, a);
    case 1024: return launch_yz_ct<T, NG, THREE, 1024>(grid, lds, st, P, a);
    case 2048: return launch_yz_ct<T, NG, THREE, 2048>(grid, lds, st, P, a);
    default: break;
  }
  *found = false;
  return hipSuccess;
}

// z chunks of the Y/z kernel: enough blocks (x groups x fields x chunks) to
// reach target_blocks, at most one chunk per z slice.
uint32_t yz_chunks(uint32_t nx, uint32_t nz, uint32_t pp, uint32_t target_blocks) {
  const uint32_t per = ceil_div(nx, pp) * kEbFields;
  uint32_t nch = target_blocks / per;
  if (nch < 1) nch = 1;
  return nch > nz ? nz : nch;
}

uint32_t yz_wg_of(uint32_t E) { return E <= 8 ? 1024u : 512u; }

}  // namespace

namespace {
std::mutex g_tune_mu;
std::map<std::tuple<uint32_t, uint32_t, uint32_t, int>, YzTuning> g_tune;
}  // namespace

YzTuning ebasis_yz_tuning(uint32_t nx, uint32_t ny, uint32_t nz, int precision) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(std::make_tuple(nx, ny, nz, precision));
  return it == g_tune.end() ? YzTuning{} : it->second;
}

void ebasis_set_yz_tuning(uint32_t nx, uint32_t ny, uint32_t nz, int precision, YzTuning t) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (t.columns == 0 && t.target_blocks == 0) g_tune.erase(std::make_tuple(nx, ny, nz, precision));
  else g_tune[std::make_tuple(nx, ny, nz, precision)] = t;
}

bool ebasis_supported(uint32_t nx, uint32_t ny, uint32_t nz, int precision) {
  StagePlan P;
  uint32_t lcm;
  const size_t cs = precision == CYR_FP64 ? 16 : 8;
  const uint32_t ecap = precision == CYR_FP64 ? 8 : 16;
  const size_t lds = device_max_lds();
  if (nx < 2 || ny < 2 || nz > 0xffffu) return false;
  if (!ct_usable(nx, precision)) {  // run-time X plan: one row per <= 256 threads
    if (!make_stage_plan(nx, P, lcm)) return false;
    const uint32_t ex = choose_e(nx, lcm, 256, ecap);
    if (!e_instantiated(ex, precision) || nx / ex > 256) return false;
    const size_t pp = 256 / (nx / ex);
    if (pp * nx * cs > lds) return false;  // row tile
  }
  // Y kernel LDS: tile + (3D) Y and z twiddles; the run-time plan can narrow to one column
  const size_t tw = nz > 1 ? ((size_t)ny + nz) * cs : 0;
  if (ct_usable(ny, precision) && (size_t)ct_yz_pp_of(ny) * ny * cs + tw <= lds) return true;
  if (!make_stage_plan(ny, P, lcm)) return false;
  const uint32_t ey = choose_e(ny, lcm, yz_wg_of(lcm), ecap);
  if (!e_instantiated(ey, precision) || ny / ey > yz_wg_of(ey)) return false;
  return (size_t)ny * cs + tw <= lds;
}

uint32_t ebasis_max_chunks(uint32_t nx, uint32_t ny, uint32_t nz, int precision) {
  // at most 16 columns per block; wider tiles give fewer groups, so more chunks
  const YzTuning t = ebasis_yz_tuning(nx, ny, nz, precision);
  const uint32_t target = t.target_blocks > kYzTargetBlocks ? t.target_blocks : kYzTargetBlocks;
  return yz_chunks(nx, nz, 16, target);
}

cyr_status launch_ebasis_bins(const EwaldGeom &g, int32_t *kzt, hipStream_t st) {
  const uint64_t P = (uint64_t)g.nx * g.ny;
  hipLaunchKernelGGL(ebasis_bins_kernel, dim3(ceil_div(P, 256)), dim3(256), 0, st, g, kzt);
  CYR_LAUNCHED();
  return CYR_OK;
}

template <typename T>
cyr_status launch_ebasis_polx(PolxArgs<T> a, const void *vox, int storage, int morph, bool hann, hipStream_t st) {
  return launch_xfields<T, kXTensor>(a, vox, storage, morph, hann, st);
}

template <typename T> cyr_status ebasis_yz_plan(YzArgs<T> &a, uint32_t target_blocks) {
  StagePlan P;
  uint32_t lcm;
  if (!make_stage_plan(a.ny, P, lcm)) return fail(CYR_E_UNSUPPORTED, "E-basis: Ny must be 2^a 3^b 5^c");
  const int prec = sizeof(T) == 8 ? CYR_FP64 : CYR_FP32;
  const size_t tw = a.nz > 1 ? ((size_t)a.nz + a.ny) * sizeof(cx<T>) : 0;
  auto lds_of = [&](uint32_t q) { return (size_t)q * a.ny * sizeof(cx<T>) + tw; };
  // explicit arguments win over the tuning table, which wins over the defaults
  const YzTuning tune = ebasis_yz_tuning(a.nx, a.ny, a.nz, prec);
  if (!target_blocks) target_blocks = tune.target_blocks ? tune.target_blocks : kYzTargetBlocks;
  const uint32_t want_pp = a.pp ? a.pp : tune.columns;
  a.ct = 0;
  if (!a.e && ct_usable(a.ny, prec)) {
    const uint32_t pp = ct_yz_pp_ok(a.ny, want_pp) ? want_pp : ct_yz_pp_of(a.ny);
    if (lds_of(pp) <= device_max_lds()) {
      a.ct = 1;
      a.e = ct_elements(a.ny);
      a.pp = pp;
      a.tpp = a.ny / a.e;
    }
  }
  if (!a.ct) {
    const uint32_t ecap = sizeof(T) == 8 ? 8 : 16;
    uint32_t pp = want_pp ? want_pp : 16;
    uint32_t E = a.e ? a.e : choose_e(a.ny, lcm, yz_wg_of(lcm) / pp, ecap);
    if (E % lcm != 0 || a.ny % E != 0 || !e_instantiated(E, prec)) E = lcm;
    const uint32_t tpp = a.ny / E;
    const uint32_t wg = yz_wg_of(E);
    while (pp > 1 && (lds_of(pp) > device_max_lds() || pp * tpp > wg)) pp /= 2;
    if (lds_of(pp) > device_max_lds() || pp * tpp > wg) return fail(CYR_E_UNSUPPORTED, "E-basis: column does not fit");
    a.e = E;
    a.tpp = tpp;
    a.pp = pp;
  }
  if (!a.zchunk) a.zchunk = ceil_div(a.nz, yz_chunks(a.nx, a.nz, a.pp, target_blocks));
  a.nchunks = ceil_div(a.nz, a.zchunk);
  return CYR_OK;
}

template <typename T> cyr_status launch_ebasis_yz(YzArgs<T> a, int ng, hipStream_t st) {
  StagePlan P;
  uint32_t lcm;
  // (the run-time plan's elements per thread are a multiple of its radix lcm; a
  // compile-time plan has its own radix sequence)
  if (!make_stage_plan(a.ny, P, lcm) || !a.e || !a.pp || !a.tpp || !a.zchunk || a.ny % a.e != 0 ||
      (!a.ct && a.e % lcm != 0) || a.tpp != a.ny / a.e)
    return fail(CYR_E_INVALID, "E-basis: Y plan missing (call ebasis_yz_plan)");
  const int prec = sizeof(T) == 8 ? CYR_FP64 : CYR_FP32;
  if (a.ct && (!ct_usable(a.ny, prec) || a.e != ct_elements(a.ny) || !ct_yz_pp_ok(a.ny, a.pp)))
    return fail(CYR_E_INVALID, "E-basis: inconsistent compile-time Y plan");
  if (!a.ct && a.pp * a.tpp > yz_wg_of(a.e)) return fail(CYR_E_INVALID, "E-basis: Y block larger than its launch bound");
  const bool three = a.nz > 1;
  if (three && (!a.twz || !a.kzt)) return fail(CYR_E_INVALID, "E-basis: z tables missing");
  if (three && a.nz > 0xffffu) return fail(CYR_E_UNSUPPORTED, "E-basis: Nz above 65535");
  if (!three && (a.nchunks != 1 || ng != 1)) return fail(CYR_E_INVALID, "E-basis: 2D plan must be one chunk");
  cyr_status s;
  a.twy = twiddles<T>(a.ny, &s);
  if (!a.twy) return s;
  const uint32_t groups = ceil_div(a.nx, a.pp);
  const dim3 grid(xcd_grid((uint64_t)groups * kEbFields * a.nchunks)), block(a.pp * a.tpp);
  const size_t lds =
      (size_t)a.pp * a.ny * sizeof(cx<T>) + (three ? ((size_t)a.nz + a.ny) * sizeof(cx<T>) : 0);
  if (lds > device_max_lds()) return fail(CYR_E_UNSUPPORTED, "E-basis: Y tile does not fit in LDS");
  bool found = false;
  hipError_t e;
  if (a.ct) {
    if (!three) e = launch_yz_ct_n<T, 1, false>(a.ny, grid, lds, st, P, a, &found);
    else if (ng == 2) e = launch_yz_ct_n<T, 2, true>(a.ny, grid, lds, st, P, a, &found);
    else e = launch_yz_ct_n<T, 1, true>(a.ny, grid, lds, st, P, a, &found);
  } else {
    if (!three) e = launch_yz_rt_e<T, 1, false>(a.e, grid, block, lds, st, P, a, &found);
    else if (ng == 2) e = launch_yz_rt_e<T, 2, true>(a.e, grid, block, lds, st, P, a, &found);
    else e = launch_yz_rt_e<T, 1, true>(a.e, grid, block, lds, st, P, a, &found);
  }
  if (e != hipSuccess) return hip_fail(e, "ebasis_yz_kernel", __FILE__, __LINE__);
  if (!found) return fail(CYR_E_UNSUPPORTED, "E-basis: no Y kernel for this length");
  return CYR_OK;
}

template <typename T> cyr_status launch_ebasis_dc(const DcArgs<T> &a, hipStream_t st) {
  if (a.nx < 2 || a.ny < 2 || !a.twy || (a.nz > 1 && !a.twz)) return fail(CYR_E_INVALID, "E-basis: DC tables missing");
  const uint32_t n = a.count ? a.count : kEbFields;
  if (n > (uint32_t)kXMaxOut) return fail(CYR_E_INVALID, "DC sums: too many volumes");
  hipLaunchKernelGGL((ebasis_dc_kernel<T>), dim3(a.nz, n), dim3(256), 0, st, a);
  CYR_LAUNCHED();
  return CYR_OK;
}

template <typename T> cyr_status launch_ebasis_angles(const AnglesArgs<T> &a, int ng, hipStream_t st) {
  const uint64_t P = (uint64_t)a.g.nx * a.g.ny;
  const dim3 grid(ceil_div(P, 256)), block(256);
  if (ng == 2) hipLaunchKernelGGL((ebasis_angles_kernel<T, 2>), grid, block, 0, st, a);
  else hipLaunchKernelGGL((ebasis_angles_kernel<T, 1>), grid, block, 0, st, a);
  CYR_LAUNCHED();
  return CYR_OK;
}

#define CYR_INST(T)                                                                                           \
  template cyr_status launch_ebasis_polx<T>(PolxArgs<T>, const void *, int, int, bool, hipStream_t);        \
  template cyr_status ebasis_yz_plan<T>(YzArgs<T> &, uint32_t);                                              \
  template cyr_status launch_ebasis_yz<T>(YzArgs<T>, int, hipStream_t);                                      \
  template cyr_status launch_ebasis_dc<T>(const DcArgs<T> &, hipStream_t);                                   \
  template cyr_status launch_ebasis_angles<T>(const AnglesArgs<T> &, int, hipStream_t);
CYR_INST(float)
CYR_INST(double)
#undef CYR_INST

}  // namespace cyr
