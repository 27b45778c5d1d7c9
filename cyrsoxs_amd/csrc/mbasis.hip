// Kernels of the material-basis executor.
//
// The E-basis tensor fields of an energy are linear in that energy's optical
// constants with energy-independent real coefficients (xfields.hpp):
//   N'00 = sum_m b_m alpha_m + c_m beta_m + d_m q00_m,   N'01 = sum_m d_m q01_m, ...
// Windowing, the 3D FFT and the DC replacement are linear, so per k vector the
// seven real basis fields of every material are transformed once and each
// energy's transformed fields are formed at its Ewald bins as
//   F_c(k) = sum over basis fields of (constant of the energy) x (basis spectrum)(k).
// Two real fields A, B share one complex volume Z = A + iB; with
// Zm(k) = Z(-k) their spectra are A^(k) = (Z + conj Zm) / 2 and
// B^(k) = (Z - conj Zm) / 2i, so every column keeps Z at its own Ewald bins
// and at the mirrored bins of the opposite column.
//
// Launches per (k, energy group): mbasis_bands_kernel (+ band width reduction),
// then per material pass xfields_kernel (pair/single mode) -> DC neighbour sums
// -> Y FFT pass -> mbasis_zband_kernel (Z FFT writing only band values; for Nz
// without a compile-time plan: Z FFT pass + mbasis_extract_kernel);
// mbasis_dcsum_kernel; then per energy mbasis_angles_kernel and the shared
// rotation / detector kernels.
#include <hip/hip_runtime.h>

#include "fft.hpp"
#include "fftpolicy.hpp"
#include "geom.hpp"
#include "kernels.hpp"
#include "xfields.hpp"

namespace cyr {

namespace {

template <typename T> __device__ inline T mb_qnan();
template <> __device__ inline float mb_qnan<float>() { return __int_as_float(0x7fc00000); }
template <> __device__ inline double mb_qnan<double>() { return __longlong_as_double(0x7ff8000000000000ll); }

// ---- band of shifted z indices per column over the energy group ---------------
// band[col] = Zlo | (width << 16) over every energy's Ewald z index (and its
// linear-interpolation partner z + 1); 0 = no energy samples the column.  Each
// block also writes the widest band and the sum of the widths it saw (no atomics:
// a second kernel reduces): bmax[kBandHead + 2b] = widest, bmax[kBandHead + 2b + 1]
// = sum of block b.
__global__ void __launch_bounds__(256) mbasis_bands_kernel(EwaldGeom g, const double *__restrict__ kmag, uint32_t ne,
                                                           int32_t *__restrict__ band, uint32_t *__restrict__ bmax) {
  __shared__ uint32_t red[2][256];
  const uint64_t P = (uint64_t)g.nx * g.ny;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t w = 0;
  if (t < P) {
    const uint32_t Y = (uint32_t)(t / g.nx), X = (uint32_t)(t - (uint64_t)Y * g.nx);
    uint32_t lo = 0xffffffffu, hi = 0;
    for (uint32_t e = 0; e < ne; e++) {
      EwaldGeom ge = g;
      ge.k = kmag[e];
      uint32_t z0 = 0;
      double f = 0;
      const int kind = ewald_geometry(ge, X, Y, &z0, &f);
      if (kind == 0) continue;
      lo = z0 < lo ? z0 : lo;
      const uint32_t top = kind == 2 ? z0 + 1 : z0;
      hi = top > hi ? top : hi;
    }
    w = lo == 0xffffffffu ? 0 : hi - lo + 1;
    const uint64_t col = (uint64_t)unshift(Y, g.ny) * g.nx + unshift(X, g.nx);
    band[col] = w ? (int32_t)(lo | (w << 16)) : 0;
  }
  red[0][threadIdx.x] = w;
  red[1][threadIdx.x] = w;
  __syncthreads();
  for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[0][threadIdx.x] = red[0][threadIdx.x + s] > red[0][threadIdx.x] ? red[0][threadIdx.x + s] : red[0][threadIdx.x];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    bmax[kBandHead + 2 * blockIdx.x] = red[0][0];
    bmax[kBandHead + 2 * blockIdx.x + 1] = red[1][0];
  }
}

// ---- columns whose transform some band reads, per row ----------------------------
// Column c = (x, y) (FFT bins) is read by the band passes when its own band or
// that of its mirror column (-x, -y) is non-empty.  span[y] = lo | (hi << 16)
// covers every such column of FFT row y by its shifted x coordinate (unshift is
// an involution: it also maps bins to shifted coordinates); lo > hi when the row
// has none.  The Y pass writes, and the Z band pass reads, only columns inside
// their row's span.  One block per row.
__global__ void __launch_bounds__(256) mbasis_span_kernel(const int32_t *__restrict__ band, uint32_t nx, uint32_t ny,
                                                          uint32_t *__restrict__ span) {
  __shared__ uint32_t red[2][256];
  const uint32_t y = blockIdx.x, my = y ? ny - y : 0;
  uint32_t lo = 0xffffu, hi = 0;
  for (uint32_t x = threadIdx.x; x < nx; x += blockDim.x) {
    const uint32_t mx = x ? nx - x : 0;
    if (band[(uint64_t)y * nx + x] != 0 || band[(uint64_t)my * nx + mx] != 0) {
      const uint32_t X = unshift(x, nx);
      lo = X < lo ? X : lo;
      hi = X > hi ? X : hi;
    }
  }
  red[0][threadIdx.x] = lo;
  red[1][threadIdx.x] = hi;
  __syncthreads();
  for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[0][threadIdx.x] = red[0][threadIdx.x + s] < red[0][threadIdx.x] ? red[0][threadIdx.x + s] : red[0][threadIdx.x];
      red[1][threadIdx.x] = red[1][threadIdx.x + s] > red[1][threadIdx.x] ? red[1][threadIdx.x + s] : red[1][threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) span[y] = red[0][0] | (red[1][0] << 16);
}

// bmax[0] = widest band, bmax[1..2] = sum of all widths (uint64, little endian)
// over the n per-block entries, bmax[3..4] = columns inside the ny row spans,
// bmax[5..6] = D4, the largest (2X - nx)^2 + (2Y - ny)^2 over the spans' end
// columns (X: shifted column, Y = unshift(y): the shifted row of FFT row y).
__global__ void __launch_bounds__(256) mbasis_bmax_kernel(uint32_t *__restrict__ bmax, uint32_t n,
                                                          const uint32_t *__restrict__ span, uint32_t nx, uint32_t ny) {
  __shared__ uint32_t red[256];
  __shared__ unsigned long long sum[3][256];
  uint32_t m = 0;
  unsigned long long s = 0, c = 0, d4 = 0;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    m = bmax[kBandHead + 2 * i] > m ? bmax[kBandHead + 2 * i] : m;
    s += bmax[kBandHead + 2 * i + 1];
  }
  for (uint32_t y = threadIdx.x; y < ny; y += blockDim.x) {
    const uint32_t lo = span[y] & 0xffffu, hi = span[y] >> 16;
    if (lo <= hi) {
      c += hi - lo + 1;
      const long long dy = 2ll * unshift(y, ny) - ny, dl = 2ll * lo - nx, dh = 2ll * hi - nx;
      const unsigned long long r = (unsigned long long)(dy * dy + (dl * dl > dh * dh ? dl * dl : dh * dh));
      d4 = r > d4 ? r : d4;
    }
  }
  red[threadIdx.x] = m;
  sum[0][threadIdx.x] = s;
  sum[1][threadIdx.x] = c;
  sum[2][threadIdx.x] = d4;
  __syncthreads();
  for (uint32_t k = blockDim.x / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) {
      red[threadIdx.x] = red[threadIdx.x + k] > red[threadIdx.x] ? red[threadIdx.x + k] : red[threadIdx.x];
      sum[0][threadIdx.x] += sum[0][threadIdx.x + k];
      sum[1][threadIdx.x] += sum[1][threadIdx.x + k];
      sum[2][threadIdx.x] = sum[2][threadIdx.x + k] > sum[2][threadIdx.x] ? sum[2][threadIdx.x + k] : sum[2][threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    bmax[0] = red[0];
    bmax[1] = (uint32_t)(sum[0][0] & 0xffffffffull);
    bmax[2] = (uint32_t)(sum[0][0] >> 32);
    bmax[3] = (uint32_t)(sum[1][0] & 0xffffffffull);
    bmax[4] = (uint32_t)(sum[1][0] >> 32);
    bmax[5] = (uint32_t)(sum[2][0] & 0xffffffffull);
    bmax[6] = (uint32_t)(sum[2][0] >> 32);
  }
}

// Whether bins x0..x1 (inclusive) of a row meet shifted coordinates lo..hi.
// The shifted coordinate m - x (x <= m = n/2), n + m - x (x > m) falls on each
// of the two parts, so each part is a range.
__device__ inline bool bins_meet_span(uint32_t x0, uint32_t x1, uint32_t n, uint32_t lo, uint32_t hi) {
  const uint32_t m = n / 2;
  if (lo > hi) return false;
  if (x0 <= m) {
    const uint32_t a = m - (x1 < m ? x1 : m), b = m - x0;
    if (a <= hi && b >= lo) return true;
  }
  if (x1 > m) {
    const uint32_t s = x0 > m + 1 ? x0 : m + 1;
    const uint32_t a = n + m - x1, b = n + m - s;
    if (a <= hi && b >= lo) return true;
  }
  return false;
}

// ---- band values of the transformed packed volumes ----------------------------
// own[v][j][col] = Z_v(col, kz(Zlo + j)),  mir[v][j][col] = Z_v(-col, -kz(Zlo + j))
// for j < width(col).  Grid: (columns, j, volume).
template <typename T>
__global__ void __launch_bounds__(256) mbasis_extract_kernel(BandArgs<T> a) {
  const uint64_t P = (uint64_t)a.nx * a.ny;
  const uint64_t col = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t j = blockIdx.y, v = blockIdx.z;
  if (col >= P) return;
  const uint32_t b = (uint32_t)a.band[col];
  if (j >= (b >> 16)) return;
  const uint32_t kz = a.nz > 1 ? unshift((b & 0xffffu) + j, a.nz) : 0;
  const uint32_t iy = (uint32_t)(col / a.nx), ix = (uint32_t)(col - (uint64_t)iy * a.nx);
  const uint32_t mx = ix ? a.nx - ix : 0, my = iy ? a.ny - iy : 0, mz = kz ? a.nz - kz : 0;
  const cx<T> *Z = a.vol[v];
  const uint64_t o = ((uint64_t)(a.vbase + v) * a.J + j) * P + col;
  a.own[o] = Z[(uint64_t)kz * P + col];
  a.mir[o] = Z[(uint64_t)mz * P + (uint64_t)my * a.nx + mx];
}

// ---- Z FFT fused with the band extraction --------------------------------------
// Same launch layout as the strided FFT pass (PP adjacent x columns of one y row
// per block, z pencils, twiddles staged in LDS, XCD-ordered 1D grid), but the
// transformed pencil is not written back: each output Z(col, kz) goes to own[]
// when kz is in the column's band and to mir[] of the opposite column when -kz
// is in that column's band.
constexpr uint32_t kZbandChunk = 16;
template <typename T, class F, int WG>
__global__ void __launch_bounds__(WG) mbasis_zband_kernel(BandArgs<T> a, const cx<T> *__restrict__ tw) {
  constexpr int E = F::kE, TPP = F::TPP, N = TPP * E, PP = WG / TPP;
  extern __shared__ unsigned char smem_raw[];
  cx<T> *lds = reinterpret_cast<cx<T> *>(smem_raw);
  const uint32_t groups = (a.nx + PP - 1) / PP, per_vol = groups * a.ny;
  // chunks of 16 tiles over the XCDs: the tiles outside the row spans leave at
  // once, and they cluster at the rows far from ky = 0 (one contiguous tile range
  // per XCD left some XCDs with little work: 426 against 396 us at 512^2 x 128)
  const uint32_t q = xcd_chunk_tile(blockIdx.x, kZbandChunk);
  if (q >= per_vol * a.count) return;  // grid padding: the whole block leaves before any barrier
  const uint32_t p = threadIdx.x % PP, t = threadIdx.x / PP;
  const uint32_t v = q / per_vol, tile = q % per_vol;
  const uint32_t g = tile % groups, y = tile / groups;
  // columns outside the row's span are neither written by the Y pass nor read
  // here (no band reads them): a tile with none inside leaves (block-uniform,
  // before any barrier), the others load only their columns inside it
  const uint32_t sp = a.span ? a.span[y] : 0xffff0000u;
  const uint32_t slo = sp & 0xffffu, shi = sp >> 16;
  if (!bins_meet_span(g * PP, (g * PP + PP < a.nx ? g * PP + PP : a.nx) - 1, a.nx, slo, shi)) return;
  const uint32_t x = g * PP + p;
  const bool active = x < a.nx;
  const uint64_t P = (uint64_t)a.nx * a.ny;
  const uint64_t col = (uint64_t)y * a.nx + (active ? x : 0);
  const uint32_t sx = unshift(active ? x : 0, a.nx);
  const bool need = active && sx >= slo && sx <= shi;
  const cx<T> *base = a.vol[v] + col;
  // pencil and band-table loads first: they stay in flight across the twiddle
  // staging barrier and the FFT
  cx<T> r[E];
#pragma unroll
  for (int i = 0; i < E; i++) r[i] = need ? base[(uint64_t)F::in_pos(t, i) * P] : mk<T>(0, 0);
  const uint32_t xx = active ? x : 0;
  const uint32_t mx = xx ? a.nx - xx : 0, my = y ? a.ny - y : 0;
  const uint64_t cm = (uint64_t)my * a.nx + mx;
  const uint32_t b = (uint32_t)a.band[col], bm = (uint32_t)a.band[cm];
  cx<T> *ltw = lds + (size_t)PP * N;
  // (with the last stage's FP32 residuals when they fit without costing occupancy)
  constexpr bool LO = tw_lo_in_lds<T>((size_t)PP * N, N, WG);
  for (uint32_t i = threadIdx.x; i < (uint32_t)(LO ? 2 * N : N); i += WG) ltw[i] = tw[i];
  lds_sync();
  F::run(r, t, p, lds, ltw, LO ? ltw + N : tw + N);
  if (!active) return;
  const uint32_t lo = b & 0xffffu, w = b >> 16;
  const uint32_t lom = bm & 0xffffu, wm = bm >> 16;
  cx<T> *own = a.own + (uint64_t)(a.vbase + v) * a.J * P;
  cx<T> *mir = a.mir + (uint64_t)(a.vbase + v) * a.J * P;
#pragma unroll
  for (int i = 0; i < E; i++) {
    const uint32_t kz = F::out_pos(t, i);
    const uint32_t j = unshift(kz, N) - lo;  // wraps above w when below the band
    if (j < w) own[(uint64_t)j * P + col] = r[i];
    const uint32_t jm = unshift(kz ? N - kz : 0, N) - lom;
    if (jm < wm) mir[(uint64_t)jm * P + cm] = r[i];
  }
}

// Columns per tile: 32 for short pencils (FP32: 256 B contiguous per z row,
// measured faster than 16 or 64 at Nz = 128, and than 16 at Nz = 256: 1024^2 x 256
// 3.50 against 3.94 ms; FP64 keeps 16 at Nz = 256, where 32 columns would need
// 132 KB of LDS per block), else as many as 1024 threads allow.
template <typename T> constexpr int zband_pp(int N) {
  return N <= 128 || (N == 256 && sizeof(T) == 4) ? 32 : (N <= 512 ? 16 : (N == 1024 ? 8 : 4));
}

template <typename T, int N>
hipError_t launch_zband_n(const BandArgs<T> &a, hipStream_t st, const cx<T> *tw) {
  constexpr int PP = zband_pp<T>(N);
  using F = typename CtPlan<T, N, false, PP>::type;
  constexpr int WG = PP * F::TPP;
  const uint64_t tiles = (uint64_t)((a.nx + PP - 1) / PP) * a.ny * a.count;
  if (tiles > 0x7ffffff0ull) return hipErrorInvalidConfiguration;
  constexpr bool LO = tw_lo_in_lds<T>((size_t)PP * N, N, WG);
  const size_t lds = ((size_t)PP * N + (LO ? 2 * N : N)) * sizeof(cx<T>);
  hipLaunchKernelGGL((mbasis_zband_kernel<T, F, WG>), dim3(xcd_chunk_grid(tiles, kZbandChunk)), dim3(WG), lds, st,
                     a, tw);
  return hipGetLastError();
}

// dcv[v] = (sum_z dcz[v][z]) / cnt, fixed-order tree (deterministic); one block per volume
__global__ void __launch_bounds__(1024) mbasis_dcsum_kernel(const double2 *__restrict__ dcz, uint32_t nz, double cnt,
                                                            double2 *__restrict__ dcv) {
  __shared__ double red[2][1024];
  const uint32_t v = blockIdx.x;
  double sr = 0, si = 0;
  for (uint32_t z = threadIdx.x; z < nz; z += blockDim.x) {
    const double2 x = dcz[(uint64_t)v * nz + z];
    sr += x.x;
    si += x.y;
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
    o.x = red[0][0] / cnt;
    o.y = red[1][0] / cnt;
    dcv[v] = o;
  }
}

template <typename T> __device__ inline void cmac(cx<T> &acc, T cr, T ci, const cx<T> &x) {
  acc.x += cr * x.x - ci * x.y;
  acc.y += cr * x.y + ci * x.x;
}

// Spectra of the two real fields packed in one volume, from Z(k) and Z(-k).
template <typename T> __device__ inline void unpack(const cx<T> &z, const cx<T> &zm, cx<T> &a, cx<T> &b) {
  const T h = (T)0.5;
  a = mk<T>(h * (z.x + zm.x), h * (z.y - zm.y));
  b = mk<T>(h * (z.y + zm.y), -h * (z.x - zm.x));
}

// Adds basis field r (alpha, beta, q00, q11, q01, q02, q12) of material constant mc.
template <typename T>
__device__ inline void add_basis(int r, const MatConst<T> &mc, const cx<T> &x, cx<T> &iso, cx<T> F[kEbFields]) {
  switch (r) {
    case 0: cmac<T>(iso, mc.br, mc.bi, x); break;
    case 1: cmac<T>(iso, mc.cr, mc.ci, x); break;
    case 2: cmac<T>(F[0], mc.dr, mc.di, x); break;
    case 3: cmac<T>(F[3], mc.dr, mc.di, x); break;
    case 4: cmac<T>(F[1], mc.dr, mc.di, x); break;
    case 5: cmac<T>(F[2], mc.dr, mc.di, x); break;
    default: cmac<T>(F[4], mc.dr, mc.di, x); break;
  }
}

// ---- per energy: the tensor fields at a pixel's Ewald bins ----------------------
// Returns 0 for a NaN pixel; else fills F[q] (bin z0 + q), the scattering
// matrices md[q] and the linear weight w1 of bin z0 + 1.
template <typename T, int NG>
__device__ inline int mb_fields(const MbAnglesArgs<T> &a, uint32_t X, uint32_t Y, cx<T> (&F)[NG][kEbFields],
                                double (&md)[NG][6], T &w1) {
  const EwaldGeom &g = a.g;
  const uint64_t P2 = (uint64_t)g.nx * g.ny;
  uint32_t z0 = 0;
  double f = 0;
  const int kind = ewald_geometry(g, X, Y, &z0, &f);
  if (kind == 0) return 0;
  const uint32_t ix = unshift(X, g.nx), iy = unshift(Y, g.ny);
  const uint64_t col = (uint64_t)iy * g.nx + ix;
  const bool three = !g.enable2d;
  const uint32_t j0 = three ? z0 - ((uint32_t)a.band[col] & 0xffffu) : 0;
  // the DC-replaced value stands in for Z(0) (its own mirror) at FFT bin (0, 0, 0)
  bool dc[NG];
#pragma unroll
  for (int q = 0; q < NG; q++) dc[q] = ix == 0 && iy == 0 && (three ? unshift(z0 + q, g.nz) == 0 : true);
  const uint64_t slab = (uint64_t)a.J * P2;
  const cx<T> *own = a.own + (uint64_t)j0 * P2 + col, *mir = a.mir + (uint64_t)j0 * P2 + col;
  auto fetch = [&](uint32_t v, int q, cx<T> &z, cx<T> &zm) {
    if (dc[q]) {
      const double2 d = a.dcv[v];
      z = zm = mk<T>((T)d.x, (T)d.y);
    } else {
      z = own[(uint64_t)v * slab + (uint64_t)q * P2];
      zm = mir[(uint64_t)v * slab + (uint64_t)q * P2];
    }
  };
  cx<T> iso[NG];
#pragma unroll
  for (int q = 0; q < NG; q++) {
    iso[q] = mk<T>(0, 0);
#pragma unroll
    for (int c = 0; c < kEbFields; c++) F[q][c] = mk<T>(0, 0);
  }
  const uint32_t pairs = a.M / 2;
  for (uint32_t p = 0; p < pairs; p++) {
    const MatConst<T> ma = a.mats[2 * p], mb = a.mats[2 * p + 1];
#pragma unroll
    for (int r = 0; r < 7; r++) {
#pragma unroll
      for (int q = 0; q < NG; q++) {
        cx<T> z, zm, xa, xb;
        fetch(7 * p + r, q, z, zm);
        unpack<T>(z, zm, xa, xb);
        add_basis<T>(r, ma, xa, iso[q], F[q]);
        add_basis<T>(r, mb, xb, iso[q], F[q]);
      }
    }
  }
  if (a.M & 1) {  // last material alone: (alpha, beta), (q00, q11), (q01, q02), (q12, -)
    const MatConst<T> mc = a.mats[a.M - 1];
    const uint32_t vb = 7 * pairs;
#pragma unroll
    for (int r = 0; r < 4; r++) {
#pragma unroll
      for (int q = 0; q < NG; q++) {
        cx<T> z, zm, xa, xb;
        fetch(vb + r, q, z, zm);
        unpack<T>(z, zm, xa, xb);
        add_basis<T>(2 * r, mc, xa, iso[q], F[q]);
        if (r < 3) add_basis<T>(2 * r + 1, mc, xb, iso[q], F[q]);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < NG; q++) {
    F[q][0].x += iso[q].x;
    F[q][0].y += iso[q].y;
    F[q][3].x += iso[q].x;
    F[q][3].y += iso[q].y;
    scatter_matrix(g, X, Y, three ? z0 + q : 0, md[q]);
  }
  w1 = (T)f;
  return kind;
}

// ---- per energy: every angle's projection ---------------------------------------
template <typename T, int NG>
__global__ void __launch_bounds__(256) mbasis_angles_kernel(MbAnglesArgs<T> a) {
  const uint64_t P2 = (uint64_t)a.g.nx * a.g.ny;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P2) return;
  const uint32_t Y = (uint32_t)(t / a.g.nx), X = (uint32_t)(t - (uint64_t)Y * a.g.nx);
  cx<T> F[NG][kEbFields];
  double md[NG][6];
  T w1;
  if (!mb_fields<T, NG>(a, X, Y, F, md, w1)) {
    for (uint32_t an = 0; an < a.nA; an++) a.proj[(uint64_t)an * P2 + t] = mb_qnan<T>();
    return;
  }
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

// ---- per energy, laboratory frame: the intensity's quadratic form ---------------
// With E(theta) = c e1 + s e2 the polarization is p = c u + s v, u = Q (F0, F1, F2),
// v = Q (F1, F3, F4), so I(theta) = c^2 |M u|^2 + s^2 |M v|^2 + 2 c s Re<M u, M v>.
// quad[pixel] = (|Mu|^2, |Mv|^2, Re<Mu, Mv>), interpolated between the bins.
template <typename T, int NG>
__device__ inline void quad_form_write(const Mat3<T> &Q, const cx<T> (&F)[NG][kEbFields], const double (&md)[NG][6],
                                       T w1, T *o) {
  T qa[NG], qb[NG], qm[NG];
#pragma unroll
  for (int q = 0; q < NG; q++) {
    cx<T> u[3], v[3], mu[3], mv[3];
    const cx<T> ue[3] = {F[q][0], F[q][1], F[q][2]}, ve[3] = {F[q][1], F[q][3], F[q][4]};
#pragma unroll
    for (int i = 0; i < 3; i++) {
      u[i] = mk<T>(Q.m[3 * i] * ue[0].x + Q.m[3 * i + 1] * ue[1].x + Q.m[3 * i + 2] * ue[2].x,
                   Q.m[3 * i] * ue[0].y + Q.m[3 * i + 1] * ue[1].y + Q.m[3 * i + 2] * ue[2].y);
      v[i] = mk<T>(Q.m[3 * i] * ve[0].x + Q.m[3 * i + 1] * ve[1].x + Q.m[3 * i + 2] * ve[2].x,
                   Q.m[3 * i] * ve[0].y + Q.m[3 * i + 1] * ve[1].y + Q.m[3 * i + 2] * ve[2].y);
    }
    const T m0 = (T)md[q][0], m1 = (T)md[q][1], m2 = (T)md[q][2], m3 = (T)md[q][3], m4 = (T)md[q][4],
            m5 = (T)md[q][5];
    const T mm[3][3] = {{m0, m1, m2}, {m1, m3, m4}, {m2, m4, m5}};
#pragma unroll
    for (int i = 0; i < 3; i++) {
      mu[i] = mk<T>(mm[i][0] * u[0].x + mm[i][1] * u[1].x + mm[i][2] * u[2].x,
                    mm[i][0] * u[0].y + mm[i][1] * u[1].y + mm[i][2] * u[2].y);
      mv[i] = mk<T>(mm[i][0] * v[0].x + mm[i][1] * v[1].x + mm[i][2] * v[2].x,
                    mm[i][0] * v[0].y + mm[i][1] * v[1].y + mm[i][2] * v[2].y);
    }
    qa[q] = mu[0].x * mu[0].x + mu[0].y * mu[0].y + mu[1].x * mu[1].x + mu[1].y * mu[1].y + mu[2].x * mu[2].x +
            mu[2].y * mu[2].y;
    qb[q] = mv[0].x * mv[0].x + mv[0].y * mv[0].y + mv[1].x * mv[1].x + mv[1].y * mv[1].y + mv[2].x * mv[2].x +
            mv[2].y * mv[2].y;
    qm[q] = mu[0].x * mv[0].x + mu[0].y * mv[0].y + mu[1].x * mv[1].x + mu[1].y * mv[1].y + mu[2].x * mv[2].x +
            mu[2].y * mv[2].y;
  }
  if constexpr (NG == 2) {
    o[0] = (1 - w1) * qa[0] + w1 * qa[1];
    o[1] = (1 - w1) * qb[0] + w1 * qb[1];
    o[2] = (1 - w1) * qm[0] + w1 * qm[1];
  } else {
    o[0] = qa[0];
    o[1] = qb[0];
    o[2] = qm[0];
  }
}

// One (pixel, energy): a holds the energy's |k| and constants.
template <typename T, int NG>
__device__ inline void quad_pixel_energy(const MbAnglesArgs<T> &a, uint32_t X, uint32_t Y, T *o) {
  cx<T> F[NG][kEbFields];
  double md[NG][6];
  T w1;
  if (!mb_fields<T, NG>(a, X, Y, F, md, w1)) {
    o[0] = o[1] = o[2] = mb_qnan<T>();
    return;
  }
  quad_form_write<T, NG>(a.Q, F, md, w1, o);
}

constexpr uint32_t kQuadChunk = 16;  // pixel blocks per XCD chunk of the quadratic-form grids
template <typename T, int NG>
__global__ void __launch_bounds__(256) mbasis_quad_kernel(MbAnglesArgs<T> a) {
  const uint64_t P2 = (uint64_t)a.g.nx * a.g.ny;
  // 1D grid in XCD order over (pixel block, energy), energy fastest: the ne
  // energies of a pixel block read the same band columns and run on one XCD
  // back to back, so all but the first find them in its L2.  Chunks of 16 pixel
  // blocks go to the XCDs in turn: the pixels outside the Ewald sphere (cheap)
  // cluster in the rows far from the beam.
  const uint32_t q = xcd_chunk_tile(blockIdx.x, kQuadChunk * a.ne);
  const uint32_t e = q % a.ne, b = q / a.ne;
  const uint64_t t = (uint64_t)b * blockDim.x + threadIdx.x;
  if (t >= P2) return;
  a.g.k = a.kmag[e];
  a.mats += (uint64_t)e * a.M;
  const uint32_t Y = (uint32_t)(t / a.g.nx), X = (uint32_t)(t - (uint64_t)Y * a.g.nx);
  quad_pixel_energy<T, NG>(a, X, Y, a.quad + ((uint64_t)e * P2 + quad_index(X, Y, a.g.nx, a.g.ny)) * kQuadRecord);
}

// Two energies (2 ep, 2 ep + 1) of one pixel per thread.  Neighbouring energies'
// first Ewald bins are usually equal or one apart; then every band value either
// energy needs lies in the NG + 1 bins from the lower one, so each volume's values
// are gathered (and unpacked) once for both: the kernel is bound by its gathers.
// Pixels whose bins lie further apart, the FFT bin (0, 0) (DC values) and NaN
// pixels take the one-energy path for each energy.
template <typename T, int NG>
__global__ void __launch_bounds__(256) mbasis_quad2_kernel(MbAnglesArgs<T> a) {
  const uint64_t P2 = (uint64_t)a.g.nx * a.g.ny;
  const uint32_t npair = (a.ne + 1) / 2;
  const uint32_t qt = xcd_chunk_tile(blockIdx.x, kQuadChunk * npair);  // (as mbasis_quad_kernel)
  const uint32_t ep = qt % npair, b = qt / npair;
  const uint64_t t = (uint64_t)b * blockDim.x + threadIdx.x;
  if (t >= P2) return;
  const uint32_t e0 = 2 * ep;
  const bool two = e0 + 1 < a.ne;
  const uint32_t Y = (uint32_t)(t / a.g.nx), X = (uint32_t)(t - (uint64_t)Y * a.g.nx);
  const uint32_t rec = quad_index(X, Y, a.g.nx, a.g.ny);
  MbAnglesArgs<T> ak[2] = {a, a};
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const uint32_t e = two ? e0 + k : e0;
    ak[k].g.k = a.kmag[e];
    ak[k].mats += (uint64_t)e * a.M;
  }
  T *o0 = a.quad + ((uint64_t)e0 * P2 + rec) * kQuadRecord, *o1 = o0 + P2 * kQuadRecord;
  uint32_t z0[2];
  double f[2];
  const int kind0 = ewald_geometry(ak[0].g, X, Y, &z0[0], &f[0]);
  const int kind1 = two ? ewald_geometry(ak[1].g, X, Y, &z0[1], &f[1]) : 0;
  const uint32_t ix = unshift(X, a.g.nx), iy = unshift(Y, a.g.ny);
  const bool three = !a.g.enable2d;
  const uint32_t gap = z0[0] > z0[1] ? z0[0] - z0[1] : z0[1] - z0[0];
  if (!kind0 || !kind1 || (ix == 0 && iy == 0) || (three && gap > 1)) {
    quad_pixel_energy<T, NG>(ak[0], X, Y, o0);
    if (two) quad_pixel_energy<T, NG>(ak[1], X, Y, o1);
    return;
  }
  const uint64_t col = (uint64_t)iy * a.g.nx + ix;
  const uint32_t zlo = three ? ((uint32_t)a.band[col] & 0xffffu) : 0;
  const uint32_t zl = z0[0] < z0[1] ? z0[0] : z0[1];
  const uint32_t jl = three ? zl - zlo : 0;
  const bool d[2] = {three && z0[0] > zl, three && z0[1] > zl};
  const bool wide = d[0] || d[1];  // the bins span NG + 1 band rows
  const uint64_t slab = (uint64_t)a.J * P2;
  const cx<T> *own = a.own + (uint64_t)jl * P2 + col, *mir = a.mir + (uint64_t)jl * P2 + col;
  cx<T> F[2][NG][kEbFields], iso[2][NG];
#pragma unroll
  for (int k = 0; k < 2; k++)
#pragma unroll
    for (int q = 0; q < NG; q++) {
      iso[k][q] = mk<T>(0, 0);
#pragma unroll
      for (int c = 0; c < kEbFields; c++) F[k][q][c] = mk<T>(0, 0);
    }
  // the packed values of volume v at the NG (+ 1) rows, unpacked
  auto gather = [&](uint32_t v, cx<T> (&xa)[NG + 1], cx<T> (&xb)[NG + 1]) {
#pragma unroll
    for (int i = 0; i <= NG; i++) {
      cx<T> z = mk<T>(0, 0), zm = mk<T>(0, 0);
      if (i < NG || wide) {
        z = own[(uint64_t)v * slab + (uint64_t)i * P2];
        zm = mir[(uint64_t)v * slab + (uint64_t)i * P2];
      }
      unpack<T>(z, zm, xa[i], xb[i]);
    }
  };
  auto sel = [&](const cx<T> (&x)[NG + 1], int k, int q) { return d[k] ? x[q + 1] : x[q]; };
  const uint32_t pairs = a.M / 2;
  for (uint32_t p = 0; p < pairs; p++) {
    const MatConst<T> ma[2] = {ak[0].mats[2 * p], ak[1].mats[2 * p]};
    const MatConst<T> mb[2] = {ak[0].mats[2 * p + 1], ak[1].mats[2 * p + 1]};
#pragma unroll
    for (int r = 0; r < 7; r++) {
      cx<T> xa[NG + 1], xb[NG + 1];
      gather(7 * p + r, xa, xb);
#pragma unroll
      for (int k = 0; k < 2; k++)
#pragma unroll
        for (int q = 0; q < NG; q++) {
          add_basis<T>(r, ma[k], sel(xa, k, q), iso[k][q], F[k][q]);
          add_basis<T>(r, mb[k], sel(xb, k, q), iso[k][q], F[k][q]);
        }
    }
  }
  if (a.M & 1) {  // last material alone: (alpha, beta), (q00, q11), (q01, q02), (q12, -)
    const MatConst<T> mc[2] = {ak[0].mats[a.M - 1], ak[1].mats[a.M - 1]};
    const uint32_t vb = 7 * pairs;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      cx<T> xa[NG + 1], xb[NG + 1];
      gather(vb + r, xa, xb);
#pragma unroll
      for (int k = 0; k < 2; k++)
#pragma unroll
        for (int q = 0; q < NG; q++) {
          add_basis<T>(2 * r, mc[k], sel(xa, k, q), iso[k][q], F[k][q]);
          if (r < 3) add_basis<T>(2 * r + 1, mc[k], sel(xb, k, q), iso[k][q], F[k][q]);
        }
    }
  }
#pragma unroll
  for (int k = 0; k < 2; k++) {
    double md[NG][6];
#pragma unroll
    for (int q = 0; q < NG; q++) {
      F[k][q][0].x += iso[k][q].x;
      F[k][q][0].y += iso[k][q].y;
      F[k][q][3].x += iso[k][q].x;
      F[k][q][3].y += iso[k][q].y;
      scatter_matrix(ak[k].g, X, Y, three ? z0[k] + q : 0, md[q]);
    }
    quad_form_write<T, NG>(a.Q, F[k], md, (T)f[k], k ? o1 : o0);
  }
}

}  // namespace

cyr_status launch_mbasis_bands(const EwaldGeom &g, const double *kmag, uint32_t ne, int32_t *band, uint32_t *bmax,
                               hipStream_t st) {
  const uint64_t P = (uint64_t)g.nx * g.ny;
  const uint32_t blocks = ceil_div(P, 256);
  if (g.nx > 0xfffe) return fail(CYR_E_UNSUPPORTED, "material basis: Nx above 65534");
  uint32_t *span = mbasis_span_table(bmax, g.nx, g.ny);
  hipLaunchKernelGGL(mbasis_bands_kernel, dim3(blocks), dim3(256), 0, st, g, kmag, ne, band, bmax);
  CYR_LAUNCHED();
  hipLaunchKernelGGL(mbasis_span_kernel, dim3(g.ny), dim3(256), 0, st, band, g.nx, g.ny, span);
  CYR_LAUNCHED();
  hipLaunchKernelGGL(mbasis_bmax_kernel, dim3(1), dim3(256), 0, st, bmax, blocks, span, g.nx, g.ny);
  CYR_LAUNCHED();
  return CYR_OK;
}

uint64_t mbasis_bmax_words(uint32_t nx, uint32_t ny) { return kBandHead + 2 * ceil_div((uint64_t)nx * ny, 256) + ny; }
uint32_t *mbasis_span_table(uint32_t *bmax, uint32_t nx, uint32_t ny) {
  return bmax + kBandHead + 2 * ceil_div((uint64_t)nx * ny, 256);
}

template <typename T> cyr_status launch_mbasis_extract(const BandArgs<T> &a, hipStream_t st) {
  if (a.count == 0) return CYR_OK;
  if (a.count > (uint32_t)kXMaxOut || a.J == 0 || a.J > 65535) return fail(CYR_E_INVALID, "band extraction: bad shape");
  const uint64_t P = (uint64_t)a.nx * a.ny;
  hipLaunchKernelGGL((mbasis_extract_kernel<T>), dim3(ceil_div(P, 256), a.J, a.count), dim3(256), 0, st, a);
  CYR_LAUNCHED();
  return CYR_OK;
}

cyr_status launch_mbasis_dcsum(const double2 *dcz, uint32_t nvol, uint32_t nz, bool three, double2 *dcv,
                               hipStream_t st) {
  if (nvol == 0) return CYR_OK;
  hipLaunchKernelGGL(mbasis_dcsum_kernel, dim3(nvol), dim3(1024), 0, st, dcz, nz, three ? 6.0 : 4.0, dcv);
  CYR_LAUNCHED();
  return CYR_OK;
}

template <typename T> cyr_status launch_mbasis_zband(const BandArgs<T> &a, hipStream_t st, bool *done) {
  *done = false;
  const int prec = sizeof(T) == 8 ? CYR_FP64 : CYR_FP32;
  if (a.count == 0 || a.nz < 2 || !ct_usable(a.nz, prec)) return CYR_OK;
  if (a.count > (uint32_t)kXMaxOut || a.J == 0) return fail(CYR_E_INVALID, "Z band pass: bad shape");
  const uint32_t pp = a.nz <= 128 ? 32 : (a.nz <= 512 ? 16 : (a.nz == 1024 ? 8 : 4));
  if (((size_t)pp * a.nz + a.nz) * sizeof(cx<T>) > device_max_lds()) return CYR_OK;
  cyr_status s;
  const cx<T> *tw = twiddles<T>(a.nz, &s);
  if (!tw) return s;
  hipError_t e = hipSuccess;
  switch (a.nz) {
    case 64: e = launch_zband_n<T, 64>(a, st, tw); break;
    case 128: e = launch_zband_n<T, 128>(a, st, tw); break;
    case 256: e = launch_zband_n<T, 256>(a, st, tw); break;
    case 384:
      if constexpr (sizeof(T) == 4) e = launch_zband_n<T, 384>(a, st, tw);
      break;
    case 512: e = launch_zband_n<T, 512>(a, st, tw); break;
    case 1024: e = launch_zband_n<T, 1024>(a, st, tw); break;
    case 2048: e = launch_zband_n<T, 2048>(a, st, tw); break;
    default: return CYR_OK;
  }
  if (e != hipSuccess) return hip_fail(e, "mbasis_zband_kernel", __FILE__, __LINE__);
  *done = true;
  return CYR_OK;
}

template <typename T> cyr_status launch_mbasis_angles(const MbAnglesArgs<T> &a, int ng, hipStream_t st) {
  const uint64_t P = (uint64_t)a.g.nx * a.g.ny;
  const dim3 grid(ceil_div(P, 256)), block(256);
  if (a.quad) {
    if (!a.kmag || a.ne == 0 || a.ne > 65535) return fail(CYR_E_INVALID, "quadratic form: energy table missing");
    // FP32: energies in pairs (one thread per pixel and pair); FP64 (whose pair
    // kernel would hold 256 registers) and single energies: one at a time
    const bool paired = a.ne >= 2 && sizeof(T) == 4;
    const uint32_t per = paired ? (a.ne + 1) / 2 : a.ne;
    const uint64_t tiles = (uint64_t)ceil_div(P, 256) * per;
    const uint64_t padded = (uint64_t)ceil_div(tiles, (uint64_t)kXcds * kQuadChunk * per) * kXcds * kQuadChunk * per;
    if (padded > 0x7ffffff0ull) return fail(CYR_E_UNSUPPORTED, "quadratic form: grid too large");
    const dim3 g2((unsigned)padded);
    bool launched = false;
    if constexpr (sizeof(T) == 4) {
      if (paired) {
        if (ng == 2) hipLaunchKernelGGL((mbasis_quad2_kernel<T, 2>), g2, block, 0, st, a);
        else hipLaunchKernelGGL((mbasis_quad2_kernel<T, 1>), g2, block, 0, st, a);
        launched = true;
      }
    }
    if (!launched) {
      if (ng == 2) hipLaunchKernelGGL((mbasis_quad_kernel<T, 2>), g2, block, 0, st, a);
      else hipLaunchKernelGGL((mbasis_quad_kernel<T, 1>), g2, block, 0, st, a);
    }
  } else {
    if (ng == 2) hipLaunchKernelGGL((mbasis_angles_kernel<T, 2>), grid, block, 0, st, a);
    else hipLaunchKernelGGL((mbasis_angles_kernel<T, 1>), grid, block, 0, st, a);
  }
  CYR_LAUNCHED();
  return CYR_OK;
}

template <typename T>
cyr_status launch_mbasis_xfields(PolxArgs<T> a, bool pair, const void *vox, int storage, int morph, bool hann,
                                 hipStream_t st) {
  if (pair) return launch_xfields<T, kXPair>(a, vox, storage, morph, hann, st);
  return launch_xfields<T, kXSingle>(a, vox, storage, morph, hann, st);
}

#define CYR_INST(T)                                                                                          \
  template cyr_status launch_mbasis_extract<T>(const BandArgs<T> &, hipStream_t);                          \
  template cyr_status launch_mbasis_zband<T>(const BandArgs<T> &, hipStream_t, bool *);                    \
  template cyr_status launch_mbasis_angles<T>(const MbAnglesArgs<T> &, int, hipStream_t);                  \
  template cyr_status launch_mbasis_xfields<T>(PolxArgs<T>, bool, const void *, int, int, bool, hipStream_t);
CYR_INST(float)
CYR_INST(double)
#undef CYR_INST

}  // namespace cyr
