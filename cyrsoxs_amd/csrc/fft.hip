// Forward complex-to-complex FFT pencils for gfx950 (replaces the reference's
// cufftExecC2C/Z2Z calls, include/cudaMain.h:184-190, src/cudaMain.cu:668-670).
//
// Contract: F[k] = sum_n x[n] exp(-2 pi i n k / N) along one axis, unnormalised,
// in place.  A 3D transform is three axis passes (X, Y, Z).
//
// Design (wave64 / LDS):
//   * A block owns `pp` whole pencils.  For the contiguous X axis the threads of a
//     pencil are consecutive lanes (row loads are unit-stride); for the strided
//     Y/Z axes the lanes run across `pp` adjacent pencils (each element load is a
//     pp*sizeof(complex) contiguous segment).
//   * Each thread holds E elements in registers; every Stockham stage is
//     radix R in {2,3,4,5,8} with E/R butterflies per thread.  Stage 0 reads
//     straight from HBM and the last stage writes straight to HBM (both in the
//     Stockham access order, which is coalesced); intermediate stages go through
//     LDS.  Each element is read once and written once per pass.
//   * Twiddles come from a per-length table generated in FP64 on the host.
#include <hip/hip_runtime.h>

#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "fft.hpp"
#include "fftpolicy.hpp"
#include "stockham.hpp"

namespace cyr {

namespace {

struct PassArgs {
  void *vol[8];
  uint32_t nvol;  // volumes of this launch (1D XCD-ordered grids)
  uint64_t elem_stride;
  uint64_t inner_stride;
  uint64_t outer_stride;
  uint32_t inner_count;
  uint32_t outer_count;
  uint32_t groups;  // ceil(inner_count / pp)
  uint32_t pp;
  uint32_t tpp;
  double2 *dcz;        // non-null: the Y pass adds the DC-neighbour partial sums (FftDc)
  const void *dc_twz;  // Z twiddles of those sums (null in 2D)
  uint32_t dc_nz;
  const uint32_t *span;  // Y pass: per output row, the shifted x range it writes (null: all)
};

// ---- one Stockham stage ------------------------------------------------------
//
// Butterfly j (0 <= j < n/R) of a stage with sub-transform length Ls reads
// x[j + r n/R], multiplies by W_{Ls R}^{r (j mod Ls)}, does an R-point DFT and
// writes y[(j / Ls) Ls R + (j mod Ls) + r Ls].

template <typename T, int E, int R, bool CONTIG>
__device__ inline void run_stage(cx<T> (&v)[E], const StagePlan &P, uint32_t s, bool first, bool last,
                                 uint32_t t, uint32_t p, const PassArgs &A, cx<T> *gbase, bool active,
                                 cx<T> *lds, uint32_t lpitch, const cx<T> *__restrict__ tw) {
  constexpr int NB = E / R;
  const uint32_t n = P.n;
  const uint32_t m = n / R;
  const uint32_t Ls = P.ls[s];
  const uint32_t tstep = P.tstep[s];
  const uint32_t tpp = A.tpp;
  // ---- gather inputs
#pragma unroll
  for (int b = 0; b < NB; b++) {
    const uint32_t j = t + b * tpp;
#pragma unroll
    for (int r = 0; r < R; r++) {
      const uint32_t idx = j + r * m;
      if (first) {
        v[b * R + r] = active ? gbase[(uint64_t)idx * A.elem_stride] : mk<T>(0, 0);
      } else {
        v[b * R + r] = CONTIG ? lds[p * lpitch + idx] : lds[idx * lpitch + p];
      }
    }
  }
  if (!first) __syncthreads();
  // ---- twiddle and butterfly
#pragma unroll
  for (int b = 0; b < NB; b++) {
    const uint32_t j = t + b * tpp;
    const uint32_t k = j % Ls;
    if (k != 0) {
      if (last && tw_blocks<T>() == 2) {  // finest twiddles: compensated (hi + lo)
#pragma unroll
        for (int r = 1; r < R; r++) v[b * R + r] = cmul_tw<T>(v[b * R + r], tw, n, r * k * tstep);
      } else {
#pragma unroll
        for (int r = 1; r < R; r++) v[b * R + r] = cmul<T>(v[b * R + r], tw[r * k * tstep]);
      }
    }
    bfly<T, R>(&v[b * R]);
  }
  // ---- scatter outputs
#pragma unroll
  for (int b = 0; b < NB; b++) {
    const uint32_t j = t + b * tpp;
    const uint32_t k = j % Ls;
    const uint32_t o0 = (j / Ls) * Ls * R + k;
#pragma unroll
    for (int r = 0; r < R; r++) {
      const uint32_t o = o0 + r * Ls;
      if (last) {
        if (active) gbase[(uint64_t)o * A.elem_stride] = v[b * R + r];
      } else {
        if (CONTIG) lds[p * lpitch + o] = v[b * R + r];
        else lds[o * lpitch + p] = v[b * R + r];
      }
    }
  }
  if (!last) __syncthreads();
}

template <typename T, int E, bool CONTIG>
__global__ void __launch_bounds__(1024) fft_pencil_kernel(StagePlan P, PassArgs A,
                                                          const cx<T> *__restrict__ tw) {
  extern __shared__ unsigned char smem_raw[];
  cx<T> *lds = reinterpret_cast<cx<T> *>(smem_raw);
  const uint32_t tid = threadIdx.x;
  uint32_t t, p;
  if (CONTIG) {
    t = tid % A.tpp;
    p = tid / A.tpp;
  } else {
    p = tid % A.pp;
    t = tid / A.pp;
  }
  const uint32_t g = blockIdx.x % A.groups;
  const uint32_t o = blockIdx.x / A.groups;
  const uint32_t pencil = g * A.pp + p;
  const bool active = pencil < A.inner_count;
  cx<T> *gbase = reinterpret_cast<cx<T> *>(A.vol[blockIdx.y]) + (uint64_t)pencil * A.inner_stride +
                 (uint64_t)o * A.outer_stride;
  const uint32_t lpitch = CONTIG ? P.n : A.pp;
  cx<T> v[E];
  for (uint32_t s = 0; s < P.nstages; s++) {
    const bool first = s == 0, last = s + 1 == P.nstages;
    switch (P.radix[s]) {
      case 2:
        if constexpr (E % 2 == 0) run_stage<T, E, 2, CONTIG>(v, P, s, first, last, t, p, A, gbase, active, lds, lpitch, tw);
        break;
      case 3:
        if constexpr (E % 3 == 0) run_stage<T, E, 3, CONTIG>(v, P, s, first, last, t, p, A, gbase, active, lds, lpitch, tw);
        break;
      case 4:
        if constexpr (E % 4 == 0) run_stage<T, E, 4, CONTIG>(v, P, s, first, last, t, p, A, gbase, active, lds, lpitch, tw);
        break;
      case 5:
        if constexpr (E % 5 == 0) run_stage<T, E, 5, CONTIG>(v, P, s, first, last, t, p, A, gbase, active, lds, lpitch, tw);
        break;
      default:
        if constexpr (E % 8 == 0) run_stage<T, E, 8, CONTIG>(v, P, s, first, last, t, p, A, gbase, active, lds, lpitch, tw);
        break;
    }
  }
}

// Compile-time codelet pass (lengths with a CtPlan): pencils held in registers,
// stages exchanged through LDS with immediate offsets.  Contiguous pencils (X)
// are padded LDS rows, several per wave when short, twiddles read from the
// global table; strided pencils (Y, Z) are PP-column tiles whose twiddles are
// staged in LDS next to the tile once per block.
// Load of one complex element marked non-temporal (streamed: each element of a
// strided pass is read once).  For the 512-point Y pass over 7 volumes of
// 512^2 x 128 it cuts the pass from 717 to 637 us; non-temporal stores add nothing.
template <typename T> __device__ inline cx<T> stream_load(const cx<T> *p) {
  if constexpr (sizeof(T) == 4) {
    const unsigned long long v = __builtin_nontemporal_load(reinterpret_cast<const unsigned long long *>(p));
    return mk<T>(__uint_as_float((unsigned)v), __uint_as_float((unsigned)(v >> 32)));
  } else {
    const double *d = reinterpret_cast<const double *>(p);
    return mk<T>(__builtin_nontemporal_load(d), __builtin_nontemporal_load(d + 1));
  }
}

template <typename T, class F, bool CONTIG, int WG>
__global__ void __launch_bounds__(WG) fft_ct_kernel(PassArgs A, const cx<T> *__restrict__ tw) {
  constexpr int E = F::kE, TPP = F::TPP, N = TPP * E;
  constexpr bool LO = !CONTIG && tw_lo_in_lds<T>((size_t)F::PE * N, N, WG);
  extern __shared__ unsigned char smem_raw[];
  cx<T> *lds = reinterpret_cast<cx<T> *>(smem_raw);
  const cx<T> *twp = tw;
  uint32_t t, p;
  if constexpr (CONTIG) {
    t = threadIdx.x % TPP;
    p = threadIdx.x / TPP;
  } else {
    if (xcd_tile(blockIdx.x, gridDim.x) >= A.groups * A.outer_count * A.nvol)
      return;  // grid padding: the whole block leaves before any barrier
    p = threadIdx.x % A.pp;
    t = threadIdx.x / A.pp;
  }
  uint32_t tile = blockIdx.x, vol = blockIdx.y;
  if constexpr (!CONTIG) {
    // 1D grid in XCD order: neighbouring column groups (sharing HBM sectors) on one XCD
    const uint32_t q = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t per_vol = A.groups * A.outer_count;
    vol = q / per_vol;
    tile = q % per_vol;
  }
  const uint32_t g = tile % A.groups;
  const uint32_t o = tile / A.groups;
  const uint32_t pencil = g * A.pp + p;
  const bool active = pencil < A.inner_count;
  cx<T> *gbase = reinterpret_cast<cx<T> *>(A.vol[vol]) + (uint64_t)(active ? pencil : 0) * A.inner_stride +
                 (uint64_t)o * A.outer_stride;
  cx<T> v[E];
#pragma unroll
  for (int i = 0; i < E; i++) {
    if constexpr (CONTIG) v[i] = active ? gbase[(uint64_t)F::in_pos(t, i) * A.elem_stride] : mk<T>(0, 0);
    else v[i] = active ? stream_load<T>(gbase + (uint64_t)F::in_pos(t, i) * A.elem_stride) : mk<T>(0, 0);
  }
  if constexpr (!CONTIG) {
    // DC-neighbour sums (Y pass of X-transformed volumes, Nx >= 3): columns 0, 1
    // and Nx - 1 of slice z = o add sum_y w G(x, y, z), w = 2 cos(2 pi y / Ny)
    // + 2 cos(2 pi z / Nz) (3D) at x = 0 and 1 otherwise; one FP64 partial per
    // (volume, z, column, wave), the lanes of a wave that hold one column (lane
    // stride pp) combined in a fixed order
    if (A.dcz && (g == 0 || g == A.groups - 1)) {  // the same for the whole block
      const bool mine = active && (pencil <= 1 || pencil == A.inner_count - 1);
      const uint32_t which = pencil <= 1 ? pencil : 2;
      double sr = 0, si = 0;
      if (mine) {
        const double cz =
            A.dc_nz > 1 ? 2.0 * tw_cos<T>(reinterpret_cast<const cx<T> *>(A.dc_twz), A.dc_nz, o) : 0.0;
#pragma unroll
        for (int i = 0; i < E; i++) {
          const double w = which == 0 ? 2.0 * tw_cos<T>(tw, N, F::in_pos(t, i)) + cz : 1.0;
          sr += w * (double)v[i].x;
          si += w * (double)v[i].y;
        }
      }
      for (uint32_t off = A.pp; off < 64; off <<= 1) {
        sr += __shfl_xor(sr, (int)off);
        si += __shfl_xor(si, (int)off);
      }
      if (mine && (threadIdx.x & 63) < A.pp)
        A.dcz[(((uint64_t)vol * A.dc_nz + o) * 3 + which) * (WG / 64) + threadIdx.x / 64] = make_double2(sr, si);
    }
  }
  if constexpr (!CONTIG) {
    // twiddles staged while the pencil loads are in flight, with the last
    // stage's FP32 residuals when they fit without costing occupancy (else they
    // are read from the global table, L2 resident)
    cx<T> *ltw = lds + (size_t)A.pp * N;
    for (uint32_t i = threadIdx.x; i < (uint32_t)(LO ? 2 * N : N); i += WG) ltw[i] = tw[i];
    if (A.span) {  // row spans beside the twiddles
      uint32_t *ls = reinterpret_cast<uint32_t *>(ltw + (LO ? 2 * N : N));
      for (uint32_t i = threadIdx.x; i < (uint32_t)N; i += WG) ls[i] = A.span[i];
    }
    lds_sync();
    twp = ltw;
  }
  // Row spans: output row y is written only where some Ewald band reads it, in
  // whole lines of the block's pp columns (a line is kept when any of its
  // columns' shifted x coordinates, m - x for bins x <= m = n/2 and n + m - x
  // above, lies in the row's lo..hi).  Skipping single columns leaves partial
  // lines, which made the pass slower than writing everything.
  uint32_t keep = (1u << E) - 1;
  if (!CONTIG && A.span) {
    const uint32_t *ls = reinterpret_cast<const uint32_t *>(lds + (size_t)A.pp * N + (LO ? 2 * N : N));
    const uint32_t n = A.inner_count, m = n / 2;
    const uint32_t x0 = g * A.pp, x1 = (x0 + A.pp < n ? x0 + A.pp : n) - 1;
    const bool lo_part = x0 <= m, hi_part = x1 > m;  // the line's bins <= m and > m
    const uint32_t a0 = m - (x1 < m ? x1 : m), a1 = m - x0;
    const uint32_t b0 = n + m - x1, b1 = n + m - (x0 > m + 1 ? x0 : m + 1);
    keep = 0;
#pragma unroll
    for (int i = 0; i < E; i++) {
      const uint32_t sp = ls[F::out_pos(t, i)], lo = sp & 0xffffu, hi = sp >> 16;
      const bool meet = lo <= hi && ((lo_part && a0 <= hi && a1 >= lo) || (hi_part && b0 <= hi && b1 >= lo));
      keep |= meet ? 1u << i : 0u;
    }
  }
  F::run(v, t, p, lds, twp, LO ? twp + N : tw + N);
  if (active) {
#pragma unroll
    for (int i = 0; i < E; i++)
      if (keep & (1u << i)) gbase[(uint64_t)F::out_pos(t, i) * A.elem_stride] = v[i];
  }
}

template <typename T, int N>
hipError_t launch_ct_n(bool contig, uint32_t nc, PassArgs &A, hipStream_t stream, const cx<T> *tw) {
  constexpr int E = N == 384 ? 12 : 8, TPP = N / E;
  if (contig) {
    using F = typename CtPlan<T, N, true, 1>::type;
    constexpr int WG = TPP <= 64 ? 64 : TPP;
    A.pp = WG / TPP;
    A.tpp = TPP;
    A.groups = (A.inner_count + A.pp - 1) / A.pp;
    const uint64_t blocks = (uint64_t)A.groups * A.outer_count;
    if (blocks > 0x7fffffffull) return hipErrorInvalidConfiguration;
    const size_t lds = (size_t)A.pp * F::kPitch * sizeof(cx<T>);
    hipLaunchKernelGGL((fft_ct_kernel<T, F, true, WG>), dim3((unsigned)blocks, nc), dim3(WG), lds, stream, A, tw);
  } else {
    constexpr int PP = N <= 512 ? 16 : (N == 1024 ? 8 : 4);
    using F = typename CtPlan<T, N, false, PP>::type;
    constexpr int WG = PP * TPP;
    A.pp = PP;
    A.tpp = TPP;
    A.groups = (A.inner_count + PP - 1) / PP;
    A.nvol = nc;
    const uint64_t tiles = (uint64_t)A.groups * A.outer_count * nc;
    if (tiles > 0x7ffffff0ull) return hipErrorInvalidConfiguration;
    constexpr bool LO = tw_lo_in_lds<T>((size_t)PP * N, N, WG);
    const size_t lds = ((size_t)PP * N + (LO ? 2 * N : N)) * sizeof(cx<T>) + (A.span ? 4 * (size_t)N : 0);
    hipLaunchKernelGGL((fft_ct_kernel<T, F, false, WG>), dim3(xcd_grid(tiles)), dim3(WG), lds, stream, A, tw);
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_ct(uint32_t n, bool contig, uint32_t nc, PassArgs &A, hipStream_t stream, const cx<T> *tw,
                     bool *found) {
  *found = true;
  switch (n) {
    case 64: return launch_ct_n<T, 64>(contig, nc, A, stream, tw);
    case 128: return launch_ct_n<T, 128>(contig, nc, A, stream, tw);
    case 256: return launch_ct_n<T, 256>(contig, nc, A, stream, tw);
    case 384:
      if constexpr (sizeof(T) == 4) return launch_ct_n<T, 384>(contig, nc, A, stream, tw);
      break;
    case 512: return launch_ct_n<T, 512>(contig, nc, A, stream, tw);
    case 1024: return launch_ct_n<T, 1024>(contig, nc, A, stream, tw);
    case 2048: return launch_ct_n<T, 2048>(contig, nc, A, stream, tw);
    default: break;
  }
  *found = false;
  return hipSuccess;
}

// Direct O(n^2) DFT of each pencil for lengths with prime factors > 5: the
// pencil is staged in LDS, then every thread produces outputs.
template <typename T>
__global__ void __launch_bounds__(256) dft_pencil_kernel(uint32_t n, PassArgs A, const cx<T> *__restrict__ tw) {
  extern __shared__ unsigned char smem_raw[];
  cx<T> *x = reinterpret_cast<cx<T> *>(smem_raw);
  const uint32_t g = blockIdx.x % A.inner_count;
  const uint32_t o = blockIdx.x / A.inner_count;
  cx<T> *base = reinterpret_cast<cx<T> *>(A.vol[blockIdx.y]) + (uint64_t)g * A.inner_stride +
                (uint64_t)o * A.outer_stride;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) x[i] = base[(uint64_t)i * A.elem_stride];
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) {
    T re = 0, im = 0;
    uint32_t e = 0;
    for (uint32_t i = 0; i < n; i++) {
      const cx<T> u = cmul_tw<T>(x[i], tw, n, e);
      re += u.x;
      im += u.y;
      e += k;
      if (e >= n) e -= n;
    }
    base[(uint64_t)k * A.elem_stride] = mk<T>(re, im);
  }
}

// ---- host-side planning ------------------------------------------------------

struct Factoring {
  bool ok = false;
  std::vector<uint32_t> radices;
  uint32_t lcm = 1;
};

Factoring factor(uint32_t n) {
  Factoring f;
  uint32_t r = n, a = 0, b = 0, c = 0;
  while (r % 2 == 0) { r /= 2; a++; }
  while (r % 3 == 0) { r /= 3; b++; }
  while (r % 5 == 0) { r /= 5; c++; }
  if (r != 1) return f;
  while (a >= 3) { f.radices.push_back(8); a -= 3; }
  uint32_t p2 = 1;
  for (uint32_t rr : f.radices) p2 = rr > p2 ? rr : p2;
  if (a == 2) { f.radices.push_back(4); p2 = p2 > 4 ? p2 : 4; }
  if (a == 1) { f.radices.push_back(2); p2 = p2 > 2 ? p2 : 2; }
  for (uint32_t i = 0; i < b; i++) f.radices.push_back(3);
  for (uint32_t i = 0; i < c; i++) f.radices.push_back(5);
  f.lcm = p2 * (b ? 3 : 1) * (c ? 5 : 1);
  f.ok = f.radices.size() <= (size_t)kMaxStages;
  return f;
}

std::mutex g_mu;
std::map<std::tuple<int, uint32_t, int>, DeviceBuffer> g_tw;  // (device, n, precision)
std::map<std::tuple<uint32_t, int, int>, FftShape> g_shapes;   // (n, contig, precision)

template <typename T> constexpr int prec_of() { return sizeof(T) == 8 ? CYR_FP64 : CYR_FP32; }

template <typename T, bool CONTIG>
hipError_t launch_e(uint32_t E, dim3 grid, dim3 block, size_t lds, hipStream_t st, const StagePlan &P,
                    const PassArgs &A, const cx<T> *tw, bool *found) {
  *found = true;
#define CYR_FFT_CASE(EV)                                                                     \
  case EV:                                                                                   \
    hipLaunchKernelGGL((fft_pencil_kernel<T, EV, CONTIG>), grid, block, lds, st, P, A, tw); \
    return hipGetLastError();
  switch (E) {
    CYR_FFT_CASE(2)
    CYR_FFT_CASE(3)
    CYR_FFT_CASE(4)
    CYR_FFT_CASE(5)
    CYR_FFT_CASE(6)
    CYR_FFT_CASE(8)
    CYR_FFT_CASE(10)
    CYR_FFT_CASE(12)
    CYR_FFT_CASE(16)
    CYR_FFT_CASE(20)
    CYR_FFT_CASE(24)
    CYR_FFT_CASE(32)
    CYR_FFT_CASE(40)
    default:
      *found = false;
      return hipSuccess;
  }
#undef CYR_FFT_CASE
}

size_t max_lds_bytes() {
  static size_t v = 0;
  if (!v) {
    int dev = 0;
    hipGetDevice(&dev);
    int b = 0;
    if (hipDeviceGetAttribute(&b, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || b <= 0)
      b = 160 * 1024;  // gfx950 (the only target this library is built for)
    v = (size_t)b;
  }
  return v;
}

}  // namespace

bool make_stage_plan(uint32_t n, StagePlan &P, uint32_t &lcm) {
  const Factoring f = factor(n);
  if (!f.ok) return false;
  P = StagePlan{};
  P.n = n;
  P.nstages = (uint32_t)f.radices.size();
  uint32_t Ls = 1;
  for (uint32_t s = 0; s < P.nstages; s++) {
    P.radix[s] = f.radices[s];
    P.ls[s] = Ls;
    P.tstep[s] = n / (Ls * f.radices[s]);
    Ls *= f.radices[s];
  }
  lcm = f.lcm;
  return true;
}

size_t device_max_lds() { return max_lds_bytes(); }

// Host image of the length-n table (tw_len<T>(n) entries).
template <typename T> void twiddle_host(uint32_t n, cx<T> *h) {
  for (uint32_t m = 0; m < n; m++) {
    // exact-angle FP64 twiddle, rounded once to the compute precision; FP32
    // tables also keep the rounding residual (block 2)
    const double a = 2.0 * M_PI * (double)m / (double)n;
    const double c = std::cos(a), sn = -std::sin(a);
    h[m].x = (T)c;
    h[m].y = (T)sn;
    if (tw_blocks<T>() == 2) {
      h[n + m].x = (T)(c - (double)h[m].x);
      h[n + m].y = (T)(sn - (double)h[m].y);
    }
  }
}

// Under an installed scratch arena the table is built into the arena for this
// call (queued on the arena's stream), never into the process-wide cache.
template <typename T> const cx<T> *arena_twiddles(ScratchArena &a, uint32_t n, cyr_status *st) {
  const auto key = std::make_pair(n, prec_of<T>());
  auto it = a.tw.find(key);
  if (it != a.tw.end()) {
    *st = CYR_OK;
    return reinterpret_cast<const cx<T> *>(it->second);
  }
  const size_t bytes = sizeof(cx<T>) * tw_len<T>(n);
  void *d = a.carve(bytes);
  if (!d) {
    *st = fail(CYR_E_NOMEM, "allocation-free call: twiddle tables exceed the scratch array");
    return nullptr;
  }
  a.keep.emplace_back(bytes);
  twiddle_host<T>(n, reinterpret_cast<cx<T> *>(a.keep.back().data()));
  if (hipMemcpyAsync(d, a.keep.back().data(), bytes, hipMemcpyHostToDevice, a.stream) != hipSuccess) {
    (void)hipGetLastError();
    *st = fail(CYR_E_DEVICE, "twiddle upload failed");
    return nullptr;
  }
  a.tw[key] = d;
  *st = CYR_OK;
  return reinterpret_cast<const cx<T> *>(d);
}

template <typename T> const cx<T> *twiddles(uint32_t n, cyr_status *st) {
  if (ScratchArena *a = scratch_arena()) return arena_twiddles<T>(*a, n, st);
  int dev = 0;
  hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(g_mu);
  auto key = std::make_tuple(dev, n, prec_of<T>());
  auto it = g_tw.find(key);
  if (it != g_tw.end()) {
    *st = CYR_OK;
    return reinterpret_cast<const cx<T> *>(it->second.ptr);
  }
  const size_t len = tw_len<T>(n);
  std::vector<cx<T>> h(len);
  twiddle_host<T>(n, h.data());
  DeviceBuffer b;
  cyr_status s = ensure_buffer(b, sizeof(cx<T>) * len);
  if (s != CYR_OK) {
    *st = s;
    return nullptr;
  }
  if (hipMemcpy(b.ptr, h.data(), sizeof(cx<T>) * len, hipMemcpyHostToDevice) != hipSuccess) {
    free_buffer(b);
    *st = fail(CYR_E_DEVICE, "twiddle upload failed");
    return nullptr;
  }
  g_tw[key] = b;
  *st = CYR_OK;
  return reinterpret_cast<const cx<T> *>(b.ptr);
}

void fft_set_shape(uint32_t n, int contig, int precision, FftShape s) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_shapes[std::make_tuple(n, contig, precision)] = s;
}

void fft_clear_shapes() {
  std::lock_guard<std::mutex> lk(g_mu);
  g_shapes.clear();
}

void fft_release() {
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto &kv : g_tw) free_buffer(kv.second);
  g_tw.clear();
}

// Whether a strided pass of length n runs the compile-time codelets.
template <typename T> bool strided_ct(uint32_t n) {
  if (!ct_length(n) || (sizeof(T) == 8 && ct_elements(n) > 8)) return false;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_shapes.count(std::make_tuple(n, 0, prec_of<T>()))) return false;
  }
  const uint32_t pp = n <= 512 ? 16 : (n == 1024 ? 8 : 4);
  return ((size_t)pp * n + n) * sizeof(cx<T>) <= max_lds_bytes();
}

// Whether the strided codelet's block of length n also holds the row-span table
// (4 bytes per row) without losing occupancy.
template <typename T> bool fft_span_fits(uint32_t n) {
  if (!strided_ct<T>(n)) return false;
  const uint32_t pp = n <= 512 ? 16 : (n == 1024 ? 8 : 4), wg = pp * (n / ct_elements(n));
  const bool lo = tw_lo_in_lds<T>((size_t)pp * n, n, wg);
  const size_t lds = ((size_t)pp * n + (lo ? 2 * n : n)) * sizeof(cx<T>), with = lds + 4 * (size_t)n;
  return with <= max_lds_bytes() && cu_blocks(with, wg) >= cu_blocks(lds, wg);
}

// (the strided codelet's block: pp pencils x n / E threads, in waves of 64)
template <typename T> uint32_t fft_dc_slots(uint32_t nx, uint32_t ny) {
  if (nx < 3 || !strided_ct<T>(ny)) return 0;
  const uint32_t pp = ny <= 512 ? 16 : (ny == 1024 ? 8 : 4);
  return 3 * (pp * (ny / ct_elements(ny)) / 64);
}

template <typename T>
cyr_status fft_axis(void *const *vols, uint32_t count, uint32_t nx, uint32_t ny, uint32_t nz, int axis,
                    hipStream_t stream, const FftDc *dc) {
  const uint32_t n = axis == 0 ? nx : axis == 1 ? ny : nz;
  if (n <= 1 || count == 0) return CYR_OK;
  cyr_status st;
  const cx<T> *tw = twiddles<T>(n, &st);
  if (!tw) return st;
  PassArgs A{};
  const bool contig = axis == 0;
  if (axis == 0) {
    A.elem_stride = 1;
    A.inner_stride = nx;
    A.inner_count = ny;
    A.outer_stride = (uint64_t)nx * ny;
    A.outer_count = nz;
  } else if (axis == 1) {
    A.elem_stride = nx;
    A.inner_stride = 1;
    A.inner_count = nx;
    A.outer_stride = (uint64_t)nx * ny;
    A.outer_count = nz;
  } else {
    A.elem_stride = (uint64_t)nx * ny;
    A.inner_stride = 1;
    A.inner_count = nx;
    A.outer_stride = nx;
    A.outer_count = ny;
  }
  const Factoring f = factor(n);
  const size_t lds_max = max_lds_bytes();
  // compile-time codelets unless a shape override asks for the run-time plan
  bool use_ct;
  if (contig) {
    use_ct = ct_length(n) && (sizeof(T) == 4 || ct_elements(n) <= 8);
    {
      std::lock_guard<std::mutex> lk(g_mu);
      if (g_shapes.count(std::make_tuple(n, 1, prec_of<T>()))) use_ct = false;
    }
    if (use_ct) {
      const size_t need = (size_t)(64 / (n / ct_elements(n)) > 1 ? 64 / (n / ct_elements(n)) : 1) *
                          ct_row_pitch(n) * sizeof(cx<T>);
      if (need > lds_max) use_ct = false;
    }
  } else {
    use_ct = strided_ct<T>(n);
  }
  // DC-neighbour sums only ride along a Y pass on the compile-time codelets
  const uint64_t dc_per_vol = dc ? (uint64_t)nz * fft_dc_slots<T>(nx, ny) : 0;
  if (dc && (axis != 1 || !use_ct || dc_per_vol == 0))
    return fail(CYR_E_INVALID, "FFT: DC sums need a Y pass on a compile-time plan and Nx >= 3");
  for (uint32_t c0 = 0; c0 < count; c0 += 8) {
    const uint32_t nc = count - c0 < 8 ? count - c0 : 8;
    for (uint32_t i = 0; i < nc; i++) A.vol[i] = vols[c0 + i];
    A.dcz = dc ? dc->dcz + c0 * dc_per_vol : nullptr;
    A.dc_twz = dc ? dc->twz : nullptr;
    A.dc_nz = nz;
    A.span = dc && dc->span && fft_span_fits<T>(ny) ? dc->span : nullptr;
    bool done = false;
    if (use_ct) {
      bool found = false;
      const hipError_t e = launch_ct<T>(n, contig, nc, A, stream, tw, &found);
      if (e != hipSuccess) return hip_fail(e, "fft_ct_kernel", __FILE__, __LINE__);
      done = found;
    }
    if (!done && f.ok) {
      // Elements per thread: the radix lcm, doubled while a pencil would need more
      // than 256 threads; pencils per block: up to 256 threads (contiguous) or a
      // 16-pencil coalescing group (strided), bounded by LDS.
      FftShape shp;
      {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_shapes.find(std::make_tuple(n, (int)contig, prec_of<T>()));
        if (it != g_shapes.end()) shp = it->second;
      }
      uint32_t E = f.lcm;
      const uint32_t tpp_cap = contig ? 256 : 32;  // strided: 16 pencils x 32 threads
      const uint32_t e_cap = sizeof(T) == 8 ? 16 : 32;
      if (shp.e && shp.e % f.lcm == 0 && n % shp.e == 0) E = shp.e;
      else
        while (n / E > tpp_cap && (E * 2) <= e_cap && n % (E * 2) == 0) E *= 2;
      if (n % E == 0 && n >= E) {
        const uint32_t tpp = n / E;
        uint32_t pp;
        if (shp.pp) pp = shp.pp;
        else if (contig) pp = tpp >= 256 ? 1 : 256 / tpp;
        else pp = 16;
        while (pp > 1 && ((size_t)pp * n * sizeof(cx<T>) > lds_max || pp * tpp > 1024)) pp /= 2;
        if (contig && pp > A.inner_count) pp = A.inner_count;
        if ((size_t)pp * n * sizeof(cx<T>) <= lds_max && pp * tpp <= 1024) {
          StagePlan P{};
          P.n = n;
          P.nstages = (uint32_t)f.radices.size();
          uint32_t Ls = 1;
          for (uint32_t s = 0; s < P.nstages; s++) {
            P.radix[s] = f.radices[s];
            P.ls[s] = Ls;
            P.tstep[s] = n / (Ls * f.radices[s]);
            Ls *= f.radices[s];
          }
          A.pp = pp;
          A.tpp = tpp;
          A.groups = (A.inner_count + pp - 1) / pp;
          const uint64_t blocks = (uint64_t)A.groups * A.outer_count;
          if (blocks > 0x7fffffffull) return fail(CYR_E_UNSUPPORTED, "FFT grid too large");
          dim3 grid((unsigned)blocks, nc), block(pp * tpp);
          const size_t lds = (size_t)pp * n * sizeof(cx<T>);
          bool found = false;
          hipError_t e = contig ? launch_e<T, true>(E, grid, block, lds, stream, P, A, tw, &found)
                                : launch_e<T, false>(E, grid, block, lds, stream, P, A, tw, &found);
          if (e != hipSuccess) return hip_fail(e, "fft_pencil_kernel", __FILE__, __LINE__);
          done = found;
        }
      }
    }
    if (!done) {
      if ((size_t)n * sizeof(cx<T>) > lds_max) return fail(CYR_E_UNSUPPORTED, "FFT length not supported");
      const uint64_t blocks = (uint64_t)A.inner_count * A.outer_count;
      if (blocks > 0x7fffffffull) return fail(CYR_E_UNSUPPORTED, "FFT grid too large");
      dim3 grid((unsigned)blocks, nc), block(256);
      hipLaunchKernelGGL((dft_pencil_kernel<T>), grid, block, (size_t)n * sizeof(cx<T>), stream, n, A, tw);
      CYR_LAUNCHED();
    }
  }
  return CYR_OK;
}

template <typename T>
cyr_status fft3d(void *const *vols, uint32_t count, uint32_t nx, uint32_t ny, uint32_t nz, hipStream_t stream) {
  CYR_TRY(fft_axis<T>(vols, count, nx, ny, nz, 0, stream));
  CYR_TRY(fft_axis<T>(vols, count, nx, ny, nz, 1, stream));
  CYR_TRY(fft_axis<T>(vols, count, nx, ny, nz, 2, stream));
  return CYR_OK;
}

template const cx<float> *twiddles<float>(uint32_t, cyr_status *);
template const cx<double> *twiddles<double>(uint32_t, cyr_status *);
template cyr_status fft_axis<float>(void *const *, uint32_t, uint32_t, uint32_t, uint32_t, int, hipStream_t,
                                    const FftDc *);
template cyr_status fft_axis<double>(void *const *, uint32_t, uint32_t, uint32_t, uint32_t, int, hipStream_t,
                                     const FftDc *);
template uint32_t fft_dc_slots<float>(uint32_t, uint32_t);
template uint32_t fft_dc_slots<double>(uint32_t, uint32_t);
template bool fft_span_fits<float>(uint32_t);
template bool fft_span_fits<double>(uint32_t);
template cyr_status fft3d<float>(void *const *, uint32_t, uint32_t, uint32_t, uint32_t, hipStream_t);
template cyr_status fft3d<double>(void *const *, uint32_t, uint32_t, uint32_t, uint32_t, hipStream_t);

}  // namespace cyr
