// Executors and the whole-simulation / stage-level C ABI.
//
// PLAIN (reference structure, src/cudaMain.cu:586-971): for every energy, k
// vector and E angle: polarization -> 3 full 3D FFTs -> DC replacement ->
// Ewald projection (shift folded into the gather) -> rotate + mask + FP64
// accumulate; then average, detector warp, output.
//
// EBASIS: p(theta) = c(theta) p1 + s(theta) p2 with p1, p2 the polarizations
// for the E basis vectors e1 = R_K x^, e2 = k^ x e1 (E(theta) = R(theta) x^ lies
// in their plane).  Every later stage is linear or a permutation, so per
// (energy, k) the pipeline computes two polarizations, their XY FFTs, and the z
// DFT only at each pixel's Ewald bin(s); all angles then become 2D work.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "fft.hpp"
#include "host_math.hpp"
#include "kernels.hpp"
#include "pipeline.hpp"

namespace cyr {

// The fused E-basis kernels need 2^a 3^b 5^c in-plane lengths and tiles that fit
// in LDS; FP64 runs them on native-precision morphologies only.  Other
// simulations asking for E-basis run the plain executor.
bool ebasis_usable(const cyr_sim_desc *d, const SimGeometry &g) {
  return ebasis_supported(g.nx, g.ny, g.nz, d->precision) &&
         !(d->precision == CYR_FP64 && d->storage != CYR_STORE_NATIVE);
}

// HBM bytes per voxel (E-basis: per energy and k vector; material basis: per k
// vector) and per pixel and energy.  The E-basis Y/z kernel is ALU-bound at about
// twice its bytes, so its reads count double.  Measured at 512^2 x 128, 4
// materials: E-basis 1.26 ms per energy, material basis 3.7 ms per k + 0.05 ms
// per energy; the model's break-even (2.8 energies) matches.
int resolved_pipeline(const cyr_sim_desc *d, const SimGeometry &g) {
  if (d->pipeline == CYR_PIPE_PLAIN || !ebasis_usable(d, g)) return CYR_PIPE_PLAIN;
  if (d->pipeline != CYR_PIPE_AUTO) return d->pipeline;
  const double cs = d->precision == CYR_FP64 ? 16.0 : 8.0;
  const double el = d->storage == CYR_STORE_NATIVE ? cs / 2 : 2.0;
  const double nv = mbasis_volumes(g.M), ng = (g.nz > 1 && d->interpolation == 1) ? 2.0 : 1.0;
  const double morph = 4.0 * el * g.M;
  const double eb = (double)g.nE * (double)g.V * (morph + 15.0 * cs);
  const double mb = (double)g.V * (morph + 4.0 * cs * nv) + (double)g.nE * (double)g.P * 4.0 * ng * nv * cs;
  return mb < eb ? CYR_PIPE_MBASIS : CYR_PIPE_EBASIS;
}

namespace {

template <typename U>
__global__ void __launch_bounds__(256) fill_kernel(U *__restrict__ p, U v, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <typename U> cyr_status fill(U *p, U v, uint64_t n, hipStream_t st) {
  if (n == 0) return CYR_OK;
  hipLaunchKernelGGL((fill_kernel<U>), dim3(ceil_div(n, 256)), dim3(256), 0, st, p, v, n);
  CYR_LAUNCHED();
  return CYR_OK;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// n = (1 - delta) + i beta  ->  n^2 - 1 = delta^2 - 2 delta - beta^2 + 2i (1 - delta) beta
void sq_minus_1(double delta, double beta, double &re, double &im) {
  re = delta * delta - 2.0 * delta - beta * beta;
  im = 2.0 * (1.0 - delta) * beta;
}

template <typename T> MatConst<T> mat_const(const double *m4) {
  const double dp = 1.0 - m4[0], bp = m4[1], dq = 1.0 - m4[2], bq = m4[3];
  double ar, ai, br, bi, cr, ci;
  sq_minus_1(dp, bp, ar, ai);
  sq_minus_1(dq, bq, br, bi);
  sq_minus_1((dp + 2.0 * dq) / 3.0, (bp + 2.0 * bq) / 3.0, cr, ci);
  MatConst<T> r;
  r.ar = (T)ar;
  r.ai = (T)ai;
  r.br = (T)br;
  r.bi = (T)bi;
  r.cr = (T)cr;
  r.ci = (T)ci;
  r.dr = (T)(ar - br);
  r.di = (T)(ai - bi);
  return r;
}

// Orbit tiles of the quarter-turn rotation kernel (0 when it does not apply).
inline uint32_t orbit_tiles(const SimGeometry &g) {
  return rotation_quarter_turns(g.nx, g.ny, g.nA, g.quarter) ? rotation_orbit_tile_count(g.nx) : 0;
}

// Per-simulation constant tables, uploaded once.
struct Tables {
  size_t mats = 0, han = 0, amat = 0, rotinv = 0, detinv = 0, kmag = 0, quad = 0, quad_t = 0, tiles = 0, otiles = 0,
         total = 0;
  std::vector<unsigned char> host;
};

// Frame of the E-basis tensor fields: columns e1, e2, e1 x e2 (= k^), row-major.
M3 ebasis_frame(const SimGeometry &g, uint32_t k) {
  const V3 a = g.e1[k], b = g.e2[k];
  const V3 c{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
  M3 q;
  const double col[3][3] = {{a.x, a.y, a.z}, {b.x, b.y, b.z}, {c.x, c.y, c.z}};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) q.m[i * 3 + j] = col[j][i];
  return q;
}

bool is_identity(const M3 &q) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      if (q.m[i * 3 + j] != (i == j ? 1.0 : 0.0)) return false;
  return true;
}

// Per (k, angle): A (3 x 5) with p~ = A F for the tensor fields F = (N'00, N'01,
// N'02, N'11, N'12): p' = N' (c, s, 0)^T, p = Q p', and p <- R^T p in the
// MATERIAL frame (include/uniaxial.h:163-165).
void ebasis_amat(const SimGeometry &g, uint32_t k, uint32_t a, bool matframe, double A[kEbAmat]) {
  const M3 Q = ebasis_frame(g, k);
  const M3 &R = g.R[(size_t)k * g.nA + a];
  double M[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double v = 0;
      if (matframe)
        for (int l = 0; l < 3; l++) v += R.m[l * 3 + i] * Q.m[l * 3 + j];
      else
        v = Q.m[i * 3 + j];
      M[i * 3 + j] = v;
    }
  const double cc = g.ebasis[((size_t)k * g.nA + a) * 2 + 0], ss = g.ebasis[((size_t)k * g.nA + a) * 2 + 1];
  const double C[3][kEbFields] = {{cc, ss, 0, 0, 0}, {0, cc, 0, ss, 0}, {0, 0, cc, 0, ss}};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < kEbFields; j++) {
      double v = 0;
      for (int l = 0; l < 3; l++) v += M[i * 3 + l] * C[l][j];
      A[i * kEbFields + j] = v;
    }
}

// Offsets of the table sections (shared by upload and resident runs).
template <typename T> void build_layout(const SimGeometry &g, Tables &tb) {
  size_t off = 0;
  tb.mats = off;
  off = align256(off + sizeof(MatConst<T>) * g.nE * g.M);
  tb.han = off;
  off = align256(off + sizeof(T) * (g.nx + g.ny + g.nz));
  tb.amat = off;
  off = align256(off + sizeof(T) * kEbAmat * g.nK * g.nA);
  tb.rotinv = off;
  off = align256(off + sizeof(double) * 6 * g.nK * g.nA);
  tb.detinv = off;
  off = align256(off + sizeof(double) * 6 * g.nK);
  tb.kmag = off;
  off = align256(off + sizeof(double) * g.nE);
  tb.quad = off;
  off = align256(off + sizeof(double) * 3 * g.nK * g.nA);
  tb.quad_t = off;
  off = align256(off + sizeof(T) * 3 * g.nK * g.nA);
  tb.tiles = off;
  off = align256(off + sizeof(uint32_t) * rotation_tile_count(g.nx, g.ny));
  tb.otiles = off;
  off = align256(off + sizeof(uint32_t) * orbit_tiles(g));
  tb.total = off;
}

template <typename T>
void build_tables(const cyr_sim_desc *d, const SimGeometry &g, const double *mats, Tables &tb) {
  build_layout<T>(g, tb);
  tb.host.assign(tb.total, 0);
  MatConst<T> *mc = reinterpret_cast<MatConst<T> *>(tb.host.data() + tb.mats);
  for (uint32_t e = 0; e < g.nE; e++)
    for (uint32_t m = 0; m < g.M; m++) mc[(size_t)e * g.M + m] = mat_const<T>(mats + 4 * ((size_t)e * g.M + m));
  // Hanning h(i, N) = 0.5 (1 - cos(2 pi i / N)) (src/cudaMain.cu:283-289); 1 along z in 2D
  T *h = reinterpret_cast<T *>(tb.host.data() + tb.han);
  for (uint32_t i = 0; i < g.nx; i++) h[i] = (T)(0.5 * (1 - std::cos(2 * M_PI * i / (g.nx))));
  for (uint32_t i = 0; i < g.ny; i++) h[g.nx + i] = (T)(0.5 * (1 - std::cos(2 * M_PI * i / (g.ny))));
  for (uint32_t i = 0; i < g.nz; i++)
    h[g.nx + g.ny + i] = g.nz == 1 ? (T)1 : (T)(0.5 * (1 - std::cos(2 * M_PI * i / (g.nz))));
  T *am = reinterpret_cast<T *>(tb.host.data() + tb.amat);
  for (uint32_t k = 0; k < g.nK; k++)
    for (uint32_t a = 0; a < g.nA; a++) {
      double A[kEbAmat];
      ebasis_amat(g, k, a, d->reference_frame == 0, A);
      for (int j = 0; j < kEbAmat; j++) am[((size_t)k * g.nA + a) * kEbAmat + j] = (T)A[j];
    }
  std::memcpy(tb.host.data() + tb.rotinv, g.rot_inv.data(), sizeof(double) * 6 * g.nK * g.nA);
  std::memcpy(tb.host.data() + tb.detinv, g.det_inv.data(), sizeof(double) * 6 * g.nK);
  std::memcpy(tb.host.data() + tb.kmag, g.kmag.data(), sizeof(double) * g.nE);
  // laboratory frame: I(theta) = c^2 q0 + s^2 q1 + 2 c s q2 with E = c e1 + s e2
  double *qc = reinterpret_cast<double *>(tb.host.data() + tb.quad);
  T *qt = reinterpret_cast<T *>(tb.host.data() + tb.quad_t);
  for (size_t i = 0; i < (size_t)g.nK * g.nA; i++) {
    const double c = g.ebasis[2 * i], sn = g.ebasis[2 * i + 1];
    qc[3 * i + 0] = c * c;
    qc[3 * i + 1] = sn * sn;
    qc[3 * i + 2] = 2.0 * c * sn;
    for (int j = 0; j < 3; j++) qt[3 * i + j] = (T)qc[3 * i + j];
  }
  rotation_tile_order(g.nx, g.ny, reinterpret_cast<uint32_t *>(tb.host.data() + tb.tiles));
  if (orbit_tiles(g)) rotation_orbit_tile_order(g.nx, reinterpret_cast<uint32_t *>(tb.host.data() + tb.otiles));
}

template <typename T> Vec3<T> v3(const V3 &v) { return Vec3<T>{(T)v.x, (T)v.y, (T)v.z}; }
template <typename T> Mat3<T> m3(const M3 &m) {
  Mat3<T> r;
  for (int i = 0; i < 9; i++) r.m[i] = (T)m.m[i];
  return r;
}

// Constant tables and per-simulation setup shared by both executors.
template <typename T> struct Setup {
  const MatConst<T> *mats;
  const T *hx, *hy, *hz, *amat;
  const double *rotinv, *kmag, *quad;
  const T *quad_t;
  const uint32_t *tiles, *otiles;  // otiles: null without quarter-turn orbits
  const cx<T> *twy, *twz;
  double vox_bytes, cxb;
};

// Device views of the tables in wb.tables (laid out by build_tables) and the
// twiddle tables; everything is created before any kernel is queued.
template <typename T>
cyr_status setup_views(const cyr_sim_desc *d, const SimGeometry &g, const Tables &tb, const void *tables,
                       Setup<T> &su) {
  if (!tables) return fail(CYR_E_INVALID, "tables not resident");
  const unsigned char *tbase = reinterpret_cast<const unsigned char *>(tables);
  su.mats = reinterpret_cast<const MatConst<T> *>(tbase + tb.mats);
  const T *han = reinterpret_cast<const T *>(tbase + tb.han);
  su.hx = han;
  su.hy = han + g.nx;
  su.hz = han + g.nx + g.ny;
  su.amat = reinterpret_cast<const T *>(tbase + tb.amat);
  su.rotinv = reinterpret_cast<const double *>(tbase + tb.rotinv);
  su.kmag = reinterpret_cast<const double *>(tbase + tb.kmag);
  su.quad = reinterpret_cast<const double *>(tbase + tb.quad);
  su.quad_t = reinterpret_cast<const T *>(tbase + tb.quad_t);
  su.tiles = reinterpret_cast<const uint32_t *>(tbase + tb.tiles);
  su.otiles = orbit_tiles(g) ? reinterpret_cast<const uint32_t *>(tbase + tb.otiles) : nullptr;
  cyr_status s;
  su.twz = g.nz > 1 ? twiddles<T>(g.nz, &s) : nullptr;
  if (g.nz > 1 && !su.twz) return s;
  su.twy = twiddles<T>(g.ny, &s);
  if (!su.twy) return s;
  if (!twiddles<T>(g.nx, &s)) return s;
  su.cxb = (double)sizeof(cx<T>);
  const double vox_el = d->storage == CYR_STORE_NATIVE ? (double)sizeof(T) : 2.0;
  su.vox_bytes = 4.0 * vox_el * g.M * g.V;
  return CYR_OK;
}

// PLAIN: the reference's per-angle structure (src/cudaMain.cu:586-971).
template <typename T>
cyr_status run_plain(const cyr_sim_desc *d, const SimGeometry &g, const void *d_vox, T *d_out, WorkBuffers &wb,
                     hipStream_t st, const Setup<T> &su) {
  const uint64_t V = g.V, P = g.P;
  const bool alg1 = d->algorithm == 1;
  const bool hann = d->windowing == 1 && !alg1;  // Algorithm 1 has no window (include/uniaxial.h:441-521)
  const bool matframe = d->reference_frame == 0;
  const bool full = d->scatter_approach == 1;
  const int ng = (g.nz > 1 && d->interpolation == 1) ? 2 : 1;
  LaneBuffers &lb = wb.lane[0];
  CYR_TRY(ensure_buffer(lb.pol, sizeof(cx<T>) * V * 3));
  cx<T> *pol = reinterpret_cast<cx<T> *>(lb.pol.ptr);
  cx<T> *vols[3];
  void *vptr[3];
  for (uint32_t c = 0; c < 3; c++) {
    vols[c] = pol + (uint64_t)c * V;
    vptr[c] = vols[c];
  }
  if (alg1) CYR_TRY(ensure_buffer(wb.nt, sizeof(cx<T>) * V * 6));
  if (full) CYR_TRY(ensure_buffer(wb.scatter, sizeof(T) * V));
  CYR_TRY(ensure_buffer(lb.proj, sizeof(T) * P));
  CYR_TRY(ensure_buffer(wb.acc, sizeof(double) * P));
  CYR_TRY(ensure_buffer(wb.count, sizeof(uint32_t) * P));
  CYR_TRY(ensure_buffer(lb.avg, sizeof(double) * P));
  T *proj = reinterpret_cast<T *>(lb.proj.ptr);
  double *acc = reinterpret_cast<double *>(wb.acc.ptr);
  uint32_t *count = reinterpret_cast<uint32_t *>(wb.count.ptr);
  double *avg = reinterpret_cast<double *>(lb.avg.ptr);
  const double cxb = su.cxb;

  PolarArgs<T> pa{};
  pa.M = g.M;
  pa.V = V;
  pa.nx = g.nx;
  pa.ny = g.ny;
  pa.nz = g.nz;
  pa.hx = su.hx;
  pa.hy = su.hy;
  pa.hz = su.hz;
  for (uint32_t c = 0; c < 3; c++) pa.out[c] = vols[c];

  for (uint32_t e = 0; e < g.nE; e++) {
    pa.mats = su.mats + (size_t)e * g.M;
    if (alg1)
      CYR_TRY(launch_nt<T>(pa, d_vox, d->storage, d->morphology_type, reinterpret_cast<cx<T> *>(wb.nt.ptr), st));
    for (uint32_t k = 0; k < g.nK; k++) {
      const double kv[3] = {g.kvec[k].x, g.kvec[k].y, g.kvec[k].z};
      const EwaldGeom geo = make_ewald_geom(g.nx, g.ny, g.nz, d->phys_size, g.kmag[e], kv, d->interpolation);
      T *out = d_out + ((uint64_t)e * g.nK + k) * P;
      CYR_TRY(fill<double>(acc, 0.0, P, st));
      CYR_TRY(fill<uint32_t>(count, 0u, P, st));
      for (uint32_t a = 0; a < g.nA; a++) {
        const M3 &R = g.R[(size_t)k * g.nA + a];
        pa.e[0] = Vec3<T>{(T)R.m[0], (T)R.m[3], (T)R.m[6]};  // E = R x^
        pa.R = m3<T>(R);
        kt_begin(st);
        if (alg1)
          CYR_TRY(launch_polarization_nt<T>(pa, reinterpret_cast<const cx<T> *>(wb.nt.ptr), 1, matframe, st));
        else
          CYR_TRY(launch_polarization<T>(pa, d_vox, d->storage, d->morphology_type, 1, hann, matframe, st));
        kt_end(alg1 ? "polarization_nt" : "polarization", st, (alg1 ? 6.0 * V * cxb : su.vox_bytes) + 3.0 * V * cxb);
        kt_begin(st);
        CYR_TRY(fft3d<T>(vptr, 3, g.nx, g.ny, g.nz, st));
        kt_end("fft3d_3vol", st, (g.nz > 1 ? 3.0 : 2.0) * 3 * 2 * V * cxb);
        kt_begin(st);
        CYR_TRY(launch_dc_replace<T>(vols, 3, g.nx, g.ny, g.nz, st));
        kt_end("dc_replace", st);
        if (full) {
          T *sc = reinterpret_cast<T *>(wb.scatter.ptr);
          CYR_TRY(launch_scatter3d<T>(geo, vols[0], vols[1], vols[2], sc, st));
          CYR_TRY(launch_ewald_full<T>(geo, sc, proj, st));
        } else {
          kt_begin(st);
          CYR_TRY(launch_ewald_partial<T>(geo, vols[0], vols[1], vols[2], proj, st));
          kt_end("ewald_partial", st, 3.0 * ng * P * cxb + (double)P * sizeof(T));
        }
        kt_begin(st);
        CYR_TRY(launch_rotate_accumulate<T>(proj, acc, count, g.nx, g.ny, &g.rot_inv[((size_t)k * g.nA + a) * 6],
                                            d->rot_mask, st));
        kt_end("rotate_accumulate", st, (double)P * (sizeof(T) + 2 * (8 + 4)));
      }
      CYR_TRY(launch_average(acc, count, avg, P, g.nA, d->rot_mask, st));
      CYR_TRY(launch_detector_warp<T>(avg, out, g.nx, g.ny, &g.det_inv[(size_t)k * 6], d->rot_mask, st));
    }
  }
  return CYR_OK;
}

// EBASIS: per (energy, k) unit the five tensor fields, their XY transforms, the
// z DFT at the Ewald bins, then every angle as 2D work.  With two lanes, unit u
// runs on stream u % 2 with its own work buffers and starts once unit u-1's
// X kernel has finished, so one unit's X kernel (HBM-heavy) overlaps the
// previous unit's Y/z chain (ALU/LDS-heavy).
template <typename T>
cyr_status run_ebasis(const cyr_sim_desc *d, const SimGeometry &g, const void *d_vox, T *d_out, WorkBuffers &wb,
                      hipStream_t st, const Setup<T> &su) {
  const uint64_t V = g.V, P = g.P;
  const bool hann = d->windowing == 1 && d->algorithm != 1;
  const int ng = (g.nz > 1 && d->interpolation == 1) ? 2 : 1;
  const uint32_t units = g.nE * g.nK;
  const uint32_t lanes = units > 1 ? std::min(energy_lanes(), wb.max_lanes) : 1;
  YzArgs<T> yz{};
  yz.nx = g.nx;
  yz.ny = g.ny;
  yz.nz = g.nz;
  CYR_TRY(ebasis_yz_plan<T>(yz, 0));
  for (uint32_t L = 0; L < lanes; L++) {
    LaneBuffers &lb = wb.lane[L];
    CYR_TRY(ensure_buffer(lb.pol, sizeof(cx<T>) * V * kEbFields));
    CYR_TRY(ensure_buffer(lb.proj, sizeof(T) * P * g.nA));
    CYR_TRY(ensure_buffer(lb.avg, sizeof(double) * P));
    CYR_TRY(ensure_buffer(lb.bins, sizeof(int32_t) * P));
    CYR_TRY(ensure_buffer(lb.part, sizeof(cx<T>) * P * ng * kEbFields * yz.nchunks));
    CYR_TRY(ensure_buffer(lb.dc, sizeof(double2) * kEbFields * g.nz));
  }
  hipStream_t lane_st[2] = {st, st};
  if (lanes > 1) {
    CYR_TRY(wb.ensure_aux());
    lane_st[1] = wb.aux;
    CYR_HIP(hipEventRecord(wb.ev_fork, st));
    CYR_HIP(hipStreamWaitEvent(wb.aux, wb.ev_fork, 0));
  }
  const double cxb = su.cxb;
  const double part_bytes = cxb * P * ng * kEbFields * yz.nchunks;
  for (uint32_t u = 0; u < units; u++) {
    const uint32_t e = u / g.nK, k = u % g.nK, L = u % lanes;
    hipStream_t s = lane_st[L];
    LaneBuffers &lb = wb.lane[L];
    cx<T> *pol = reinterpret_cast<cx<T> *>(lb.pol.ptr);
    int32_t *kzt = reinterpret_cast<int32_t *>(lb.bins.ptr);
    cx<T> *part = reinterpret_cast<cx<T> *>(lb.part.ptr);
    double2 *dcz = reinterpret_cast<double2 *>(lb.dc.ptr);
    T *proj = reinterpret_cast<T *>(lb.proj.ptr);
    double *avg = reinterpret_cast<double *>(lb.avg.ptr);
    if (lanes > 1 && u > 0) CYR_HIP(hipStreamWaitEvent(s, wb.ev_polx[(u - 1) % lanes], 0));
    const double kv[3] = {g.kvec[k].x, g.kvec[k].y, g.kvec[k].z};
    const EwaldGeom geo = make_ewald_geom(g.nx, g.ny, g.nz, d->phys_size, g.kmag[e], kv, d->interpolation);
    T *out = d_out + ((uint64_t)e * g.nK + k) * P;
    // bins -> tensor fields + X FFT -> DC neighbours -> Y FFT + z bins -> angles -> rotation
    kt_begin(s);
    CYR_TRY(launch_ebasis_bins(geo, kzt, s));
    kt_end("ebasis_bins", s, (double)P * 4);
    PolxArgs<T> xa{};
    xa.mats = su.mats + (size_t)e * g.M;
    xa.M = g.M;
    xa.V = V;
    xa.nx = g.nx;
    xa.ny = g.ny;
    xa.nz = g.nz;
    xa.rows = g.ny * g.nz;
    xa.hx = su.hx;
    xa.hy = su.hy;
    xa.hz = su.hz;
    const M3 Q = ebasis_frame(g, k);
    xa.Q = m3<T>(Q);
    xa.rot = is_identity(Q) ? 0 : 1;
    for (int c = 0; c < kEbFields; c++) xa.out[c] = pol + (uint64_t)c * V;
    kt_begin(s);
    CYR_TRY(launch_ebasis_polx<T>(xa, d_vox, d->storage, d->morphology_type, hann, s));
    kt_end("xfields", s, su.vox_bytes + kEbFields * (double)V * cxb);
    if (lanes > 1) CYR_HIP(hipEventRecord(wb.ev_polx[L], s));
    DcArgs<T> da{};
    for (int c = 0; c < kEbFields; c++) da.vol[c] = pol + (uint64_t)c * V;
    da.twy = su.twy;
    da.twz = su.twz;
    da.nx = g.nx;
    da.ny = g.ny;
    da.nz = g.nz;
    da.dcz = dcz;
    kt_begin(s);
    CYR_TRY(launch_ebasis_dc<T>(da, s));
    kt_end("ebasis_dc", s, 3.0 * kEbFields * (double)g.ny * g.nz * cxb);
    YzArgs<T> ya = yz;
    for (int c = 0; c < kEbFields; c++) ya.vol[c] = pol + (uint64_t)c * V;
    ya.part = part;
    ya.kzt = g.nz > 1 ? kzt : nullptr;
    ya.twz = su.twz;
    kt_begin(s);
    CYR_TRY(launch_ebasis_yz<T>(ya, ng, s));
    kt_end("ebasis_yz", s, kEbFields * (double)V * cxb + (g.nz > 1 ? (double)P * 4 : 0.0) + part_bytes);
    AnglesArgs<T> aa{};
    aa.g = geo;
    aa.part = part;
    aa.nchunks = yz.nchunks;
    aa.dcz = dcz;
    aa.amat = su.amat + (size_t)kEbAmat * k * g.nA;
    aa.nA = g.nA;
    aa.proj = proj;
    kt_begin(s);
    CYR_TRY(launch_ebasis_angles<T>(aa, ng, s));
    kt_end("ebasis_angles", s, part_bytes + (double)g.nA * P * sizeof(T));
    kt_begin(s);
    CYR_TRY(launch_rotate_all<T>(proj, su.rotinv + (size_t)6 * k * g.nA, g.nA, g.nx, g.ny, d->rot_mask, avg,
                                 su.tiles, -1.0, s));
    kt_end("rotate_all", s, (double)(g.nA + 1) * P * sizeof(T));
    CYR_TRY(launch_detector_warp<T>(avg, out, g.nx, g.ny, &g.det_inv[(size_t)k * 6], d->rot_mask, s));
  }
  if (lanes > 1) {
    CYR_HIP(hipEventRecord(wb.ev_join, wb.aux));
    CYR_HIP(hipStreamWaitEvent(st, wb.ev_join, 0));
  }
  return CYR_OK;
}

// Widest Ewald band and the sum of the columns' band widths over energies
// [e0, e0 + ne) for k, by (grid, k, |k| list): known after one eager run, so graph
// capture can size its launches.
std::mutex g_band_mu;
// per band geometry: widest band, sum of the band widths, columns inside the row
// spans, D4 (no finite plane value beyond sqrt(D4) / 2 pixels of the centre)
struct BandWidths {
  uint32_t J;
  uint64_t wsum, ssum, d4;
};
std::map<std::vector<double>, BandWidths> g_band_widths;

// Band table of energies [e0, e0 + ne) for k vector k into `band`, its widest band
// in *J and the sum of the widths in *wsum (read back once per distinct geometry,
// which waits for the stream; a stream under graph capture needs them cached by
// an earlier run).
cyr_status mbasis_band(const cyr_sim_desc *d, const SimGeometry &g, uint32_t k, uint32_t e0, uint32_t ne,
                       const double *d_kmag, int32_t *band, uint32_t *bmax, hipStream_t s, uint32_t *J,
                       uint64_t *wsum, uint64_t *ssum, uint64_t *d4) {
  const double kv[3] = {g.kvec[k].x, g.kvec[k].y, g.kvec[k].z};
  const EwaldGeom geo = make_ewald_geom(g.nx, g.ny, g.nz, d->phys_size, g.kmag[e0], kv, d->interpolation);
  CYR_TRY(launch_mbasis_bands(geo, d_kmag + e0, ne, band, bmax, s));
  std::vector<double> key = {(double)g.nx, (double)g.ny, (double)g.nz, d->phys_size, (double)d->interpolation,
                             kv[0], kv[1], kv[2]};
  key.insert(key.end(), g.kmag.begin() + e0, g.kmag.begin() + e0 + ne);
  {
    std::lock_guard<std::mutex> lk(g_band_mu);
    auto it = g_band_widths.find(key);
    if (it != g_band_widths.end()) {
      *J = it->second.J;
      *wsum = it->second.wsum;
      *ssum = it->second.ssum;
      *d4 = it->second.d4;
      return CYR_OK;
    }
  }
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  CYR_HIP(hipStreamIsCapturing(s, &cap));
  if (cap != hipStreamCaptureStatusNone)
    return fail(CYR_E_INVALID, "material basis: Ewald band widths unknown under capture (run the simulation once first)");
  uint32_t h[kBandHead] = {};
  CYR_HIP(hipMemcpyAsync(h, bmax, sizeof(h), hipMemcpyDeviceToHost, s));
  CYR_HIP(hipStreamSynchronize(s));
  const BandWidths bw{h[0], (uint64_t)h[1] | ((uint64_t)h[2] << 32), (uint64_t)h[3] | ((uint64_t)h[4] << 32),
                      (uint64_t)h[5] | ((uint64_t)h[6] << 32)};
  std::lock_guard<std::mutex> lk(g_band_mu);
  if (g_band_widths.size() >= 4096) g_band_widths.clear();  // bounded: entries are cheap to recompute
  g_band_widths[key] = bw;
  *J = bw.J;
  *wsum = bw.wsum;
  *ssum = bw.ssum;
  *d4 = bw.d4;
  return CYR_OK;
}

// Energies whose quadratic forms and averages are batched per launch (laboratory
// frame): up to 32, and no more than ~1 GiB of planes.
uint32_t mbasis_energy_batch(const SimGeometry &g, size_t real_bytes) {
  const uint64_t per = (uint64_t)g.P * (4 * real_bytes + 8);
  uint64_t nb = (1ull << 30) / (per ? per : 1);
  if (nb < 1) nb = 1;
  return (uint32_t)(nb > 32 ? 32 : nb);
}

// Largest band kept per energy group: the band values then take no more memory
// than the seven transformed volumes of a material pass.
uint32_t mbasis_band_cap(const SimGeometry &g) {
  const uint32_t nv = mbasis_volumes(g.M);
  const uint32_t cap = (uint32_t)((7ull * g.nz) / (2ull * nv));
  const uint32_t floor_cap = g.nz < 8 ? g.nz : 8;
  return cap > floor_cap ? cap : floor_cap;
}

// Whether the quadratic-form coefficients (c^2, s^2, 2cs) of every angle a of the
// first quarter turn equal, rounded to the compute precision, those of a + 2
// quarter turns (and a + A4 those of a + 3 quarter turns): E turned by 180 degrees
// gives the same intensity, so the orbit kernel forms one value for both pairs.
template <typename T> bool quad_pairs_share(const SimGeometry &g, uint32_t k) {
  const uint32_t A4 = g.quarter;
  if (A4 == 0 || g.nA != 4 * A4) return false;
  for (uint32_t a = 0; a < 2 * A4; a++) {
    const double *x = &g.ebasis[2 * ((size_t)k * g.nA + a)], *y = &g.ebasis[2 * ((size_t)k * g.nA + a + 2 * A4)];
    const double cx = x[0], sx = x[1], cy = y[0], sy = y[1];
    if ((T)(cx * cx) != (T)(cy * cy) || (T)(sx * sx) != (T)(sy * sy) || (T)(2.0 * cx * sx) != (T)(2.0 * cy * sy))
      return false;
  }
  return true;
}

// MBASIS: per (k, energy group) the seven real basis fields of every material
// (packed two per complex volume), their 3D FFTs, and their values in the band
// of Ewald bins the group samples (plus the mirrored bins); then per energy the
// transformed tensor fields at its bins as combinations with its constants, and
// every angle as 2D work.
template <typename T>
cyr_status run_mbasis(const cyr_sim_desc *d, const SimGeometry &g, const void *d_vox, T *d_out, WorkBuffers &wb,
                      hipStream_t s, const Setup<T> &su) {
  const uint64_t V = g.V, P = g.P;
  const bool hann = d->windowing == 1 && d->algorithm != 1;
  const bool three = g.nz > 1;
  const int ng = (three && d->interpolation == 1) ? 2 : 1;
  const uint32_t nvol = mbasis_volumes(g.M), pass_max = nvol < 7 ? nvol : 7;
  // laboratory frame: one quadratic form per pixel instead of every angle's projection
  const bool lab = d->reference_frame != 0;
  const uint32_t nb = lab ? std::min(mbasis_energy_batch(g, sizeof(T)), g.nE) : 1;
  LaneBuffers &lb = wb.lane[0];
  CYR_TRY(ensure_buffer(lb.pol, sizeof(cx<T>) * V * pass_max));
  CYR_TRY(ensure_buffer(lb.proj, sizeof(T) * P * (lab ? 4 * (uint64_t)nb : g.nA)));
  CYR_TRY(ensure_buffer(lb.avg, sizeof(double) * P * nb));
  CYR_TRY(ensure_buffer(lb.bins, sizeof(int32_t) * P + sizeof(uint32_t) * mbasis_bmax_words(g.nx, g.ny)));
  // DC-neighbour sums: taken by the Y pass when it runs on a compile-time plan
  // (dc_slots partials per z slice), else by their own kernel (one per slice)
  const uint32_t dc_slots = fft_dc_slots<T>(g.nx, g.ny);
  const uint64_t dc_per = (uint64_t)g.nz * (dc_slots ? dc_slots : 1);
  CYR_TRY(ensure_buffer(lb.dc, sizeof(double2) * nvol * (dc_per + 1)));
  cx<T> *pol = reinterpret_cast<cx<T> *>(lb.pol.ptr);
  int32_t *band = reinterpret_cast<int32_t *>(lb.bins.ptr);
  uint32_t *bmax = reinterpret_cast<uint32_t *>(band + P);
  const uint32_t *span = mbasis_span_table(bmax, g.nx, g.ny);
  double2 *dcz = reinterpret_cast<double2 *>(lb.dc.ptr);
  double2 *dcv = dcz + (uint64_t)nvol * dc_per;
  T *proj = reinterpret_cast<T *>(lb.proj.ptr);
  double *avg = reinterpret_cast<double *>(lb.avg.ptr);
  const double cxb = su.cxb;
  const uint32_t cap = mbasis_band_cap(g);
  const double vel = d->storage == CYR_STORE_NATIVE ? (double)sizeof(T) : 2.0;

  for (uint32_t k = 0; k < g.nK; k++) {
    const M3 Q = ebasis_frame(g, k);
    for (uint32_t e0 = 0; e0 < g.nE;) {
      uint32_t ne = g.nE - e0, J = 0;
      uint64_t wsum = 0;  // sum over columns of the band widths
      uint64_t ssum = 0;  // columns inside the row spans (those some band reads)
      uint64_t d4 = 0;    // (2 x the largest distance of a finite plane value from the centre)^2
      for (;;) {
        kt_begin(s);
        CYR_TRY(mbasis_band(d, g, k, e0, ne, su.kmag, band, bmax, s, &J, &wsum, &ssum, &d4));
        kt_end("mbasis_bands", s, (double)P * 4);
        if (J <= cap || ne == 1) break;
        ne = (ne + 1) / 2;
      }
      if (J == 0) J = 1;  // no pixel sampled: the angles kernel writes NaN only
      CYR_TRY(ensure_buffer(lb.part, sizeof(cx<T>) * 2 * (uint64_t)nvol * J * P));
      cx<T> *own = reinterpret_cast<cx<T> *>(lb.part.ptr);
      cx<T> *mir = own + (uint64_t)nvol * J * P;
      // ---- basis spectra in the band: one pass per material pair (or single)
      for (uint32_t m0 = 0, vb = 0; m0 < g.M;) {
        const bool pair = m0 + 1 < g.M;
        const uint32_t nv = pair ? 7 : 4;
        PolxArgs<T> xa{};
        xa.M = g.M;
        xa.m0 = m0;
        xa.V = V;
        xa.nx = g.nx;
        xa.ny = g.ny;
        xa.nz = g.nz;
        xa.rows = g.ny * g.nz;
        xa.hx = su.hx;
        xa.hy = su.hy;
        xa.hz = su.hz;
        xa.Q = m3<T>(Q);
        xa.rot = is_identity(Q) ? 0 : 1;
        void *vptr[kXMaxOut];
        for (uint32_t c = 0; c < nv; c++) {
          xa.out[c] = pol + (uint64_t)c * V;
          vptr[c] = xa.out[c];
        }
        kt_begin(s);
        CYR_TRY(launch_mbasis_xfields<T>(xa, pair, d_vox, d->storage, d->morphology_type, hann, s));
        kt_end("xfields", s, (pair ? 2.0 : 1.0) * 4.0 * vel * V + nv * (double)V * cxb);
        FftDc fdc{dcz + (uint64_t)vb * dc_per, su.twz, span};
        if (!dc_slots) {
          DcArgs<T> da{};
          for (uint32_t c = 0; c < nv; c++) da.vol[c] = xa.out[c];
          da.count = nv;
          da.twy = su.twy;
          da.twz = su.twz;
          da.nx = g.nx;
          da.ny = g.ny;
          da.nz = g.nz;
          da.dcz = fdc.dcz;
          kt_begin(s);
          CYR_TRY(launch_ebasis_dc<T>(da, s));
          kt_end("ebasis_dc", s, 3.0 * nv * (double)g.ny * g.nz * cxb);
        }
        kt_begin(s);
        CYR_TRY(fft_axis<T>(vptr, nv, g.nx, g.ny, g.nz, 1, s, dc_slots ? &fdc : nullptr));
        // reads every volume; writes the columns inside the row spans when it honours
        // them (algorithmic bytes: the lines at the span edges are written whole)
        const double ycols = dc_slots && fft_span_fits<T>(g.ny) ? (double)ssum : (double)P;
        kt_end("fft_y", s, nv * (double)V * cxb + nv * ycols * g.nz * cxb);
        BandArgs<T> ba{};
        for (uint32_t c = 0; c < nv; c++) ba.vol[c] = xa.out[c];
        ba.count = nv;
        ba.vbase = vb;
        ba.band = band;
        ba.span = span;  // (columns outside it are never read, written by the Y pass or not)
        ba.own = own;
        ba.mir = mir;
        ba.J = J;
        ba.nx = g.nx;
        ba.ny = g.ny;
        ba.nz = g.nz;
        bool fused_z = false;
        if (three) {
          kt_begin(s);
          CYR_TRY(launch_mbasis_zband<T>(ba, s, &fused_z));
          // reads the volumes' columns inside the row spans; writes each column's band
          // values and those of its mirror column (own + mir)
          const double zcols = ba.span ? (double)ssum : (double)P;
          if (fused_z) kt_end("mbasis_zband", s, nv * zcols * g.nz * cxb + 2.0 * nv * (double)wsum * cxb);
          if (!fused_z) {
            kt_begin(s);
            CYR_TRY(fft_axis<T>(vptr, nv, g.nx, g.ny, g.nz, 2, s));
            kt_end("fft_z", s, 2.0 * nv * (double)V * cxb);
          }
        }
        if (!fused_z) {
          kt_begin(s);
          CYR_TRY(launch_mbasis_extract<T>(ba, s));
          kt_end("mbasis_extract", s);
        }
        m0 += pair ? 2 : 1;
        vb += nv;
      }
      CYR_TRY(launch_mbasis_dcsum(dcz, nvol, (uint32_t)dc_per, three, dcv, s));
      // ---- per energy: fields at the Ewald bins, angles, rotation
      const double kv[3] = {g.kvec[k].x, g.kvec[k].y, g.kvec[k].z};
      const double *inv = su.rotinv + (size_t)6 * k * g.nA;
      MbAnglesArgs<T> aa{};
      aa.band = band;
      aa.own = own;
      aa.mir = mir;
      aa.J = J;
      aa.dcv = dcv;
      aa.M = g.M;
      aa.amat = su.amat + (size_t)kEbAmat * k * g.nA;
      aa.nA = g.nA;
      aa.Q = m3<T>(Q);
      if (lab) {
        // batches of nb energies: quadratic forms, rotation + average, detector warp
        for (uint32_t b0 = e0; b0 < e0 + ne; b0 += nb) {
          const uint32_t n = std::min(nb, e0 + ne - b0);
          aa.g = make_ewald_geom(g.nx, g.ny, g.nz, d->phys_size, g.kmag[b0], kv, d->interpolation);
          aa.kmag = su.kmag + b0;
          aa.ne = n;
          aa.mats = su.mats + (size_t)b0 * g.M;
          aa.quad = proj;
          kt_begin(s);
          CYR_TRY(launch_mbasis_angles<T>(aa, ng, s));
          // unique inputs: the group's band values (own + mirror, each column's band
          // width, shared by its energies) and the band table; outputs: n
          // quadratic-form planes
          kt_end("mbasis_quad", s, 2.0 * nvol * (double)wsum * cxb + 4.0 * P + n * (double)kQuadRecord * P * sizeof(T));
          kt_begin(s);
          CYR_TRY(launch_rotate_quad<T>(proj, inv, su.quad + (size_t)3 * k * g.nA, su.quad_t + (size_t)3 * k * g.nA,
                                        g.nA, g.nx, g.ny, d->rot_mask, avg, n, su.tiles, g.quarter, su.otiles,
                                        quad_pairs_share<T>(g, k), 0.5 * std::sqrt((double)d4), s));
          kt_end("rotate_quad", s, n * ((double)kQuadRecord * P * sizeof(T) + 8.0 * P));
          CYR_TRY(launch_detector_warp_n<T>(avg, d_out + ((uint64_t)b0 * g.nK + k) * P, (uint64_t)g.nK * P, n, g.nx,
                                            g.ny, &g.det_inv[(size_t)k * 6], d->rot_mask, s));
        }
      } else {
        for (uint32_t e = e0; e < e0 + ne; e++) {
          aa.g = make_ewald_geom(g.nx, g.ny, g.nz, d->phys_size, g.kmag[e], kv, d->interpolation);
          aa.mats = su.mats + (size_t)e * g.M;
          aa.proj = proj;
          kt_begin(s);
          CYR_TRY(launch_mbasis_angles<T>(aa, ng, s));
          kt_end("mbasis_angles", s, 2.0 * ng * nvol * (double)P * cxb + (double)g.nA * P * sizeof(T));
          kt_begin(s);
          CYR_TRY(launch_rotate_all<T>(proj, inv, g.nA, g.nx, g.ny, d->rot_mask, avg, su.tiles,
                                       0.5 * std::sqrt((double)d4), s));
          kt_end("rotate_all", s, (double)(g.nA + 1) * P * sizeof(T));
          T *out = d_out + ((uint64_t)e * g.nK + k) * P;
          CYR_TRY(launch_detector_warp<T>(avg, out, g.nx, g.ny, &g.det_inv[(size_t)k * 6], d->rot_mask, s));
        }
      }
      e0 += ne;
    }
  }
  return CYR_OK;
}

template <typename T>
cyr_status run_impl(const cyr_sim_desc *d, const SimGeometry &g, const void *d_vox, const double *mats, T *d_out,
                    WorkBuffers &wb, hipStream_t st, RunMode mode, const void *resident_tables) {
  Tables tb;
  const void *tables = resident_tables;
  if (mode == kRunResident) {
    build_layout<T>(g, tb);
  } else {
    build_tables<T>(d, g, mats, tb);
    CYR_TRY(ensure_buffer(wb.tables, tb.total));
    CYR_TRY(wb.upload(wb.tables.ptr, tb.host.data(), tb.total, st));
    tables = wb.tables.ptr;
  }
  Setup<T> su{};
  CYR_TRY(setup_views<T>(d, g, tb, tables, su));
  const int pipe = resolved_pipeline(d, g);
  const cyr_status s = pipe == CYR_PIPE_PLAIN ? run_plain<T>(d, g, d_vox, d_out, wb, st, su)
                       : pipe == CYR_PIPE_MBASIS ? run_mbasis<T>(d, g, d_vox, d_out, wb, st, su)
                                                 : run_ebasis<T>(d, g, d_vox, d_out, wb, st, su);
  if (mode != kRunSync) return s;
  // drain the stream on errors too, so no queued launch outlives the call's buffers
  const hipError_t se = hipStreamSynchronize(st);
  kt_collect();
  if (s != CYR_OK) return s;
  if (se != hipSuccess) return hip_fail(se, "hipStreamSynchronize", __FILE__, __LINE__);
  return CYR_OK;
}

// Per-device state of the plain C ABI: cached work buffers and a library stream.
struct DeviceState {
  std::mutex mu;
  WorkBuffers wb;
  DeviceBuffer vox, out;
  hipStream_t stream = nullptr;
};

std::mutex g_states_mu;
std::map<int, DeviceState *> g_states;

cyr_status device_state(DeviceState **out) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
    (void)hipGetLastError();
    return fail(CYR_E_NODEVICE, "no HIP device found");
  }
  int dev = 0;
  CYR_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_states_mu);
  auto it = g_states.find(dev);
  if (it == g_states.end()) {
    DeviceState *s = new DeviceState();
    CYR_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    it = g_states.emplace(dev, s).first;
  }
  *out = it->second;
  return CYR_OK;
}

}  // namespace

void LaneBuffers::release() {
  free_buffer(pol);
  free_buffer(proj);
  free_buffer(avg);
  free_buffer(bins);
  free_buffer(part);
  free_buffer(dc);
}

size_t LaneBuffers::bytes() const { return pol.bytes + proj.bytes + avg.bytes + bins.bytes + part.bytes + dc.bytes; }

cyr_status WorkBuffers::ensure_aux() {
  if (aux) return CYR_OK;
  CYR_HIP(hipStreamCreateWithFlags(&aux, hipStreamNonBlocking));  // default priority
  CYR_HIP(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
  CYR_HIP(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
  for (auto &ev : ev_polx) CYR_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  return CYR_OK;
}

cyr_status WorkBuffers::upload(void *dst, const void *src, size_t n, hipStream_t st) {
  if (!stage_done) CYR_HIP(hipEventCreateWithFlags(&stage_done, hipEventDisableTiming));
  if (stage_pending) {  // the previous upload must have left the staging buffer
    CYR_HIP(hipEventSynchronize(stage_done));
    stage_pending = false;
  }
  if (stage_bytes < n) {
    if (stage) (void)hipHostFree(stage);
    stage = nullptr;
    stage_bytes = 0;
    CYR_HIP(hipHostMalloc(&stage, n, hipHostMallocDefault));
    stage_bytes = n;
  }
  std::memcpy(stage, src, n);
  CYR_HIP(hipMemcpyAsync(dst, stage, n, hipMemcpyHostToDevice, st));
  CYR_HIP(hipEventRecord(stage_done, st));
  stage_pending = true;
  return CYR_OK;
}

void WorkBuffers::release() {
  if (stage_pending) (void)hipEventSynchronize(stage_done);
  stage_pending = false;
  if (stage) (void)hipHostFree(stage);
  stage = nullptr;
  stage_bytes = 0;
  if (stage_done) (void)hipEventDestroy(stage_done);
  stage_done = nullptr;
  for (auto &l : lane) l.release();
  free_buffer(nt);
  free_buffer(scatter);
  free_buffer(acc);
  free_buffer(count);
  free_buffer(tables);
  if (aux) {
    (void)hipStreamSynchronize(aux);
    (void)hipStreamDestroy(aux);
    aux = nullptr;
    (void)hipEventDestroy(ev_fork);
    (void)hipEventDestroy(ev_join);
    for (auto &ev : ev_polx) (void)hipEventDestroy(ev);
  }
}

size_t WorkBuffers::bytes() const {
  return lane[0].bytes() + lane[1].bytes() + nt.bytes + scatter.bytes + acc.bytes + count.bytes + tables.bytes;
}

namespace {
std::atomic<uint32_t> g_lanes{1};
}  // namespace

uint32_t energy_lanes() { return g_lanes.load(std::memory_order_relaxed); }
void set_energy_lanes(uint32_t n) { g_lanes.store(n < 1 ? 1 : (n > 2 ? 2 : n)); }

uint64_t morphology_bytes(const cyr_sim_desc *d) {
  const uint64_t V = (uint64_t)d->dims[0] * d->dims[1] * d->dims[2];
  const uint64_t el = d->storage == CYR_STORE_NATIVE ? (d->precision == CYR_FP64 ? 8 : 4) : 2;
  return V * d->num_materials * 4 * el;
}

uint64_t output_bytes(const cyr_sim_desc *d, const SimGeometry &g) {
  return (uint64_t)g.nE * g.nK * g.P * (d->precision == CYR_FP64 ? 8 : 4);
}

// DC-neighbour partials per volume of the material basis (see run_mbasis)
static uint64_t mbasis_dc_per(const cyr_sim_desc *d, const SimGeometry &g) {
  const uint32_t slots = d->precision == CYR_FP64 ? fft_dc_slots<double>(g.nx, g.ny) : fft_dc_slots<float>(g.nx, g.ny);
  return (uint64_t)g.nz * (slots ? slots : 1);
}

uint64_t work_bytes(const cyr_sim_desc *d, const SimGeometry &g) {
  const uint64_t cs = d->precision == CYR_FP64 ? 16 : 8, rs = cs / 2;
  const int pipe = resolved_pipeline(d, g);
  uint64_t b = 0;
  if (pipe == CYR_PIPE_MBASIS) {
    // one pass of packed volumes, band values (own + mirror, at most the band cap),
    // band table, DC sums, per-angle projections, average
    const uint64_t nv = mbasis_volumes(g.M), pass = nv < 7 ? nv : 7;
    const uint64_t J = mbasis_band_cap(g) > 2 ? mbasis_band_cap(g) : 2;
    const uint64_t nb = d->reference_frame != 0 ? std::min(mbasis_energy_batch(g, rs), g.nE) : 1;
    const uint64_t planes = d->reference_frame != 0 ? 4 * nb : g.nA;
    b += cs * g.V * pass + cs * 2 * nv * J * g.P + 4 * g.P + 4 * mbasis_bmax_words(g.nx, g.ny) +
         16 * nv * (mbasis_dc_per(d, g) + 1) + rs * g.P * planes + 8 * g.P * nb;
  } else if (pipe == CYR_PIPE_EBASIS) {  // per lane: fields, bin table, Y/z partial sums (chunks x 2 bins x fields), DC sums,
                 // per-angle projections, average
    const uint64_t nch = ebasis_max_chunks(g.nx, g.ny, g.nz, d->precision);
    const uint64_t lanes = (uint64_t)g.nE * g.nK > 1 ? energy_lanes() : 1;
    b += lanes * (cs * g.V * kEbFields + 4 * g.P + cs * g.P * 2 * kEbFields * nch + 16 * kEbFields * g.nz +
                  rs * g.P * g.nA + 8 * g.P);
  } else {
    b += cs * g.V * 3 + rs * g.P + 8 * g.P * 2 + 4 * g.P;
    if (d->algorithm == 1) b += cs * g.V * 6;
    if (d->scatter_approach == 1) b += rs * g.V;
  }
  b += 64 * 1024 + (uint64_t)g.nK * g.nA * (16 + 72 + 48) + (uint64_t)g.nE * g.M * 64;
  return b;
}

cyr_status run_simulation(const cyr_sim_desc *d, const SimGeometry &g, const void *d_vox, const double *mats,
                          void *d_out, WorkBuffers &wb, hipStream_t st, RunMode mode, const void *resident_tables) {
  if (d->precision == CYR_FP64)
    return run_impl<double>(d, g, d_vox, mats, reinterpret_cast<double *>(d_out), wb, st, mode, resident_tables);
  return run_impl<float>(d, g, d_vox, mats, reinterpret_cast<float *>(d_out), wb, st, mode, resident_tables);
}

size_t table_bytes(const cyr_sim_desc *d, const SimGeometry &g) {
  Tables tb;
  if (d->precision == CYR_FP64) build_layout<double>(g, tb);
  else build_layout<float>(g, tb);
  return tb.total;
}

void fill_tables(const cyr_sim_desc *d, const SimGeometry &g, const double *mats, void *host) {
  Tables tb;
  if (d->precision == CYR_FP64) build_tables<double>(d, g, mats, tb);
  else build_tables<float>(d, g, mats, tb);
  std::memcpy(host, tb.host.data(), tb.total);
}


void convert_storage(const cyr_sim_desc *d, const void *host_vox, std::vector<unsigned char> &out) {
  const uint64_t n = (uint64_t)d->dims[0] * d->dims[1] * d->dims[2] * d->num_materials * 4;
  out.resize(n * 2);
  uint16_t *o = reinterpret_cast<uint16_t *>(out.data());
  for (uint64_t i = 0; i < n; i++) {
    const float v = d->precision == CYR_FP64 ? (float)reinterpret_cast<const double *>(host_vox)[i]
                                             : reinterpret_cast<const float *>(host_vox)[i];
    if (d->storage == CYR_STORE_FP16) {
      __half h = __float2half(v);
      std::memcpy(&o[i], &h, 2);
    } else {
      // round to nearest even bfloat16
      uint32_t u;
      std::memcpy(&u, &v, 4);
      if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) o[i] = (uint16_t)((u >> 16) | 0x40);
      else o[i] = (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
    }
  }
}

// Autotuner of the E-basis Y/z kernel: times every candidate launch shape
// (columns per work group x target grid size) on the descriptor's grid with
// zero-filled fields and the first energy's Ewald bins, and records the fastest
// in the tuning table.  The choice does not change results beyond the order of
// the partial sums over z chunks.
template <typename T>
cyr_status autotune_impl(const cyr_sim_desc *d, const SimGeometry &g, WorkBuffers &wb, hipStream_t st,
                         YzTuning *chosen) {
  *chosen = YzTuning{};
  const bool ebasis = ebasis_usable(d, g);
  if (!ebasis || g.nz < 2) return CYR_OK;  // 2D: one slice, nothing to split
  const uint64_t V = g.V, P = g.P;
  const int ng = d->interpolation == 1 ? 2 : 1;
  LaneBuffers &lb = wb.lane[0];
  CYR_TRY(ensure_buffer(lb.pol, sizeof(cx<T>) * V * kEbFields));
  CYR_HIP(hipMemsetAsync(lb.pol.ptr, 0, sizeof(cx<T>) * V * kEbFields, st));
  CYR_TRY(ensure_buffer(lb.bins, sizeof(int32_t) * P));
  const double kv[3] = {g.kvec[0].x, g.kvec[0].y, g.kvec[0].z};
  const EwaldGeom geo = make_ewald_geom(g.nx, g.ny, g.nz, d->phys_size, g.kmag[0], kv, d->interpolation);
  CYR_TRY(launch_ebasis_bins(geo, reinterpret_cast<int32_t *>(lb.bins.ptr), st));
  cyr_status s;
  const cx<T> *twz = twiddles<T>(g.nz, &s);
  if (!twz) return s;
  hipEvent_t e0, e1;
  CYR_HIP(hipEventCreate(&e0));
  CYR_HIP(hipEventCreate(&e1));
  float best = -1;
  YzTuning pick{};
  const uint32_t cols[3] = {4, 8, 16};
  const uint32_t targets[4] = {256, 512, 1024, 2048};
  cyr_status rc = CYR_OK;
  for (uint32_t ci = 0; ci < 3 && rc == CYR_OK; ci++)
    for (uint32_t ti = 0; ti < 4 && rc == CYR_OK; ti++) {
      YzArgs<T> ya{};
      ya.nx = g.nx;
      ya.ny = g.ny;
      ya.nz = g.nz;
      ya.pp = cols[ci];
      if (ebasis_yz_plan<T>(ya, targets[ti]) != CYR_OK || ya.pp != cols[ci]) continue;
      rc = ensure_buffer(lb.part, sizeof(cx<T>) * P * ng * kEbFields * ya.nchunks);
      if (rc != CYR_OK) break;
      for (int c = 0; c < kEbFields; c++) ya.vol[c] = reinterpret_cast<cx<T> *>(lb.pol.ptr) + (uint64_t)c * V;
      ya.part = reinterpret_cast<cx<T> *>(lb.part.ptr);
      ya.kzt = reinterpret_cast<int32_t *>(lb.bins.ptr);
      ya.twz = twz;
      rc = launch_ebasis_yz<T>(ya, ng, st);  // warm-up
      if (rc != CYR_OK) break;
      constexpr int reps = 3;
      if (hipEventRecord(e0, st) != hipSuccess) { rc = fail(CYR_E_DEVICE, "event record"); break; }
      for (int r = 0; r < reps && rc == CYR_OK; r++) rc = launch_ebasis_yz<T>(ya, ng, st);
      if (rc != CYR_OK) break;
      float ms = 0;
      if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        rc = fail(CYR_E_DEVICE, "autotune timing");
        break;
      }
      if (best < 0 || ms < best) {
        best = ms;
        pick = YzTuning{cols[ci], targets[ti]};
      }
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != CYR_OK) return rc;
  if (best >= 0) ebasis_set_yz_tuning(g.nx, g.ny, g.nz, d->precision, pick);
  *chosen = pick;
  return CYR_OK;
}

cyr_status autotune_shape(const cyr_sim_desc *d, const SimGeometry &g, WorkBuffers &wb, hipStream_t st,
                          YzTuning *chosen) {
  if (d->precision == CYR_FP64) return autotune_impl<double>(d, g, wb, st, chosen);
  return autotune_impl<float>(d, g, wb, st, chosen);
}

}  // namespace cyr


using namespace cyr;

extern "C" {

cyr_status cyr_set_tuning(const uint32_t dims[3], int32_t precision, const cyr_yz_tuning *t) {
  if (!dims || !t) return fail(CYR_E_INVALID, "null argument");
  if (t->columns != 0 && t->columns != 1 && t->columns != 2 && t->columns != 4 && t->columns != 8 &&
      t->columns != 16)
    return fail(CYR_E_INVALID, "columns must be 0 or a power of two <= 16");
  ebasis_set_yz_tuning(dims[0], dims[1], dims[2], precision, YzTuning{t->columns, t->target_blocks});
  return CYR_OK;
}

cyr_status cyr_get_tuning(const uint32_t dims[3], int32_t precision, cyr_yz_tuning *t) {
  if (!dims || !t) return fail(CYR_E_INVALID, "null argument");
  const YzTuning y = ebasis_yz_tuning(dims[0], dims[1], dims[2], precision);
  t->columns = y.columns;
  t->target_blocks = y.target_blocks;
  return CYR_OK;
}

cyr_status cyr_set_energy_lanes(uint32_t lanes) {
  if (lanes < 1 || lanes > 2) return fail(CYR_E_INVALID, "energy lanes must be 1 or 2");
  set_energy_lanes(lanes);
  return CYR_OK;
}

uint32_t cyr_energy_lanes(void) { return energy_lanes(); }

cyr_status cyr_autotune(const cyr_sim_desc *desc, cyr_yz_tuning *chosen) {
  SimGeometry g;
  CYR_TRY(build_geometry(desc, g));
  auto ws = workspace_use();
  DeviceState *ds;
  CYR_TRY(device_state(&ds));
  std::lock_guard<std::mutex> lk(ds->mu);
  YzTuning y{};
  const cyr_status s = desc->precision == CYR_FP64 ? autotune_impl<double>(desc, g, ds->wb, ds->stream, &y)
                                                   : autotune_impl<float>(desc, g, ds->wb, ds->stream, &y);
  if (s != CYR_OK) return s;
  if (hipStreamSynchronize(ds->stream) != hipSuccess) return fail(CYR_E_DEVICE, "autotune sync");
  if (chosen) {
    chosen->columns = y.columns;
    chosen->target_blocks = y.target_blocks;
  }
  return CYR_OK;
}


uint64_t cyr_footprint_bytes(const cyr_sim_desc *desc) {
  SimGeometry g;
  if (build_geometry(desc, g) != CYR_OK) return 0;
  return morphology_bytes(desc) + work_bytes(desc, g) + output_bytes(desc, g);
}

int32_t cyr_resolved_pipeline(const cyr_sim_desc *desc) {
  SimGeometry g;
  if (build_geometry(desc, g) != CYR_OK) return -1;
  return resolved_pipeline(desc, g);
}

cyr_status cyr_simulate_device(const cyr_sim_desc *desc, const void *d_voxels, const double *mats, void *d_out,
                               void *stream) {
  SimGeometry g;
  CYR_TRY(build_geometry(desc, g));
  if (!d_voxels || !mats || !d_out) return fail(CYR_E_INVALID, "null buffer");
  auto ws = workspace_use();
  DeviceState *ds;
  CYR_TRY(device_state(&ds));
  std::lock_guard<std::mutex> lk(ds->mu);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  return run_simulation(desc, g, d_voxels, mats, d_out, ds->wb, st);
}

cyr_status cyr_simulate(const cyr_sim_desc *desc, const void *voxels, const double *mats, void *out) {
  SimGeometry g;
  CYR_TRY(build_geometry(desc, g));
  if (!voxels || !mats || !out) return fail(CYR_E_INVALID, "null buffer");
  auto ws = workspace_use();
  DeviceState *ds;
  CYR_TRY(device_state(&ds));
  std::lock_guard<std::mutex> lk(ds->mu);
  const uint64_t vb = morphology_bytes(desc), ob = output_bytes(desc, g);
  CYR_TRY(ensure_buffer(ds->vox, vb));
  CYR_TRY(ensure_buffer(ds->out, ob));
  if (desc->storage == CYR_STORE_NATIVE) {
    CYR_TRY(upload_host(ds->vox.ptr, voxels, vb, ds->stream));
  } else {
    std::vector<unsigned char> conv;
    convert_storage(desc, voxels, conv);
    CYR_TRY(upload_host(ds->vox.ptr, conv.data(), vb, ds->stream));
    CYR_HIP(hipStreamSynchronize(ds->stream));
  }
  CYR_TRY(run_simulation(desc, g, ds->vox.ptr, mats, ds->out.ptr, ds->wb, ds->stream));
  CYR_HIP(hipMemcpyAsync(out, ds->out.ptr, ob, hipMemcpyDeviceToHost, ds->stream));
  CYR_HIP(hipStreamSynchronize(ds->stream));
  return CYR_OK;
}

cyr_status cyr_polarization_device(const cyr_sim_desc *desc, const void *d_voxels, const double *mats_e,
                                   double eangle_rad, void *d_px, void *d_py, void *d_pz, void *stream) {
  SimGeometry g;
  CYR_TRY(build_geometry(desc, g));
  if (desc->case_type != 0)  // src/cudaMain.cu:1677-1680
    return fail(CYR_E_UNSUPPORTED, "Only implemented for Case Type = 0");
  if (!d_voxels || !mats_e || !d_px || !d_py || !d_pz) return fail(CYR_E_INVALID, "null buffer");
  auto ws = workspace_use();
  DeviceState *ds;
  CYR_TRY(device_state(&ds));
  std::lock_guard<std::mutex> lk(ds->mu);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const M3 R = rotation_matrix_e(g.kvec[0], g.RK[0], eangle_rad);
  const bool hann = desc->windowing == 1;
  auto go = [&](auto tag) -> cyr_status {
    using T = decltype(tag);
    Tables tb;
    std::vector<double> m(4 * (size_t)g.M * g.nE);
    for (uint32_t e = 0; e < g.nE; e++) std::memcpy(&m[4 * (size_t)e * g.M], mats_e, sizeof(double) * 4 * g.M);
    build_tables<T>(desc, g, m.data(), tb);
    CYR_TRY(ensure_buffer(ds->wb.tables, tb.total));
    CYR_TRY(ds->wb.upload(ds->wb.tables.ptr, tb.host.data(), tb.total, st));
    unsigned char *tbase = reinterpret_cast<unsigned char *>(ds->wb.tables.ptr);
    PolarArgs<T> pa{};
    pa.mats = reinterpret_cast<const MatConst<T> *>(tbase + tb.mats);
    pa.M = g.M;
    pa.V = g.V;
    pa.nx = g.nx;
    pa.ny = g.ny;
    pa.nz = g.nz;
    const T *h = reinterpret_cast<const T *>(tbase + tb.han);
    pa.hx = h;
    pa.hy = h + g.nx;
    pa.hz = h + g.nx + g.ny;
    pa.e[0] = Vec3<T>{(T)R.m[0], (T)R.m[3], (T)R.m[6]};
    pa.R = m3<T>(R);
    pa.out[0] = reinterpret_cast<cx<T> *>(d_px);
    pa.out[1] = reinterpret_cast<cx<T> *>(d_py);
    pa.out[2] = reinterpret_cast<cx<T> *>(d_pz);
    CYR_TRY(launch_polarization<T>(pa, d_voxels, desc->storage, desc->morphology_type, 1, hann,
                                   desc->reference_frame == 0, st));
    CYR_HIP(hipStreamSynchronize(st));
    return CYR_OK;
  };
  if (desc->precision == CYR_FP64) return go(double{});
  return go(float{});
}

cyr_status cyr_polarization(const cyr_sim_desc *desc, const void *voxels, const double *mats, uint32_t energy_id,
                            double eangle_rad, void *px, void *py, void *pz) {
  SimGeometry g;
  CYR_TRY(build_geometry(desc, g));
  if (energy_id >= g.nE) return fail(CYR_E_INVALID, "energy_id out of range");
  if (!voxels || !mats || !px || !py || !pz) return fail(CYR_E_INVALID, "null buffer");
  DeviceState *ds;
  CYR_TRY(device_state(&ds));
  const uint64_t vb = morphology_bytes(desc);
  const uint64_t cb = g.V * (desc->precision == CYR_FP64 ? 16 : 8);
  DeviceBuffer dv, dp;
  CYR_TRY(ensure_buffer(dv, vb));
  cyr_status s = ensure_buffer(dp, 3 * cb);
  if (s != CYR_OK) {
    free_buffer(dv);
    return s;
  }
  hipStream_t st = ds->stream;
  if (desc->storage == CYR_STORE_NATIVE) {
    s = hipMemcpy(dv.ptr, voxels, vb, hipMemcpyHostToDevice) == hipSuccess ? CYR_OK
                                                                            : fail(CYR_E_DEVICE, "upload failed");
  } else {
    std::vector<unsigned char> conv;
    convert_storage(desc, voxels, conv);
    s = hipMemcpy(dv.ptr, conv.data(), vb, hipMemcpyHostToDevice) == hipSuccess ? CYR_OK
                                                                                : fail(CYR_E_DEVICE, "upload failed");
  }
  unsigned char *b = reinterpret_cast<unsigned char *>(dp.ptr);
  if (s == CYR_OK)
    s = cyr_polarization_device(desc, dv.ptr, mats + 4 * (size_t)energy_id * g.M, eangle_rad, b, b + cb, b + 2 * cb,
                                st);
  if (s == CYR_OK) {
    if (hipMemcpy(px, b, cb, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(py, b + cb, cb, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(pz, b + 2 * cb, cb, hipMemcpyDeviceToHost) != hipSuccess)
      s = fail(CYR_E_DEVICE, "download failed");
  }
  free_buffer(dv);
  free_buffer(dp);
  return s;
}

cyr_status cyr_fft3d_device(uint32_t nx, uint32_t ny, uint32_t nz, int32_t precision, void *const *d_volumes,
                            uint32_t count, void *stream) {
  if (!d_volumes || nx < 1 || ny < 1 || nz < 1) return fail(CYR_E_INVALID, "bad fft arguments");
  for (uint32_t i = 0; i < count; i++)
    if (!d_volumes[i]) return fail(CYR_E_INVALID, "null volume");
  auto ws = workspace_use();
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (precision == CYR_FP64) CYR_TRY(fft3d<double>(d_volumes, count, nx, ny, nz, st));
  else if (precision == CYR_FP32) CYR_TRY(fft3d<float>(d_volumes, count, nx, ny, nz, st));
  else return fail(CYR_E_INVALID, "precision out of range");
  CYR_HIP(hipStreamSynchronize(st));
  return CYR_OK;
}

cyr_status cyr_fft_axis_device(uint32_t nx, uint32_t ny, uint32_t nz, int32_t precision, void *const *d_volumes,
                               uint32_t count, int32_t axis, void *stream) {
  if (!d_volumes || nx < 1 || ny < 1 || nz < 1 || axis < 0 || axis > 2) return fail(CYR_E_INVALID, "bad fft arguments");
  for (uint32_t i = 0; i < count; i++)
    if (!d_volumes[i]) return fail(CYR_E_INVALID, "null volume");
  auto ws = workspace_use();
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (precision == CYR_FP64) return fft_axis<double>(d_volumes, count, nx, ny, nz, axis, st);
  if (precision == CYR_FP32) return fft_axis<float>(d_volumes, count, nx, ny, nz, axis, st);
  return fail(CYR_E_INVALID, "precision out of range");
}

cyr_status cyr_dc_shift_device(uint32_t nx, uint32_t ny, uint32_t nz, int32_t precision, void *d_volume,
                               int32_t do_dc, int32_t do_shift, void *stream) {
  if (!d_volume || nx < 2 || ny < 2 || nz < 1) return fail(CYR_E_INVALID, "bad dc/shift arguments");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (precision == CYR_FP64) {
    cx<double> *v = reinterpret_cast<cx<double> *>(d_volume);
    if (do_dc) CYR_TRY(launch_dc_replace<double>(&v, 1, nx, ny, nz, st));
    if (do_shift) CYR_TRY(launch_igor_shift<double>(v, nx, ny, nz, st));
  } else if (precision == CYR_FP32) {
    cx<float> *v = reinterpret_cast<cx<float> *>(d_volume);
    if (do_dc) CYR_TRY(launch_dc_replace<float>(&v, 1, nx, ny, nz, st));
    if (do_shift) CYR_TRY(launch_igor_shift<float>(v, nx, ny, nz, st));
  } else {
    return fail(CYR_E_INVALID, "precision out of range");
  }
  CYR_HIP(hipStreamSynchronize(st));
  return CYR_OK;
}

cyr_status cyr_ewald_device(const cyr_sim_desc *desc, uint32_t energy_id, uint32_t k_id, const void *d_px,
                            const void *d_py, const void *d_pz, void *d_proj, void *stream) {
  SimGeometry g;
  CYR_TRY(build_geometry(desc, g));
  if (energy_id >= g.nE || k_id >= g.nK) return fail(CYR_E_INVALID, "energy_id/k_id out of range");
  if (!d_px || !d_py || !d_pz || !d_proj) return fail(CYR_E_INVALID, "null buffer");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const double kv[3] = {g.kvec[k_id].x, g.kvec[k_id].y, g.kvec[k_id].z};
  const EwaldGeom geo = make_ewald_geom(g.nx, g.ny, g.nz, desc->phys_size, g.kmag[energy_id], kv, desc->interpolation);
  auto go = [&](auto tag) -> cyr_status {
    using T = decltype(tag);
    const cx<T> *px = reinterpret_cast<const cx<T> *>(d_px), *py = reinterpret_cast<const cx<T> *>(d_py),
                *pz = reinterpret_cast<const cx<T> *>(d_pz);
    T *proj = reinterpret_cast<T *>(d_proj);
    if (desc->scatter_approach == 1) {
      DeviceBuffer sc;
      CYR_TRY(ensure_buffer(sc, sizeof(T) * g.V));
      cyr_status s = launch_scatter3d<T>(geo, px, py, pz, reinterpret_cast<T *>(sc.ptr), st);
      if (s == CYR_OK) s = launch_ewald_full<T>(geo, reinterpret_cast<T *>(sc.ptr), proj, st);
      if (s == CYR_OK && hipStreamSynchronize(st) != hipSuccess) s = fail(CYR_E_DEVICE, "sync failed");
      free_buffer(sc);
      return s;
    }
    CYR_TRY(launch_ewald_partial<T>(geo, px, py, pz, proj, st));
    CYR_HIP(hipStreamSynchronize(st));
    return CYR_OK;
  };
  if (desc->precision == CYR_FP64) return go(double{});
  return go(float{});
}

cyr_status cyr_warp_affine_device(uint32_t nx, uint32_t ny, int32_t precision, const void *d_src, void *d_dst,
                                  const double coeffs[6], void *stream) {
  if (!d_src || !d_dst || !coeffs || nx < 1 || ny < 1) return fail(CYR_E_INVALID, "bad warp arguments");
  const double det = coeffs[0] * coeffs[4] - coeffs[1] * coeffs[3];
  if (!(std::fabs(det) > 0) || !std::isfinite(det)) return fail(CYR_E_INVALID, "singular warp");
  double inv[6];
  affine_inverse(coeffs, inv);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (precision == CYR_FP64)
    CYR_TRY(launch_warp<double>(reinterpret_cast<const double *>(d_src), reinterpret_cast<double *>(d_dst), nx, ny,
                                inv, st));
  else if (precision == CYR_FP32)
    CYR_TRY(launch_warp<float>(reinterpret_cast<const float *>(d_src), reinterpret_cast<float *>(d_dst), nx, ny, inv,
                               st));
  else
    return fail(CYR_E_INVALID, "precision out of range");
  CYR_HIP(hipStreamSynchronize(st));
  return CYR_OK;
}

cyr_status cyr_release_workspace(void) {
  auto ws = workspace_exclusive();  // waits for calls that use the caches
  if (live_handles() > 0)
    return fail(CYR_E_INVALID, "a batch plan or serving engine is alive: destroy it before releasing the workspace");
  {
    std::lock_guard<std::mutex> lb(g_band_mu);
    g_band_widths.clear();
  }
  std::lock_guard<std::mutex> lk(g_states_mu);
  int cur = 0;
  hipGetDevice(&cur);
  for (auto &kv : g_states) {
    hipSetDevice(kv.first);
    std::lock_guard<std::mutex> lk2(kv.second->mu);
    kv.second->wb.release();
    free_buffer(kv.second->vox);
    free_buffer(kv.second->out);
  }
  hipSetDevice(cur);
  fft_release();
  release_batch_buffers();
  release_upload_staging();
  return CYR_OK;
}

}  // extern "C"
