// Kernel argument blocks and launchers shared by the executors.
#pragma once

#include <cstdint>

#include "common.hpp"

namespace cyr {

// ---- polarization (polarization.hip) ----------------------------------------
template <typename T> struct PolarArgs {
  const MatConst<T> *mats;  // this energy's M constants (device)
  uint32_t M;
  uint64_t V;
  uint32_t nx, ny, nz;
  uint32_t rows, block, blocks_per_row;  // filled by the launcher
  const T *hx, *hy, *hz;                 // Hanning tables (device) or null
  Vec3<T> e[2];                          // E basis vectors
  Mat3<T> R;                             // E rotation (MATERIAL frame)
  cx<T> *out[6];                         // [basis][component] volumes
};

template <typename T>
cyr_status launch_polarization(PolarArgs<T> a, const void *vox, int storage, int morph, int nb, bool hann,
                               bool matframe, hipStream_t st);
template <typename T>
cyr_status launch_nt(PolarArgs<T> a, const void *vox, int storage, int morph, cx<T> *nt, hipStream_t st);
template <typename T>
cyr_status launch_polarization_nt(PolarArgs<T> a, const cx<T> *nt, int nb, bool matframe, hipStream_t st);

// ---- Ewald projection (ewald.hip) -------------------------------------------
// Everything geometric is FP64: q grid, Ewald validity, z index and weight.
struct EwaldGeom {
  uint32_t nx, ny, nz;
  double start, dx, dy, dz;  // q grid (include/uniaxial.h:842-856)
  double k;                  // |k| for this energy
  double kx, ky, kz;         // unit k vector
  int interpolation;         // 0 NN, 1 linear
  int enable2d;
};
EwaldGeom make_ewald_geom(uint32_t nx, uint32_t ny, uint32_t nz, double phys_size, double k,
                          const double kvec[3], int interpolation);

// Plain path: projection of one E angle from the three UNshifted FFT volumes.
template <typename T>
cyr_status launch_ewald_partial(const EwaldGeom &g, const cx<T> *px, const cx<T> *py, const cx<T> *pz, T *proj,
                                hipStream_t st);
// Full approach: I over the whole shifted volume, then the projection.
template <typename T>
cyr_status launch_scatter3d(const EwaldGeom &g, const cx<T> *px, const cx<T> *py, const cx<T> *pz, T *scatter,
                            hipStream_t st);
template <typename T>
cyr_status launch_ewald_full(const EwaldGeom &g, const T *scatter, T *proj, hipStream_t st);
// DC replacement of F[0] in place for `count` volumes (src/cudaMain.cu:123-176).
template <typename T>
cyr_status launch_dc_replace(cx<T> *const *vols, uint32_t count, uint32_t nx, uint32_t ny, uint32_t nz,
                             hipStream_t st);
// Stand-alone Igor shift (in-place involution, include/uniaxial.h:570-613).
template <typename T>
cyr_status launch_igor_shift(cx<T> *vol, uint32_t nx, uint32_t ny, uint32_t nz, hipStream_t st);

// ---- E-basis executor (ebasis.hip) -----------------------------------------
// Tensor fields transformed per (energy, k): N'00, N'01, N'02, N'11, N'12 of
// N' = Q^T N Q, Q = [e1 e2 k^]; per-angle coefficient matrix A (3 x 5) with
// p~ = A F.
constexpr int kEbFields = 5;
constexpr int kEbAmat = 3 * kEbFields;
// True when the grid has Stockham plans and LDS tiles that fit the fused kernels
// (other grids run the plain executor).
bool ebasis_supported(uint32_t nx, uint32_t ny, uint32_t nz, int precision);
// Grid size the Y/z kernel's z chunking aims for, and the most chunks any plan
// for (nx, nz) uses (sizes the partial-sum buffer).
constexpr uint32_t kYzTargetBlocks = 1024;
uint32_t ebasis_max_chunks(uint32_t nx, uint32_t ny, uint32_t nz, int precision);
// Pixel -> FFT z bin of its (kx, ky) column: kzt[ky][kx] (-1 = not sampled); with
// linear interpolation the upper bin is kzt - 1 (mod Nz).
cyr_status launch_ebasis_bins(const EwaldGeom &g, int32_t *kzt, hipStream_t st);

// Outputs of the fields + X FFT kernel: 5 tensor fields or up to 7 packed basis volumes.
constexpr int kXMaxOut = 8;
template <typename T> struct PolxArgs {
  const MatConst<T> *mats;         // tensor mode: this energy's M constants
  uint32_t M;
  uint32_t m0;                     // basis modes: first material of the pass
  uint64_t V;
  uint32_t nx, ny, nz, rows;       // rows = ny * nz
  const T *hx, *hy, *hz;           // Hanning tables
  Mat3<T> Q;                       // columns e1, e2, k^ (row-major)
  int32_t rot;                     // 0 when Q is the identity
  const cx<T> *twx;                // filled by the launcher
  uint32_t tpp, pp, pitch;         // filled by the launcher
  cx<T> *out[kXMaxOut];            // X-transformed fields
};
template <typename T>
cyr_status launch_ebasis_polx(PolxArgs<T> a, const void *vox, int storage, int morph, bool hann, hipStream_t st);

template <typename T> struct YzArgs {
  const cx<T> *vol[kEbFields];     // X-transformed fields
  cx<T> *part;                     // [nchunks][NG][kEbFields][Ny][Nx] partial sums (2D: [kEbFields][Ny][Nx])
  const int32_t *kzt;              // [Ny][Nx] (3D)
  const cx<T> *twy, *twz;          // twy filled by the launcher; twz null in 2D
  uint32_t nx, ny, nz;
  uint32_t e, pp, tpp, zchunk, nchunks;  // 0 = automatic (ebasis_yz_plan fills)
  uint32_t ct;                     // 1: compile-time FFT plan for Ny (set by ebasis_yz_plan)
};
// Launch choices of the Y/z kernel for one (Nx, Ny, Nz, precision) shape, set by
// the autotuner: columns per block and the grid size the z chunking aims for
// (0 = built-in default).
struct YzTuning {
  uint32_t columns = 0;
  uint32_t target_blocks = 0;
};
YzTuning ebasis_yz_tuning(uint32_t nx, uint32_t ny, uint32_t nz, int precision);
void ebasis_set_yz_tuning(uint32_t nx, uint32_t ny, uint32_t nz, int precision, YzTuning t);
// Fills e, pp, tpp, zchunk, nchunks (fields left 0 are chosen automatically;
// target_blocks is the grid size the z chunking aims for, 0 = tuned/default).
template <typename T> cyr_status ebasis_yz_plan(YzArgs<T> &a, uint32_t target_blocks);
template <typename T> cyr_status launch_ebasis_yz(YzArgs<T> a, int ng, hipStream_t st);

// DC-replacement neighbour sums per (field, z) slice, FP64.
template <typename T> struct DcArgs {
  const cx<T> *vol[kXMaxOut];      // X-transformed fields
  uint32_t count;                  // volumes (0 = kEbFields)
  const cx<T> *twy, *twz;          // twz null in 2D
  uint32_t nx, ny, nz;
  double2 *dcz;                    // [count][Nz]
};
template <typename T> cyr_status launch_ebasis_dc(const DcArgs<T> &a, hipStream_t st);

template <typename T> struct AnglesArgs {
  EwaldGeom g;
  const cx<T> *part;
  uint32_t nchunks;
  const double2 *dcz;
  const T *amat;                   // [nA][kEbAmat]
  uint32_t nA;
  T *proj;                         // [nA][Ny][Nx]
};
template <typename T> cyr_status launch_ebasis_angles(const AnglesArgs<T> &a, int ng, hipStream_t st);

// ---- material-basis executor (mbasis.hip) ----------------------------------
// Packed basis volumes per k: 7 per material pair, 4 for an odd last material.
inline uint32_t mbasis_volumes(uint32_t M) { return 7 * (M / 2) + 4 * (M % 2); }
// Per column (FFT-index order): Zlo | (width << 16) of the shifted z indices the
// energies [0, ne) sample (kmag[e] = |k|); bmax[0] = widest band, bmax[1..2] = sum
// of the columns' widths (uint64), bmax[3..4] = columns inside the row spans
// (uint64), bmax[5..6] = D4 (uint64): the largest (2X - Nx)^2 + (2Y - Ny)^2 over the
// span endpoints (X shifted column, Y shifted row), so no detector pixel farther
// than sqrt(D4) / 2 from the plane's centre has a finite value in any energy of
// the group; then per-block scratch and the ny row spans (mbasis_bmax_words in all;
// mbasis_span_table points at the spans).  Row span: lo | (hi << 16), the shifted
// x coordinates that cover every column whose own or mirrored band is non-empty
// (lo > hi: none); the Y pass writes and the Z band pass reads only those.
constexpr uint32_t kBandHead = 7;
cyr_status launch_mbasis_bands(const EwaldGeom &g, const double *kmag, uint32_t ne, int32_t *band, uint32_t *bmax,
                               hipStream_t st);
uint64_t mbasis_bmax_words(uint32_t nx, uint32_t ny);
uint32_t *mbasis_span_table(uint32_t *bmax, uint32_t nx, uint32_t ny);
// Fields + X FFT of one material pass: a pair (m0, m0 + 1, 7 outputs) or a single
// material (m0, 4 outputs).
template <typename T>
cyr_status launch_mbasis_xfields(PolxArgs<T> a, bool pair, const void *vox, int storage, int morph, bool hann,
                                 hipStream_t st);
template <typename T> struct BandArgs {
  const cx<T> *vol[kXMaxOut];      // transformed packed volumes of this pass
  uint32_t count, vbase;           // volumes, global index of vol[0]
  const int32_t *band;             // [Ny][Nx] band table
  const uint32_t *span;            // [Ny] row spans (null: every column)
  cx<T> *own, *mir;                // [volumes][J][Ny][Nx]
  uint32_t J, nx, ny, nz;
};
template <typename T> cyr_status launch_mbasis_extract(const BandArgs<T> &a, hipStream_t st);
// Z FFT of the Y-transformed volumes writing only the band values (vol[] is read,
// not overwritten); *done = false when Nz has no compile-time plan (then run the
// Z FFT pass and launch_mbasis_extract).
template <typename T> cyr_status launch_mbasis_zband(const BandArgs<T> &a, hipStream_t st, bool *done);
// DC-replacement values of the packed volumes from their per-slice neighbour sums.
cyr_status launch_mbasis_dcsum(const double2 *dcz, uint32_t nvol, uint32_t nz, bool three, double2 *dcv,
                               hipStream_t st);
template <typename T> struct MbAnglesArgs {
  EwaldGeom g;                     // this energy's geometry
  const int32_t *band;
  const cx<T> *own, *mir;
  uint32_t J;
  const double2 *dcv;              // [volumes]
  const MatConst<T> *mats;         // this energy's M constants
  uint32_t M;
  const T *amat;                   // [nA][kEbAmat]
  uint32_t nA;
  T *proj;                         // [nA][Ny][Nx] (per-angle projections), or
  T *quad;                         // [ne][P][4] in quad_index order: laboratory frame, the intensity's quadratic form
  Mat3<T> Q;                       // frame e1, e2, k^ (quadratic form)
  const double *kmag;              // quadratic form: |k| of the ne energies (mats then holds ne x M)
  uint32_t ne;
};
// quad set: one quadratic form per pixel for each of ne energies in one launch
// (laboratory frame), else every angle's projection of one energy.
template <typename T> cyr_status launch_mbasis_angles(const MbAnglesArgs<T> &a, int ng, hipStream_t st);

// Values per record of a plane of quadratic forms (q0, q1, q2 of one pixel, packed
// without padding: a gather reads 12 of every 12 bytes it brings in, not 12 of 16).
constexpr uint32_t kQuadRecord = 3;

// Record of pixel (x, y) in a plane of quadratic forms: 4 x 4 pixel blocks,
// row-major inside the block and over blocks, when both sides are multiples of 4;
// row-major otherwise.  A wave gathering the bilinear taps of a rotated 8 x 8
// pixel footprint then touches fewer cache lines than with row-major records
// (measured: thin-film rotation 6.9 -> 5.5 ms per 10 energies).
__host__ __device__ inline uint32_t quad_index(uint32_t x, uint32_t y, uint32_t nx, uint32_t ny) {
  if ((nx & 3u) == 0 && (ny & 3u) == 0) return ((y >> 2) * (nx >> 2) + (x >> 2)) * 16u + ((y & 3u) << 2) + (x & 3u);
  return y * nx + x;
}

// ---- rotation / averaging (rotate.hip) --------------------------------------
// Plain path: rotate one projection by the E angle and accumulate (FP64).
template <typename T>
cyr_status launch_rotate_accumulate(const T *proj, double *acc, uint32_t *count, uint32_t nx, uint32_t ny,
                                    const double inv[6], int rot_mask, hipStream_t st);
// Average over angles into avg (FP64): mask division or 1/nA.
cyr_status launch_average(const double *acc, const uint32_t *count, double *avg, uint64_t P, uint32_t nA,
                          int rot_mask, hipStream_t st);
// Output tiles (16 x 16) of the rotation kernels, ordered by distance from the
// rotation centre (ty << 16 | tx); the kernels give XCD x the x-th annulus.
uint32_t rotation_tile_count(uint32_t nx, uint32_t ny);
// Tiles of rotation_tile_order whose centres lie within r pixels of the plane's
// centre (a prefix of that order).
uint32_t rotation_tiles_within(uint32_t nx, uint32_t ny, double r);
void rotation_tile_order(uint32_t nx, uint32_t ny, uint32_t *out);
// E-basis path: rotate every angle's projection, accumulate and average in one pass
// (tiles: rotation_tile_order on the device).
template <typename T>
cyr_status launch_rotate_all(const T *proj, const double *inv_tab, uint32_t nA, uint32_t nx, uint32_t ny,
                             int rot_mask, double *avg, const uint32_t *tiles, double rfin, hipStream_t st);
// Quarter-turn orbits of a square N x N plane (N even): orbit representatives
// in 16 x 16 tiles (ty << 16 | tx) ordered by distance from the rotation centre.
uint32_t rotation_orbit_tile_count(uint32_t n);
// Orbit tiles whose centres lie within r pixels of the plane's centre (a prefix
// of rotation_orbit_tile_order).
uint32_t rotation_orbit_tiles_within(uint32_t n, double r);
void rotation_orbit_tile_order(uint32_t n, uint32_t *out);
// Quarter turns Q (2..4) that the orbit kernel can share taps over, given
// `quarter` angle steps per 90 degrees (0: none); 0 when it does not apply.
uint32_t rotation_quarter_turns(uint32_t nx, uint32_t ny, uint32_t nA, uint32_t quarter);
// Laboratory frame: the same for ne energies from their per-pixel quadratic forms
// quad[ne][P][4] (quad_index order; I(theta) = c^2 q0 + s^2 q1 + 2 c s q2) with
// coef[a] = (c^2, s^2, 2 c s) (FP64; coef_t: the same in the compute precision)
// into avg[ne][Ny][Nx]; each angle's taps serve several energies.  quarter (angle
// steps per 90 degrees, 0 = none) with otiles (orbit tile order) lets each tap
// cell serve the angles a quarter turn apart; share: coef_t of angles a and
// a + 2 quarter turns are equal (then one intensity serves both pairs).
// rfin >= 0: no plane value farther than rfin pixels from the centre is finite
// (the tiles beyond it are then filled without sampling: every tap there is NaN,
// so their averages are NaN, or 0 with the rotation mask).
template <typename T>
cyr_status launch_rotate_quad(const T *quad, const double *inv_tab, const double *coef, const T *coef_t, uint32_t nA,
                              uint32_t nx, uint32_t ny, int rot_mask, double *avg, uint32_t ne, const uint32_t *tiles,
                              uint32_t quarter, const uint32_t *otiles, bool share, double rfin, hipStream_t st);
// Detector warp of the averaged pattern into the output (prefill 0 or NaN).
template <typename T>
cyr_status launch_detector_warp(const double *avg, T *out, uint32_t nx, uint32_t ny, const double inv[6],
                                int rot_mask, hipStream_t st);
// The same for ne averages avg[ne][Ny][Nx] into out + e * out_stride.
template <typename T>
cyr_status launch_detector_warp_n(const double *avg, T *out, uint64_t out_stride, uint32_t ne, uint32_t nx,
                                  uint32_t ny, const double inv[6], int rot_mask, hipStream_t st);
// Generic bilinear warp (stage-level entry point).
template <typename T>
cyr_status launch_warp(const T *src, T *dst, uint32_t nx, uint32_t ny, const double inv[6], hipStream_t st);

}  // namespace cyr
