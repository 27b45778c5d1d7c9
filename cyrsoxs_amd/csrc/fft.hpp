// Hand-written forward C2C FFT for gfx950: Stockham autosort pencils staged in
// LDS, one launch per axis for a batch of volumes.
#pragma once

#include <cstdint>

#include "common.hpp"
#include "stockham.hpp"

namespace cyr {

// DC-neighbour sums of X-transformed volumes taken by their Y pass (see
// fft_axis): dcz receives count x nz x fft_dc_slots(nx, ny) FP64 partials, in
// [volume][z][slot] order, whose fixed-order sum over one volume is
//   sum_{y,z} G(1,y,z) + G(Nx-1,y,z) + G(0,y,z) (2 cos(2 pi y/Ny) + 2 cos(2 pi z/Nz)),
// the z term only in 3D (twz: the Z twiddle table, null in 2D).  With span (Ny
// entries, lo | (hi << 16)), when fft_span_fits(Ny), the pass writes output row y
// only in the lines of adjacent columns (16 for Ny <= 512, 8 for 1024, 4 above)
// holding a column whose shifted x coordinate (the bin -> shifted map m - x,
// Nx + m - x, m = Nx/2) lies in lo..hi; the other columns keep their input.
struct FftDc {
  double2 *dcz;
  const void *twz;
  const uint32_t *span = nullptr;
};
// Partials per (volume, z) of a Y pass that can take the DC sums; 0 when it
// cannot (Nx < 3, or no compile-time plan for Ny).
template <typename T> uint32_t fft_dc_slots(uint32_t nx, uint32_t ny);
// Whether a Y pass of length ny honours FftDc::span.
template <typename T> bool fft_span_fits(uint32_t ny);

// One axis pass over `count` volumes laid out X-fastest with dims (nx, ny, nz).
// axis 0 = X (contiguous pencils), 1 = Y, 2 = Z (strided pencils).  With dc
// (axis 1 only, fft_dc_slots > 0) the pass also writes the DC-neighbour partials.
template <typename T>
cyr_status fft_axis(void *const *vols, uint32_t count, uint32_t nx, uint32_t ny, uint32_t nz,
                    int axis, hipStream_t stream, const FftDc *dc = nullptr);

// Full forward unnormalised 3D FFT (X, then Y, then Z) of `count` volumes.
template <typename T>
cyr_status fft3d(void *const *vols, uint32_t count, uint32_t nx, uint32_t ny, uint32_t nz,
                 hipStream_t stream);

// Twiddle table W_n[m] = exp(-2 pi i m / n), m < n, generated in FP64 on the
// host and cached per device (stored in the compute precision; FP32 tables add
// the block of rounding residuals, tw_len<T>(n) entries in all).
template <typename T> const cx<T> *twiddles(uint32_t n, cyr_status *st);

// Launch-shape override for the tuner: elements per thread and pencils per block
// for an axis length (0 = automatic choice).
struct FftShape { uint32_t e = 0, pp = 0; };
void fft_set_shape(uint32_t n, int contig, int precision, FftShape s);
void fft_clear_shapes();

void fft_release();

// Stockham stage plan of an axis length made of factors 2, 3, 5 (false otherwise);
// lcm = least common multiple of the radices (the minimum elements per thread).
bool make_stage_plan(uint32_t n, StagePlan &P, uint32_t &lcm);
size_t device_max_lds();

}  // namespace cyr
