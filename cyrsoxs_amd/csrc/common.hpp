// Shared types, error plumbing and device helpers of the CyRSoXS HIP library.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/cyrsoxs.h"

namespace cyr {

// ---------------------------------------------------------------------------
// Error plumbing: every entry point returns a cyr_status; the detail string is
// kept per thread for cyr_last_error().
// ---------------------------------------------------------------------------
void set_error(const std::string &msg);
cyr_status fail(cyr_status s, const std::string &msg);
cyr_status hip_fail(hipError_t e, const char *what, const char *file, int line);

#define CYR_HIP(expr)                                                   \
  do {                                                                  \
    hipError_t _e = (expr);                                             \
    if (_e != hipSuccess) return ::cyr::hip_fail(_e, #expr, __FILE__, __LINE__); \
  } while (0)

#define CYR_TRY(expr)                     \
  do {                                    \
    cyr_status _s = (expr);               \
    if (_s != CYR_OK) return _s;          \
  } while (0)

// Checks that the last kernel launch was accepted.
#define CYR_LAUNCHED() CYR_HIP(hipGetLastError())

// ---------------------------------------------------------------------------
// Complex types: interleaved (re, im) pairs of the compute precision.
// ---------------------------------------------------------------------------
template <typename T> struct CplxOf;
template <> struct CplxOf<float> { using type = float2; };
template <> struct CplxOf<double> { using type = double2; };
template <typename T> using cx = typename CplxOf<T>::type;

template <typename T> __host__ __device__ inline cx<T> mk(T re, T im) {
  cx<T> r;
  r.x = re;
  r.y = im;
  return r;
}
template <typename T> __device__ inline cx<T> cadd(cx<T> a, cx<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T> __device__ inline cx<T> csub(cx<T> a, cx<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T> __device__ inline cx<T> cmul(cx<T> a, cx<T> b) {
  return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T> __device__ inline cx<T> cscale(cx<T> a, T s) { return mk<T>(a.x * s, a.y * s); }

// Twiddle tables of length n hold w^m = exp(-2 pi i m / n), m < n, computed in
// FP64 and rounded once.  An FP32 table has a second block of n entries right
// after it: the residuals lo[m] = w^m - hi[m], rounded to FP32, so that hi + lo
// carries the FP64 twiddle to about 48 bits.  tw_len is the table's entry count
// (what an LDS copy of the table must hold).
template <typename T> __host__ __device__ constexpr uint32_t tw_blocks() { return sizeof(T) == 4 ? 2u : 1u; }
template <typename T> __host__ __device__ constexpr size_t tw_len(size_t n) { return n * tw_blocks<T>(); }

// Blocks of wg threads using `bytes` of LDS that one CU holds at once
// (160 KiB of LDS and 32 waves per CU).
__host__ __device__ constexpr unsigned cu_blocks(size_t bytes, unsigned wg) {
  const unsigned by_waves = 32u / ((wg + 63u) / 64u);
  const size_t by_lds = bytes ? (size_t)163840 / bytes : by_waves;
  return by_lds < by_waves ? (unsigned)by_lds : by_waves;
}
// Whether a kernel with `used` LDS entries besides a staged length-n twiddle
// copy can also stage the FP32 residual block without losing occupancy
// (otherwise the residuals are read from the global table).
template <typename T> __host__ __device__ constexpr bool tw_lo_in_lds(size_t used, size_t n, unsigned wg) {
  return tw_blocks<T>() == 2 &&
         cu_blocks((used + 2 * n) * sizeof(cx<T>), wg) >= cu_blocks((used + n) * sizeof(cx<T>), wg);
}

// Compensated twiddle product v (hi + lo): the small product v lo is formed
// first and folded into the fused multiply-adds of v hi, so the twiddle's own
// rounding error (up to 2^-24 relative) drops out and only the arithmetic's
// remains.  FP64 tables have no residual block and take the plain product.
template <typename T> __device__ inline cx<T> cmul_comp(cx<T> v, cx<T> hi, cx<T> lo) {
  if constexpr (sizeof(T) == 4) {
    const float cr = v.x * lo.x - v.y * lo.y, ci = v.x * lo.y + v.y * lo.x;
    return mk<T>(__builtin_fmaf(v.x, hi.x, __builtin_fmaf(-v.y, hi.y, cr)),
                 __builtin_fmaf(v.x, hi.y, __builtin_fmaf(v.y, hi.x, ci)));
  } else {
    (void)lo;
    return cmul<T>(v, hi);
  }
}
// Re w^e of a length-n table in FP64 (hi + lo for FP32 tables).
template <typename T> __device__ inline double tw_cos(const cx<T> *tw, uint32_t n, uint32_t e) {
  if constexpr (sizeof(T) == 4) return (double)tw[e].x + (double)tw[n + e].x;
  else return (double)tw[e].x;
}
// v * w^e from a table of length n (tw: hi block; the residual block follows at tw + n).
template <typename T> __device__ inline cx<T> cmul_tw(cx<T> v, const cx<T> *tw, uint32_t n, uint32_t e) {
  if constexpr (sizeof(T) == 4) return cmul_comp<T>(v, tw[e], tw[n + e]);
  else return cmul<T>(v, tw[e]);
}
// Compile-time loop: fn(std::integral_constant<int, I>) for I in [I0, N).  Used
// where a #pragma unroll may be declined (large bodies), which would turn
// register arrays indexed by the loop variable into scratch memory.
template <int I, int N, class Fn> __device__ inline void static_for(Fn &&fn) {
  if constexpr (I < N) {
    fn(std::integral_constant<int, I>{});
    static_for<I + 1, N>(fn);
  }
}
// multiply by -i
template <typename T> __device__ inline cx<T> cmul_mi(cx<T> a) { return mk<T>(a.y, -a.x); }

// ---------------------------------------------------------------------------
// Morphology storage: 4 values per voxel per material.
// ---------------------------------------------------------------------------
struct alignas(8) half4 { __half x, y, z, w; };
struct alignas(8) bf164 { __hip_bfloat16 x, y, z, w; };

template <typename T, typename S> struct VoxLoad;
template <typename T> struct VoxLoad<T, float4> {
  __device__ static inline void get(const float4 *p, uint64_t i, T v[4]) {
    float4 a = p[i];
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
  }
};
template <typename T> struct VoxLoad<T, double4> {
  __device__ static inline void get(const double4 *p, uint64_t i, T v[4]) {
    double4 a = p[i];
    v[0] = (T)a.x; v[1] = (T)a.y; v[2] = (T)a.z; v[3] = (T)a.w;
  }
};
template <typename T> struct VoxLoad<T, half4> {
  __device__ static inline void get(const half4 *p, uint64_t i, T v[4]) {
    half4 a = p[i];
    v[0] = (T)__half2float(a.x); v[1] = (T)__half2float(a.y);
    v[2] = (T)__half2float(a.z); v[3] = (T)__half2float(a.w);
  }
};
template <typename T> struct VoxLoad<T, bf164> {
  __device__ static inline void get(const bf164 *p, uint64_t i, T v[4]) {
    bf164 a = p[i];
    v[0] = (T)__bfloat162float(a.x); v[1] = (T)__bfloat162float(a.y);
    v[2] = (T)__bfloat162float(a.z); v[3] = (T)__bfloat162float(a.w);
  }
};

// ---------------------------------------------------------------------------
// Per-material optical constants for one energy, derived on the host in FP64:
//   a = npara^2 - 1, b = nperp^2 - 1, c = (npara + 2 nperp)^2 / 9 - 1,
//   d = npara^2 - nperp^2   (complex, stored re/im).
// The vacuum term (the "- Vfrac" of include/uniaxial.h:106-150) is folded in
// analytically so FP32 does not lose three digits to 1 - n^2 cancellation.
// ---------------------------------------------------------------------------
template <typename T> struct MatConst {
  T ar, ai, br, bi, cr, ci, dr, di;
};

// Device-side 3x3 real matrix and 3-vector (row-major).
template <typename T> struct Mat3 { T m[9]; };
template <typename T> struct Vec3 { T x, y, z; };

inline unsigned ceil_div(uint64_t a, uint64_t b) { return (unsigned)((a + b - 1) / b); }

// XCD-aware block order.  The dispatcher hands work group w to XCD w mod 8 (MI355X
// has 8 XCDs, each with its own L2); kernels whose neighbouring tiles share HBM
// sectors launch a 1D grid padded to a multiple of 8 and map block w to tile
//   q = (w mod 8) * (grid / 8) + w / 8,
// so every XCD walks one contiguous range of tiles and neighbours meet in its L2.
constexpr unsigned kXcds = 8;
inline unsigned xcd_grid(uint64_t tiles) { return ceil_div(tiles, kXcds) * kXcds; }
__device__ inline uint32_t xcd_tile(uint32_t w, uint32_t grid) {
  return (w % kXcds) * (grid / kXcds) + w / kXcds;
}
// Chunked form for grids whose work varies along the tile range (tiles that
// leave at once): XCD x walks chunks x, x + 8, x + 16, ... of C consecutive
// tiles, q = ((i / C) * 8 + x) * C + i mod C with i = w / 8, so neighbours still
// meet in one L2 and every XCD gets its share of the busy tiles.  The grid is
// padded to a multiple of 8 C; blocks past the last tile leave.
inline unsigned xcd_chunk_grid(uint64_t tiles, unsigned C) { return ceil_div(tiles, kXcds * C) * kXcds * C; }
__device__ inline uint32_t xcd_chunk_tile(uint32_t w, uint32_t C) {
  const uint32_t i = w / kXcds;
  return ((i / C) * kXcds + w % kXcds) * C + i % C;
}

// ---------------------------------------------------------------------------
// Per-device cached work buffers (grown on demand, freed by
// cyr_release_workspace or at process exit).
// ---------------------------------------------------------------------------
struct DeviceBuffer {
  void *ptr = nullptr;
  size_t bytes = 0;
};
cyr_status ensure_buffer(DeviceBuffer &b, size_t bytes);
void free_buffer(DeviceBuffer &b);

// A caller-provided device region that replaces allocation on the calling thread
// while it is installed (the allocation-free entry point): ensure_buffer carves
// 256-byte-aligned slices from it with a bump pointer (growing the topmost slice
// in place) and fails with CYR_E_NOMEM instead of allocating; free_buffer forgets
// slices without freeing them; twiddle tables are built into it per call, their
// host images kept alive in `keep` until the caller has drained `stream`.
struct ScratchArena {
  unsigned char *base = nullptr;
  size_t cap = 0, top = 0;
  size_t last_off = 0;
  void *last_ptr = nullptr;
  hipStream_t stream = nullptr;
  std::vector<std::vector<unsigned char>> keep;
  std::map<std::pair<uint32_t, int>, void *> tw;  // (length, precision) -> table
  // 256-byte-aligned slice of `bytes`, or nullptr when the region is exhausted
  void *carve(size_t bytes);
  bool owns(const void *p) const {
    return p >= (const void *)base && p < (const void *)(base + cap);
  }
};
void set_scratch_arena(ScratchArena *a);  // nullptr uninstalls
ScratchArena *scratch_arena();

// Process-wide device caches (twiddle tables, per-device work buffers) are shared
// by every call that may use them (workspace_use(), held for the call) and freed
// only by cyr_release_workspace under workspace_exclusive().  Objects that keep
// pointers into them beyond one call (batch plans, serving engines) are counted
// as live handles; release is refused while any exists.
std::shared_lock<std::shared_mutex> workspace_use();
std::unique_lock<std::shared_mutex> workspace_exclusive();
void live_handles_add(int delta);
int live_handles();
// Host -> device copy of a caller's (pageable) buffer.  Large buffers go through
// per-device pinned staging with several host threads and return once the copy
// has landed (after the work already queued on st); small ones are one copy
// queued on st.
cyr_status upload_host(void *dst, const void *src, size_t n, hipStream_t st);
void release_upload_staging();

// ---------------------------------------------------------------------------
// Opt-in per-kernel event timing (cyr_kernel_timing).  kt_begin/kt_end bracket
// one launch on `st`; `bytes` is the launch's algorithmic HBM traffic (each
// input read once, each output written once).  kt_collect (after the stream has
// drained) folds the recorded pairs into the per-name totals.
// ---------------------------------------------------------------------------
bool kt_enabled();
void kt_begin(hipStream_t st);
void kt_end(const char *name, hipStream_t st, double bytes = 0.0);
void kt_collect();

}  // namespace cyr
