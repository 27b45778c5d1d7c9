// Register/LDS FFT policies shared by the fused E-basis kernels.
//
// A policy transforms one pencil held by tpp threads, E elements each:
//   in_pos(t, i)   pencil index loaded into register slot i before run()
//   out_pos(t, i)  pencil index held by register slot i after run()
//   run(v, t, p, lds, tw)  Stockham stages, exchanging through LDS
// Layouts: CONTIG pencils are rows of the LDS tile (element e of pencil p at
// p*pitch + e); otherwise pencils are columns (e*pp + p), which keeps a wave's
// accesses to consecutive pencils in consecutive LDS words.
//
// RtFft runs a plan chosen at run time (any 2^a 3^b 5^c length); CtFft is the
// same algorithm with the radix sequence, the pencil length and the LDS pitch
// fixed at compile time, so every LDS offset and most index arithmetic fold
// into immediates and the register footprint shrinks.
#pragma once

#include <hip/hip_runtime.h>

#include "stockham.hpp"

namespace cyr {

template <typename T, int E> struct RtFft {
  static constexpr int kE = E;
  // the kernel argument itself (constant memory): a copy indexed by run-time
  // stage numbers would be placed in scratch
  const StagePlan &P;
  uint32_t tpp;
  LdsMap L;
  __device__ RtFft(const StagePlan &plan, uint32_t tpp_, uint32_t pitch, bool contig)
      : P(plan), tpp(tpp_), L{pitch, contig} {}
  __device__ uint32_t in_pos(uint32_t t, int i) const {
    const uint32_t R0 = P.radix[0];
    return t + ((uint32_t)i / R0) * tpp + ((uint32_t)i % R0) * (P.n / R0);
  }
  __device__ uint32_t out_pos(uint32_t t, int i) const {
    return final_index(t, tpp, P.n, P.radix[P.nstages - 1], i);
  }
  // in_pos(t, i) - in_pos(t, 0): the same for every thread
  __device__ uint32_t in_delta(int i) const {
    const uint32_t R0 = P.radix[0];
    return ((uint32_t)i / R0) * tpp + ((uint32_t)i % R0) * (P.n / R0);
  }
  __device__ bool exchanges() const { return P.nstages > 1; }
  // tw: the length-n twiddle block (global or an LDS copy); twlo: its FP32
  // residual block (may sit elsewhere; unused for FP64).  COMP = false keeps
  // plain products in every stage (twlo unused).
  template <bool COMP = true>
  __device__ void run(cx<T> (&v)[E], uint32_t t, uint32_t p, cx<T> *lds, const cx<T> *__restrict__ tw,
                      const cx<T> *twlo) const {
    stockham_from_regs<T, E>(v, P, t, tpp, p, lds, L, tw, COMP ? twlo : nullptr);
  }
};

template <int A, int... B> struct FirstOf { static constexpr int value = A; };
template <int A, int... B> struct LastOf { static constexpr int value = LastOf<B...>::value; };
template <int A> struct LastOf<A> { static constexpr int value = A; };

// LDS addressing of the compile-time stages.  Column tiles (pencils are
// columns, PE = pencils per tile) place element e at e * PE; row tiles (PE = 1)
// pad one slot per 8 elements, e -> e + e / 8, so the radix-R scatter of a
// Stockham stage (lane stride R elements) spreads over the LDS banks.
__host__ __device__ constexpr uint32_t pad8(uint32_t e) { return e + (e >> 3); }

// Largest (A mod 8) over the stage's runtime store bases A = (t / LS) LS R + t mod LS.
template <int LS, int R> constexpr uint32_t max_base_mod8() {
  uint32_t m = 0;
  for (uint32_t t = 0; t < 8u * LS; t++) {
    const uint32_t a = ((t / LS) * (LS * R) + (t % LS)) % 8;
    m = a > m ? a : m;
  }
  return m;
}

// LDS slot of element A + B (A runtime, B compile-time).  For row tiles the
// padding splits when the low bits of A and B cannot carry into bit 3.
template <int PE, int LS, int R, uint32_t B> __device__ inline uint32_t lds_slot(uint32_t A) {
  if constexpr (PE != 1) {
    return (A + B) * PE;
  } else if constexpr (max_base_mod8<LS, R>() + (B % 8) <= 7) {
    return pad8(A) + (B / 8) * 9 + (B % 8);
  } else {
    return pad8(A + B);
  }
}

// Register slot i = b R + r -> LDS element o0 + (b TPP / LS) LS R + r LS (stage outputs).
template <typename T, int E, int PE, int LS, int R, int TPP, int I>
__device__ inline void ct_store(const cx<T> (&v)[E], uint32_t o0, cx<T> *lb) {
  if constexpr (I < E) {
    constexpr int b = I / R, r = I % R;
    constexpr uint32_t B = (uint32_t)((b * TPP / LS) * (LS * R) + r * LS);
    lb[lds_slot<PE, LS, R, B>(o0)] = v[I];
    ct_store<T, E, PE, LS, R, TPP, I + 1>(v, o0, lb);
  }
}

// Register slot i = b R2 + r <- LDS element t + b TPP + r M2 (next stage's inputs).
template <typename T, int E, int PE, int TPP, int M2, int R2, int I>
__device__ inline void ct_load(cx<T> (&v)[E], uint32_t t, const cx<T> *lb) {
  if constexpr (I < E) {
    constexpr int b = I / R2, r = I % R2;
    constexpr uint32_t B = (uint32_t)(b * TPP + r * M2);
    static_assert(PE != 1 || B % 8 == 0, "row-tile loads need 8-aligned offsets");
    if constexpr (PE != 1) v[I] = lb[(t + B) * PE];
    else v[I] = lb[pad8(t) + (B / 8) * 9];
    ct_load<T, E, PE, TPP, M2, R2, I + 1>(v, t, lb);
  }
}

// Stage with sub-transform length LS and radix R on registers, then (unless it
// is the last) the LDS exchange into the next stage's input order.  lb is the
// pencil's LDS base, PE the LDS stride between consecutive pencil elements
// (PE = 1: padded row, see lds_slot).
template <typename T, int N, int E, int PE, bool CP, int LS, int R, int... Rest>
__device__ inline void ct_stages(cx<T> (&v)[E], uint32_t t, cx<T> *lb, const cx<T> *__restrict__ tw,
                                 const cx<T> *__restrict__ twlo) {
  constexpr int TPP = N / E, NB = E / R, TSTEP = N / (LS * R);
  // with CP the last stage applies the finest twiddles as compensated products (FP32 hi + lo)
  constexpr bool COMP = CP && sizeof...(Rest) == 0 && tw_blocks<T>() == 2;
  static_assert(E % R == 0, "elements per thread must be a multiple of every radix");
  if constexpr (LS > 1) {
    if constexpr (TPP % LS == 0) {
      // every butterfly of this thread has the same twiddle exponent k = t mod LS
      const uint32_t k = t % LS;
      if (k != 0) {
        cx<T> w[R], wl[R];
#pragma unroll
        for (int r = 1; r < R; r++) {
          w[r] = tw[r * k * TSTEP];
          if constexpr (COMP) wl[r] = twlo[r * k * TSTEP];
        }
#pragma unroll
        for (int b = 0; b < NB; b++)
#pragma unroll
          for (int r = 1; r < R; r++) {
            if constexpr (COMP) v[b * R + r] = cmul_comp<T>(v[b * R + r], w[r], wl[r]);
            else v[b * R + r] = cmul<T>(v[b * R + r], w[r]);
          }
      }
    } else {
#pragma unroll
      for (int b = 0; b < NB; b++) {
        const uint32_t k = (t + b * TPP) % LS;
        if (k != 0) {
#pragma unroll
          for (int r = 1; r < R; r++) {
            if constexpr (COMP) v[b * R + r] = cmul_comp<T>(v[b * R + r], tw[r * k * TSTEP], twlo[r * k * TSTEP]);
            else v[b * R + r] = cmul<T>(v[b * R + r], tw[r * k * TSTEP]);
          }
        }
      }
    }
  }
#pragma unroll
  for (int b = 0; b < NB; b++) bfly<T, R>(&v[b * R]);
  if constexpr (sizeof...(Rest) > 0) {
    if constexpr (TPP % LS == 0) {
      const uint32_t o0 = (t / LS) * (LS * R) + (t % LS);
      ct_store<T, E, PE, LS, R, TPP, 0>(v, o0, lb);
    } else {
#pragma unroll
      for (int b = 0; b < NB; b++) {
        const uint32_t j = t + b * TPP;
        const uint32_t o0 = (j / LS) * (LS * R) + (j % LS);
#pragma unroll
        for (int r = 0; r < R; r++) lb[PE == 1 ? pad8(o0 + r * LS) : (o0 + r * LS) * PE] = v[b * R + r];
      }
    }
    lds_sync();
    constexpr int R2 = FirstOf<Rest...>::value;
    static_assert(E % R2 == 0, "elements per thread must be a multiple of every radix");
    ct_load<T, E, PE, TPP, N / R2, R2, 0>(v, t, lb);
    lds_sync();
    ct_stages<T, N, E, PE, CP, LS * R, Rest...>(v, t, lb, tw, twlo);
  }
}

// CONTIG: pencil p at lds + p * pitch(N) as a padded row; otherwise lds + p with
// element stride PP.
template <typename T, int N, int E, bool CONTIG, int PP, int... R> struct CtFft {
  static constexpr uint32_t kPitch = N + N / 8;  // padded row length (CONTIG)
  static constexpr int kE = E, TPP = N / E;
  static constexpr int R0 = FirstOf<R...>::value, RL = LastOf<R...>::value;
  static constexpr int PE = CONTIG ? 1 : PP;
  __device__ CtFft(const StagePlan &, uint32_t, uint32_t, bool) {}
  __device__ static uint32_t in_pos(uint32_t t, int i) { return t + (i / R0) * TPP + (i % R0) * (N / R0); }
  __device__ static uint32_t out_pos(uint32_t t, int i) { return t + (i / RL) * TPP + (i % RL) * (N / RL); }
  __device__ static constexpr uint32_t in_delta(int i) { return (i / R0) * TPP + (i % R0) * (N / R0); }
  __device__ static constexpr bool exchanges() { return sizeof...(R) > 1; }
  template <bool COMP = true>
  __device__ static void run(cx<T> (&v)[E], uint32_t t, uint32_t p, cx<T> *lds, const cx<T> *__restrict__ tw,
                             const cx<T> *__restrict__ twlo) {
    cx<T> *lb = CONTIG ? lds + (size_t)p * kPitch : lds + p;
    ct_stages<T, N, E, PE, COMP, 1, R...>(v, t, lb, tw, twlo);
  }
};

// Compile-time plans of the lengths the workloads use: elements per thread and
// radix sequence (radix-8 first; E = lcm of the radices).
template <typename T, int N, bool CONTIG, int PP> struct CtPlan;
template <typename T, bool C, int PP> struct CtPlan<T, 64, C, PP> { using type = CtFft<T, 64, 8, C, PP, 8, 8>; };
template <typename T, bool C, int PP> struct CtPlan<T, 128, C, PP> { using type = CtFft<T, 128, 8, C, PP, 8, 8, 2>; };
template <typename T, bool C, int PP> struct CtPlan<T, 256, C, PP> { using type = CtFft<T, 256, 8, C, PP, 8, 8, 4>; };
template <typename T, bool C, int PP> struct CtPlan<T, 384, C, PP> {
  using type = CtFft<T, 384, 12, C, PP, 4, 4, 4, 2, 3>;
};
template <typename T, bool C, int PP> struct CtPlan<T, 512, C, PP> { using type = CtFft<T, 512, 8, C, PP, 8, 8, 8>; };
template <typename T, bool C, int PP> struct CtPlan<T, 1024, C, PP> {
  using type = CtFft<T, 1024, 8, C, PP, 8, 8, 8, 2>;
};
template <typename T, bool C, int PP> struct CtPlan<T, 2048, C, PP> {
  using type = CtFft<T, 2048, 8, C, PP, 8, 8, 8, 4>;
};

// Lengths with a compile-time plan, its elements per thread, and the LDS
// elements one padded row of it takes.
inline bool ct_length(uint32_t n) {
  return n == 64 || n == 128 || n == 256 || n == 384 || n == 512 || n == 1024 || n == 2048;
}
inline uint32_t ct_elements(uint32_t n) { return n == 384 ? 12u : 8u; }
inline uint32_t ct_row_pitch(uint32_t n) { return n + n / 8; }

}  // namespace cyr
