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
  StagePlan P;
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
  __device__ void run(cx<T> (&v)[E], uint32_t t, uint32_t p, cx<T> *lds, const cx<T> *__restrict__ tw) const {
    stockham_from_regs<T, E>(v, P, t, tpp, p, lds, L, tw);
  }
};

template <int A, int... B> struct FirstOf { static constexpr int value = A; };
template <int A, int... B> struct LastOf { static constexpr int value = LastOf<B...>::value; };
template <int A> struct LastOf<A> { static constexpr int value = A; };

// Stage with sub-transform length LS and radix R on registers, then (unless it
// is the last) the LDS exchange into the next stage's input order.  lb is the
// pencil's LDS base, PE the LDS stride between consecutive pencil elements.
template <typename T, int N, int E, int PE, int LS, int R, int... Rest>
__device__ inline void ct_stages(cx<T> (&v)[E], uint32_t t, cx<T> *lb, const cx<T> *__restrict__ tw) {
  constexpr int TPP = N / E, NB = E / R, TSTEP = N / (LS * R);
  static_assert(E % R == 0, "elements per thread must be a multiple of every radix");
  if constexpr (LS > 1) {
    if constexpr (TPP % LS == 0) {
      // every butterfly of this thread has the same twiddle exponent k = t mod LS
      const uint32_t k = t % LS;
      if (k != 0) {
        cx<T> w[R];
#pragma unroll
        for (int r = 1; r < R; r++) w[r] = tw[r * k * TSTEP];
#pragma unroll
        for (int b = 0; b < NB; b++)
#pragma unroll
          for (int r = 1; r < R; r++) v[b * R + r] = cmul<T>(v[b * R + r], w[r]);
      }
    } else {
#pragma unroll
      for (int b = 0; b < NB; b++) {
        const uint32_t k = (t + b * TPP) % LS;
        if (k != 0) {
#pragma unroll
          for (int r = 1; r < R; r++) v[b * R + r] = cmul<T>(v[b * R + r], tw[r * k * TSTEP]);
        }
      }
    }
  }
#pragma unroll
  for (int b = 0; b < NB; b++) bfly<T, R>(&v[b * R]);
  if constexpr (sizeof...(Rest) > 0) {
    if constexpr (TPP % LS == 0) {
      const uint32_t o0 = (t / LS) * (LS * R) + (t % LS);
#pragma unroll
      for (int b = 0; b < NB; b++)
#pragma unroll
        for (int r = 0; r < R; r++) lb[(o0 + (b * TPP / LS) * (LS * R) + r * LS) * PE] = v[b * R + r];
    } else {
#pragma unroll
      for (int b = 0; b < NB; b++) {
        const uint32_t j = t + b * TPP;
        const uint32_t o0 = (j / LS) * (LS * R) + (j % LS);
#pragma unroll
        for (int r = 0; r < R; r++) lb[(o0 + r * LS) * PE] = v[b * R + r];
      }
    }
    lds_sync();
    constexpr int R2 = FirstOf<Rest...>::value;
    constexpr int NB2 = E / R2, M2 = N / R2;
    static_assert(E % R2 == 0, "elements per thread must be a multiple of every radix");
#pragma unroll
    for (int b = 0; b < NB2; b++)
#pragma unroll
      for (int r = 0; r < R2; r++) v[b * R2 + r] = lb[(t + b * TPP + r * M2) * PE];
    lds_sync();
    ct_stages<T, N, E, PE, LS * R, Rest...>(v, t, lb, tw);
  }
}

// CONTIG: pencil p at lds + p * N; otherwise lds + p with element stride PP.
template <typename T, int N, int E, bool CONTIG, int PP, int... R> struct CtFft {
  static constexpr int kE = E, TPP = N / E;
  static constexpr int R0 = FirstOf<R...>::value, RL = LastOf<R...>::value;
  static constexpr int PE = CONTIG ? 1 : PP;
  __device__ CtFft(const StagePlan &, uint32_t, uint32_t, bool) {}
  __device__ static uint32_t in_pos(uint32_t t, int i) { return t + (i / R0) * TPP + (i % R0) * (N / R0); }
  __device__ static uint32_t out_pos(uint32_t t, int i) { return t + (i / RL) * TPP + (i % RL) * (N / RL); }
  __device__ static constexpr uint32_t in_delta(int i) { return (i / R0) * TPP + (i % R0) * (N / R0); }
  __device__ static constexpr bool exchanges() { return sizeof...(R) > 1; }
  __device__ static void run(cx<T> (&v)[E], uint32_t t, uint32_t p, cx<T> *lds, const cx<T> *__restrict__ tw) {
    cx<T> *lb = CONTIG ? lds + (size_t)p * N : lds + p;
    ct_stages<T, N, E, PE, 1, R...>(v, t, lb, tw);
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

// Lengths with a compile-time plan, and its elements per thread.
inline bool ct_length(uint32_t n) {
  return n == 64 || n == 128 || n == 256 || n == 384 || n == 512 || n == 1024 || n == 2048;
}
inline uint32_t ct_elements(uint32_t n) { return n == 384 ? 12u : 8u; }

}  // namespace cyr
