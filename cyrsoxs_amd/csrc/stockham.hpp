// Stockham autosort stages shared by the FFT pencil kernels and the fused
// E-basis kernels.
//
// Butterfly j (0 <= j < n/R) of a stage with sub-transform length Ls reads
// x[j + r n/R], multiplies by W_{Ls R}^{r (j mod Ls)}, does an R-point DFT and
// writes y[(j / Ls) Ls R + (j mod Ls) + r Ls].  A thread holds E elements
// (E/R butterflies, j = t + b tpp).  Stage 0 reads from HBM (or from a caller
// functor), intermediate stages exchange through LDS, and the last stage either
// writes HBM or leaves its outputs in registers: after the last stage element
// v[b R + r] is output index j_b + r n/R.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace cyr {

constexpr int kMaxStages = 16;

struct StagePlan {
  uint32_t n;
  uint32_t nstages;
  uint32_t radix[kMaxStages];
  uint32_t ls[kMaxStages];
  uint32_t tstep[kMaxStages];
};

// Work-group barrier that orders LDS accesses only: outstanding global loads and
// stores stay in flight across it (a prefetched slice, or the previous field's
// stores, overlap the next stage).  In a single-wave work group the compiler
// drops the s_barrier and keeps the LDS wait.
__device__ inline void lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// ---- radix butterflies (forward, W = exp(-2 pi i / R)) ----------------------

template <typename T> __device__ inline void bfly2(cx<T> *a) {
  cx<T> t = a[0];
  a[0] = cadd<T>(t, a[1]);
  a[1] = csub<T>(t, a[1]);
}

template <typename T> __device__ inline void bfly4(cx<T> *a) {
  cx<T> t0 = cadd<T>(a[0], a[2]), t1 = csub<T>(a[0], a[2]);
  cx<T> t2 = cadd<T>(a[1], a[3]), t3 = cmul_mi<T>(csub<T>(a[1], a[3]));
  a[0] = cadd<T>(t0, t2);
  a[1] = cadd<T>(t1, t3);
  a[2] = csub<T>(t0, t2);
  a[3] = csub<T>(t1, t3);
}

template <typename T> __device__ inline void bfly8(cx<T> *a) {
  const T r = (T)0.70710678118654752440084436210484903928;
  cx<T> e[4] = {a[0], a[2], a[4], a[6]};
  cx<T> o[4] = {a[1], a[3], a[5], a[7]};
  bfly4<T>(e);
  bfly4<T>(o);
  // W8^1 = (1 - i)/sqrt2, W8^2 = -i, W8^3 = -(1 + i)/sqrt2
  cx<T> o1 = mk<T>((o[1].x + o[1].y) * r, (o[1].y - o[1].x) * r);
  cx<T> o2 = cmul_mi<T>(o[2]);
  cx<T> o3 = mk<T>((o[3].y - o[3].x) * r, -(o[3].x + o[3].y) * r);
  a[0] = cadd<T>(e[0], o[0]);
  a[4] = csub<T>(e[0], o[0]);
  a[1] = cadd<T>(e[1], o1);
  a[5] = csub<T>(e[1], o1);
  a[2] = cadd<T>(e[2], o2);
  a[6] = csub<T>(e[2], o2);
  a[3] = cadd<T>(e[3], o3);
  a[7] = csub<T>(e[3], o3);
}

template <typename T> __device__ inline void bfly3(cx<T> *a) {
  const T s = (T)0.86602540378443864676372317075293618347;  // sin(2 pi/3)
  cx<T> t1 = cadd<T>(a[1], a[2]);
  cx<T> t2 = csub<T>(a[1], a[2]);
  cx<T> m = mk<T>(a[0].x - (T)0.5 * t1.x, a[0].y - (T)0.5 * t1.y);
  cx<T> u = mk<T>(s * t2.y, -s * t2.x);  // -i s t2
  a[0] = cadd<T>(a[0], t1);
  a[1] = cadd<T>(m, u);
  a[2] = csub<T>(m, u);
}

template <typename T> __device__ inline void bfly5(cx<T> *a) {
  const T c1 = (T)0.30901699437494742410229341718281905886;   // cos(2 pi/5)
  const T c2 = (T)-0.80901699437494742410229341718281905886;  // cos(4 pi/5)
  const T s1 = (T)0.95105651629515357211643933337938214340;   // sin(2 pi/5)
  const T s2 = (T)0.58778525229247312916870595463907276860;   // sin(4 pi/5)
  cx<T> b1 = cadd<T>(a[1], a[4]), b2 = cadd<T>(a[2], a[3]);
  cx<T> d1 = csub<T>(a[1], a[4]), d2 = csub<T>(a[2], a[3]);
  cx<T> m1 = mk<T>(a[0].x + c1 * b1.x + c2 * b2.x, a[0].y + c1 * b1.y + c2 * b2.y);
  cx<T> m2 = mk<T>(a[0].x + c2 * b1.x + c1 * b2.x, a[0].y + c2 * b1.y + c1 * b2.y);
  cx<T> n1 = mk<T>(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y);
  cx<T> n2 = mk<T>(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y);
  a[0] = mk<T>(a[0].x + b1.x + b2.x, a[0].y + b1.y + b2.y);
  // y = m -/+ i n
  a[1] = mk<T>(m1.x + n1.y, m1.y - n1.x);
  a[4] = mk<T>(m1.x - n1.y, m1.y + n1.x);
  a[2] = mk<T>(m2.x + n2.y, m2.y - n2.x);
  a[3] = mk<T>(m2.x - n2.y, m2.y + n2.x);
}

template <typename T, int R> __device__ inline void bfly(cx<T> *a) {
  if constexpr (R == 2) bfly2<T>(a);
  else if constexpr (R == 3) bfly3<T>(a);
  else if constexpr (R == 4) bfly4<T>(a);
  else if constexpr (R == 5) bfly5<T>(a);
  else bfly8<T>(a);
}

// Twiddles and butterflies of one stage on registers.  twlo (the residual
// block of an FP32 table, or null) turns the twiddles into compensated products;
// the caller passes it for the last stage only.
template <typename T, int E, int R>
__device__ inline void stage_compute(cx<T> (&v)[E], uint32_t t, uint32_t tpp, uint32_t Ls, uint32_t tstep,
                                     const cx<T> *__restrict__ tw, const cx<T> *twlo = nullptr) {
  constexpr int NB = E / R;
#pragma unroll
  for (int b = 0; b < NB; b++) {
    const uint32_t j = t + b * tpp;
    const uint32_t k = j % Ls;
    if (k != 0) {
      if (tw_blocks<T>() == 2 && twlo) {
#pragma unroll
        for (int r = 1; r < R; r++) v[b * R + r] = cmul_comp<T>(v[b * R + r], tw[r * k * tstep], twlo[r * k * tstep]);
      } else {
#pragma unroll
        for (int r = 1; r < R; r++) v[b * R + r] = cmul<T>(v[b * R + r], tw[r * k * tstep]);
      }
    }
    bfly<T, R>(&v[b * R]);
  }
}

// LDS addressing of element e of the block's pencil p.
struct LdsMap {
  uint32_t pitch;
  bool contig;
  __device__ inline uint32_t operator()(uint32_t p, uint32_t e) const {
    return contig ? p * pitch + e : e * pitch + p;
  }
};

// Stage input from LDS in Stockham order.
template <typename T, int E, int R>
__device__ inline void stage_load_lds(cx<T> (&v)[E], uint32_t t, uint32_t tpp, uint32_t m, uint32_t p,
                                      const cx<T> *lds, LdsMap L) {
  constexpr int NB = E / R;
#pragma unroll
  for (int b = 0; b < NB; b++) {
    const uint32_t j = t + b * tpp;
#pragma unroll
    for (int r = 0; r < R; r++) v[b * R + r] = lds[L(p, j + r * m)];
  }
}

// Stage output to LDS in Stockham order.
template <typename T, int E, int R>
__device__ inline void stage_store_lds(const cx<T> (&v)[E], uint32_t t, uint32_t tpp, uint32_t Ls, uint32_t p,
                                       cx<T> *lds, LdsMap L) {
  constexpr int NB = E / R;
#pragma unroll
  for (int b = 0; b < NB; b++) {
    const uint32_t j = t + b * tpp;
    const uint32_t o0 = (j / Ls) * Ls * R + (j % Ls);
#pragma unroll
    for (int r = 0; r < R; r++) lds[L(p, o0 + r * Ls)] = v[b * R + r];
  }
}

// Input position of register slot (b, r) in stage 0 (and of any stage with m = n/R).
__device__ inline uint32_t stage_in_index(uint32_t t, uint32_t tpp, uint32_t m, int b, int r) {
  return t + b * tpp + r * m;
}

// Runs stages [1, nstages) on registers that already hold the stage-0 inputs:
// stage 0 compute, then LDS exchange for every later stage.  After return the
// registers hold the final outputs in last-stage order (v[b R + r] = output
// j_b + r n/R_last).  Uniform control flow: every thread of the block calls it.
// The last stage's twiddles are compensated products with the residual block
// twlo of an FP32 table.
template <typename T, int E>
__device__ inline void stockham_from_regs(cx<T> (&v)[E], const StagePlan &P, uint32_t t, uint32_t tpp, uint32_t p,
                                          cx<T> *lds, LdsMap L, const cx<T> *__restrict__ tw,
                                          const cx<T> *__restrict__ twlo) {
  for (uint32_t s = 0; s < P.nstages; s++) {
    const uint32_t R = P.radix[s], Ls = P.ls[s], ts = P.tstep[s], m = P.n / R;
    if (s > 0) {
      switch (R) {
        case 2: if constexpr (E % 2 == 0) stage_load_lds<T, E, 2>(v, t, tpp, m, p, lds, L); break;
        case 3: if constexpr (E % 3 == 0) stage_load_lds<T, E, 3>(v, t, tpp, m, p, lds, L); break;
        case 4: if constexpr (E % 4 == 0) stage_load_lds<T, E, 4>(v, t, tpp, m, p, lds, L); break;
        case 5: if constexpr (E % 5 == 0) stage_load_lds<T, E, 5>(v, t, tpp, m, p, lds, L); break;
        default: if constexpr (E % 8 == 0) stage_load_lds<T, E, 8>(v, t, tpp, m, p, lds, L); break;
      }
      lds_sync();
    }
    const cx<T> *lo = s + 1 == P.nstages && tw_blocks<T>() == 2 ? twlo : nullptr;
    switch (R) {
      case 2: if constexpr (E % 2 == 0) stage_compute<T, E, 2>(v, t, tpp, Ls, ts, tw, lo); break;
      case 3: if constexpr (E % 3 == 0) stage_compute<T, E, 3>(v, t, tpp, Ls, ts, tw, lo); break;
      case 4: if constexpr (E % 4 == 0) stage_compute<T, E, 4>(v, t, tpp, Ls, ts, tw, lo); break;
      case 5: if constexpr (E % 5 == 0) stage_compute<T, E, 5>(v, t, tpp, Ls, ts, tw, lo); break;
      default: if constexpr (E % 8 == 0) stage_compute<T, E, 8>(v, t, tpp, Ls, ts, tw, lo); break;
    }
    if (s + 1 < P.nstages) {
      switch (R) {
        case 2: if constexpr (E % 2 == 0) stage_store_lds<T, E, 2>(v, t, tpp, Ls, p, lds, L); break;
        case 3: if constexpr (E % 3 == 0) stage_store_lds<T, E, 3>(v, t, tpp, Ls, p, lds, L); break;
        case 4: if constexpr (E % 4 == 0) stage_store_lds<T, E, 4>(v, t, tpp, Ls, p, lds, L); break;
        case 5: if constexpr (E % 5 == 0) stage_store_lds<T, E, 5>(v, t, tpp, Ls, p, lds, L); break;
        default: if constexpr (E % 8 == 0) stage_store_lds<T, E, 8>(v, t, tpp, Ls, p, lds, L); break;
      }
      lds_sync();
    }
  }
}

// Output index held by register slot i after stockham_from_regs: slot i = b R + r
// of the last stage with radix R holds output (t + b tpp) + r n/R.
__device__ inline uint32_t final_index(uint32_t t, uint32_t tpp, uint32_t n, uint32_t R_last, int i) {
  const uint32_t b = (uint32_t)i / R_last, r = (uint32_t)i % R_last;
  return t + b * tpp + r * (n / R_last);
}

}  // namespace cyr
