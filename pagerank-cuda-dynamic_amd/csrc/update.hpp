// update.hpp -- the per-vertex update shared by every plan of the rank update (rank.hip, binned.hip).
//
// Per vertex v (pagerankUpdateRanks{Thread,Block}CukU, pagerankCuda.hxx:87-109, :162-188), given its in-sum
// cv = sum_{u in in(v)} r[u]/deg[u]:
//   a[v] = C0 + P*cv                      -- or the DF-P closed form (C0 + P*(cv - r[v]/d))/(1 - P/d)
//   prune / frontier flags                 -- vaff[v] = 0 if |a-r|/max(r,a) <= C; naff[v] = (|a-r|/max > D)
// Unaffected rows of the dynamic variants are carried over unchanged.
#pragma once
#include "prh_device.hpp"
#include "prh_internal.hpp"

namespace prh {

template <int VARIANT>
struct Traits {
  static constexpr bool weighted = VARIANT >= kWeightedForward;
  static constexpr bool dynamic = !weighted && VARIANT >= kTraversal;
  static constexpr bool frontier = !weighted && VARIANT >= kFrontier;
  static constexpr bool prune = VARIANT == kPrune;
};

template <int VARIANT>
__device__ __forceinline__ bool affected(const RankArgs &A, uint32_t v) {
  if constexpr (Traits<VARIANT>::dynamic) return A.vaff[v] != 0;
  else return true;
}

// Flags a frontier variant leaves on a row: still affected (not pruned), and in the frontier.
struct RowFlags {
  bool affected = true, frontier = false;
};

// CONTRIB: the plan keeps contributions (ca = a/d for the next gathers); the propagation-blocked plan computes them
// from the ranks where it reads them, and passes crv = r/d itself (the same value, bit for bit).
template <int VARIANT, bool CONTRIB = true>
__device__ __forceinline__ double finish_vertex(const RankArgs &A, uint32_t v, double cv, uint32_t d, double rv,
                                                double crv, RowFlags &f) {
  using T = Traits<VARIANT>;
  double av;
  if constexpr (T::prune) av = fma(A.P, cv - crv, A.C0) * (1.0 / (1.0 - A.P / d));
  else av = fma(A.P, cv, A.C0);
  A.a[v] = av;
  if constexpr (CONTRIB) A.ca[v] = av / d;
  const double ev = fabs(av - rv);
  if constexpr (T::frontier) {
    const double rel = ev / fmax(rv, av);
    if (T::prune && rel <= A.C) {
      A.vaff[v] = 0;
      f.affected = false;
    }
    f.frontier = rel > A.D;
    A.naff[v] = f.frontier ? 1 : 0;
  }
  return ev;
}

template <int VARIANT>
__device__ __forceinline__ double finish_vertex(const RankArgs &A, uint32_t v, double cv, uint32_t d, double rv,
                                                double crv) {
  RowFlags f;
  return finish_vertex<VARIANT>(A, v, cv, d, rv, crv, f);
}

// Per-lane sums of the extension's two costs (frontier variants with A.stats set).
struct CostAcc {
  double pull = 0.0, push = 0.0;
  __device__ void add(const RowFlags &f, uint32_t in_deg, uint32_t out_deg) {
    if (!f.affected) pull += in_deg;
    if (f.frontier) push += out_deg;
  }
};

// Workgroup totals of the extension costs to A.stats[2 slot], [2 slot + 1] (frontier variants with A.stats), for a
// workgroup of WAVES waves. Every thread must call it.
template <int VARIANT, int WAVES>
__device__ __forceinline__ void wg_store_costs(const RankArgs &A, const CostAcc &c, uint32_t slot, double *lds) {
  if constexpr (Traits<VARIANT>::frontier) {
    if (!A.stats) return;
    const int w = threadIdx.x / kWave;
    const double pull = group_sum<kWave>(c.pull), push = group_sum<kWave>(c.push);
    if ((threadIdx.x & (kWave - 1)) == 0) {
      lds[w] = pull;
      lds[WAVES + w] = push;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double a = 0.0, b = 0.0;
      for (int i = 0; i < WAVES; ++i) {
        a += lds[i];
        b += lds[WAVES + i];
      }
      A.stats[2 * size_t(slot)] = a;
      A.stats[2 * size_t(slot) + 1] = b;
    }
  }
}

// Workgroup-wide max over WAVES waves, valid in thread 0. Every thread must call it.
template <int WAVES>
__device__ __forceinline__ double wg_max(double x, double *lds) {
  x = group_max<kWave>(x);
  const int w = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) lds[w] = x;
  __syncthreads();
  double m = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < WAVES; ++i) m = fmax(m, lds[i]);
  return m;
}

template <int VARIANT>
__device__ __forceinline__ double finish_vertex(const RankArgs &A, uint32_t v, double cv) {
  return finish_vertex<VARIANT>(A, v, cv, A.deg[v], A.r[v], Traits<VARIANT>::prune ? A.cr[v] : 0.0);
}

// Unaffected row of a dynamic variant: carried over unchanged (rank and, where the plan keeps them, contribution).
template <int VARIANT, bool CONTRIB = true>
__device__ __forceinline__ void copy_vertex(const RankArgs &A, uint32_t v, double rv, double crv) {
  A.a[v] = rv;
  if constexpr (CONTRIB) A.ca[v] = crv;
  if constexpr (Traits<VARIANT>::frontier) A.naff[v] = 0;
}

template <int VARIANT>
__device__ __forceinline__ void copy_vertex(const RankArgs &A, uint32_t v) {
  A.a[v] = A.r[v];
  A.ca[v] = A.cr[v];
  if constexpr (Traits<VARIANT>::frontier) A.naff[v] = 0;
}

}  // namespace prh
