// frontier.hip -- affected-vertex marking for the dynamic solvers.
//
//   mark_partial: naff[u] = 1 for every deletion or insertion source u, vaff[v] = 1 for every deletion target v
//                 (pagerankAffectedFrontierPartialCukU, pagerankCuda.hxx:272-281).
//   extend:       vaff[w] = 1 for every out-neighbour w of a vertex u with naff[u]
//                 (pagerankAffectedExtend{Thread,Block}CukU + pagerankMarkNeighborsCudU, :55-63, :313-366),
//                 split by out-degree: one lane per low-out-degree vertex, one wave per high-out-degree vertex.
//                 With `fresh` set it is one breadth-first level: only newly marked vertices enter `fresh`
//                 and `changed` records whether any vertex was newly marked (the DT closure, :438-456).
// Concurrent writers only ever store 1, so the flags are the same for any interleaving.
#include "prh_device.hpp"
#include "prh_internal.hpp"

namespace prh {
namespace {

__global__ __launch_bounds__(kBlock) void mark_partial_kernel(uint8_t *naff, uint8_t *vaff, const uint32_t *delu,
                                                              const uint32_t *delv, size_t ND, const uint32_t *insu,
                                                              size_t NI) {
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < ND + NI; i += size_t(gridDim.x) * kBlock) {
    const uint32_t u = i < ND ? delu[i] : insu[i - ND];
    naff[u] = 1;
    if (i < ND) vaff[delv[i]] = 1;
  }
}

// Marks w. Plain marking stores unconditionally: a byte store does not hold up the wave, a test before it would
// (measured slower). Traversal levels test first, since only newly marked vertices join the next level.
__device__ __forceinline__ void mark(const ExtendArgs &A, uint32_t w) {
  if (!A.fresh) {
    A.vaff[w] = 1;
  } else if (!A.vaff[w]) {
    A.vaff[w] = 1;
    A.fresh[w] = 1;
    *A.changed = 1;
  }
}

constexpr int kLowLanes = 8;                     // lanes per low-out-degree vertex
constexpr int kLowPerLane = 64 / kLowLanes;      // edges per lane (out-degree <= 64 in the low class)

// High out-degree vertices: one wave per vertex. Low ones (out-degree <= 64): kLowLanes lanes per vertex, each
// lane loading all of its (at most kLowPerLane) targets before marking any, so that the loads overlap.
__global__ __launch_bounds__(kBlock) void extend_kernel(ExtendArgs A, uint32_t nb_high, uint32_t nb_low, LoopCtl ctl,
                                                        int use_ctl) {
  if (use_ctl && !iteration_continues(ctl)) return;
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  if (blockIdx.x < nb_high) {
    const uint32_t n = A.sched.n_high + A.sched.n_huge;
    const uint32_t *order = A.sched.order + A.sched.n_low;
    for (uint32_t i = blockIdx.x * kWavesPerBlock + w; i < n; i += nb_high * kWavesPerBlock) {
      const uint32_t u = order[i];
      if (!A.naff[u]) continue;
      const uint32_t beg = A.off[u], end = A.off[u + 1];
      for (uint32_t e = beg + lane; e < end; e += kWave) mark(A, A.dst[e]);
    }
  } else {
    const uint32_t n = A.sched.n_low;
    constexpr uint32_t kGroups = kBlock / kLowLanes;
    const uint32_t g = threadIdx.x / kLowLanes, j = threadIdx.x % kLowLanes;
    for (uint32_t i = (blockIdx.x - nb_high) * kGroups + g; i < n; i += nb_low * kGroups) {
      const uint32_t u = A.sched.order[i];
      if (!A.naff[u]) continue;
      const uint32_t beg = A.off[u], end = A.off[u + 1];
      uint32_t t[kLowPerLane];
#pragma unroll
      for (int k = 0; k < kLowPerLane; ++k) {
        const uint32_t e = beg + j + uint32_t(k) * kLowLanes;
        t[k] = e < end ? A.dst[e] : 0xFFFFFFFFu;
      }
#pragma unroll
      for (int k = 0; k < kLowPerLane; ++k)
        if (t[k] != 0xFFFFFFFFu) mark(A, t[k]);
      // Only reached if the schedule's low class was split above 64.
      for (uint32_t e = beg + kLowLanes * kLowPerLane + j; e < end; e += kLowLanes) mark(A, A.dst[e]);
    }
  }
}

inline uint32_t clamp_blocks(size_t n, size_t per, uint32_t cap) {
  const size_t g = (n + per - 1) / per;
  return uint32_t(g > cap ? cap : g);
}

}  // namespace

hipError_t launch_mark_partial(uint8_t *naff, uint8_t *vaff, const uint32_t *delu, const uint32_t *delv, size_t ND,
                               const uint32_t *insu, size_t NI, hipStream_t s) {
  if (ND + NI == 0) return hipSuccess;
  mark_partial_kernel<<<clamp_blocks(ND + NI, kBlock, 4096), kBlock, 0, s>>>(naff, vaff, delu, delv, ND, insu, NI);
  return hipGetLastError();
}

hipError_t launch_extend(const ExtendArgs &A, unsigned long long *err, int *done, int l, double tol, hipStream_t s) {
  const uint32_t nb_high = clamp_blocks(size_t(A.sched.n_high) + A.sched.n_huge, kWavesPerBlock, 4096);
  const uint32_t nb_low = clamp_blocks(A.sched.n_low, kBlock / kLowLanes, 8192);
  if (nb_high + nb_low == 0) return hipSuccess;
  const LoopCtl ctl{err, done, l, tol};
  extend_kernel<<<nb_high + nb_low, kBlock, 0, s>>>(A, nb_high, nb_low, ctl, err ? 1 : 0);
  return hipGetLastError();
}

}  // namespace prh
