// frontier.hip -- affected-vertex marking for the dynamic solvers.
//
//   mark_partial: naff[u] = 1 for every deletion or insertion source u, vaff[v] = 1 for every deletion target v
//                 (pagerankAffectedFrontierPartialCukU, pagerankCuda.hxx:272-281).
//   extend:       vaff[w] = 1 for every out-neighbour w of a vertex u with naff[u]
//                 (pagerankAffectedExtend{Thread,Block}CukU + pagerankMarkNeighborsCudU, :55-63, :313-366),
//                 split by out-degree: 8 lanes per low-out-degree vertex, one wave per high-out-degree vertex.
//                 With `fresh` set it is one breadth-first level: only newly marked vertices enter `fresh`
//                 and `changed` records whether any vertex was newly marked (the DT closure, :438-456).
//   extend_batch  the same marking for the initial extension, whose flagged vertices are exactly the batch's
//                 sources: driven by the batch list instead of a scan of every vertex, with each source's edges
//                 cut into chunks of kBatchChunk spread over all the waves (a hub's edges are not left to one).
//   direction /   the same extension in pull form (the direction-optimizing idea of Beamer et al., SC 2012):
//   extend_pull   every unaffected row w takes vaff[w] = 1 when an in-neighbour is in the frontier. The rank
//                 launch leaves per-workgroup costs of both forms (in-edges of unaffected rows, out-edges of
//                 frontier rows); `direction` totals them and the cheaper form runs. Both give the same flags.
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
// lane loading all of its (at most kLowPerLane) targets before marking any, so that the loads overlap. Huge ones,
// when the chunk table is given: one wave per chunk of kExtendChunk edges (blocks past nb_high + nb_low).
__global__ __launch_bounds__(kBlock) void extend_kernel(ExtendArgs A, uint32_t nb_high, uint32_t nb_low, LoopCtl ctl,
                                                        int use_ctl) {
  if (use_ctl && !iteration_continues(ctl)) return;
  if (A.direction && *A.direction != 0) return;
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  if (blockIdx.x >= nb_high + nb_low) {
    const uint32_t nw = (gridDim.x - nb_high - nb_low) * kWavesPerBlock;
    for (uint32_t k = (blockIdx.x - nb_high - nb_low) * kWavesPerBlock + w; k < A.n_chunks; k += nw) {
      const uint32_t u = A.chunk_row[k];
      if (!A.naff[u]) continue;
      for (uint32_t e = A.chunk_beg[k] + lane; e < A.chunk_end[k]; e += kWave) mark(A, A.dst[e]);
    }
  } else if (blockIdx.x < nb_high) {
    const uint32_t n = A.sched.n_high + (A.chunk_row ? 0u : A.sched.n_huge);
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

constexpr uint32_t kBatchChunk = 1024;  // edges of one source per wave-sized work unit of extend_batch

__device__ __forceinline__ uint32_t batch_source(const uint32_t *delu, size_t ND, const uint32_t *insu, size_t i) {
  return i < ND ? delu[i] : insu[i - ND];
}

// Claims source u for one batch entry: naff[u] goes from 1 (set by mark_partial) to 2 by a compare-and-swap on
// the aligned 32-bit word holding it (naff holds a multiple of 4 bytes); true for the one entry that made the change.
// Both values read as "in the frontier" wherever naff is used.
__device__ __forceinline__ bool claim_source(uint8_t *naff, uint32_t u) {
  uint32_t *word = reinterpret_cast<uint32_t *>(naff + (u & ~3u));
  const uint32_t shift = (u & 3u) * 8;
  uint32_t old = __hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (((old >> shift) & 0xFFu) == 1u) {
    const uint32_t prev = atomicCAS(word, old, (old & ~(0xFFu << shift)) | (2u << shift));
    if (prev == old) return true;
    old = prev;
  }
  return false;
}

// cnt[i] = the number of kBatchChunk-edge chunks of batch source i's out-edges for the first entry of each source
// (a source repeated in the batch is extended once), 0 for the others; cnt[n] = 0.
__global__ __launch_bounds__(kBlock) void batch_chunks_kernel(const uint32_t *delu, size_t ND, const uint32_t *insu,
                                                              size_t NI, const uint32_t *off, uint8_t *naff,
                                                              uint32_t *cnt) {
  const size_t n = ND + NI;
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i <= n; i += size_t(gridDim.x) * kBlock) {
    if (i == n) {
      cnt[n] = 0;
    } else {
      const uint32_t u = batch_source(delu, ND, insu, i);
      cnt[i] = claim_source(naff, u) ? (off[u + 1] - off[u] + kBatchChunk - 1) / kBatchChunk : 0u;
    }
  }
}

// Chunk t (of start[n] in all, start = exclusive scan of cnt) belongs to the source i with the largest start[i] <= t
// (sources without out-edges own no chunk and are skipped by that rule); one wave marks its edges.
__global__ __launch_bounds__(kBlock) void extend_batch_kernel(const uint32_t *delu, size_t ND, const uint32_t *insu,
                                                              size_t NI, const uint32_t *off, const uint32_t *dst,
                                                              const uint32_t *start, uint8_t *vaff) {
  const uint32_t n = uint32_t(ND + NI);
  const uint32_t total = __builtin_amdgcn_readfirstlane(start[n]);
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t nw = gridDim.x * kWavesPerBlock;
  for (uint32_t t = __builtin_amdgcn_readfirstlane(blockIdx.x * kWavesPerBlock + threadIdx.x / kWave); t < total;
       t += nw) {
    uint32_t lo = 0, hi = n;  // start[n] = total > t, so the answer is below n
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) / 2;
      if (__builtin_amdgcn_readfirstlane(start[mid]) <= t) lo = mid;
      else hi = mid - 1;
    }
    const uint32_t u = __builtin_amdgcn_readfirstlane(batch_source(delu, ND, insu, lo));
    const uint32_t e0 = off[u] + (t - start[lo]) * kBatchChunk;
    const uint32_t e1 = min(off[u + 1], e0 + kBatchChunk);
    for (uint32_t e = e0 + lane; e < e1; e += kWave) vaff[dst[e]] = 1;
  }
}

// The same two steps over a tidied batch held as out-CSR keys (source << 32 | target), as the streaming session keeps
// it on the device: deletions dk[0, *nd) in entries [0, cap) and insertions ik[0, *ni) in entries [cap, 2 cap); an
// entry past its side's count has no source and owns no chunk. Sizes come from the capacity, so the launches can be
// captured once and replayed for every batch.
__device__ __forceinline__ bool key_source(const unsigned long long *dk, const uint32_t *nd,
                                           const unsigned long long *ik, const uint32_t *ni, uint32_t cap, uint32_t i,
                                           uint32_t &u) {
  if (i < cap) {
    if (i >= *nd) return false;
    u = uint32_t(dk[i] >> 32);
    return true;
  }
  i -= cap;
  if (i >= *ni) return false;
  u = uint32_t(ik[i] >> 32);
  return true;
}

__global__ __launch_bounds__(kBlock) void batch_chunks_keys_kernel(const unsigned long long *dk, const uint32_t *nd,
                                                                   const unsigned long long *ik, const uint32_t *ni,
                                                                   uint32_t cap, const uint32_t *off, uint8_t *naff,
                                                                   uint32_t *cnt) {
  const uint32_t n = 2 * cap;
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i <= n; i += gridDim.x * kBlock) {
    uint32_t u = 0;
    cnt[i] = i < n && key_source(dk, nd, ik, ni, cap, i, u) && claim_source(naff, u)
                 ? (off[u + 1] - off[u] + kBatchChunk - 1) / kBatchChunk
                 : 0u;
  }
}

__global__ __launch_bounds__(kBlock) void extend_batch_keys_kernel(const unsigned long long *dk, const uint32_t *nd,
                                                                   const unsigned long long *ik, const uint32_t *ni,
                                                                   uint32_t cap, const uint32_t *off,
                                                                   const uint32_t *dst, const uint32_t *start,
                                                                   uint8_t *vaff) {
  const uint32_t n = 2 * cap;
  const uint32_t total = __builtin_amdgcn_readfirstlane(start[n]);
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t nw = gridDim.x * kWavesPerBlock;
  for (uint32_t t = __builtin_amdgcn_readfirstlane(blockIdx.x * kWavesPerBlock + threadIdx.x / kWave); t < total;
       t += nw) {
    uint32_t lo = 0, hi = n;  // start[n] = total > t, so the owner is below n
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) / 2;
      if (__builtin_amdgcn_readfirstlane(start[mid]) <= t) lo = mid;
      else hi = mid - 1;
    }
    uint32_t u = 0;
    key_source(dk, nd, ik, ni, cap, lo, u);  // entry lo owns chunk t, so it has a source
    u = __builtin_amdgcn_readfirstlane(u);
    const uint32_t e0 = off[u] + (t - start[lo]) * kBatchChunk;
    const uint32_t e1 = min(off[u + 1], e0 + kBatchChunk);
    for (uint32_t e = e0 + lane; e < e1; e += kWave) vaff[dst[e]] = 1;
  }
}

constexpr int kDirThreads = 1024;

// One workgroup: totals of the per-workgroup costs, then the choice (pull only when strictly cheaper).
__global__ __launch_bounds__(kDirThreads) void direction_kernel(const double *stats, uint32_t n, int *direction,
                                                                LoopCtl ctl) {
  if (!iteration_continues(ctl)) return;
  __shared__ double lds[2 * (kDirThreads / kWave)];
  double pull = 0.0, push = 0.0;
  for (uint32_t i = threadIdx.x; i < n; i += kDirThreads) {
    pull += stats[2 * size_t(i)];
    push += stats[2 * size_t(i) + 1];
  }
  pull = group_sum<kWave>(pull);
  push = group_sum<kWave>(push);
  const int w = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    lds[w] = pull;
    lds[kDirThreads / kWave + w] = push;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int k = 0; k < kDirThreads / kWave; ++k) {
      a += lds[k];
      b += lds[kDirThreads / kWave + k];
    }
    *direction = a < b ? 1 : 0;
  }
}

constexpr int kPullLanes = 8;

// Pull form of the extension over the in-CSR rows (rows sorted by in-degree, descending): a wave per row for
// rows [0, n_wave), kPullLanes lanes per row for [n_wave, n_group), one lane per row after that. Only rows not yet
// affected look at their in-neighbours, and stop at the first one in the frontier.
__global__ __launch_bounds__(kBlock) void extend_pull_kernel(PullArgs A, uint32_t nb_wave, uint32_t nb_group,
                                                             uint32_t nb_one, LoopCtl ctl) {
  if (!iteration_continues(ctl)) return;
  if (*A.direction != 1) return;
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  if (blockIdx.x < nb_wave) {
    for (uint32_t v = blockIdx.x * kWavesPerBlock + w; v < A.n_wave; v += nb_wave * kWavesPerBlock) {
      if (A.vaff[v]) continue;  // wave-uniform
      const uint32_t b = A.off[v], e = A.off[v + 1];
      for (uint32_t k = b; k < e; k += kWave) {
        const bool hit = k + lane < e && A.naff[A.src[k + lane]];
        if (__ballot(hit)) {
          if (lane == 0) A.vaff[v] = 1;
          break;
        }
      }
    }
  } else if (blockIdx.x < nb_wave + nb_group) {
    const uint32_t g = threadIdx.x / kPullLanes, j = threadIdx.x % kPullLanes;
    constexpr uint32_t kGroups = kBlock / kPullLanes;
    const uint32_t first = A.n_wave;
    for (uint32_t r0 = first + (blockIdx.x - nb_wave) * kGroups; r0 < A.n_group; r0 += nb_group * kGroups) {
      const uint32_t v = r0 + g;  // rows r0 .. r0 + kGroups - 1 in this round (block-uniform loop)
      bool hit = false;
      if (v < A.n_group && !A.vaff[v]) {
        const uint32_t b = A.off[v], e = A.off[v + 1];
        for (uint32_t k = b + j; k < e && !hit; k += kPullLanes) hit = A.naff[A.src[k]] != 0;
      }
      const unsigned long long bal = __ballot(hit);
      const unsigned group_bits = unsigned(bal >> ((lane / kPullLanes) * kPullLanes)) & ((1u << kPullLanes) - 1);
      if (j == 0 && group_bits) A.vaff[v] = 1;
    }
  } else {
    for (uint32_t v = A.n_group + (blockIdx.x - nb_wave - nb_group) * kBlock + threadIdx.x; v < A.N;
         v += nb_one * kBlock) {
      if (A.vaff[v]) continue;
      const uint32_t b = A.off[v], e = A.off[v + 1];
      for (uint32_t k = b; k < e; ++k)
        if (A.naff[A.src[k]]) {
          A.vaff[v] = 1;
          break;
        }
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
  const uint32_t nb_high =
      clamp_blocks(size_t(A.sched.n_high) + (A.chunk_row ? 0u : A.sched.n_huge), kWavesPerBlock, 4096);
  const uint32_t nb_low = clamp_blocks(A.sched.n_low, kBlock / kLowLanes, 8192);
  const uint32_t nb_chunk = A.chunk_row ? clamp_blocks(A.n_chunks, kWavesPerBlock, 4096) : 0u;
  if (nb_high + nb_low + nb_chunk == 0) return hipSuccess;
  const LoopCtl ctl{err, done, l, tol};
  extend_kernel<<<nb_high + nb_low + nb_chunk, kBlock, 0, s>>>(A, nb_high, nb_low, ctl, err ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_extend_batch(uint8_t *vaff, uint8_t *naff, const uint32_t *delu, size_t ND, const uint32_t *insu,
                               size_t NI, const uint32_t *off, const uint32_t *dst, uint32_t *cnt, uint32_t *start,
                               void *tmp, size_t &tmp_bytes, hipStream_t s) {
  const size_t n = ND + NI;
  if (!tmp) return scan_u32(nullptr, tmp_bytes, cnt, start, n + 1, s);
  if (!n) return hipSuccess;
  batch_chunks_kernel<<<clamp_blocks(n + 1, kBlock, 1024), kBlock, 0, s>>>(delu, ND, insu, NI, off, naff, cnt);
  hipError_t e = scan_u32(tmp, tmp_bytes, cnt, start, n + 1, s);
  if (e != hipSuccess) return e;
  extend_batch_kernel<<<1024, kBlock, 0, s>>>(delu, ND, insu, NI, off, dst, start, vaff);
  return hipGetLastError();
}

hipError_t launch_extend_batch_keys(uint8_t *vaff, uint8_t *naff, const unsigned long long *dk, const uint32_t *nd,
                                    const unsigned long long *ik, const uint32_t *ni, uint32_t cap,
                                    const uint32_t *off, const uint32_t *dst, uint32_t *cnt, uint32_t *start,
                                    void *tmp, size_t &tmp_bytes, hipStream_t s) {
  const size_t n = 2 * size_t(cap);
  if (!tmp) return scan_u32(nullptr, tmp_bytes, cnt, start, n + 1, s);
  batch_chunks_keys_kernel<<<clamp_blocks(n + 1, kBlock, 1024), kBlock, 0, s>>>(dk, nd, ik, ni, cap, off, naff, cnt);
  hipError_t e = scan_u32(tmp, tmp_bytes, cnt, start, n + 1, s);
  if (e != hipSuccess) return e;
  extend_batch_keys_kernel<<<1024, kBlock, 0, s>>>(dk, nd, ik, ni, cap, off, dst, start, vaff);
  return hipGetLastError();
}

hipError_t launch_direction(const double *stats, uint32_t n, int *direction, unsigned long long *err, int *done, int l,
                            double tol, hipStream_t s) {
  const LoopCtl ctl{err, done, l, tol};
  direction_kernel<<<1, kDirThreads, 0, s>>>(stats, n, direction, ctl);
  return hipGetLastError();
}

hipError_t launch_extend_pull(const PullArgs &A, unsigned long long *err, int *done, int l, double tol,
                              hipStream_t s) {
  const uint32_t nb_wave = clamp_blocks(A.n_wave, kWavesPerBlock, 2048);
  const uint32_t nb_group = clamp_blocks(size_t(A.n_group) - A.n_wave, kBlock / kPullLanes, 4096);
  const uint32_t nb_one = clamp_blocks(size_t(A.N) - A.n_group, kBlock, 4096);
  if (nb_wave + nb_group + nb_one == 0) return hipSuccess;
  const LoopCtl ctl{err, done, l, tol};
  extend_pull_kernel<<<nb_wave + nb_group + nb_one, kBlock, 0, s>>>(A, nb_wave, nb_group, nb_one, ctl);
  return hipGetLastError();
}

}  // namespace prh
