// rank.hip -- the pull-based, atomic-free rank update with the fused L-inf reduction, plus small map kernels.
//
// Row plan (every variant). One launch per iteration covers all vertices of the degree-ordered layout (rows
// sorted by in-degree, descending; layout.hip) through a table of sections, heaviest first so that long rows
// start early:
//   chunked rows  (in-degree above the chunk threshold): one workgroup per fixed-size edge chunk, each writing its
//                 partial sum; finish_chunked_kernel (the next launch) adds a row's partials in chunk order;
//   wave rows     one wave per row, DPP/swizzle wave sum;
//   group rows    W lanes per row (W = 1..32), one band per W over a contiguous row range, DPP group sum.
//
// Sliced plan (Static / Naive-dynamic; see SlicedPlan): sliced_gather_kernel sums each row's sources slice by
// slice, slice x on the workgroups that land on XCD x, and sliced_finish_kernel adds a row's slice partials.
//
// Per vertex v (pagerankUpdateRanks{Thread,Block}CukU, pagerankCuda.hxx:87-109, :162-188):
//   cv = sum_{u in in(v)} r[u]/deg[u]     -- gathered from the precomputed contributions cr[u] = r[u]/deg[u],
//                                            the same FP64 value the reference divides per edge
//   a[v] = C0 + P*cv                      -- or the DF-P closed form (C0 + P*(cv - r[v]/d))/(1 - P/d)
//   prune / frontier flags                 -- vaff[v] = 0 if |a-r|/max(r,a) <= C; naff[v] = (|a-r|/max > D)
// Each vertex has exactly one writer, so the update needs no atomics on ranks; the only atomic is one 64-bit max
// per workgroup for the L-inf change. Every sum is taken in a fixed order, so results repeat bit for bit.
#include <algorithm>
#include <type_traits>

#include "prh_device.hpp"
#include "prh_internal.hpp"
#include "update.hpp"

namespace prh {
namespace {

// Streamed arrays (source ids, offsets) are read with the non-temporal hint so that they do not displace the
// gathered contributions from L2; the gathers themselves use the default policy.
__device__ __forceinline__ uint32_t stream_u32(const uint32_t *p) { return __builtin_nontemporal_load(p); }

// Everything a row's update reads besides its gathers, loaded in one batch before them so that the loads overlap
// instead of adding round trips after the gathers.
struct RowOperands {
  uint32_t b, e;  // edge range
  uint32_t d;     // out-degree
  double rv;      // old rank
  double crv;     // old contribution (DF-P's self term)
  bool upd;       // the row is updated (always, except unaffected rows of the dynamic variants)
};

template <int VARIANT>
__device__ __forceinline__ RowOperands row_operands(const RankArgs &A, uint32_t v) {
  RowOperands o;
  o.b = stream_u32(A.off + v);
  o.e = stream_u32(A.off + v + 1);
  o.d = A.deg[v];
  o.rv = A.r[v];
  o.crv = Traits<VARIANT>::dynamic ? A.cr[v] : 0.0;
  o.upd = affected<VARIANT>(A, v);
  return o;
}

// Workgroup totals of the costs to A.stats[2 slot], [2 slot + 1]. Every thread must call it.
template <int VARIANT>
__device__ __forceinline__ void store_costs(const RankArgs &A, CostAcc c, uint32_t slot, double *lds8) {
  if constexpr (Traits<VARIANT>::frontier) {
    if (!A.stats) return;
    const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    const double pull = group_sum<kWave>(c.pull), push = group_sum<kWave>(c.push);
    if (lane == 0) {
      lds8[w] = pull;
      lds8[kWavesPerBlock + w] = push;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      A.stats[2 * size_t(slot)] = (lds8[0] + lds8[1]) + (lds8[2] + lds8[3]);
      A.stats[2 * size_t(slot) + 1] = (lds8[4] + lds8[5]) + (lds8[6] + lds8[7]);
    }
  }
}

// Sum of cr[src[e]] for e = first, first + STEP, ... < end, four gathers in flight per lane.
template <int STEP>
__device__ __forceinline__ double gather_sum(const uint32_t *src, const double *cr, uint32_t e, uint32_t end) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  for (; e + 3 * STEP < end; e += 4 * STEP) {
    const uint32_t u0 = stream_u32(src + e), u1 = stream_u32(src + e + STEP);
    const uint32_t u2 = stream_u32(src + e + 2 * STEP), u3 = stream_u32(src + e + 3 * STEP);
    s0 += cr[u0];
    s1 += cr[u1];
    s2 += cr[u2];
    s3 += cr[u3];
  }
  for (; e < end; e += STEP) s0 += cr[stream_u32(src + e)];
  return (s0 + s1) + (s2 + s3);
}

template <int STEP>
__device__ __forceinline__ double gather_sum(const RankArgs &A, uint32_t e, uint32_t end) {
  return gather_sum<STEP>(A.src, A.cr, e, end);
}

// Sum of a workgroup-strided gather, valid in thread 0. Every thread must call it.
__device__ __forceinline__ double block_gather_sum(const uint32_t *src, const double *cr, uint32_t b, uint32_t e,
                                                   double *lds) {
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  double acc = group_sum<kWave>(gather_sum<kBlock>(src, cr, b + threadIdx.x, e));
  if (lane == 0) lds[w] = acc;
  __syncthreads();
  const double t = (lds[0] + lds[1]) + (lds[2] + lds[3]);
  __syncthreads();
  return t;
}

// One workgroup per chunk item: its partial sum goes to the item's slot; rows are finished by
// finish_chunked_kernel after this launch (the kernel boundary makes every chunk's partial visible).
template <int VARIANT>
__device__ __forceinline__ void chunked_rows(const RankArgs &A, const RankPlan &P, const RankSection &S, uint32_t blk,
                                             double *lds) {
  const uint32_t ib = S.begin + blk * S.per_block, ie = min(S.end, ib + S.per_block);
  for (uint32_t it = ib; it < ie; ++it) {  // block-uniform
    const uint32_t v = P.item_row[it];
    if (!affected<VARIANT>(A, v)) continue;
    const double t = block_gather_sum(A.src, A.cr, P.item_beg[it], P.item_end[it], lds);
    if (threadIdx.x == 0) P.partial[it] = t;
  }
}

template <int VARIANT>
__device__ __forceinline__ void wave_rows(const RankArgs &A, const RankSection &S, uint32_t blk, double &emax,
                                          CostAcc &cost) {
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  const uint32_t rb = S.begin + blk * S.per_block, re = min(S.end, rb + S.per_block);
  for (uint32_t v = rb + w; v < re; v += kWavesPerBlock) {  // wave-uniform
    const RowOperands o = row_operands<VARIANT>(A, v);
    RowFlags f;
    if (!o.upd) {
      if (lane == 0) {
        copy_vertex<VARIANT>(A, v, o.rv, o.crv);
        f.affected = false;
        cost.add(f, o.e - o.b, o.d);
      }
      continue;
    }
    double acc = gather_sum<kWave>(A, o.b + lane, o.e);
    acc = group_sum<kWave>(acc);
    if (lane == 0) {
      emax = fmax(emax, finish_vertex<VARIANT>(A, v, acc, o.d, o.rv, o.crv, f));
      cost.add(f, o.e - o.b, o.d);
    }
  }
}

template <int VARIANT, int W>
__device__ __forceinline__ void group_rows(const RankArgs &A, const RankSection &S, uint32_t blk, double &emax,
                                           CostAcc &cost) {
  constexpr int G = kWave / W;  // rows per wave per round
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  const int g = lane / W, j = lane % W;
  const uint32_t rb = S.begin + blk * S.per_block, re = min(S.end, rb + S.per_block);
  for (uint32_t r0 = rb + w * G; r0 < re; r0 += kWavesPerBlock * G) {  // wave-uniform
    const uint32_t v = r0 + g;
    const bool valid = v < re;
    RowOperands o{};
    double acc = 0.0;
    if (valid) {
      o = row_operands<VARIANT>(A, v);
      if (o.upd) acc = gather_sum<W>(A, o.b + j, o.e);
    }
    acc = group_sum<W>(acc);  // all lanes active again here
    if (valid && j == 0) {
      RowFlags f;
      if (o.upd) {
        emax = fmax(emax, finish_vertex<VARIANT>(A, v, acc, o.d, o.rv, o.crv, f));
      } else {
        copy_vertex<VARIANT>(A, v, o.rv, o.crv);
        f.affected = false;
      }
      cost.add(f, o.e - o.b, o.d);
    }
  }
}

template <int VARIANT>
__global__ __launch_bounds__(kBlock) void rank_update_kernel(RankArgs A, RankPlan P, LoopCtl ctl) {
  if (!iteration_begin(ctl)) return;
  __shared__ double lds_sum[kWavesPerBlock];
  __shared__ double lds_max[kWavesPerBlock];
  int s = 0;
  while (s + 1 < P.nsec && blockIdx.x >= P.sec[s + 1].block_begin) ++s;
  const RankSection &S = P.sec[s];
  const uint32_t blk = blockIdx.x - S.block_begin;
  double emax = 0.0;
  CostAcc cost;
  if (S.kind == kChunkedRows) {
    chunked_rows<VARIANT>(A, P, S, blk, lds_sum);
  } else if (S.kind == kWaveRows) {
    wave_rows<VARIANT>(A, S, blk, emax, cost);
  } else {
    switch (S.width) {
      case 1: group_rows<VARIANT, 1>(A, S, blk, emax, cost); break;
      case 2: group_rows<VARIANT, 2>(A, S, blk, emax, cost); break;
      case 4: group_rows<VARIANT, 4>(A, S, blk, emax, cost); break;
      case 8: group_rows<VARIANT, 8>(A, S, blk, emax, cost); break;
      case 16: group_rows<VARIANT, 16>(A, S, blk, emax, cost); break;
      default: group_rows<VARIANT, 32>(A, S, blk, emax, cost); break;
    }
  }
  __shared__ double lds_cost[2 * kWavesPerBlock];
  store_costs<VARIANT>(A, cost, blockIdx.x, lds_cost);
  emax = block_max(emax, lds_max);
  if (threadIdx.x == 0) atomic_max_nonneg(&ctl.err[iter_index(ctl)], emax);
}

// Chunked rows, after the rank kernel of the same iteration: one wave per row adds the row's chunk partials (in a
// fixed order) and finishes the row, or copies it through when unaffected. Runs only if that kernel ran.
template <int VARIANT>
__global__ __launch_bounds__(kBlock) void finish_chunked_kernel(RankArgs A, RankPlan P, uint32_t n_rows, LoopCtl ctl) {
  if (!iteration_begin(ctl)) return;  // same outcome as the rank kernel's entry test of this iteration
  __shared__ double lds_max[kWavesPerBlock];
  __shared__ double lds_cost[2 * kWavesPerBlock];
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  double emax = 0.0;
  CostAcc cost;
  for (uint32_t v = blockIdx.x * kWavesPerBlock + w; v < n_rows; v += gridDim.x * kWavesPerBlock) {
    const uint32_t in_deg = A.off[v + 1] - A.off[v], d = A.deg[v];
    RowFlags f;
    if (!affected<VARIANT>(A, v)) {
      if (lane == 0) {
        copy_vertex<VARIANT>(A, v);
        f.affected = false;
        cost.add(f, in_deg, d);
      }
      continue;
    }
    double s = 0.0;
    for (uint32_t k = P.row_first[v] + lane; k < P.row_first[v + 1]; k += kWave) s += P.partial[k];
    s = group_sum<kWave>(s);
    if (lane == 0) {
      emax = fmax(emax, finish_vertex<VARIANT>(A, v, s, d, A.r[v], Traits<VARIANT>::prune ? A.cr[v] : 0.0, f));
      cost.add(f, in_deg, d);
    }
  }
  store_costs<VARIANT>(A, cost, P.blocks + blockIdx.x, lds_cost);
  emax = block_max(emax, lds_max);
  if (threadIdx.x == 0) atomic_max_nonneg(&ctl.err[iter_index(ctl)], emax);
}

template <int VARIANT>
hipError_t launch_variant(const RankArgs &A, const RankPlan &P, const LoopCtl &ctl, hipStream_t s, hipEvent_t mark) {
  rank_update_kernel<VARIANT><<<P.blocks, kBlock, 0, s>>>(A, P, ctl);
  if (mark) {
    const hipError_t e = hipEventRecord(mark, s);
    if (e != hipSuccess) return e;
  }
  if (P.n_chunked) finish_chunked_kernel<VARIANT><<<finish_chunked_grid(P.n_chunked), kBlock, 0, s>>>(A, P, P.n_chunked, ctl);
  return hipGetLastError();
}

// ---- sliced plan ----

// Gather launch: workgroup b works on slice x = b % slices (its XCD with 8 slices) and on block b / slices of that
// slice's section table. Items are (row, slice) runs of sources; each leaves one partial sum.
__global__ __launch_bounds__(kBlock) void sliced_gather_kernel(SlicedPlan P, const double *cr, LoopCtl ctl) {
  if (!iteration_begin(ctl)) return;
  __shared__ double lds_sum[kWavesPerBlock];
  const uint32_t x = blockIdx.x & uint32_t(P.slices - 1);
  const uint32_t j = blockIdx.x / uint32_t(P.slices);
  const int ns = P.nsec[x];
  if (ns == 0) return;
  int s = 0;
  while (s + 1 < ns && j >= P.sec[x][s + 1].block_begin) ++s;
  const RankSection &S = P.sec[x][s];
  const uint64_t first = uint64_t(j - S.block_begin) * S.per_block;
  if (j < S.block_begin || first >= uint64_t(S.end - S.begin)) return;  // padding block of a shorter slice
  const uint32_t ib = S.begin + uint32_t(first), ie = min(S.end, ib + S.per_block);
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  if (S.kind == kChunkedRows) {
    for (uint32_t i = ib; i < ie; ++i) {  // block-uniform
      const uint32_t c = P.chunk_list[i];
      const double t = block_gather_sum(P.src, cr, P.chunk_beg[c], P.chunk_end[c], lds_sum);
      if (threadIdx.x == 0) P.chunk_partial[c] = t;
    }
  } else if (S.kind == kWaveRows) {
    for (uint32_t k = ib + w; k < ie; k += kWavesPerBlock) {  // wave-uniform
      const double t = group_sum<kWave>(
          gather_sum<kWave>(P.src, cr, stream_u32(P.item_off + k) + lane, stream_u32(P.item_off + k + 1)));
      if (lane == 0) P.partial[k] = t;
    }
  } else {
    auto items = [&](auto width) {
      constexpr int W = decltype(width)::value, G = kWave / W;
      const int g = lane / W, q = lane % W;
      for (uint32_t k0 = ib + w * G; k0 < ie; k0 += kWavesPerBlock * G) {  // wave-uniform
        const uint32_t k = k0 + g;
        double acc = 0.0;
        if (k < ie) acc = gather_sum<W>(P.src, cr, stream_u32(P.item_off + k) + q, stream_u32(P.item_off + k + 1));
        acc = group_sum<W>(acc);
        if (k < ie && q == 0) P.partial[k] = acc;
      }
    };
    switch (S.width) {
      case 1: items(std::integral_constant<int, 1>{}); break;
      case 2: items(std::integral_constant<int, 2>{}); break;
      case 4: items(std::integral_constant<int, 4>{}); break;
      case 8: items(std::integral_constant<int, 8>{}); break;
      case 16: items(std::integral_constant<int, 16>{}); break;
      default: items(std::integral_constant<int, 32>{}); break;
    }
  }
}

// Finish launch: the first `huge_blocks` workgroups finish the chunked rows (a wave per row over its chunk
// partials); the rest take one row per lane from row n_huge on, adding the row's slice partials in slice order.
// A lane finds its partial in slice x from the 64-row group's first item id plus the number of lanes below it
// whose rows also have an item in slice x (ballot + mbcnt). Every lane loads one value per slice, so that the S
// loads are in flight together: a lane without an item in slice x reads slot 0 (always allocated) and drops it.
template <int S>
__global__ __launch_bounds__(kBlock) void sliced_finish_kernel(RankArgs A, SlicedPlan P, uint32_t huge_blocks,
                                                               LoopCtl ctl) {
  if (!iteration_begin(ctl)) return;
  __shared__ double lds_max[kWavesPerBlock];
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  double emax = 0.0;
  if (blockIdx.x < huge_blocks) {
    const uint32_t v = blockIdx.x * kWavesPerBlock + w;  // wave-uniform
    if (v < P.n_huge) {
      double t = 0.0;
      for (uint32_t k = P.row_chunk_first[v] + lane; k < P.row_chunk_first[v + 1]; k += kWave) t += P.chunk_partial[k];
      t = group_sum<kWave>(t);
      if (lane == 0) emax = finish_vertex<kPlain>(A, v, t);
    }
  } else {
    const uint32_t rows = P.N - P.n_huge;
    const uint32_t stride = (gridDim.x - huge_blocks) * kBlock;
    // Grid-stride over rows from n_huge, in 256-row tiles (a bounded grid: short workgroups by the ten thousand
    // would be limited by workgroup dispatch). Waves stay aligned to 64-row groups.
    for (uint32_t i0 = (blockIdx.x - huge_blocks) * kBlock; i0 < rows; i0 += stride) {
      const uint32_t i = i0 + threadIdx.x;  // row offset from n_huge
      const uint32_t v = P.n_huge + i;
      const bool valid = v < P.N;
      const uint32_t vv = valid ? v : P.N - 1;
      const unsigned m = valid ? P.mask[vv] : 0u;
      const uint32_t d = A.deg[vv];
      const double rv = A.r[vv];
      const uint32_t grp = __builtin_amdgcn_readfirstlane(i / kWave);
      // The wave's 64-row group exists when its first row does; otherwise every m is 0 and no base is needed.
      const bool group_ok = P.n_huge + grp * kWave < P.N;
      const uint32_t *wb = P.wbase + size_t(group_ok ? grp : 0) * S;
      uint32_t base[S];
#pragma unroll
      for (int x = 0; x < S; ++x) base[x] = wb[x];
      double part[S];
#pragma unroll
      for (int x = 0; x < S; ++x) {
        const bool has = (m >> x) & 1u;
        const uint32_t k = has ? base[x] + lanes_below(__ballot(has)) : 0u;
        const double t = P.partial[k];
        part[x] = has ? t : 0.0;
      }
      double cv = 0.0;
#pragma unroll
      for (int x = 0; x < S; ++x) cv += part[x];  // slice order; absent slices add +0.0
      if (valid) emax = fmax(emax, finish_vertex<kPlain>(A, v, cv, d, rv, 0.0));
    }
  }
  emax = block_max(emax, lds_max);
  if (threadIdx.x == 0) atomic_max_nonneg(&ctl.err[iter_index(ctl)], emax);
}

// r = a = 1/N (or keep r as loaded and copy it to a); contributions cr = ca = r/deg.
// pagerankInitializeRanksCukW (pagerankCuda.hxx:232-239) and the warm-start copy (:817).
__global__ __launch_bounds__(kBlock) void init_ranks_kernel(double *r, double *a, double *cr, double *ca,
                                                            const uint32_t *deg, uint32_t N, int from_r) {
  for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < N; v += gridDim.x * kBlock) {
    const double rv = from_r ? r[v] : 1.0 / N;
    if (!from_r) r[v] = rv;
    a[v] = rv;
    if (cr) {  // null: the plan computes contributions from the ranks (propagation blocking) and keeps none
      const double c = rv / deg[v];
      cr[v] = c;
      ca[v] = c;
    }
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void fill_kernel(T *p, T v, size_t n) {
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += size_t(gridDim.x) * kBlock) p[i] = v;
}

__global__ __launch_bounds__(kBlock) void copy_f64_kernel(double *dst, const double *src, size_t n) {
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += size_t(gridDim.x) * kBlock) dst[i] = src[i];
}

// Graph checks; flag bits: 1 offsets not monotone / not 0..M, 2 out-degree 0 (dead end), 4 out-degree disagrees
// with the out-CSR, 8 in-CSR id out of range, 16 out-CSR id out of range, 32 out-CSR offsets not 0..M.
__global__ __launch_bounds__(kBlock) void validate_graph_kernel(const uint32_t *in_off, const uint32_t *in_src,
                                                                const uint32_t *outdeg, const uint32_t *out_off,
                                                                const uint32_t *out_dst, uint32_t N, uint64_t M,
                                                                int *flags) {
  int bits = 0;
  const size_t tid = size_t(blockIdx.x) * kBlock + threadIdx.x, stride = size_t(gridDim.x) * kBlock;
  if (tid == 0) {
    if (in_off[0] != 0 || in_off[N] != M) bits |= 1;
    if (out_off && (out_off[0] != 0 || out_off[N] != M)) bits |= 32;
  }
  for (size_t v = tid; v < N; v += stride) {
    if (in_off[v + 1] < in_off[v]) bits |= 1;
    if (outdeg[v] == 0) bits |= 2;
    if (out_off) {
      if (out_off[v + 1] < out_off[v]) bits |= 32;
      else if (out_off[v + 1] - out_off[v] != outdeg[v]) bits |= 4;
    }
  }
  for (size_t e = tid; e < M; e += stride) {
    if (in_src[e] >= N) bits |= 8;
    if (out_dst && out_dst[e] >= N) bits |= 16;
  }
  if (bits) atomicOr(flags, bits);
}

__global__ __launch_bounds__(kBlock) void validate_ids_kernel(const uint32_t *ids, size_t n, uint32_t N, int *flags,
                                                              int bit) {
  bool bad = false;
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += size_t(gridDim.x) * kBlock)
    bad |= ids[i] >= N;
  if (bad) atomicOr(flags, bit);
}

__global__ __launch_bounds__(kBlock) void check_finite_kernel(const double *x, size_t n, int *flags, int bit) {
  bool bad = false;
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += size_t(gridDim.x) * kBlock)
    bad |= !isfinite(x[i]);
  if (bad) atomicOr(flags, bit);
}

__global__ __launch_bounds__(kBlock) void gather_ranges_kernel(const uint32_t *off, const uint32_t *ids, size_t n,
                                                               uint32_t *beg, uint32_t *end) {
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += size_t(gridDim.x) * kBlock) {
    const uint32_t v = ids[i];
    beg[i] = off[v];
    end[i] = off[v + 1];
  }
}

inline uint32_t map_grid(size_t n) {
  const size_t g = (n + kBlock - 1) / kBlock;
  return uint32_t(g < 1 ? 1 : (g > 8192 ? 8192 : g));
}

}  // namespace

hipError_t launch_rank_update(int variant, const RankArgs &A, const RankPlan &P, unsigned long long *err, int *done,
                              int l, double tol, hipStream_t s, hipEvent_t mark) {
  if (P.blocks == 0) return mark ? hipEventRecord(mark, s) : hipSuccess;
  const LoopCtl ctl{err, done, l, tol};
  switch (variant) {
    case kPlain: return launch_variant<kPlain>(A, P, ctl, s, mark);
    case kTraversal: return launch_variant<kTraversal>(A, P, ctl, s, mark);
    case kFrontier: return launch_variant<kFrontier>(A, P, ctl, s, mark);
    case kPrune: return launch_variant<kPrune>(A, P, ctl, s, mark);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_sliced_update(const RankArgs &A, const SlicedPlan &P, unsigned long long *err, int *done, int l,
                                double tol, hipStream_t s, hipEvent_t mark) {
  const LoopCtl ctl{err, done, l, tol};
  const uint32_t gather_blocks = P.blocks_per_slice * uint32_t(P.slices);
  if (gather_blocks) sliced_gather_kernel<<<gather_blocks, kBlock, 0, s>>>(P, A.cr, ctl);
  if (mark) {
    const hipError_t e = hipEventRecord(mark, s);
    if (e != hipSuccess) return e;
  }
  const uint32_t huge_blocks = (P.n_huge + kWavesPerBlock - 1) / kWavesPerBlock;
  const uint32_t row_blocks = std::min<uint32_t>((P.N - P.n_huge + kBlock - 1) / kBlock, 4096);
  const uint32_t g = huge_blocks + row_blocks;
  if (g) {
    switch (P.slices) {
      case 1: sliced_finish_kernel<1><<<g, kBlock, 0, s>>>(A, P, huge_blocks, ctl); break;
      case 2: sliced_finish_kernel<2><<<g, kBlock, 0, s>>>(A, P, huge_blocks, ctl); break;
      case 4: sliced_finish_kernel<4><<<g, kBlock, 0, s>>>(A, P, huge_blocks, ctl); break;
      case 8: sliced_finish_kernel<8><<<g, kBlock, 0, s>>>(A, P, huge_blocks, ctl); break;
      default: return hipErrorInvalidValue;
    }
  }
  return hipGetLastError();
}

hipError_t launch_init_ranks(double *r, double *a, double *cr, double *ca, const uint32_t *deg, uint32_t N,
                             bool from_r, hipStream_t s) {
  if (N == 0) return hipSuccess;
  init_ranks_kernel<<<map_grid(N), kBlock, 0, s>>>(r, a, cr, ca, deg, N, from_r ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_fill_u8(uint8_t *p, uint8_t v, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  fill_kernel<uint8_t><<<map_grid(n), kBlock, 0, s>>>(p, v, n);
  return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  fill_kernel<uint32_t><<<map_grid(n), kBlock, 0, s>>>(p, v, n);
  return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long *p, unsigned long long v, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  fill_kernel<unsigned long long><<<map_grid(n), kBlock, 0, s>>>(p, v, n);
  return hipGetLastError();
}

hipError_t launch_fill_i32(int *p, int v, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  fill_kernel<int><<<map_grid(n), kBlock, 0, s>>>(p, v, n);
  return hipGetLastError();
}

hipError_t launch_copy_f64(double *dst, const double *src, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  copy_f64_kernel<<<map_grid(n), kBlock, 0, s>>>(dst, src, n);
  return hipGetLastError();
}

hipError_t launch_validate_graph(const uint32_t *in_off, const uint32_t *in_src, const uint32_t *outdeg,
                                 const uint32_t *out_off, const uint32_t *out_dst, uint32_t N, uint64_t M,
                                 int *flags, hipStream_t s) {
  validate_graph_kernel<<<map_grid(M > N ? M : N), kBlock, 0, s>>>(in_off, in_src, outdeg, out_off, out_dst, N, M,
                                                                   flags);
  return hipGetLastError();
}

hipError_t launch_validate_ids(const uint32_t *ids, size_t n, uint32_t N, int *flags, int bit, hipStream_t s) {
  if (n == 0) return hipSuccess;
  validate_ids_kernel<<<map_grid(n), kBlock, 0, s>>>(ids, n, N, flags, bit);
  return hipGetLastError();
}

hipError_t launch_gather_ranges(const uint32_t *off, const uint32_t *ids, size_t n, uint32_t *beg, uint32_t *end,
                                hipStream_t s) {
  if (n == 0) return hipSuccess;
  gather_ranges_kernel<<<map_grid(n), kBlock, 0, s>>>(off, ids, n, beg, end);
  return hipGetLastError();
}

hipError_t launch_check_finite(const double *x, size_t n, int *flags, int bit, hipStream_t s) {
  if (n == 0) return hipSuccess;
  check_finite_kernel<<<map_grid(n), kBlock, 0, s>>>(x, n, flags, bit);
  return hipGetLastError();
}

}  // namespace prh
