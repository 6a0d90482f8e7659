// rank.hip -- the pull-based, atomic-free rank update with the fused L-inf reduction, plus small map kernels.
//
// One launch per iteration covers all vertices of the degree-ordered layout (rows sorted by in-degree,
// descending; layout.hip) through a table of sections, heaviest first so that long rows start early:
//   chunked rows  (in-degree above the chunk threshold): one workgroup per fixed-size edge chunk; each chunk's
//                 sum goes to a partial slot and the workgroup that arrives last for the row adds the row's
//                 partials in chunk order and finishes it (deterministic; the arrival count only picks who adds);
//   wave rows     one wave per row, DPP/swizzle wave sum;
//   group rows    W lanes per row (W = 1..32), one band per W over a contiguous row range, DPP group sum.
//
// Per vertex v (pagerankUpdateRanks{Thread,Block}CukU, pagerankCuda.hxx:87-109, :162-188):
//   cv = sum_{u in in(v)} r[u]/deg[u]     -- gathered from the precomputed contributions cr[u] = r[u]/deg[u],
//                                            the same FP64 value the reference divides per edge
//   a[v] = C0 + P*cv                      -- or the DF-P closed form (C0 + P*(cv - r[v]/d))/(1 - P/d)
//   prune / frontier flags                 -- vaff[v] = 0 if |a-r|/max(r,a) <= C; naff[v] = (|a-r|/max > D)
// Each vertex has exactly one writer, so the update needs no atomics on ranks; the atomics are one 64-bit max
// per workgroup for the L-inf change and one arrival count per chunk of a chunked row.
#include "prh_device.hpp"
#include "prh_internal.hpp"

namespace prh {
namespace {

template <int VARIANT>
struct Traits {
  static constexpr bool dynamic = VARIANT >= kTraversal;
  static constexpr bool frontier = VARIANT >= kFrontier;
  static constexpr bool prune = VARIANT == kPrune;
};

template <int VARIANT>
__device__ __forceinline__ double finish_vertex(const RankArgs &A, uint32_t v, double cv) {
  using T = Traits<VARIANT>;
  const uint32_t d = A.deg[v];
  const double rv = A.r[v];
  double av;
  if constexpr (T::prune) av = fma(A.P, cv - A.cr[v], A.C0) * (1.0 / (1.0 - A.P / d));
  else av = fma(A.P, cv, A.C0);
  A.a[v] = av;
  A.ca[v] = av / d;
  const double ev = fabs(av - rv);
  if constexpr (T::frontier) {
    const double rel = ev / fmax(rv, av);
    if (T::prune && rel <= A.C) A.vaff[v] = 0;
    A.naff[v] = rel > A.D ? 1 : 0;
  }
  return ev;
}

template <int VARIANT>
__device__ __forceinline__ void copy_vertex(const RankArgs &A, uint32_t v) {
  A.a[v] = A.r[v];
  A.ca[v] = A.cr[v];
  if constexpr (Traits<VARIANT>::frontier) A.naff[v] = 0;
}

template <int VARIANT>
__device__ __forceinline__ bool affected(const RankArgs &A, uint32_t v) {
  if constexpr (Traits<VARIANT>::dynamic) return A.vaff[v] != 0;
  else return true;
}

// Streamed arrays (source ids, offsets) are read with the non-temporal hint so that they do not displace the
// gathered contributions from L2; the gathers themselves use the default policy.
__device__ __forceinline__ uint32_t stream_u32(const uint32_t *p) { return __builtin_nontemporal_load(p); }

// Sum of cr[src[e]] for e = first, first + STEP, ... < end, four gathers in flight per lane.
template <int STEP>
__device__ __forceinline__ double gather_sum(const RankArgs &A, uint32_t e, uint32_t end) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  for (; e + 3 * STEP < end; e += 4 * STEP) {
    const uint32_t u0 = stream_u32(A.src + e), u1 = stream_u32(A.src + e + STEP);
    const uint32_t u2 = stream_u32(A.src + e + 2 * STEP), u3 = stream_u32(A.src + e + 3 * STEP);
    s0 += A.cr[u0];
    s1 += A.cr[u1];
    s2 += A.cr[u2];
    s3 += A.cr[u3];
  }
  for (; e < end; e += STEP) s0 += A.cr[stream_u32(A.src + e)];
  return (s0 + s1) + (s2 + s3);
}

// One workgroup per chunk item: its partial sum goes to the item's slot; rows are finished by
// finish_chunked_kernel after this launch (the kernel boundary makes every chunk's partial visible).
template <int VARIANT>
__device__ __forceinline__ void chunked_rows(const RankArgs &A, const RankPlan &P, const RankSection &S, uint32_t blk,
                                             double *lds) {
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  const uint32_t ib = S.begin + blk * S.per_block, ie = min(S.end, ib + S.per_block);
  for (uint32_t it = ib; it < ie; ++it) {  // block-uniform
    const uint32_t v = P.item_row[it];
    if (!affected<VARIANT>(A, v)) continue;
    double acc = gather_sum<kBlock>(A, P.item_beg[it] + threadIdx.x, P.item_end[it]);
    acc = group_sum<kWave>(acc);
    if (lane == 0) lds[w] = acc;
    __syncthreads();
    if (threadIdx.x == 0) P.partial[it] = (lds[0] + lds[1]) + (lds[2] + lds[3]);
    __syncthreads();
  }
}

template <int VARIANT>
__device__ __forceinline__ void wave_rows(const RankArgs &A, const RankSection &S, uint32_t blk, double &emax) {
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  const uint32_t rb = S.begin + blk * S.per_block, re = min(S.end, rb + S.per_block);
  for (uint32_t v = rb + w; v < re; v += kWavesPerBlock) {  // wave-uniform
    if (!affected<VARIANT>(A, v)) {
      if (lane == 0) copy_vertex<VARIANT>(A, v);
      continue;
    }
    double acc = gather_sum<kWave>(A, stream_u32(A.off + v) + lane, stream_u32(A.off + v + 1));
    acc = group_sum<kWave>(acc);
    if (lane == 0) emax = fmax(emax, finish_vertex<VARIANT>(A, v, acc));
  }
}

template <int VARIANT, int W>
__device__ __forceinline__ void group_rows(const RankArgs &A, const RankSection &S, uint32_t blk, double &emax) {
  constexpr int G = kWave / W;  // rows per wave per round
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  const int g = lane / W, j = lane % W;
  const uint32_t rb = S.begin + blk * S.per_block, re = min(S.end, rb + S.per_block);
  for (uint32_t r0 = rb + w * G; r0 < re; r0 += kWavesPerBlock * G) {  // wave-uniform
    const uint32_t v = r0 + g;
    const bool valid = v < re;
    bool upd = false;
    double acc = 0.0;
    if (valid) {
      upd = affected<VARIANT>(A, v);
      if (upd) acc = gather_sum<W>(A, stream_u32(A.off + v) + j, stream_u32(A.off + v + 1));
    }
    acc = group_sum<W>(acc);  // all lanes active again here
    if (valid && j == 0) {
      if (upd) emax = fmax(emax, finish_vertex<VARIANT>(A, v, acc));
      else copy_vertex<VARIANT>(A, v);
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
  if (S.kind == kChunkedRows) {
    chunked_rows<VARIANT>(A, P, S, blk, lds_sum);
  } else if (S.kind == kWaveRows) {
    wave_rows<VARIANT>(A, S, blk, emax);
  } else {
    switch (S.width) {
      case 1: group_rows<VARIANT, 1>(A, S, blk, emax); break;
      case 2: group_rows<VARIANT, 2>(A, S, blk, emax); break;
      case 4: group_rows<VARIANT, 4>(A, S, blk, emax); break;
      case 8: group_rows<VARIANT, 8>(A, S, blk, emax); break;
      case 16: group_rows<VARIANT, 16>(A, S, blk, emax); break;
      default: group_rows<VARIANT, 32>(A, S, blk, emax); break;
    }
  }
  emax = block_max(emax, lds_max);
  if (threadIdx.x == 0) atomic_max_nonneg(&ctl.err[ctl.l], emax);
}

// Chunked rows, after the rank kernel of the same iteration: one wave per row adds the row's chunk partials (in a
// fixed order) and finishes the row, or copies it through when unaffected. Runs only if that kernel ran.
template <int VARIANT>
__global__ __launch_bounds__(kBlock) void finish_chunked_kernel(RankArgs A, RankPlan P, uint32_t n_rows, LoopCtl ctl) {
  if (!iteration_begin(ctl)) return;  // same outcome as the rank kernel's entry test of this iteration
  __shared__ double lds_max[kWavesPerBlock];
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  double emax = 0.0;
  for (uint32_t v = blockIdx.x * kWavesPerBlock + w; v < n_rows; v += gridDim.x * kWavesPerBlock) {
    if (!affected<VARIANT>(A, v)) {
      if (lane == 0) copy_vertex<VARIANT>(A, v);
      continue;
    }
    double s = 0.0;
    for (uint32_t k = P.row_first[v] + lane; k < P.row_first[v + 1]; k += kWave) s += P.partial[k];
    s = group_sum<kWave>(s);
    if (lane == 0) emax = fmax(emax, finish_vertex<VARIANT>(A, v, s));
  }
  emax = block_max(emax, lds_max);
  if (threadIdx.x == 0) atomic_max_nonneg(&ctl.err[ctl.l], emax);
}

template <int VARIANT>
hipError_t launch_variant(const RankArgs &A, const RankPlan &P, const LoopCtl &ctl, hipStream_t s, hipEvent_t mark) {
  rank_update_kernel<VARIANT><<<P.blocks, kBlock, 0, s>>>(A, P, ctl);
  if (mark) {
    const hipError_t e = hipEventRecord(mark, s);
    if (e != hipSuccess) return e;
  }
  if (P.n_chunked) {
    const uint32_t g = (P.n_chunked + kWavesPerBlock - 1) / kWavesPerBlock;
    finish_chunked_kernel<VARIANT><<<g > 4096 ? 4096 : g, kBlock, 0, s>>>(A, P, P.n_chunked, ctl);
  }
  return hipGetLastError();
}

// r = a = 1/N (or keep r as loaded and copy it to a); contributions cr = ca = r/deg.
// pagerankInitializeRanksCukW (pagerankCuda.hxx:232-239) and the warm-start copy (:817).
__global__ __launch_bounds__(kBlock) void init_ranks_kernel(double *r, double *a, double *cr, double *ca,
                                                            const uint32_t *deg, uint32_t N, int from_r) {
  for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < N; v += gridDim.x * kBlock) {
    const double rv = from_r ? r[v] : 1.0 / N;
    const double c = rv / deg[v];
    if (!from_r) r[v] = rv;
    a[v] = rv;
    cr[v] = c;
    ca[v] = c;
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
