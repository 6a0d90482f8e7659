// ppr.hip -- batched personalized PageRank: up to 64 teleport columns solved together.
//
// Column j: r_j <- (1-P) s_j + P * sum_{u in in(v)} r_j[u]/deg[u], from r_j = s_j, stopping at the first iteration
// whose L-inf change of column j is below the tolerance (the rule of pagerankLoopCuU, pagerankCuda.hxx:680-690,
// applied per column). Layout: ranks and contributions are [V][64] FP64 rows, so a wave handles one vertex with
// lane j = column j and every gathered neighbour is one fully used 512-byte row (four 128-byte lines).
// Converged columns are masked: their lanes issue no loads and write nothing, and once every column has
// converged the kernels exit at entry. Each column's result stays in the buffer its last iteration wrote.
//
// In-degree is split at a chunk size: a light vertex is summed by one wave (ascending edge order, one
// accumulator, as a single-column solve) and finished at once; a heavy vertex is summed in fixed chunks by many
// waves into a partial-sum table, and a second kernel adds its chunks in order and finishes it (no atomics on
// ranks). The iteration index lives on the device (PprState::iter) and every iteration issues identical launches,
// so a run of iterations can be captured once in a HIP graph and replayed.
#include "prh_device.hpp"
#include "prh_internal.hpp"

namespace prh {
namespace {

constexpr int kCols = kPprCols;

// Per-column activity at iteration l; with `record`, block 0 stores l as the iteration count of the columns that
// converged at l-1. Returns the wave's active lane mask (lane j = column j). Uniform over the grid.
__device__ __forceinline__ unsigned long long ppr_active(const PprArgs &A, int l, bool record) {
  const int j = threadIdx.x & (kWave - 1);
  bool a = false;
  if (j < A.k && l < A.max_iter) {
    a = *(volatile int *)&A.state->done[j] == 0;
    if (a && l > 0 && __longlong_as_double((long long)A.err[size_t(l - 1) * kCols + j]) < A.tol) {
      a = false;
      if (record && blockIdx.x == 0 && threadIdx.x < kWave) *(volatile int *)&A.state->done[j] = l;
    }
  }
  return __ballot(a);
}

// Sum of contrib[src[i]][j] over i in [b, e); the edge index is wave-uniform, four rows in flight.
// ORDERED keeps a single accumulator in ascending edge order.
template <bool ORDERED>
__device__ __forceinline__ double ppr_sum(const PprArgs &A, const double *c, uint32_t b, uint32_t e) {
  const int j = threadIdx.x & (kWave - 1);
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  uint32_t i = b;
  for (; i + 3 < e; i += 4) {
    const uint32_t u0 = A.src[i], u1 = A.src[i + 1], u2 = A.src[i + 2], u3 = A.src[i + 3];
    const double x0 = c[size_t(u0) * kCols + j], x1 = c[size_t(u1) * kCols + j];
    const double x2 = c[size_t(u2) * kCols + j], x3 = c[size_t(u3) * kCols + j];
    if constexpr (ORDERED) {
      s0 += x0;
      s0 += x1;
      s0 += x2;
      s0 += x3;
    } else {
      s0 += x0;
      s1 += x1;
      s2 += x2;
      s3 += x3;
    }
  }
  for (; i < e; ++i) s0 += c[size_t(A.src[i]) * kCols + j];
  return ORDERED ? s0 : (s0 + s1) + (s2 + s3);
}

// a = (1-P) s + P cv, contribution a/deg, and the column's change |a - r|.
__device__ __forceinline__ double ppr_finish(const PprArgs &A, int l, uint32_t v, double cv) {
  const int j = threadIdx.x & (kWave - 1);
  const double *r = A.rank[l & 1];
  double *a = A.rank[(l + 1) & 1];
  double *ca = A.contrib[(l + 1) & 1];
  const int32_t row = A.seed_row[v];
  const double s = row >= 0 ? A.seed_mass[size_t(row) * kCols + j] : 0.0;
  const size_t x = size_t(v) * kCols + j;
  const double av = fma(A.P, cv, (1 - A.P) * s);
  a[x] = av;
  ca[x] = av / A.deg[v];
  return fabs(av - r[x]);
}

__device__ __forceinline__ void ppr_reduce_err(const PprArgs &A, int l, double e, double *lds) {
  const int j = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  lds[w * kCols + j] = e;
  __syncthreads();
  if (threadIdx.x < kCols) {
    const double m = fmax(fmax(lds[j], lds[kCols + j]), fmax(lds[2 * kCols + j], lds[3 * kCols + j]));
    if (j < A.k) atomic_max_nonneg(&A.err[size_t(l) * kCols + j], m);
  }
}

// Light vertices (finished here) and heavy-vertex chunks (partial sums), one wave per item.
__global__ __launch_bounds__(kBlock) void ppr_gather_kernel(PprArgs A, uint32_t nb_chunks) {
  __shared__ double lds[kWavesPerBlock * kCols];
  const int l = *(volatile int *)&A.state->iter;
  const unsigned long long mask = ppr_active(A, l, true);
  if (!mask) return;
  const int j = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const bool on = (mask >> j) & 1;
  const double *c = A.contrib[l & 1];
  double emax = 0.0;
  if (blockIdx.x < nb_chunks) {
    for (uint32_t i = blockIdx.x * kWavesPerBlock + w; i < A.n_chunks; i += nb_chunks * kWavesPerBlock) {
      if (!on) continue;
      A.partial[size_t(i) * kCols + j] = ppr_sum<false>(A, c, A.chunk_beg[i], A.chunk_end[i]);
    }
  } else {
    const uint32_t nb = gridDim.x - nb_chunks;
    for (uint32_t i = (blockIdx.x - nb_chunks) * kWavesPerBlock + w; i < A.n_light; i += nb * kWavesPerBlock) {
      if (!on) continue;
      const uint32_t v = A.light[i];
      emax = fmax(emax, ppr_finish(A, l, v, ppr_sum<true>(A, c, A.off[v], A.off[v + 1])));
    }
  }
  ppr_reduce_err(A, l, emax, lds);
}

// Heavy vertices: add the chunk partials in chunk order and finish.
__global__ __launch_bounds__(kBlock) void ppr_combine_kernel(PprArgs A) {
  __shared__ double lds[kWavesPerBlock * kCols];
  const int l = *(volatile int *)&A.state->iter;
  const unsigned long long mask = ppr_active(A, l, false);
  if (!mask) return;
  const int j = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const bool on = (mask >> j) & 1;
  double emax = 0.0;
  for (uint32_t h = blockIdx.x * kWavesPerBlock + w; h < A.n_heavy; h += gridDim.x * kWavesPerBlock) {
    if (!on) continue;
    double s = 0.0;
    for (uint32_t i = A.heavy_first[h]; i < A.heavy_first[h + 1]; ++i) s += A.partial[size_t(i) * kCols + j];
    emax = fmax(emax, ppr_finish(A, l, A.heavy_v[h], s));
  }
  ppr_reduce_err(A, l, emax, lds);
}

__global__ void ppr_advance_kernel(PprState *S) {
  if (threadIdx.x == 0) {
    volatile int *it = &S->iter;
    *it = *it + 1;
  }
}

__global__ __launch_bounds__(kBlock) void ppr_reset_kernel(PprArgs A) {
  const size_t n = size_t(A.max_iter) * kCols;
  for (size_t x = size_t(blockIdx.x) * kBlock + threadIdx.x; x < n; x += size_t(gridDim.x) * kBlock) A.err[x] = 0ull;
  if (blockIdx.x == 0) {
    if (threadIdx.x < kCols) A.state->done[threadIdx.x] = 0;
    if (threadIdx.x == 0) A.state->iter = 0;
  }
}

// r0 = s, c0 = s/deg for all columns (zero outside the seeds and for unused columns).
__global__ __launch_bounds__(kBlock) void ppr_init_kernel(PprArgs A) {
  const size_t n = size_t(A.N) * kCols;
  for (size_t x = size_t(blockIdx.x) * kBlock + threadIdx.x; x < n; x += size_t(gridDim.x) * kBlock) {
    const uint32_t v = uint32_t(x / kCols);
    const int j = int(x % kCols);
    const int32_t row = A.seed_row[v];
    const double s = (row >= 0 && j < A.k) ? A.seed_mass[size_t(row) * kCols + j] : 0.0;
    A.rank[0][x] = s;
    A.contrib[0][x] = s / A.deg[v];
  }
}

// The ranks of column j are in rank[n_j & 1] (n_j its iteration count); gather them into out[v*k + j].
__global__ __launch_bounds__(kBlock) void ppr_collect_kernel(PprArgs A, double *out) {
  const size_t n = size_t(A.N) * A.k;
  for (size_t x = size_t(blockIdx.x) * kBlock + threadIdx.x; x < n; x += size_t(gridDim.x) * kBlock) {
    const uint32_t v = uint32_t(x / A.k);
    const int j = int(x % A.k);
    int d = A.state->done[j];
    if (d == 0) d = A.max_iter;
    out[x] = A.rank[d & 1][size_t(v) * kCols + j];
  }
}

inline uint32_t grid_for(size_t n, size_t per, uint32_t cap) {
  const size_t g = (n + per - 1) / per;
  return uint32_t(g < 1 ? 1 : (g > cap ? cap : g));
}

}  // namespace

hipError_t launch_ppr_reset(const PprArgs &A, hipStream_t s) {
  ppr_reset_kernel<<<grid_for(size_t(A.max_iter) * kCols, kBlock, 1024), kBlock, 0, s>>>(A);
  return hipGetLastError();
}

hipError_t launch_ppr_init(const PprArgs &A, hipStream_t s) {
  ppr_init_kernel<<<grid_for(size_t(A.N) * kCols, kBlock, 16384), kBlock, 0, s>>>(A);
  return hipGetLastError();
}

hipError_t launch_ppr_iteration(const PprArgs &A, hipStream_t s) {
  const uint32_t nb_chunks = A.n_chunks ? grid_for(A.n_chunks, kWavesPerBlock, 8192) : 0;
  const uint32_t nb_light = A.n_light ? grid_for(A.n_light, kWavesPerBlock, 16384) : 0;
  if (nb_chunks + nb_light) ppr_gather_kernel<<<nb_chunks + nb_light, kBlock, 0, s>>>(A, nb_chunks);
  if (A.n_heavy) ppr_combine_kernel<<<grid_for(A.n_heavy, kWavesPerBlock, 4096), kBlock, 0, s>>>(A);
  ppr_advance_kernel<<<1, 64, 0, s>>>(A.state);
  return hipGetLastError();
}

hipError_t launch_ppr_collect(const PprArgs &A, double *out, hipStream_t s) {
  ppr_collect_kernel<<<grid_for(size_t(A.N) * A.k, kBlock, 16384), kBlock, 0, s>>>(A, out);
  return hipGetLastError();
}

}  // namespace prh
