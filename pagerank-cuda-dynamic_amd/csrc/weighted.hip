// weighted.hip -- weighted PageRank, its adjoint solve and the gradient assembly (no counterpart in the reference).
//
// Transition probabilities P[u,v] = w_uv / W_u with FP32 edge weights w and FP64 out-weight sums W_u.
//   forward  (pull over the in-CSR):  r_v   <- (1-a) s_v + a * sum_{u in in(v)}  w_uv * (r_u / W_u)
//   adjoint  (pull over the out-CSR): lam_u <- g_u + (a / W_u) * sum_{v in out(u)} w_uv * lam_v
// i.e. (I - a P^T) r = (1-a) s and its transpose (I - a P) lam = g, both by the same Jacobi pull with the convergence
// control of the rank update (fused L-inf, device-side stop). The solves run on the degree-ordered layout (its in-CSR
// forward, its out-CSR for the adjoint), with the weights carried to layout slots through layout_edge_map_kernel, so
// that the most-gathered values share cache lines as in the rank update. The gradient of L(r) follows without storing
// per-edge terms: dL/ds = (1-a) lam and dL/dw_uv = (a r_u / W_u) (lam_v - (lam_u - g_u)/a), where (lam_u - g_u)/a =
// sum_v P[u,v] lam_v is recomputed from the adjoint fixed point.
#include "prh_device.hpp"
#include "prh_internal.hpp"

namespace prh {
namespace {

template <bool ADJOINT>
__device__ __forceinline__ double wfinish(const WeightedArgs &A, uint32_t i, double sum) {
  double a;
  if constexpr (ADJOINT) {
    a = fma(A.P / A.W[i], sum, A.term[i]);
    A.z_new[i] = a;
  } else {
    a = fma(A.P, sum, (1 - A.P) * A.term[i]);
    A.z_new[i] = a / A.W[i];
  }
  A.x_new[i] = a;
  return fabs(a - A.x_old[i]);
}

template <int STEP>
__device__ __forceinline__ double wgather(const WeightedArgs &A, uint32_t e, uint32_t end) {
  double s0 = 0.0, s1 = 0.0;
  for (; e + STEP < end; e += 2 * STEP) {
    const uint32_t u0 = A.idx[e], u1 = A.idx[e + STEP];
    const float w0 = A.w[e], w1 = A.w[e + STEP];
    s0 += double(w0) * A.z[u0];
    s1 += double(w1) * A.z[u1];
  }
  for (; e < end; e += STEP) s0 += double(A.w[e]) * A.z[A.idx[e]];
  return s0 + s1;
}

// Rows order[rb .. re) of one band, W lanes per row, 64/W rows per wave per round.
template <bool ADJOINT, int W>
__device__ __forceinline__ void band_rows(const WeightedArgs &A, uint32_t rb, uint32_t re, double &emax) {
  constexpr int G = kWave / W;
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  const int g = lane / W, j = lane % W;
  for (uint32_t i0 = rb + w * G; i0 < re; i0 += kWavesPerBlock * G) {  // wave-uniform
    const uint32_t i = i0 + g;
    double s = 0.0;
    uint32_t v = 0;
    if (i < re) {
      v = A.sched.order[i];
      s = wgather<W>(A, A.off[v] + j, A.off[v + 1]);
    }
    s = group_sum<W>(s);
    if (i < re && j == 0) emax = fmax(emax, wfinish<ADJOINT>(A, v, s));
  }
}

// One Jacobi step. Blocks [0, n_chunks) sum one chunk of a heavy row each (its partial); the next nb_high blocks
// take the rows of degree in (64, kWeightedChunk], a wave per row; the rest the low-row bands. Fused L-inf for the
// rows finished here.
template <bool ADJOINT>
__global__ __launch_bounds__(kBlock) void weighted_step_kernel(WeightedArgs A, uint32_t nb_high, LoopCtl ctl) {
  if (!iteration_begin(ctl)) return;
  __shared__ double lds[kWavesPerBlock];
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  if (blockIdx.x < A.n_chunks) {  // block-uniform
    double s = group_sum<kWave>(wgather<kBlock>(A, A.chunk_beg[blockIdx.x] + threadIdx.x, A.chunk_end[blockIdx.x]));
    if (lane == 0) lds[w] = s;
    __syncthreads();
    if (threadIdx.x == 0) A.partial[blockIdx.x] = (lds[0] + lds[1]) + (lds[2] + lds[3]);
    return;
  }
  const uint32_t b = blockIdx.x - A.n_chunks;
  double emax = 0.0;
  if (b < nb_high) {
    const uint32_t *order = A.sched.order + A.sched.n_low;
    const uint32_t n = A.sched.n_high;
    for (uint32_t i = b * kWavesPerBlock + w; i < n; i += nb_high * kWavesPerBlock) {
      const uint32_t v = order[i];
      double s = wgather<kWave>(A, A.off[v] + lane, A.off[v + 1]);
      s = group_sum<kWave>(s);
      if (lane == 0) emax = fmax(emax, wfinish<ADJOINT>(A, v, s));
    }
  } else {
    const uint32_t lb = b - nb_high;
    int k = 0;
    while (k + 1 < A.n_bands && lb >= A.band[k + 1].block_begin) ++k;
    const RankSection &S = A.band[k];
    const uint32_t rb = S.begin + (lb - S.block_begin) * S.per_block;
    const uint32_t re = min(S.end, rb + S.per_block);
    switch (S.width) {
      case 32: band_rows<ADJOINT, 32>(A, rb, re, emax); break;
      case 16: band_rows<ADJOINT, 16>(A, rb, re, emax); break;
      case 8: band_rows<ADJOINT, 8>(A, rb, re, emax); break;
      case 4: band_rows<ADJOINT, 4>(A, rb, re, emax); break;
      case 2: band_rows<ADJOINT, 2>(A, rb, re, emax); break;
      default: band_rows<ADJOINT, 1>(A, rb, re, emax); break;
    }
  }
  emax = block_max(emax, lds);
  if (threadIdx.x == 0) atomic_max_nonneg(&ctl.err[iter_index(ctl)], emax);
}

// Heavy rows, after the step of the same iteration: a wave per row adds its chunk sums in chunk order.
template <bool ADJOINT>
__global__ __launch_bounds__(kBlock) void weighted_heavy_kernel(WeightedArgs A, LoopCtl ctl) {
  if (!iteration_begin(ctl)) return;
  __shared__ double lds[kWavesPerBlock];
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  const uint32_t *order = A.sched.order + A.sched.n_low + A.sched.n_high;
  double emax = 0.0;
  for (uint32_t h = blockIdx.x * kWavesPerBlock + w; h < A.sched.n_huge; h += gridDim.x * kWavesPerBlock) {
    double s = 0.0;
    for (uint32_t k = A.heavy_first[h] + lane; k < A.heavy_first[h + 1]; k += kWave) s += A.partial[k];
    s = group_sum<kWave>(s);
    if (lane == 0) emax = fmax(emax, wfinish<ADJOINT>(A, order[h], s));
  }
  emax = block_max(emax, lds);
  if (threadIdx.x == 0) atomic_max_nonneg(&ctl.err[iter_index(ctl)], emax);
}

// W_u = sum of u's out-weights (FP64), one wave per row.
__global__ __launch_bounds__(kBlock) void out_weight_sum_kernel(const uint32_t *out_off, const float *w, uint32_t N,
                                                                double *W) {
  const int wv = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  for (uint32_t u = blockIdx.x * kWavesPerBlock + wv; u < N; u += gridDim.x * kWavesPerBlock) {
    double s = 0.0;
    for (uint32_t e = out_off[u] + lane; e < out_off[u + 1]; e += kWave) s += double(w[e]);
    s = group_sum<kWave>(s);
    if (lane == 0) W[u] = s;
  }
}

// For a CSR of the degree-ordered layout (rows and ids in new numbering, perm: new -> old): eid[k] = the raw
// out-CSR slot of the same edge as slot k, by binary search in the source's ascending out-row. transpose: rows are
// destinations (the layout in-CSR), else sources (its out-CSR).
__global__ __launch_bounds__(kBlock) void layout_edge_map_kernel(const uint32_t *l_off, const uint32_t *l_idx,
                                                                 const uint32_t *perm, const uint32_t *out_off,
                                                                 const uint32_t *out_dst, uint32_t N, int transpose,
                                                                 uint32_t *eid, int *missing) {
  const int wv = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  bool bad = false;
  for (uint32_t i = blockIdx.x * kWavesPerBlock + wv; i < N; i += gridDim.x * kWavesPerBlock) {
    const uint32_t a = perm[i];
    for (uint32_t k = l_off[i] + lane; k < l_off[i + 1]; k += kWave) {
      const uint32_t b = perm[l_idx[k]];
      const uint32_t u = transpose ? b : a, v = transpose ? a : b;
      uint32_t lo = out_off[u], hi = out_off[u + 1];
      while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (out_dst[mid] < v) lo = mid + 1;
        else hi = mid;
      }
      const bool found = lo < out_off[u + 1] && out_dst[lo] == v;
      bad |= !found;
      eid[k] = found ? lo : 0u;
    }
  }
  if (bad) atomicOr(missing, 1);
}

__global__ __launch_bounds__(kBlock) void permute_weights_kernel(const float *w_out, const uint32_t *in_eid, size_t M,
                                                                 float *w_in) {
  for (size_t k = size_t(blockIdx.x) * kBlock + threadIdx.x; k < M; k += size_t(gridDim.x) * kBlock)
    w_in[k] = w_out[in_eid[k]];
}

// x = z = start (adjoint: lam0 = g) or x = s, z = s / W (forward: r0 = s).
__global__ __launch_bounds__(kBlock) void weighted_init_kernel(const double *start, const double *W, uint32_t N,
                                                               int adjoint, double *x, double *z) {
  for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < N; v += gridDim.x * kBlock) {
    const double s = start[v];
    x[v] = s;
    z[v] = adjoint ? s : s / W[v];
  }
}

// dL/dw (out-CSR order) and dL/ds, one wave per source row.
__global__ __launch_bounds__(kBlock) void weighted_grad_kernel(const uint32_t *out_off, const uint32_t *out_dst,
                                                               const double *W, const double *r, const double *lam,
                                                               const double *g, double P, uint32_t N,
                                                               double *grad_w, double *grad_s) {
  const int wv = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  for (uint32_t u = blockIdx.x * kWavesPerBlock + wv; u < N; u += gridDim.x * kWavesPerBlock) {
    const double lu = lam[u];
    const double su = (lu - g[u]) / P;
    const double fu = P * r[u] / W[u];
    for (uint32_t e = out_off[u] + lane; e < out_off[u + 1]; e += kWave) grad_w[e] = fu * (lam[out_dst[e]] - su);
    if (lane == 0) grad_s[u] = (1 - P) * lu;
  }
}

inline uint32_t blocks(size_t n, size_t per, uint32_t cap) {
  const size_t g = (n + per - 1) / per;
  return uint32_t(g < 1 ? 1 : (g > cap ? cap : g));
}

// key[i] = 64 - degree of row order[i] (low rows only: degree <= 64), for the degree sort of the low rows.
__global__ __launch_bounds__(kBlock) void degree_keys_kernel(const uint32_t *off, const uint32_t *order, uint32_t n,
                                                             uint32_t *key) {
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const uint32_t v = order[i];
    key[i] = 64u - (off[v + 1] - off[v]);
  }
}

template <bool ADJOINT>
hipError_t launch_step(const WeightedArgs &A, const LoopCtl &ctl, hipStream_t s) {
  const uint32_t nb_high = A.sched.n_high ? blocks(A.sched.n_high, kWavesPerBlock, 16384) : 0;
  uint32_t nb_low = 0;
  for (int k = 0; k < A.n_bands; ++k)
    nb_low = max(nb_low, A.band[k].block_begin +
                             (A.band[k].end - A.band[k].begin + A.band[k].per_block - 1) / A.band[k].per_block);
  const uint32_t grid = A.n_chunks + nb_high + nb_low;
  if (grid) weighted_step_kernel<ADJOINT><<<grid, kBlock, 0, s>>>(A, nb_high, ctl);
  if (A.sched.n_huge)
    weighted_heavy_kernel<ADJOINT><<<blocks(A.sched.n_huge, kWavesPerBlock, 4096), kBlock, 0, s>>>(A, ctl);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_weighted_step(bool adjoint, const WeightedArgs &A, unsigned long long *err, int *done, int l,
                                double tol, hipStream_t s) {
  const LoopCtl ctl{err, done, l, tol};
  return adjoint ? launch_step<true>(A, ctl, s) : launch_step<false>(A, ctl, s);
}

hipError_t launch_degree_keys(const uint32_t *off, const uint32_t *order, uint32_t n, uint32_t *key, hipStream_t s) {
  if (!n) return hipSuccess;
  degree_keys_kernel<<<blocks(n, kBlock, 8192), kBlock, 0, s>>>(off, order, n, key);
  return hipGetLastError();
}

hipError_t launch_out_weight_sum(const uint32_t *out_off, const float *w, uint32_t N, double *W, hipStream_t s) {
  if (!N) return hipSuccess;
  out_weight_sum_kernel<<<blocks(N, kWavesPerBlock, 16384), kBlock, 0, s>>>(out_off, w, N, W);
  return hipGetLastError();
}

hipError_t launch_layout_edge_map(const uint32_t *l_off, const uint32_t *l_idx, const uint32_t *perm,
                                  const uint32_t *out_off, const uint32_t *out_dst, uint32_t N, bool transpose,
                                  uint32_t *eid, int *missing, hipStream_t s) {
  if (!N) return hipSuccess;
  layout_edge_map_kernel<<<blocks(N, kWavesPerBlock, 16384), kBlock, 0, s>>>(l_off, l_idx, perm, out_off, out_dst, N,
                                                                              transpose ? 1 : 0, eid, missing);
  return hipGetLastError();
}

hipError_t launch_permute_weights(const float *w_out, const uint32_t *in_eid, size_t M, float *w_in, hipStream_t s) {
  if (!M) return hipSuccess;
  permute_weights_kernel<<<blocks(M, kBlock, 16384), kBlock, 0, s>>>(w_out, in_eid, M, w_in);
  return hipGetLastError();
}

hipError_t launch_weighted_init(const double *start, const double *W, uint32_t N, bool adjoint, double *x, double *z,
                                hipStream_t s) {
  if (!N) return hipSuccess;
  weighted_init_kernel<<<blocks(N, kBlock, 8192), kBlock, 0, s>>>(start, W, N, adjoint ? 1 : 0, x, z);
  return hipGetLastError();
}

hipError_t launch_weighted_grad(const uint32_t *out_off, const uint32_t *out_dst, const double *W, const double *r,
                                const double *lam, const double *g, double P, uint32_t N, double *grad_w,
                                double *grad_s, hipStream_t s) {
  if (!N) return hipSuccess;
  weighted_grad_kernel<<<blocks(N, kWavesPerBlock, 16384), kBlock, 0, s>>>(out_off, out_dst, W, r, lam, g, P, N,
                                                                           grad_w, grad_s);
  return hipGetLastError();
}

}  // namespace prh
