// weighted.hip -- weighted PageRank, its adjoint solve and the gradient assembly (no counterpart in the reference).
//
// Transition probabilities P[u,v] = w_uv / W_u with FP32 edge weights w and FP64 out-weight sums W_u.
//   forward  (pull over the in-CSR):  r_v   <- (1-a) s_v + a * sum_{u in in(v)}  w_uv * (r_u / W_u)
//   adjoint  (pull over the out-CSR): lam_u <- g_u + (a / W_u) * sum_{v in out(u)} w_uv * lam_v
// i.e. (I - a P^T) r = (1-a) s and its transpose (I - a P) lam = g, both by the same Jacobi pull with the
// convergence control of the rank update (fused L-inf, device-side stop). The solves run on the degree-ordered
// layout (its in-CSR forward, its out-CSR for the adjoint), with the weights carried to layout slots through
// layout_edge_map_kernel, so that the most-gathered values share cache lines as in the rank update. The gradient of L(r) follows without
// storing per-edge terms: dL/ds = (1-a) lam and dL/dw_uv = (a r_u / W_u) (lam_v - (lam_u - g_u)/a), where
// (lam_u - g_u)/a = sum_v P[u,v] lam_v is recomputed from the adjoint fixed point.
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

// One Jacobi step: low-degree rows with W lanes each, the rest one wave per row; fused L-inf.
template <bool ADJOINT, int W>
__global__ __launch_bounds__(kBlock) void weighted_step_kernel(WeightedArgs A, uint32_t nb_high, LoopCtl ctl) {
  if (!iteration_begin(ctl)) return;
  __shared__ double lds[kWavesPerBlock];
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  double emax = 0.0;
  if (blockIdx.x < nb_high) {
    const uint32_t *order = A.sched.order + A.sched.n_low;
    const uint32_t n = A.sched.n_high + A.sched.n_huge;
    for (uint32_t i = blockIdx.x * kWavesPerBlock + w; i < n; i += nb_high * kWavesPerBlock) {
      const uint32_t v = order[i];
      double s = wgather<kWave>(A, A.off[v] + lane, A.off[v + 1]);
      s = group_sum<kWave>(s);
      if (lane == 0) emax = fmax(emax, wfinish<ADJOINT>(A, v, s));
    }
  } else {
    constexpr int G = kWave / W;
    const int g = lane / W, j = lane % W;
    const uint32_t nb = gridDim.x - nb_high, n = A.sched.n_low;
    for (uint32_t i0 = ((blockIdx.x - nb_high) * kWavesPerBlock + w) * G; i0 < n; i0 += nb * kWavesPerBlock * G) {
      const uint32_t i = i0 + g;
      double s = 0.0;
      uint32_t v = 0;
      if (i < n) {
        v = A.sched.order[i];
        s = wgather<W>(A, A.off[v] + j, A.off[v + 1]);
      }
      s = group_sum<W>(s);
      if (i < n && j == 0) emax = fmax(emax, wfinish<ADJOINT>(A, v, s));
    }
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

// in_eid[k] = out-CSR position of in-CSR slot k: for row v and source u = in_src[k], the index of v in u's
// ascending out-row (binary search). One wave per in-row.
__global__ __launch_bounds__(kBlock) void edge_map_kernel(const uint32_t *in_off, const uint32_t *in_src,
                                                          const uint32_t *out_off, const uint32_t *out_dst, uint32_t N,
                                                          uint32_t *in_eid, int *missing) {
  const int wv = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  bool bad = false;
  for (uint32_t v = blockIdx.x * kWavesPerBlock + wv; v < N; v += gridDim.x * kWavesPerBlock) {
    for (uint32_t k = in_off[v] + lane; k < in_off[v + 1]; k += kWave) {
      const uint32_t u = in_src[k];
      uint32_t lo = out_off[u], hi = out_off[u + 1];
      while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (out_dst[mid] < v) lo = mid + 1;
        else hi = mid;
      }
      const bool found = lo < out_off[u + 1] && out_dst[lo] == v;
      bad |= !found;
      in_eid[k] = found ? lo : 0u;
    }
  }
  if (bad) atomicOr(missing, 1);
}

// The same map for a CSR of the degree-ordered layout (rows and ids in new numbering, perm: new -> old): eid[k] =
// raw out-CSR slot of slot k. transpose: rows are destinations (the layout in-CSR), else sources (its out-CSR).
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

template <bool ADJOINT>
hipError_t launch_step(const WeightedArgs &A, int width, const LoopCtl &ctl, hipStream_t s) {
  const uint32_t nb_high = A.sched.n_high + A.sched.n_huge
                               ? blocks(size_t(A.sched.n_high) + A.sched.n_huge, kWavesPerBlock, 16384) : 0;
  const uint32_t nb_low = A.sched.n_low ? blocks(A.sched.n_low, size_t(kWavesPerBlock) * (kWave / width), 32768) : 0;
  const uint32_t grid = nb_high + nb_low;
  if (!grid) return hipSuccess;
  switch (width) {
    case 1: weighted_step_kernel<ADJOINT, 1><<<grid, kBlock, 0, s>>>(A, nb_high, ctl); break;
    case 2: weighted_step_kernel<ADJOINT, 2><<<grid, kBlock, 0, s>>>(A, nb_high, ctl); break;
    case 4: weighted_step_kernel<ADJOINT, 4><<<grid, kBlock, 0, s>>>(A, nb_high, ctl); break;
    case 8: weighted_step_kernel<ADJOINT, 8><<<grid, kBlock, 0, s>>>(A, nb_high, ctl); break;
    case 16: weighted_step_kernel<ADJOINT, 16><<<grid, kBlock, 0, s>>>(A, nb_high, ctl); break;
    case 32: weighted_step_kernel<ADJOINT, 32><<<grid, kBlock, 0, s>>>(A, nb_high, ctl); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_weighted_step(bool adjoint, const WeightedArgs &A, int width, unsigned long long *err, int *done,
                                int l, double tol, hipStream_t s) {
  const LoopCtl ctl{err, done, l, tol};
  return adjoint ? launch_step<true>(A, width, ctl, s) : launch_step<false>(A, width, ctl, s);
}

hipError_t launch_out_weight_sum(const uint32_t *out_off, const float *w, uint32_t N, double *W, hipStream_t s) {
  if (!N) return hipSuccess;
  out_weight_sum_kernel<<<blocks(N, kWavesPerBlock, 16384), kBlock, 0, s>>>(out_off, w, N, W);
  return hipGetLastError();
}

hipError_t launch_edge_map(const uint32_t *in_off, const uint32_t *in_src, const uint32_t *out_off,
                           const uint32_t *out_dst, uint32_t N, uint32_t *in_eid, int *missing, hipStream_t s) {
  if (!N) return hipSuccess;
  edge_map_kernel<<<blocks(N, kWavesPerBlock, 16384), kBlock, 0, s>>>(in_off, in_src, out_off, out_dst, N, in_eid,
                                                                      missing);
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
