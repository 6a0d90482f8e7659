// layout.hip -- the degree-ordered device layout of a graph, and the permutations at its boundary.
//
// Vertices are renumbered by in-degree, descending (stable in id), so that
//   * the rank kernel's degree classes are contiguous row ranges (no schedule indirection, coalesced per-vertex
//     reads and writes), and
//   * the most-gathered vertices (in RMAT-like graphs in- and out-degree go together) share cache lines, which
//     cuts the 128-byte line fills that random FP64 gathers cost.
// Each in-row of the relabelled transpose is sorted by new source id, so a wave's gathers walk the hot region in
// order. The permutation (new -> old) is computed on the host by a counting sort over degrees; the device builds
// the inverse, the relabelled rows and sorts them.
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_segmented_radix_sort.hpp>

#include "prh_device.hpp"
#include "prh_internal.hpp"

namespace prh {
namespace {

inline uint32_t grid(size_t n, size_t per, uint32_t cap) {
  const size_t g = (n + per - 1) / per;
  return uint32_t(g < 1 ? 1 : (g > cap ? cap : g));
}

__global__ __launch_bounds__(kBlock) void invert_kernel(const uint32_t *perm, uint32_t N, uint32_t *inv) {
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < N; i += gridDim.x * kBlock) inv[perm[i]] = i;
}

// Row i of the output is row perm[i] of the input with every id mapped through inv; one wave per row.
__global__ __launch_bounds__(kBlock) void relabel_rows_kernel(const uint32_t *off, const uint32_t *idx,
                                                              const uint32_t *perm, const uint32_t *inv,
                                                              const uint32_t *new_off, uint32_t N, uint32_t *out) {
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  for (uint32_t i = blockIdx.x * kWavesPerBlock + w; i < N; i += gridDim.x * kWavesPerBlock) {
    const uint32_t v = perm[i], b = off[v], d = off[v + 1] - b, o = new_off[i];
    for (uint32_t k = lane; k < d; k += kWave) out[o + k] = inv[idx[b + k]];
  }
}

__global__ __launch_bounds__(kBlock) void iota_kernel(uint32_t *p, uint32_t n) {
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) p[i] = i;
}

template <class T>
__global__ __launch_bounds__(kBlock) void gather_kernel(const T *src, const uint32_t *idx, size_t n, T *dst) {
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += size_t(gridDim.x) * kBlock)
    dst[i] = src[idx[i]];
}

template <class T>
__global__ __launch_bounds__(kBlock) void scatter_kernel(const T *src, const uint32_t *idx, size_t n, T *dst) {
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += size_t(gridDim.x) * kBlock)
    dst[idx[i]] = src[i];
}

__global__ __launch_bounds__(kBlock) void map_ids_kernel(uint32_t *ids, size_t n, const uint32_t *inv) {
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += size_t(gridDim.x) * kBlock)
    ids[i] = inv[ids[i]];
}

// ---- the source-sliced layout (see SlicedPlan) ----

__device__ __forceinline__ uint32_t slice_of(uint32_t u, int S) { return (u >> kSliceLineShift) & uint32_t(S - 1); }

// One wave per row counts the row's sources per slice, 64 edges per step by ballots (wave-uniform counts).
__global__ __launch_bounds__(kBlock) void slice_count_kernel(const uint32_t *off, const uint32_t *src, uint32_t N,
                                                             int S, uint32_t *cnt) {
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  for (uint32_t v = blockIdx.x * kWavesPerBlock + w; v < N; v += gridDim.x * kWavesPerBlock) {
    const uint32_t b = off[v], e = off[v + 1];
    uint32_t c[kMaxSlices] = {};
    for (uint32_t k = b; k < e; k += kWave) {
      const uint32_t x = k + lane < e ? slice_of(src[k + lane], S) : 0xFFFFFFFFu;
#pragma unroll
      for (int y = 0; y < kMaxSlices; ++y) c[y] += uint32_t(__popcll(__ballot(x == uint32_t(y))));
    }
    if (lane == 0) {
#pragma unroll
      for (int y = 0; y < kMaxSlices; ++y)
        if (y < S) cnt[size_t(y) * N + v] = c[y];
    }
  }
}

__global__ __launch_bounds__(kBlock) void slice_flags_kernel(const uint32_t *cnt_x, uint32_t N, uint32_t n_huge,
                                                             uint32_t *has, uint32_t *ecnt) {
  for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v <= N; v += gridDim.x * kBlock) {
    const uint32_t c = v < N && v >= n_huge ? cnt_x[v] : 0u;
    has[v] = c ? 1u : 0u;
    ecnt[v] = c;
  }
}

// One wave per row: the row's sources go to their slices' runs in the slice-major array, keeping ascending order
// within each run (ballot + mbcnt ranks, 64 edges per step). Chunked rows (v < n_huge) keep their edge range and
// are only regrouped by slice inside it; other rows' runs start at item_off of their (row, slice) items.
__global__ __launch_bounds__(kBlock) void slice_scatter_kernel(const uint32_t *off, const uint32_t *src, uint32_t N,
                                                               uint32_t n_huge, int S, const uint32_t *cnt,
                                                               const uint32_t *item, const uint32_t *eoff,
                                                               SliceBases B, uint32_t *s_src, uint32_t *item_off,
                                                               uint8_t *mask) {
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  for (uint32_t v = blockIdx.x * kWavesPerBlock + w; v < N; v += gridDim.x * kWavesPerBlock) {
    uint32_t base[kMaxSlices] = {};
    unsigned m = 0;
    if (v < n_huge) {
      uint32_t acc = off[v];
#pragma unroll
      for (int y = 0; y < kMaxSlices; ++y)
        if (y < S) {
          base[y] = acc;
          acc += cnt[size_t(y) * N + v];
        }
    } else {
#pragma unroll
      for (int y = 0; y < kMaxSlices; ++y)
        if (y < S) {
          base[y] = B.edge[y] + eoff[size_t(y) * (N + 1) + v];
          if (cnt[size_t(y) * N + v]) {
            m |= 1u << y;
            if (lane == 0) item_off[B.item[y] + item[size_t(y) * (N + 1) + v]] = base[y];
          }
        }
    }
    if (lane == 0) mask[v] = uint8_t(m);
    const uint32_t b = off[v], e = off[v + 1];
    for (uint32_t k = b; k < e; k += kWave) {
      const bool in = k + lane < e;
      const uint32_t u = in ? src[k + lane] : 0u;
      const uint32_t x = in ? slice_of(u, S) : 0xFFFFFFFFu;
#pragma unroll
      for (int y = 0; y < kMaxSlices; ++y) {
        const unsigned long long bal = __ballot(x == uint32_t(y));
        if (x == uint32_t(y)) s_src[base[y] + lanes_below(bal)] = u;
        base[y] += uint32_t(__popcll(bal));
      }
    }
  }
}

// wbase[g * S + y]: item id of the first slice-y item at or after row n_huge + 64 g.
__global__ __launch_bounds__(kBlock) void slice_wbase_kernel(const uint32_t *item, uint32_t N, uint32_t n_huge, int S,
                                                             SliceBases B, uint32_t *wbase) {
  const uint32_t groups = (N - n_huge + kWave - 1) / kWave;
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < groups * uint32_t(S); i += gridDim.x * kBlock) {
    const uint32_t g = i / uint32_t(S), y = i % uint32_t(S);
    wbase[i] = B.item[y] + item[size_t(y) * (N + 1) + n_huge + size_t(g) * kWave];
  }
}

}  // namespace

hipError_t launch_invert(const uint32_t *perm, uint32_t N, uint32_t *inv, hipStream_t s) {
  if (!N) return hipSuccess;
  invert_kernel<<<grid(N, kBlock, 8192), kBlock, 0, s>>>(perm, N, inv);
  return hipGetLastError();
}

hipError_t launch_relabel_rows(const uint32_t *off, const uint32_t *idx, const uint32_t *perm, const uint32_t *inv,
                               const uint32_t *new_off, uint32_t N, uint32_t *out, hipStream_t s) {
  if (!N) return hipSuccess;
  relabel_rows_kernel<<<grid(N, kWavesPerBlock, 16384), kBlock, 0, s>>>(off, idx, perm, inv, new_off, N, out);
  return hipGetLastError();
}

hipError_t sort_rows(void *tmp, size_t &tmp_bytes, const uint32_t *in, uint32_t *out, size_t M, uint32_t N,
                     const uint32_t *off, hipStream_t s) {
  return rocprim::segmented_radix_sort_keys(tmp, tmp_bytes, in, out, M, N, off, off + 1, 0, 32, s);
}

hipError_t launch_iota_u32(uint32_t *p, uint32_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  iota_kernel<<<grid(n, kBlock, 8192), kBlock, 0, s>>>(p, n);
  return hipGetLastError();
}

hipError_t launch_gather_u32(const uint32_t *src, const uint32_t *idx, size_t n, uint32_t *dst, hipStream_t s) {
  if (!n) return hipSuccess;
  gather_kernel<uint32_t><<<grid(n, kBlock, 8192), kBlock, 0, s>>>(src, idx, n, dst);
  return hipGetLastError();
}

hipError_t launch_gather_f64(const double *src, const uint32_t *idx, size_t n, double *dst, hipStream_t s) {
  if (!n) return hipSuccess;
  gather_kernel<double><<<grid(n, kBlock, 8192), kBlock, 0, s>>>(src, idx, n, dst);
  return hipGetLastError();
}

hipError_t launch_scatter_f64(const double *src, const uint32_t *idx, size_t n, double *dst, hipStream_t s) {
  if (!n) return hipSuccess;
  scatter_kernel<double><<<grid(n, kBlock, 8192), kBlock, 0, s>>>(src, idx, n, dst);
  return hipGetLastError();
}

hipError_t launch_scatter_u8(const uint8_t *src, const uint32_t *idx, size_t n, uint8_t *dst, hipStream_t s) {
  if (!n) return hipSuccess;
  scatter_kernel<uint8_t><<<grid(n, kBlock, 8192), kBlock, 0, s>>>(src, idx, n, dst);
  return hipGetLastError();
}

hipError_t launch_map_ids(uint32_t *ids, size_t n, const uint32_t *inv, hipStream_t s) {
  if (!n) return hipSuccess;
  map_ids_kernel<<<grid(n, kBlock, 4096), kBlock, 0, s>>>(ids, n, inv);
  return hipGetLastError();
}

hipError_t launch_slice_count(const uint32_t *off, const uint32_t *src, uint32_t N, int slices, uint32_t *cnt,
                              hipStream_t s) {
  if (!N) return hipSuccess;
  slice_count_kernel<<<grid(N, kWavesPerBlock, 16384), kBlock, 0, s>>>(off, src, N, slices, cnt);
  return hipGetLastError();
}

hipError_t launch_slice_flags(const uint32_t *cnt_x, uint32_t N, uint32_t n_huge, uint32_t *has, uint32_t *ecnt,
                              hipStream_t s) {
  slice_flags_kernel<<<grid(size_t(N) + 1, kBlock, 8192), kBlock, 0, s>>>(cnt_x, N, n_huge, has, ecnt);
  return hipGetLastError();
}

hipError_t scan_u32(void *tmp, size_t &tmp_bytes, const uint32_t *in, uint32_t *out, size_t n, hipStream_t s) {
  return rocprim::exclusive_scan(tmp, tmp_bytes, in, out, 0u, n, rocprim::plus<uint32_t>(), s);
}

hipError_t launch_slice_scatter(const uint32_t *off, const uint32_t *src, uint32_t N, uint32_t n_huge, int slices,
                                const uint32_t *cnt, const uint32_t *item, const uint32_t *eoff, const SliceBases &b,
                                uint32_t *s_src, uint32_t *item_off, uint8_t *mask, hipStream_t s) {
  if (!N) return hipSuccess;
  slice_scatter_kernel<<<grid(N, kWavesPerBlock, 16384), kBlock, 0, s>>>(off, src, N, n_huge, slices, cnt, item, eoff,
                                                                         b, s_src, item_off, mask);
  return hipGetLastError();
}

hipError_t launch_slice_wbase(const uint32_t *item, uint32_t N, uint32_t n_huge, int slices, const SliceBases &b,
                              uint32_t *wbase, hipStream_t s) {
  const size_t n = size_t((N - n_huge + kWave - 1) / kWave) * slices;
  if (!n) return hipSuccess;
  slice_wbase_kernel<<<grid(n, kBlock, 4096), kBlock, 0, s>>>(item, N, n_huge, slices, b, wbase);
  return hipGetLastError();
}

}  // namespace prh
