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

}  // namespace prh
