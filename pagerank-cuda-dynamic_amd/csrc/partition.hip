// partition.hip -- stable degree partition of the vertices of a CSR (count / scan / scatter).
//
// Semantics follow pagerankPartitionVerticesCuW (pagerankCuda.hxx:625-638): vertices with degree <= t1 in
// ascending id order, followed by the rest in ascending id order. Here the rest is optionally split once more
// at t2 (t1 < d <= t2, then d > t2) so the rank kernel can give the heaviest rows a whole workgroup; with
// t2 = UINT32_MAX the output is exactly the reference's two-way partition and counts[0] its partition point.
//
// Three passes over tiles of 2048 vertices: per-tile class counts (wave ballots), one workgroup scanning the
// tile counts into tile offsets, and a scatter that ranks each vertex inside its tile with ballot + mbcnt.
#include "prh_device.hpp"
#include "prh_internal.hpp"

namespace prh {
namespace {

constexpr int kPartIpt = 8;                      // vertices per thread per tile
constexpr uint32_t kTile = kBlock * kPartIpt;    // 2048 vertices per tile
static_assert(kTile == kPartitionTile, "tile size is part of the scratch contract");
constexpr int kScanThreads = 1024;

__device__ __forceinline__ int degree_class(const uint32_t *off, uint32_t v, uint32_t t1, uint32_t t2) {
  const uint32_t d = off[v + 1] - off[v];
  return d <= t1 ? 0 : (d <= t2 ? 1 : 2);
}

__global__ __launch_bounds__(kBlock) void partition_count_kernel(const uint32_t *off, uint32_t N, uint32_t t1,
                                                                 uint32_t t2, uint32_t *tiles) {
  __shared__ uint32_t lds[2][kWavesPerBlock];
  const uint32_t base = blockIdx.x * kTile;
  const int w = threadIdx.x / kWave;
  uint32_t c0 = 0, c1 = 0;  // wave-uniform
  for (int j = 0; j < kPartIpt; ++j) {
    const uint32_t v = base + j * kBlock + threadIdx.x;
    const int k = v < N ? degree_class(off, v, t1, t2) : 3;
    c0 += __popcll(__ballot(k == 0));
    c1 += __popcll(__ballot(k == 1));
  }
  if ((threadIdx.x & (kWave - 1)) == 0) {
    lds[0][w] = c0;
    lds[1][w] = c1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    tiles[3 * blockIdx.x + 0] = lds[0][0] + lds[0][1] + lds[0][2] + lds[0][3];
    tiles[3 * blockIdx.x + 1] = lds[1][0] + lds[1][1] + lds[1][2] + lds[1][3];
  }
}

// One workgroup: turns per-tile class counts into per-tile output offsets of the three classes.
__global__ __launch_bounds__(kScanThreads) void partition_scan_kernel(uint32_t *tiles, uint32_t ntiles, uint32_t N,
                                                                      uint32_t *counts) {
  __shared__ uint32_t s0[kScanThreads], s1[kScanThreads];
  __shared__ uint32_t tot[2];
  const uint32_t t = threadIdx.x;
  const uint32_t per = (ntiles + kScanThreads - 1) / kScanThreads;
  const uint32_t tb = min(ntiles, t * per), te = min(ntiles, tb + per);
  uint32_t a0 = 0, a1 = 0;
  for (uint32_t k = tb; k < te; ++k) {
    a0 += tiles[3 * k + 0];
    a1 += tiles[3 * k + 1];
  }
  s0[t] = a0;
  s1[t] = a1;
  __syncthreads();
  // Inclusive Hillis-Steele scan of the per-thread sums.
  for (uint32_t d = 1; d < kScanThreads; d <<= 1) {
    const uint32_t x0 = t >= d ? s0[t - d] : 0, x1 = t >= d ? s1[t - d] : 0;
    __syncthreads();
    s0[t] += x0;
    s1[t] += x1;
    __syncthreads();
  }
  if (t == kScanThreads - 1) {
    tot[0] = s0[t];
    tot[1] = s1[t];
  }
  __syncthreads();
  const uint32_t n0 = tot[0], n1 = tot[1];
  uint32_t p0 = s0[t] - a0, p1 = s1[t] - a1;  // exclusive prefixes
  for (uint32_t k = tb; k < te; ++k) {
    const uint32_t c0 = tiles[3 * k + 0], c1 = tiles[3 * k + 1];
    tiles[3 * k + 0] = p0;
    tiles[3 * k + 1] = n0 + p1;
    tiles[3 * k + 2] = n0 + n1 + (k * kTile - p0 - p1);
    p0 += c0;
    p1 += c1;
  }
  if (t == 0) {
    counts[0] = n0;
    counts[1] = n1;
    counts[2] = N - n0 - n1;
  }
}

__global__ __launch_bounds__(kBlock) void partition_scatter_kernel(const uint32_t *off, uint32_t N, uint32_t t1,
                                                                   uint32_t t2, const uint32_t *tiles,
                                                                   uint32_t *order) {
  __shared__ uint32_t wc[3][kWavesPerBlock];
  const uint32_t base = blockIdx.x * kTile;
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  uint32_t run0 = tiles[3 * blockIdx.x + 0], run1 = tiles[3 * blockIdx.x + 1], run2 = tiles[3 * blockIdx.x + 2];
  for (int j = 0; j < kPartIpt; ++j) {
    const uint32_t v = base + j * kBlock + threadIdx.x;
    const int k = v < N ? degree_class(off, v, t1, t2) : 3;
    const unsigned long long m0 = __ballot(k == 0), m1 = __ballot(k == 1), m2 = __ballot(k == 2);
    if (lane == 0) {
      wc[0][w] = __popcll(m0);
      wc[1][w] = __popcll(m1);
      wc[2][w] = __popcll(m2);
    }
    __syncthreads();
    if (k < 3) {
      const unsigned long long m = k == 0 ? m0 : (k == 1 ? m1 : m2);
      uint32_t pos = k == 0 ? run0 : (k == 1 ? run1 : run2);
      for (int x = 0; x < w; ++x) pos += wc[k][x];
      order[pos + lanes_below(m)] = v;
    }
    run0 += wc[0][0] + wc[0][1] + wc[0][2] + wc[0][3];
    run1 += wc[1][0] + wc[1][1] + wc[1][2] + wc[1][3];
    run2 += wc[2][0] + wc[2][1] + wc[2][2] + wc[2][3];
    __syncthreads();
  }
}

}  // namespace

uint32_t partition_tiles(uint32_t N) { return (N + kTile - 1) / kTile; }

hipError_t launch_partition(const uint32_t *off, uint32_t N, uint32_t t1, uint32_t t2, uint32_t *order,
                            uint32_t *tiles, uint32_t *counts_dev, hipStream_t s) {
  const uint32_t nt = partition_tiles(N);
  if (nt == 0) return hipSuccess;
  partition_count_kernel<<<nt, kBlock, 0, s>>>(off, N, t1, t2, tiles);
  partition_scan_kernel<<<1, kScanThreads, 0, s>>>(tiles, nt, N, counts_dev);
  partition_scatter_kernel<<<nt, kBlock, 0, s>>>(off, N, t1, t2, tiles, order);
  return hipGetLastError();
}

}  // namespace prh
