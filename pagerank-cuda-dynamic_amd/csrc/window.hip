// window.hip -- the banded rank update: rows in the caller's vertex order, each workgroup sweeping a contiguous range
// of them chunk by chunk through a window of contributions held in LDS.
//
// On a graph whose edges join vertices close together in the caller's order (a road network, a mesh, a grid: every
// in-edge of row v comes from a vertex within H of v), the contributions that a chunk of R rows gathers all lie in
// the window [c0 - H, c0 + R + H). A workgroup keeps that window in a ring of S >= R + 2H FP64 entries in LDS; the
// next chunk brings only R new entries into it. Each contribution r/deg is thus computed from the ranks and read from
// HBM about once per iteration (plus 2H entries per workgroup's range), and every gather is an LDS read. Per
// iteration and vertex: 4 bytes of offset, 4 per in-edge (its source id), 12 per window entry (rank and out-degree),
// 8 for the old rank (an L2 hit: it entered the window a few chunks earlier) and 8 for the new one -- against the
// 8 bytes per edge that the pull model charges for a gather. Sources outside the window (edges an in-place update
// inserted, or a graph whose spread exceeds what the ring holds) are gathered from the ranks directly: same values,
// only slower. The plan keeps no contribution array.
//
// Pipelining: a thread owns one row of each chunk (two rows a thread, with their loads in flight, took 90 VGPRs and
// left one workgroup per CU: slower). While it sums chunk k, the loads of chunk k + 1 are in flight
// (its row's first source ids, old rank, out-degree and affected flag; one new window entry) and so are the offsets
// of chunk k + 2, so that each step of the sweep waits for one round trip to memory, not three.
//
// Every vertex gets its full in-sum, taken in edge order, and the update from it is the other plans' (update.hpp):
// the dynamic variants carry unaffected vertices over and set their prune and frontier flags as the row kernel does;
// the frontier's extension runs in the separate extension kernels, as with the row plan.
#include "prh_device.hpp"
#include "prh_internal.hpp"
#include "update.hpp"

namespace prh {
namespace {

constexpr uint32_t kWinRows = kBinBlock;  // rows per chunk, one per thread
constexpr int kWinWaves = kBinBlock / kWave;
constexpr int kWinIds = 6;  // source ids a row loads ahead of its chunk; further ones are read when it is summed

struct WinRow {
  uint32_t b, e;         // edge range
  uint32_t d;            // out-degree (loaded where the variant reads it)
  double rv;             // old rank
  bool upd;              // updated (else carried over)
  uint32_t id[kWinIds];  // the first sources
};

template <int VARIANT>
__device__ __forceinline__ WinRow win_row(const RankArgs &A, uint32_t v, uint32_t b, uint32_t e, bool valid) {
  WinRow o;
  o.b = b;
  o.e = e;
  o.d = 1;
  o.rv = 0.0;
  o.upd = false;
#pragma unroll
  for (int j = 0; j < kWinIds; ++j) o.id[j] = 0;
  if (valid) {
    o.rv = A.r[v];
    o.upd = affected<VARIANT>(A, v);
    if constexpr (Traits<VARIANT>::prune || Traits<VARIANT>::frontier) o.d = A.deg[v];
    // Plain loads: a wave's rows take consecutive ids, so the kWinIds loads of a wave share cache lines (with the
    // non-temporal hint each one fetched its lines again: 290 against 247 us per update on the 6000 x 4000 grid).
#pragma unroll
    for (int j = 0; j < kWinIds; ++j)
      if (b + j < e) o.id[j] = A.src[b + j];
  }
  return o;
}

template <int VARIANT, int S>
__global__ __launch_bounds__(kBinBlock) void window_update_kernel(RankArgs A, WindowPlan W, LoopCtl ctl) {
  if (!iteration_begin(ctl)) return;
  static_assert(S >= int(kWinRows) + 2, "ring smaller than a chunk");
  constexpr uint32_t R = kWinRows;
  __shared__ double ring[S];
  __shared__ double lds_max[kWinWaves];
  __shared__ double lds_cost[2 * kWinWaves];
  const uint32_t N = W.N, H = W.H, t = threadIdx.x;
  const uint32_t a = uint32_t(uint64_t(blockIdx.x) * N / gridDim.x);
  const uint32_t z = uint32_t(uint64_t(blockIdx.x + 1) * N / gridDim.x);  // this workgroup's rows [a, z)
  const uint32_t zh = min(N, z + H);  // entries at or above zh are never loaded: rows end below z
  double emax = 0.0;
  CostAcc cost;
  if (a < z) {
    // The window of the first chunk, then its rows' operands and the offsets of the second chunk.
    const uint32_t w0 = a > H ? a - H : 0u, w1 = min(zh, a + R + H);
    for (uint32_t x = w0 + t; x < w1; x += R) ring[x % S] = A.r[x] / A.deg[x];
    const uint32_t v0 = a + t;
    const bool ok0 = v0 < z;
    WinRow cur = win_row<VARIANT>(A, v0, ok0 ? A.off[v0] : 0u, ok0 ? A.off[v0 + 1] : 0u, ok0);
    const uint32_t v1 = a + R + t;
    uint32_t nb = v1 < z ? A.off[v1] : 0u, ne = v1 < z ? A.off[v1 + 1] : 0u;
    __syncthreads();
    for (uint32_t c0 = a; c0 < z; c0 += R) {  // block-uniform
      const uint32_t c1 = c0 + R;
      // In flight during this chunk: the next chunk's rows (their offsets were loaded one chunk earlier), its new
      // window entry [c1 + H, c1 + R + H), and the offsets of the chunk after it.
      const uint32_t vn = c1 + t;
      const WinRow nxt = win_row<VARIANT>(A, vn, nb, ne, vn < z);
      const uint32_t x = c1 + H + t;
      const bool xin = c1 < z && x < zh;
      double xr = 0.0;
      uint32_t xd = 1;
      if (xin) {
        xr = A.r[x];
        xd = A.deg[x];
      }
      const uint32_t v2 = c1 + R + t;
      const uint32_t nb2 = v2 < z ? A.off[v2] : 0u, ne2 = v2 < z ? A.off[v2 + 1] : 0u;
      // This chunk: row v sums its sources from the window [lo, hi) (loaded, and not yet overwritten since
      // S >= R + 2H), or from the ranks outside it.
      const uint32_t v = c0 + t;
      if (v < z) {
        // Entry u sits at ring slot u mod S = (u - lo) + (lo mod S), less S once past the end (span <= S).
        const uint32_t lo = c0 > H ? c0 - H : 0u, span = min(zh, c1 + H) - lo, base = lo % S;
        auto take = [&](uint32_t u) {
          const uint32_t k = u - lo;
          if (k < span) {
            const uint32_t q = k + base;
            return ring[q >= S ? q - S : q];
          }
          return A.r[u] / A.deg[u];
        };
        RowFlags f;
        if (cur.upd) {
          const uint32_t n = cur.e - cur.b;
          double s0 = 0.0, s1 = 0.0;
#pragma unroll
          for (int j = 0; j < kWinIds; j += 2) {
            if (uint32_t(j) < n) s0 += take(cur.id[j]);
            if (uint32_t(j + 1) < n) s1 += take(cur.id[j + 1]);
          }
          for (uint32_t e = cur.b + kWinIds; e < cur.e; ++e) s0 += take(A.src[e]);
          const double crv = Traits<VARIANT>::prune ? ring[v % S] : 0.0;  // r/d of v itself, as gathered
          emax = fmax(emax, finish_vertex<VARIANT, false>(A, v, s0 + s1, cur.d, cur.rv, crv, f));
        } else {
          copy_vertex<VARIANT, false>(A, v, cur.rv, 0.0);
          f.affected = false;
        }
        if constexpr (Traits<VARIANT>::frontier)
          if (A.stats) cost.add(f, cur.e - cur.b, cur.d);
      }
      __syncthreads();  // the window's reads for this chunk are done; entry x - S is no longer needed
      if (xin) ring[x % S] = xr / xd;
      __syncthreads();
      cur = nxt;
      nb = nb2;
      ne = ne2;
    }
  }
  wg_store_costs<VARIANT, kWinWaves>(A, cost, blockIdx.x, lds_cost);
  emax = wg_max<kWinWaves>(emax, lds_max);
  if (threadIdx.x == 0) atomic_max_nonneg(&ctl.err[iter_index(ctl)], emax);
}

template <int S>
hipError_t window_variant(int variant, const RankArgs &A, const WindowPlan &W, const LoopCtl &ctl, hipStream_t s) {
  const dim3 grid(W.wgs), block(kBinBlock);
  switch (variant) {
    case kPlain: window_update_kernel<kPlain, S><<<grid, block, 0, s>>>(A, W, ctl); break;
    case kTraversal: window_update_kernel<kTraversal, S><<<grid, block, 0, s>>>(A, W, ctl); break;
    case kFrontier: window_update_kernel<kFrontier, S><<<grid, block, 0, s>>>(A, W, ctl); break;
    case kPrune: window_update_kernel<kPrune, S><<<grid, block, 0, s>>>(A, W, ctl); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Largest |source - destination| over the in-edges of rows [0, N) (a thread per row), into *out (zeroed first).
__global__ __launch_bounds__(kBlock) void max_spread_kernel(const uint32_t *off, const uint32_t *src, uint32_t N,
                                                            uint32_t *out) {
  uint32_t m = 0;
  for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < N; v += gridDim.x * kBlock)
    for (uint32_t e = off[v]; e < off[v + 1]; ++e) {
      const uint32_t u = src[e];
      m = max(m, u > v ? u - v : v - u);
    }
  for (int k = kWave / 2; k > 0; k >>= 1) m = max(m, uint32_t(__shfl_xor(int(m), k)));
  if ((threadIdx.x & (kWave - 1)) == 0 && m) atomicMax(out, m);
}

}  // namespace

hipError_t launch_window_update(int variant, const RankArgs &A, const WindowPlan &W, unsigned long long *err,
                                int *done, int l, double tol, hipStream_t s) {
  const LoopCtl ctl{err, done, l, tol};
  switch (W.S) {
    case kWindowRingSmall: return window_variant<kWindowRingSmall>(variant, A, W, ctl, s);
    case kWindowRingLarge: return window_variant<kWindowRingLarge>(variant, A, W, ctl, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_max_spread(const uint32_t *off, const uint32_t *src, uint32_t N, uint32_t *out, hipStream_t s) {
  const hipError_t e = launch_fill_u32(out, 0u, 1, s);
  if (e != hipSuccess) return e;
  if (N == 0) return hipSuccess;
  const uint32_t g = std::min<uint32_t>((N + kBlock - 1) / kBlock, 4096);
  max_spread_kernel<<<g, kBlock, 0, s>>>(off, src, N, out);
  return hipGetLastError();
}

}  // namespace prh
