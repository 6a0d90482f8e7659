// stream.hip -- applying an edge batch to a device-resident graph (the streaming dynamic session).
//
// The session keeps the degree-ordered layout of its base graph (the vertex permutation stays fixed) and updates
// both CSRs of that layout in place of the host's tidy-and-apply (tidyBatchUpdateU, applyBatchUpdateU,
// batch.hxx:152-247): deletions keep only existing edges, insertions only absent ones, both sorted and unique;
// deletions are removed and insertions added. Every row stays sorted by column.
//
// Per CSR (rows = targets for the in-CSR, sources for the out-CSR), with key = row << 32 | column:
//   batch_keys     ids mapped to layout ids, keys built (entries past the batch's size or with a bad id: ~0)
//   (sort)         rocPRIM merge sort of the keys, at the session's fixed capacity
//   batch_valid    an entry survives if it is the first of its key and its edge exists (deletions) or is
//                  absent (insertions) in the current rows (binary search)
//   (scan)         exclusive scan of the survivor flags
//   batch_compact  survivors packed in key order, their count to a device word
//   row_ranges     per changed row: its deletions and insertions in the compacted keys, degree change, flag
//   new_degrees    old degree + change (the out-CSR also flags rows left with no out-edge)
//   (scan)         new offsets
//   (scan)         positions of the changed rows among all changed rows
//   units          per changed row its old and new offsets: the runs of unchanged rows between changed rows
//   rewrite        every old edge to its new place: unchanged rows move by their run's shift, changed rows skip
//                  deleted columns and make room for inserted ones (ranks found by binary search in the row's
//                  deletions/insertions)
//   insert         every inserted column to its place in its row
// Sizes are read from device words, so one captured sequence serves every batch up to the capacity.
#include <rocprim/device/device_merge.hpp>
#include <rocprim/device/device_merge_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "prh_device.hpp"
#include "prh_internal.hpp"

namespace prh {
namespace {

constexpr unsigned long long kNoKey = ~0ull;
constexpr unsigned long long kRowMask = 0xFFFFFFFF00000000ull;

inline uint32_t grid_for(size_t n, size_t per, uint32_t cap) {
  const size_t g = (n + per - 1) / per;
  return uint32_t(g < 1 ? 1 : (g > cap ? cap : g));
}

// First index in a[lo, hi) with a[i] >= x.
template <class T>
__device__ __forceinline__ uint32_t lower_bound(const T *a, uint32_t lo, uint32_t hi, T x) {
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (a[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(kBlock) void batch_keys_kernel(const uint32_t *bu, const uint32_t *bv,
                                                            const uint32_t *count, uint32_t cap, const uint32_t *inv,
                                                            uint32_t N, unsigned long long *in_keys,
                                                            unsigned long long *out_keys, int *flags, int bad_bit) {
  const uint32_t n = *count;
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < cap; i += gridDim.x * kBlock) {
    unsigned long long ki = kNoKey, ko = kNoKey;
    if (i < n) {
      const uint32_t u = bu[i], v = bv[i];
      if (u < N && v < N) {
        const unsigned long long lu = inv[u], lv = inv[v];
        ki = lv << 32 | lu;
        ko = lu << 32 | lv;
      } else {
        atomicOr(flags, bad_bit);
      }
    }
    in_keys[i] = ki;
    out_keys[i] = ko;
  }
}

// valid[j] = 1 when a[j] does not occur in b (both ascending).
__global__ __launch_bounds__(kBlock) void absent_from_kernel(const unsigned long long *a, uint32_t na,
                                                             const unsigned long long *b, uint32_t nb,
                                                             uint32_t *valid) {
  for (uint32_t j = blockIdx.x * kBlock + threadIdx.x; j < na; j += gridDim.x * kBlock) {
    const unsigned long long k = a[j];
    const uint32_t at = lower_bound(b, 0u, nb, k);
    valid[j] = at < nb && b[at] == k ? 0u : 1u;
  }
}

// In-CSR keys of a tidied batch from its out-CSR keys (source << 32 | target becomes target << 32 | source); the
// entries past the count become kNoKey, so that a sort of all cap entries puts the count keys first.
__global__ __launch_bounds__(kBlock) void transpose_keys_kernel(const unsigned long long *in, const uint32_t *count,
                                                                uint32_t cap, unsigned long long *out) {
  const uint32_t n = *count;
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < cap; i += gridDim.x * kBlock) {
    unsigned long long k = kNoKey;
    if (i < n) {
      const unsigned long long o = in[i];
      k = o << 32 | o >> 32;
    }
    out[i] = k;
  }
}

__global__ __launch_bounds__(kBlock) void batch_valid_kernel(const unsigned long long *keys, uint32_t cap,
                                                             const uint32_t *off, const uint32_t *idx, int want_exists,
                                                             uint32_t *valid) {
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < cap; i += gridDim.x * kBlock) {
    const unsigned long long k = keys[i];
    bool ok = k != kNoKey && (i == 0 || keys[i - 1] != k);
    if (ok) {
      const uint32_t row = uint32_t(k >> 32), col = uint32_t(k);
      const uint32_t b = off[row], e = off[row + 1];
      const uint32_t at = lower_bound(idx, b, e, col);
      const bool exists = at < e && idx[at] == col;
      ok = exists == (want_exists != 0);
    }
    valid[i] = ok ? 1u : 0u;
  }
}

__global__ __launch_bounds__(kBlock) void batch_compact_kernel(const unsigned long long *keys, const uint32_t *valid,
                                                               const uint32_t *pos, uint32_t cap,
                                                               unsigned long long *out, uint32_t *count) {
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < cap; i += gridDim.x * kBlock) {
    if (valid[i]) out[pos[i]] = keys[i];
    if (i == cap - 1) *count = pos[i] + valid[i];
  }
}

// The first entry of a row on either side finds the row's deletions [x, y) and insertions [z, w) in the two
// compacted key arrays and writes them with the row's degree change and its "changed" flag. A row with entries on
// both sides is written twice with the same values; rows without entries are left alone (delta and changed are
// cleared beforehand, the ranges are read only where changed is set).
__global__ __launch_bounds__(kBlock) void row_ranges_kernel(const unsigned long long *dkeys, const uint32_t *nd,
                                                            const unsigned long long *ikeys, const uint32_t *ni,
                                                            int *delta, uint8_t *changed, uint4 *ranges) {
  const uint32_t a = *nd, b = *ni;
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < a + b; i += gridDim.x * kBlock) {
    const bool del = i < a;
    const unsigned long long *k = del ? dkeys : ikeys;
    const uint32_t j = del ? i : i - a;
    const unsigned long long rk = k[j] & kRowMask;
    if (j > 0 && (k[j - 1] & kRowMask) == rk) continue;
    const unsigned long long next = rk + (1ull << 32);  // rows are below 2^32 - 1
    uint4 r;
    r.x = lower_bound(dkeys, 0u, a, rk);
    r.y = lower_bound(dkeys, r.x, a, next);
    r.z = lower_bound(ikeys, 0u, b, rk);
    r.w = lower_bound(ikeys, r.z, b, next);
    const uint32_t row = uint32_t(rk >> 32);
    ranges[row] = r;
    delta[row] = int(r.w - r.z) - int(r.y - r.x);
    changed[row] = 1;
  }
}

__global__ __launch_bounds__(kBlock) void new_degrees_kernel(const uint32_t *off, const int *delta, uint32_t N,
                                                             uint32_t *deg, int *flags, int empty_bit) {
  for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v <= N; v += gridDim.x * kBlock) {
    if (v == N) {
      deg[v] = 0;
      continue;
    }
    const uint32_t d = uint32_t(int(off[v + 1] - off[v]) + delta[v]);
    deg[v] = d;
    if (empty_bit && d == 0) atomicOr(flags, empty_bit);
  }
}

constexpr uint32_t kRewriteTile = 16384;  // old edges per workgroup
constexpr int kRewriteLoads = 16;         // loads a thread keeps in flight
constexpr uint32_t kTileUnits = 8;        // a tile over more units than this moves edge by edge
constexpr uint32_t kRowKeys = kBlock;     // a changed row's deletions and insertions staged in LDS, up to this many
constexpr uint32_t kStageUnits = kBlock;  // a tile's units staged in LDS for the edge-by-edge move, up to this many

// Units of the rewrite: unit i (i < C) is the run of unchanged rows before the i-th changed row R_i, then R_i
// itself; unit C runs from the last changed row's end to the last edge. Per unit: ue = old end of R_i (its edges are
// [ub, ue)), ub = old start of R_i, unb = new start of R_i, urow = R_i (N for the closing unit, whose ub = ue = M).
// Every unchanged row of unit i moves by unb - ub (rows between two changed rows keep their degrees).
struct RewriteUnits {
  uint32_t *ue, *ub, *unb, *urow;
  uint32_t *count;  // C
};

// pos = exclusive scan of changed; the i-th changed row writes unit i, row N - 1 also the closing unit and C.
__global__ __launch_bounds__(kBlock) void units_kernel(const uint32_t *old_off, const uint32_t *new_off,
                                                       const uint8_t *changed, const uint32_t *pos, uint32_t N,
                                                       RewriteUnits U) {
  for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < N; v += gridDim.x * kBlock) {
    const bool ch = changed[v] != 0;
    if (ch) {
      const uint32_t i = pos[v];
      U.ue[i] = old_off[v + 1];
      U.ub[i] = old_off[v];
      U.unb[i] = new_off[v];
      U.urow[i] = v;
    }
    if (v == N - 1) {
      const uint32_t C = pos[v] + (ch ? 1u : 0u);
      U.ue[C] = U.ub[C] = old_off[N];
      U.unb[C] = new_off[N];
      U.urow[C] = N;
      *U.count = C;
    }
  }
}

// First unit i in [lo, hi] whose old end ue[i] is above e (the closing unit's end is M, above every edge).
__device__ __forceinline__ uint32_t unit_of(const uint32_t *ue, uint32_t lo, uint32_t hi, uint32_t e) {
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (ue[mid] <= e) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// The same search by the whole workgroup (every thread calls it with the same arguments): each round tests kBlock
// evenly spaced units at once and keeps the interval after the last one whose end is at or below e, so a search over
// C units takes about log_kBlock(C) rounds of one load each. ue is non-decreasing (empty changed rows repeat an end)
// and the answer lies in [lo, hi].
__device__ uint32_t block_unit_of(const uint32_t *ue, uint32_t lo, uint32_t hi, uint32_t e) {
  while (lo < hi) {
    const uint32_t step = (hi - lo + kBlock - 1) / kBlock;
    const uint32_t p = lo + threadIdx.x * step;
    const uint32_t below = uint32_t(__syncthreads_count(p < hi && ue[p] <= e));
    if (below == 0) return lo;
    const uint32_t last = lo + (below - 1) * step;  // ue[last] <= e; unit last + step, if below hi, ends above e
    lo = last + 1;
    hi = min(hi, last + step);
  }
  return lo;
}

// A store that stays inside the new index array whatever position it is given.
__device__ __forceinline__ void store_in(uint32_t *__restrict__ dst, int64_t p, uint64_t cap, uint32_t v) {
  if (uint64_t(p) < cap) dst[p] = v;
}

// Old edges [a, b) of unchanged rows to their new places, shifted by `shift`; each thread keeps kRewriteLoads loads
// in flight before it stores.
__device__ __forceinline__ void move_run(const uint32_t *__restrict__ src, uint64_t a, uint64_t b, int64_t shift,
                                         uint32_t *__restrict__ dst, uint64_t cap) {
  for (uint64_t e = a + threadIdx.x; e < b; e += uint64_t(kRewriteLoads) * kBlock) {
    uint32_t v[kRewriteLoads];
#pragma unroll
    for (int k = 0; k < kRewriteLoads; ++k)
      if (e + uint64_t(k) * kBlock < b) v[k] = src[e + uint64_t(k) * kBlock];
#pragma unroll
    for (int k = 0; k < kRewriteLoads; ++k)
      if (e + uint64_t(k) * kBlock < b) store_in(dst, int64_t(e + uint64_t(k) * kBlock) + shift, cap, v[k]);
  }
}

// Old edges [a, b) of changed row `row` (old start ub, new start nb; its deletions dkeys[r.x, r.y) and insertions
// ikeys[r.z, r.w), keys row << 32 | column, sorted): deleted columns are dropped, the others move past the row's
// earlier deletions and insertions. The row's columns of both kinds are staged in LDS when they fit; all threads call
// this (it synchronises).
__device__ void move_changed_row(const uint32_t *__restrict__ src, uint64_t a, uint64_t b, uint32_t row, uint32_t ub,
                                 uint32_t nb, uint4 r, const unsigned long long *__restrict__ dkeys,
                                 const unsigned long long *__restrict__ ikeys, uint32_t *sdel, uint32_t *sins,
                                 uint32_t *__restrict__ dst, uint64_t cap) {
  const uint32_t nd = r.y - r.x, ni = r.w - r.z;
  const bool staged = nd <= kRowKeys && ni <= kRowKeys;
  if (staged) {
    if (threadIdx.x < nd) sdel[threadIdx.x] = uint32_t(dkeys[r.x + threadIdx.x]);
    if (threadIdx.x < ni) sins[threadIdx.x] = uint32_t(ikeys[r.z + threadIdx.x]);
    __syncthreads();
  }
  for (uint64_t e = a + threadIdx.x; e < b; e += uint64_t(kRewriteLoads) * kBlock) {
    uint32_t v[kRewriteLoads];
#pragma unroll
    for (int k = 0; k < kRewriteLoads; ++k)
      if (e + uint64_t(k) * kBlock < b) v[k] = src[e + uint64_t(k) * kBlock];
#pragma unroll
    for (int k = 0; k < kRewriteLoads; ++k) {
      const uint64_t ek = e + uint64_t(k) * kBlock;
      if (ek >= b) continue;
      const uint32_t col = v[k];
      uint32_t d_at, i_at;
      if (staged) {
        d_at = lower_bound(sdel, 0u, nd, col);
        if (d_at < nd && sdel[d_at] == col) continue;  // deleted
        i_at = lower_bound(sins, 0u, ni, col);
      } else {
        const unsigned long long key = (unsigned long long)row << 32 | col;
        d_at = lower_bound(dkeys, r.x, r.y, key);
        if (d_at < r.y && dkeys[d_at] == key) continue;  // deleted
        d_at -= r.x;
        i_at = lower_bound(ikeys, r.z, r.w, key) - r.z;
      }
      store_in(dst, int64_t(nb) + int64_t(ek - ub) - int64_t(d_at) + int64_t(i_at), cap, col);
    }
  }
  __syncthreads();  // the staged columns are read before the next row overwrites them
}

// Workgroup b moves old edges [b * tile, (b + 1) * tile) into the new index array (new_cap entries). The tile's
// first and last units are found by the whole workgroup. A tile over a few units moves unit by unit: the unchanged
// rows as one shifted run, the changed row against its staged deletions and insertions. A tile over many units (many
// short changed rows) moves edge by edge, each finding its unit by binary search, in LDS when the tile's units (at
// most kStageUnits) and their rows' ranges are staged there.
__global__ __launch_bounds__(kBlock) void rewrite_kernel(const uint32_t *__restrict__ old_off,
                                                         const uint32_t *__restrict__ old_idx, RewriteUnits U,
                                                         const uint4 *__restrict__ ranges,
                                                         const unsigned long long *__restrict__ dkeys,
                                                         const unsigned long long *__restrict__ ikeys, uint32_t N,
                                                         uint64_t new_cap, uint32_t *__restrict__ new_idx) {
  __shared__ uint32_t sdel[kRowKeys], sins[kRowKeys];
  __shared__ uint32_t s_ue[kStageUnits], s_ub[kStageUnits], s_unb[kStageUnits], s_row[kStageUnits];
  __shared__ uint4 s_rng[kStageUnits];
  const uint32_t M = old_off[N];
  const uint64_t e0 = uint64_t(blockIdx.x) * kRewriteTile;
  if (e0 >= M) return;
  const uint64_t e1 = min(uint64_t(M), e0 + kRewriteTile);
  const uint32_t C = *U.count;
  const uint32_t i0 = block_unit_of(U.ue, 0u, C, uint32_t(e0));
  const uint32_t i1 = block_unit_of(U.ue, i0, C, uint32_t(e1 - 1));
  if (i1 - i0 < kTileUnits) {
    for (uint32_t i = i0; i <= i1; ++i) {
      const uint32_t ub = U.ub[i], ue = U.ue[i], nb = U.unb[i];
      const uint64_t lo = i == i0 ? e0 : uint64_t(U.ue[i - 1]);  // unit i starts where unit i - 1 ends
      const uint64_t hi = min(e1, uint64_t(ue));
      const uint64_t mid = max(lo, min(hi, uint64_t(ub)));
      if (lo < mid) move_run(old_idx, lo, mid, int64_t(nb) - int64_t(ub), new_idx, new_cap);
      if (mid < hi) {
        const uint32_t row = U.urow[i];
        move_changed_row(old_idx, mid, hi, row, ub, nb, ranges[row], dkeys, ikeys, sdel, sins, new_idx, new_cap);
      }
    }
    return;
  }
  // Units i0..i1 staged: s_*[j] is unit i0 + j; the closing unit (i = C, row N) has no changed edges and no range.
  const uint32_t nu = i1 - i0 + 1;
  const bool staged = nu <= kStageUnits;
  if (staged) {
    if (threadIdx.x < nu) {
      const uint32_t i = i0 + threadIdx.x;
      s_ue[threadIdx.x] = U.ue[i];
      s_ub[threadIdx.x] = U.ub[i];
      s_unb[threadIdx.x] = U.unb[i];
      const uint32_t row = U.urow[i];
      s_row[threadIdx.x] = row;
      if (i < C) s_rng[threadIdx.x] = ranges[row];
    }
    __syncthreads();
  }
  for (uint64_t e = e0 + threadIdx.x; e < e1; e += uint64_t(kRewriteLoads) * kBlock) {
    uint32_t v[kRewriteLoads];
#pragma unroll
    for (int k = 0; k < kRewriteLoads; ++k)
      if (e + uint64_t(k) * kBlock < e1) v[k] = old_idx[e + uint64_t(k) * kBlock];
#pragma unroll
    for (int k = 0; k < kRewriteLoads; ++k) {
      const uint64_t ek = e + uint64_t(k) * kBlock;
      if (ek >= e1) continue;
      const uint32_t col = v[k];
      uint32_t b, nb, row;
      uint4 r;
      if (staged) {
        const uint32_t j = unit_of(s_ue, 0u, nu - 1, uint32_t(ek));
        b = s_ub[j];
        nb = s_unb[j];
        if (ek < b) {
          store_in(new_idx, int64_t(ek) + int64_t(nb) - int64_t(b), new_cap, col);
          continue;
        }
        row = s_row[j];
        r = s_rng[j];
      } else {
        const uint32_t i = unit_of(U.ue, i0, i1, uint32_t(ek));
        b = U.ub[i];
        nb = U.unb[i];
        if (ek < b) {
          store_in(new_idx, int64_t(ek) + int64_t(nb) - int64_t(b), new_cap, col);
          continue;
        }
        row = U.urow[i];
        r = ranges[row];
      }
      const unsigned long long key = (unsigned long long)row << 32 | col;
      const uint32_t d_at = lower_bound(dkeys, r.x, r.y, key);
      if (d_at < r.y && dkeys[d_at] == key) continue;  // deleted
      const uint32_t i_at = lower_bound(ikeys, r.z, r.w, key);
      store_in(new_idx, int64_t(nb) + int64_t(ek - b) - int64_t(d_at - r.x) + int64_t(i_at - r.z), new_cap, col);
    }
  }
}

__global__ __launch_bounds__(kBlock) void insert_kernel(const uint32_t *old_off, const uint32_t *old_idx,
                                                        const uint32_t *new_off, const uint4 *ranges,
                                                        const unsigned long long *dkeys,
                                                        const unsigned long long *ikeys, const uint32_t *ni,
                                                        uint32_t *new_idx) {
  const uint32_t I = *ni;
  for (uint32_t j = blockIdx.x * kBlock + threadIdx.x; j < I; j += gridDim.x * kBlock) {
    const unsigned long long key = ikeys[j];
    const uint32_t row = uint32_t(key >> 32), col = uint32_t(key);
    const uint4 r = ranges[row];
    const uint32_t b = old_off[row], e = old_off[row + 1];
    const uint32_t o_at = lower_bound(old_idx, b, e, col) - b;
    const uint32_t d_at = lower_bound(dkeys, r.x, r.y, key);
    new_idx[new_off[row] + o_at - (d_at - r.x) + (j - r.z)] = col;
  }
}

__global__ __launch_bounds__(kBlock) void degrees_kernel(const uint32_t *off, uint32_t N, uint32_t *deg) {
  for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < N; v += gridDim.x * kBlock) deg[v] = off[v + 1] - off[v];
}

// Item i of chunked row r = item_row[i], the k-th of the row's items, covers [off[r] + k chunk, ...); the row's
// last item also takes any remainder, so a row that grew past its items is still covered.
__global__ __launch_bounds__(kBlock) void refresh_chunks_kernel(const uint32_t *item_row, const uint32_t *row_first,
                                                                uint32_t n_items, const uint32_t *off, uint32_t chunk,
                                                                uint32_t *item_beg, uint32_t *item_end) {
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n_items; i += gridDim.x * kBlock) {
    const uint32_t r = item_row[i], k = i - row_first[r];
    const bool last = i + 1 == row_first[r + 1];
    const uint32_t b = off[r], e = off[r + 1];
    const uint64_t lo = uint64_t(b) + uint64_t(k) * chunk;
    item_beg[i] = uint32_t(lo < e ? lo : e);
    item_end[i] = last ? e : uint32_t(lo + chunk < e ? lo + chunk : e);
  }
}

// DF initial marking from the compacted batch in layout ids (out-CSR keys: source << 32 | target):
// naff[u] for every deletion and insertion source, vaff[v] for every deletion target (pagerankCuda.hxx:272-281).
__global__ __launch_bounds__(kBlock) void mark_batch_kernel(uint8_t *naff, uint8_t *vaff,
                                                            const unsigned long long *dkeys, const uint32_t *nd,
                                                            const unsigned long long *ikeys, const uint32_t *ni) {
  const uint32_t a = *nd, b = *ni;
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < a + b; i += gridDim.x * kBlock) {
    const unsigned long long k = i < a ? dkeys[i] : ikeys[i - a];
    naff[uint32_t(k >> 32)] = 1;
    if (i < a) vaff[uint32_t(k)] = 1;
  }
}

// Traversal closure levels (lvl: [0] level, [1] stop, [2] and [3] changed flags of even and odd levels). Before
// level l: stop if level l - 1 marked nothing new, else clear level l's flag.
__global__ void closure_step_kernel(int *lvl) {
  if (lvl[1]) return;
  const int l = lvl[0];
  if (l > 0 && lvl[2 + ((l - 1) & 1)] == 0) {
    lvl[1] = 1;
    return;
  }
  lvl[2 + (l & 1)] = 0;
}

__global__ void closure_advance_kernel(int *lvl) {
  if (!lvl[1]) ++lvl[0];
}

__global__ __launch_bounds__(kBlock) void copy_u8_kernel(uint8_t *dst, const uint8_t *src, size_t n) {
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += size_t(gridDim.x) * kBlock) dst[i] = src[i];
}

// After iteration l's launches: stop when its L-inf change is below tolerance (done = l + 1 iterations) or when
// the iteration cap is reached; otherwise move to iteration l + 1.
__global__ void advance_kernel(int *l, int *done, const unsigned long long *err, double tol, int max_iter) {
  if (*done) return;
  const int i = *l;
  const double e = __longlong_as_double((long long)err[i]);
  if (e < tol || i + 1 >= max_iter) *done = i + 1;
  else *l = i + 1;
}

}  // namespace

hipError_t launch_batch_keys(const uint32_t *bu, const uint32_t *bv, const uint32_t *count, uint32_t cap,
                             const uint32_t *inv, uint32_t N, unsigned long long *in_keys,
                             unsigned long long *out_keys, int *flags, int bad_bit, hipStream_t s) {
  batch_keys_kernel<<<grid_for(cap, kBlock, 1024), kBlock, 0, s>>>(bu, bv, count, cap, inv, N, in_keys, out_keys,
                                                                    flags, bad_bit);
  return hipGetLastError();
}

// Merge sort: its launches are kernels only (rocPRIM's radix sort clears state with hipMemsetAsync, which must not
// be recorded into the captured graphs that replay this).
hipError_t sort_keys_u64(void *tmp, size_t &tmp_bytes, const unsigned long long *in, unsigned long long *out,
                         uint32_t n, hipStream_t s) {
  return rocprim::merge_sort(tmp, tmp_bytes, in, out, n, rocprim::less<unsigned long long>(), s);
}

hipError_t merge_keys_u64(void *tmp, size_t &tmp_bytes, const unsigned long long *a, uint32_t na,
                          const unsigned long long *b, uint32_t nb, unsigned long long *out, hipStream_t s) {
  return rocprim::merge(tmp, tmp_bytes, a, b, out, na, nb, rocprim::less<unsigned long long>(), s);
}

hipError_t launch_absent_from(const unsigned long long *a, uint32_t na, const unsigned long long *b, uint32_t nb,
                              uint32_t *valid, hipStream_t s) {
  absent_from_kernel<<<grid_for(na, kBlock, 4096), kBlock, 0, s>>>(a, na, b, nb, valid);
  return hipGetLastError();
}

hipError_t launch_transpose_keys(const unsigned long long *in, const uint32_t *count, uint32_t cap,
                                unsigned long long *out, hipStream_t s) {
  transpose_keys_kernel<<<grid_for(cap, kBlock, 1024), kBlock, 0, s>>>(in, count, cap, out);
  return hipGetLastError();
}

hipError_t launch_batch_valid(const unsigned long long *keys, uint32_t cap, const uint32_t *off, const uint32_t *idx,
                              bool want_exists, uint32_t *valid, hipStream_t s) {
  batch_valid_kernel<<<grid_for(cap, kBlock, 1024), kBlock, 0, s>>>(keys, cap, off, idx, want_exists ? 1 : 0, valid);
  return hipGetLastError();
}

hipError_t launch_batch_compact(const unsigned long long *keys, const uint32_t *valid, const uint32_t *pos,
                                uint32_t cap, unsigned long long *out, uint32_t *count, hipStream_t s) {
  batch_compact_kernel<<<grid_for(cap, kBlock, 1024), kBlock, 0, s>>>(keys, valid, pos, cap, out, count);
  return hipGetLastError();
}

hipError_t launch_row_ranges(const unsigned long long *dkeys, const uint32_t *nd, const unsigned long long *ikeys,
                             const uint32_t *ni, uint32_t cap_total, int *delta, uint8_t *changed, uint4 *ranges,
                             hipStream_t s) {
  row_ranges_kernel<<<grid_for(cap_total, kBlock, 1024), kBlock, 0, s>>>(dkeys, nd, ikeys, ni, delta, changed, ranges);
  return hipGetLastError();
}

hipError_t launch_new_degrees(const uint32_t *off, const int *delta, uint32_t N, uint32_t *deg, int *flags,
                              int empty_bit, hipStream_t s) {
  new_degrees_kernel<<<grid_for(size_t(N) + 1, kBlock, 8192), kBlock, 0, s>>>(off, delta, N, deg, flags, empty_bit);
  return hipGetLastError();
}

hipError_t scan_u8(void *tmp, size_t &tmp_bytes, const uint8_t *in, uint32_t *out, size_t n, hipStream_t s) {
  return rocprim::exclusive_scan(tmp, tmp_bytes, in, out, 0u, n, rocprim::plus<uint32_t>(), s);
}

hipError_t launch_rewrite(const uint32_t *old_off, const uint32_t *old_idx, const uint32_t *new_off,
                          const uint8_t *changed, const uint32_t *pos, uint32_t *units, const uint4 *ranges,
                          const unsigned long long *dkeys, const unsigned long long *ikeys, uint32_t N,
                          uint32_t max_units, uint64_t max_edges, uint64_t new_cap, uint32_t *new_idx,
                          hipStream_t s) {
  if (N == 0) return hipSuccess;
  const RewriteUnits U{units, units + max_units, units + 2 * size_t(max_units), units + 3 * size_t(max_units),
                       units + 4 * size_t(max_units)};
  units_kernel<<<grid_for(N, kBlock, 16384), kBlock, 0, s>>>(old_off, new_off, changed, pos, N, U);
  const uint64_t g = (max_edges + kRewriteTile - 1) / kRewriteTile;
  rewrite_kernel<<<uint32_t(g ? g : 1), kBlock, 0, s>>>(old_off, old_idx, U, ranges, dkeys, ikeys, N, new_cap,
                                                        new_idx);
  return hipGetLastError();
}

hipError_t launch_insert(const uint32_t *old_off, const uint32_t *old_idx, const uint32_t *new_off,
                         const uint4 *ranges, const unsigned long long *dkeys, const unsigned long long *ikeys,
                         const uint32_t *ni, uint32_t cap, uint32_t *new_idx, hipStream_t s) {
  insert_kernel<<<grid_for(cap, kBlock, 1024), kBlock, 0, s>>>(old_off, old_idx, new_off, ranges, dkeys, ikeys, ni,
                                                               new_idx);
  return hipGetLastError();
}

hipError_t launch_degrees(const uint32_t *off, uint32_t N, uint32_t *deg, hipStream_t s) {
  degrees_kernel<<<grid_for(N, kBlock, 8192), kBlock, 0, s>>>(off, N, deg);
  return hipGetLastError();
}

hipError_t launch_refresh_chunks(const uint32_t *item_row, const uint32_t *row_first, uint32_t n_items,
                                 const uint32_t *off, uint32_t chunk, uint32_t *item_beg, uint32_t *item_end,
                                 hipStream_t s) {
  if (!n_items) return hipSuccess;
  refresh_chunks_kernel<<<grid_for(n_items, kBlock, 1024), kBlock, 0, s>>>(item_row, row_first, n_items, off, chunk,
                                                                           item_beg, item_end);
  return hipGetLastError();
}

hipError_t launch_mark_batch(uint8_t *naff, uint8_t *vaff, const unsigned long long *dkeys, const uint32_t *nd,
                             const unsigned long long *ikeys, const uint32_t *ni, uint32_t cap_total, hipStream_t s) {
  mark_batch_kernel<<<grid_for(cap_total, kBlock, 1024), kBlock, 0, s>>>(naff, vaff, dkeys, nd, ikeys, ni);
  return hipGetLastError();
}

hipError_t launch_closure_step(int *lvl, int parity, hipStream_t s) {
  (void)parity;  // level l uses flag l & 1; replays start at even levels, so this is the capture slot's parity
  closure_step_kernel<<<1, 1, 0, s>>>(lvl);
  return hipGetLastError();
}

hipError_t launch_closure_advance(int *lvl, hipStream_t s) {
  closure_advance_kernel<<<1, 1, 0, s>>>(lvl);
  return hipGetLastError();
}

hipError_t launch_copy_u8(uint8_t *dst, const uint8_t *src, size_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  copy_u8_kernel<<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(dst, src, n);
  return hipGetLastError();
}

hipError_t launch_advance(int *l, int *done, const unsigned long long *err, double tol, int max_iter, hipStream_t s) {
  advance_kernel<<<1, 1, 0, s>>>(l, done, err, tol, max_iter);
  return hipGetLastError();
}

}  // namespace prh
