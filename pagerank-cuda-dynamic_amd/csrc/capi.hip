// capi.hip -- the C ABI (include/pagerank_hip.h): contexts, graph upload and the solver driver.
//
// The driver restates pagerankInvokeCuda + pagerankLoopCuU (pagerankCuda.hxx:675-692, :713-849) for one HIP
// stream: per repeat, partition (and initial ranks) = initialization time, affected marking = marking time,
// iterations = computation time, each measured with HIP events. Iterations run without a host round trip per
// iteration: convergence is detected on the device (LoopCtl), and the host polls the device flag only every
// check_interval iterations, so the iteration count and final ranks equal those of the per-iteration check.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pagerank_hip.h"
#include "prh_internal.hpp"

namespace {

constexpr size_t kPinnedBytes = 4096;

template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t cap = 0;  // elements
};

}  // namespace

struct prh_context {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string error;
  size_t limit = 0, used = 0;
  prh_tuning tuning{};
  DevBuf<double> rank[2], contrib[2];
  DevBuf<uint8_t> vaff, naff, buff;
  DevBuf<uint32_t> sched_in, sched_out, tiles, counts, batch, scratch_u32;
  DevBuf<unsigned long long> loop;  // [L] error bits, then one slot holding `done`
  DevBuf<int> flags;                // [0] validation bits, [1] traversal "changed", [2] extension direction
  DevBuf<double> ext_stats;         // per rank workgroup: pull and push costs of the next extension
  DevBuf<uint32_t> ext_chunks;      // initial extension: chunk counts and their scan per batch source
  DevBuf<uint8_t> ext_tmp;          // its scan's temporary storage
  // batched personalized PageRank
  DevBuf<double> ppr_rank[2], ppr_contrib[2], ppr_partial, ppr_mass, ppr_out;
  DevBuf<int32_t> ppr_row;
  DevBuf<uint32_t> ppr_tables;
  DevBuf<unsigned long long> ppr_err;
  DevBuf<prh::PprState> ppr_state;
  // weighted forward / adjoint
  // weighted forward / adjoint, on the degree-ordered layout ("l": layout order, else the caller's order)
  DevBuf<double> wt_x[2], wt_z[2], wt_W, wt_Wl, wt_term, wt_terml, wt_r, wt_rl, wt_gwl, wt_gsl, wt_outM, wt_outN;
  DevBuf<float> wt_wout, wt_wl;
  DevBuf<uint32_t> wt_chunks;       // heavy-row chunk tables of the weighted pull
  DevBuf<uint32_t> wt_sort, wt_sort_tmp;  // degree sort of its low rows
  DevBuf<double> wt_partial;        // their sums
  uint64_t *pinned = nullptr;       // small host-pinned readback area (kPinnedBytes)
  hipEvent_t ev[4] = {};
  bool ktiming = false;
  std::vector<hipEvent_t> kev;  // start/stop pairs per iteration
  uint64_t k_launches = 0;
  double k_ms = 0.0;
};

struct BinnedCache {
  int B = 0;
  bool valid = false;
  bool slot = false;  // the plan keeps the X-entry -> CSR-slot map (weighted solves)
  prh::BinnedPlan plan{};
  std::vector<std::pair<void *, size_t>> bufs;  // its device allocations
};

struct prh_graph {
  prh_context *ctx = nullptr;
  uint32_t N = 0;
  uint64_t M = 0;
  // As uploaded (caller's ids): used by the batched and weighted solvers.
  uint32_t *in_off = nullptr, *in_src = nullptr, *outdeg = nullptr, *out_off = nullptr, *out_dst = nullptr;
  // Raw out-CSR slot of every slot of the layout's in-CSR and out-CSR, built on first weighted use.
  uint32_t *l_in_eid = nullptr, *l_out_eid = nullptr;
  // Degree-ordered layout used by the five solvers: layout row i is vertex perm[i] (in-degree descending, stable),
  // inv is the inverse; in-rows sorted by new source id.
  uint32_t *perm = nullptr, *inv = nullptr;
  uint32_t *l_in_off = nullptr, *l_in_src = nullptr, *l_outdeg = nullptr, *l_out_off = nullptr, *l_out_dst = nullptr;
  std::vector<uint32_t> sorted_deg;  // host: in-degree of layout row i
  float prep_ms = 0;
  // Rank-kernel plan for the current tuning: section table and the chunk tables of the heaviest rows.
  prh_tuning plan_tuning{};
  bool plan_valid = false;
  prh::RankPlan plan{};
  uint32_t *chunk_tables = nullptr;  // item_row, item_beg, item_end, row_first
  size_t chunk_tables_bytes = 0;
  double *chunk_partial = nullptr;
  size_t chunk_partial_bytes = 0;
  uint32_t n_chunked = 0;
  // Source-sliced plan of the plain update, built on first use for the current slices / chunk threshold.
  int sliced_slices = 0;
  uint32_t sliced_huge = 0;
  uint32_t sliced_block_edges = 0;
  bool sliced_valid = false;
  prh::SlicedPlan sliced{};
  std::vector<std::pair<void *, size_t>> sliced_bufs;  // its device allocations
  // Out-degree schedule of the layout's out-CSR for the extension kernels (built on first use): the stable
  // partition [<= 64 | <= kExtendChunk | heavier] and the heavier rows cut into chunks (beg, end, row).
  bool out_valid = false;
  uint32_t *out_order = nullptr, *out_chunks = nullptr;
  uint32_t out_n[3] = {0, 0, 0}, out_n_chunks = 0;
  // Propagation-blocked plans, built on first use for the current block width: of the layout in-CSR (the rank
  // update, the weighted forward) and of its out-CSR (the weighted adjoint).
  BinnedCache bin, bin_t;
};

namespace {

int fail(prh_context *c, int status, const char *fmt, ...) {
  if (c) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    c->error = buf;
  }
  return status;
}

#define PRH_HIP(ctx, expr)                                                                              \
  do {                                                                                                  \
    hipError_t e_ = (expr);                                                                             \
    if (e_ != hipSuccess) return fail(ctx, PRH_ERROR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

#define PRH_CHECK(expr)      \
  do {                       \
    int s_ = (expr);         \
    if (s_ != PRH_SUCCESS) return s_; \
  } while (0)

int dev_alloc(prh_context *c, void **p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (c->limit && c->used + bytes > c->limit)
    return fail(c, PRH_ERROR_OUT_OF_MEMORY, "device memory limit: %zu bytes in use, %zu more requested, limit %zu",
                c->used, bytes, c->limit);
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(c, PRH_ERROR_OUT_OF_MEMORY, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  }
  c->used += bytes;
  return PRH_SUCCESS;
}

void dev_free(prh_context *c, void *p, size_t bytes) {
  if (!p) return;
  (void)hipFree(p);
  c->used -= std::min(c->used, bytes ? bytes : 16);
}

template <class T>
int ensure(prh_context *c, DevBuf<T> &b, size_t n) {
  if (b.cap >= n && b.p) return PRH_SUCCESS;
  if (b.p) {
    PRH_HIP(c, hipStreamSynchronize(c->stream));
    dev_free(c, b.p, b.cap * sizeof(T));
    b.p = nullptr;
    b.cap = 0;
  }
  void *p = nullptr;
  PRH_CHECK(dev_alloc(c, &p, n * sizeof(T)));
  b.p = static_cast<T *>(p);
  b.cap = n;
  return PRH_SUCCESS;
}

template <class T>
void release(prh_context *c, DevBuf<T> &b) {
  dev_free(c, b.p, b.cap * sizeof(T));
  b.p = nullptr;
  b.cap = 0;
}

bool valid_method(int m) { return m >= PRH_STATIC && m <= PRH_PRUNE_DYNAMIC_FRONTIER; }

// Number of rows with in-degree > x in the descending degree list.
uint32_t rows_above(const std::vector<uint32_t> &deg, uint32_t x) {
  return uint32_t(std::lower_bound(deg.begin(), deg.end(), x, [](uint32_t a, uint32_t b) { return a > b; }) -
                  deg.begin());
}

void free_sliced(prh_context *c, prh_graph *g) {
  for (auto &b : g->sliced_bufs) dev_free(c, b.first, b.second);
  g->sliced_bufs.clear();
  g->sliced_valid = false;
}

// Source-sliced plan (prh::SlicedPlan) for the context's tuning, from the degree-ordered layout:
//   1. per row and slice, the number of sources (slice_count);
//   2. per slice, exclusive scans over rows of "has an item" and of the item's edge count (rows >= n_huge);
//   3. the slice-major source array, item offsets, row masks and per-64-row item bases (slice_scatter, wbase);
//   4. chunk tables for the rows above the chunk threshold, whose sources stay in their own edge range grouped
//      by slice, in chunks of huge_degree edges;
//   5. per slice, a section table: its chunks, then bands of items by row in-degree d: a wave per item for
//      d > 64 S, W lanes for d in (W S, 2 W S], W = 32 .. 1 (items hold about d / S sources).
int build_sliced_plan(prh_context *c, prh_graph *g) {
  const prh_tuning &t = c->tuning;
  const int S = t.source_slices;
  if (g->sliced_valid && g->sliced_slices == S && g->sliced_huge == t.huge_degree &&
      g->sliced_block_edges == t.block_edges)
    return PRH_SUCCESS;
  hipStream_t s = c->stream;
  PRH_HIP(c, hipStreamSynchronize(s));
  free_sliced(c, g);
  const uint32_t N = g->N;
  const uint64_t M = g->M;
  const std::vector<uint32_t> &deg = g->sorted_deg;
  const uint32_t n_huge = rows_above(deg, t.huge_degree);
  const uint32_t chunk = std::max<uint32_t>(t.huge_degree, 64);
  const uint64_t kEdgesPerBlock = t.block_edges;
  auto keep = [&](void **p, size_t bytes) -> int {
    PRH_CHECK(dev_alloc(c, p, bytes));
    g->sliced_bufs.emplace_back(*p, bytes ? bytes : 16);
    return PRH_SUCCESS;
  };
  // Temporaries, freed at the end (also on error, by the guard).
  struct Temps {
    prh_context *c;
    std::vector<std::pair<void *, size_t>> v;
    ~Temps() {
      for (auto &b : v) dev_free(c, b.first, b.second);
    }
  } tmp{c, {}};
  auto temp = [&](void **p, size_t bytes) -> int {
    PRH_CHECK(dev_alloc(c, p, bytes));
    tmp.v.emplace_back(*p, bytes ? bytes : 16);
    return PRH_SUCCESS;
  };
  const size_t N1 = size_t(N) + 1;
  uint32_t *cnt, *item, *eoff, *has, *ecnt;
  PRH_CHECK(temp(reinterpret_cast<void **>(&cnt), size_t(S) * N * 4));
  PRH_CHECK(temp(reinterpret_cast<void **>(&item), size_t(S) * N1 * 4));
  PRH_CHECK(temp(reinterpret_cast<void **>(&eoff), size_t(S) * N1 * 4));
  PRH_CHECK(temp(reinterpret_cast<void **>(&has), N1 * 4));
  PRH_CHECK(temp(reinterpret_cast<void **>(&ecnt), N1 * 4));
  size_t scan_bytes = 0;
  PRH_HIP(c, prh::scan_u32(nullptr, scan_bytes, has, item, N1, s));
  void *scan_tmp;
  PRH_CHECK(temp(&scan_tmp, scan_bytes));
  PRH_HIP(c, prh::launch_slice_count(g->l_in_off, g->l_in_src, N, S, cnt, s));
  for (int x = 0; x < S; ++x) {
    PRH_HIP(c, prh::launch_slice_flags(cnt + size_t(x) * N, N, n_huge, has, ecnt, s));
    PRH_HIP(c, prh::scan_u32(scan_tmp, scan_bytes, has, item + size_t(x) * N1, N1, s));
    PRH_HIP(c, prh::scan_u32(scan_tmp, scan_bytes, ecnt, eoff + size_t(x) * N1, N1, s));
  }
  // Values of the scans at given rows, read back: at(a, x, r) = a[x][r].
  auto at = [&](const uint32_t *a, int x, uint32_t r, uint32_t &out) -> int {
    PRH_HIP(c, hipMemcpyAsync(c->pinned, a + size_t(x) * N1 + r, 4, hipMemcpyDeviceToHost, s));
    PRH_HIP(c, hipStreamSynchronize(s));
    out = *reinterpret_cast<uint32_t *>(c->pinned);
    return PRH_SUCCESS;
  };
  uint64_t huge_edges = 0;
  for (uint32_t r = 0; r < n_huge; ++r) huge_edges += deg[r];
  prh::SliceBases B{};
  uint64_t n_items = 0, edge = huge_edges;
  for (int x = 0; x < S; ++x) {
    uint32_t k, e;
    PRH_CHECK(at(item, x, N, k));
    PRH_CHECK(at(eoff, x, N, e));
    B.item[x] = uint32_t(n_items);
    B.edge[x] = uint32_t(edge);
    n_items += k;
    edge += e;
  }
  if (edge != M) return fail(c, PRH_ERROR_DEVICE, "sliced layout: %llu edges placed of %llu",
                             (unsigned long long)edge, (unsigned long long)M);
  prh::SlicedPlan P{};
  P.slices = S;
  P.N = N;
  P.n_huge = n_huge;
  const uint32_t groups = (N - n_huge + prh::kWave - 1) / prh::kWave;
  void *p;
  PRH_CHECK(keep(&p, M * 4));
  uint32_t *s_src = static_cast<uint32_t *>(p);
  PRH_CHECK(keep(&p, (n_items + 1) * 4));
  uint32_t *item_off = static_cast<uint32_t *>(p);
  PRH_CHECK(keep(&p, n_items * 8));
  P.partial = static_cast<double *>(p);
  PRH_CHECK(keep(&p, N));
  uint8_t *mask = static_cast<uint8_t *>(p);
  PRH_CHECK(keep(&p, size_t(groups) * S * 4));
  uint32_t *wbase = static_cast<uint32_t *>(p);
  PRH_HIP(c, prh::launch_slice_scatter(g->l_in_off, g->l_in_src, N, n_huge, S, cnt, item, eoff, B, s_src, item_off,
                                       mask, s));
  *reinterpret_cast<uint32_t *>(c->pinned) = uint32_t(M);
  PRH_HIP(c, hipMemcpyAsync(item_off + n_items, c->pinned, 4, hipMemcpyHostToDevice, s));
  PRH_HIP(c, prh::launch_slice_wbase(item, N, n_huge, S, B, wbase, s));
  P.src = s_src;
  P.item_off = item_off;
  P.mask = mask;
  P.wbase = wbase;
  // Chunk tables: chunk ids in (row, slice, part) order; chunk_list groups them by slice.
  std::vector<uint32_t> hcnt(size_t(S) * n_huge);
  for (int x = 0; x < S && n_huge; ++x)
    PRH_HIP(c, hipMemcpy(hcnt.data() + size_t(x) * n_huge, cnt + size_t(x) * N, size_t(n_huge) * 4,
                         hipMemcpyDeviceToHost));
  std::vector<uint32_t> cbeg, cend, cslice, row_first(size_t(n_huge) + 1);
  uint64_t off = 0;
  for (uint32_t r = 0; r < n_huge; ++r) {
    row_first[r] = uint32_t(cbeg.size());
    for (int x = 0; x < S; ++x) {
      const uint32_t len = hcnt[size_t(x) * n_huge + r];
      for (uint32_t e = 0; e < len; e += chunk) {
        cbeg.push_back(uint32_t(off + e));
        cend.push_back(uint32_t(off + std::min(len, e + chunk)));
        cslice.push_back(uint32_t(x));
      }
      off += len;
    }
  }
  row_first[n_huge] = uint32_t(cbeg.size());
  const uint32_t n_chunks = uint32_t(cbeg.size());
  std::vector<uint32_t> clist;
  std::vector<uint32_t> cfirst(size_t(S) + 1);
  for (int x = 0; x < S; ++x) {
    cfirst[x] = uint32_t(clist.size());
    for (uint32_t i = 0; i < n_chunks; ++i)
      if (cslice[i] == uint32_t(x)) clist.push_back(i);
  }
  cfirst[S] = uint32_t(clist.size());
  PRH_CHECK(keep(&p, (size_t(n_chunks) * 3 + n_huge + 1) * 4));
  uint32_t *ct = static_cast<uint32_t *>(p);
  if (n_chunks) {
    PRH_HIP(c, hipMemcpy(ct, clist.data(), size_t(n_chunks) * 4, hipMemcpyHostToDevice));
    PRH_HIP(c, hipMemcpy(ct + n_chunks, cbeg.data(), size_t(n_chunks) * 4, hipMemcpyHostToDevice));
    PRH_HIP(c, hipMemcpy(ct + 2 * size_t(n_chunks), cend.data(), size_t(n_chunks) * 4, hipMemcpyHostToDevice));
  }
  PRH_HIP(c, hipMemcpy(ct + 3 * size_t(n_chunks), row_first.data(), (size_t(n_huge) + 1) * 4, hipMemcpyHostToDevice));
  PRH_CHECK(keep(&p, size_t(n_chunks) * 8));
  P.chunk_partial = static_cast<double *>(p);
  P.chunk_list = ct;
  P.chunk_beg = ct + n_chunks;
  P.chunk_end = ct + 2 * size_t(n_chunks);
  P.row_chunk_first = ct + 3 * size_t(n_chunks);
  // Section tables. Band b covers rows [rb[b], rb[b+1]); the first band takes a wave per item.
  std::vector<uint32_t> rb{n_huge};
  std::vector<int> width{64};
  for (int w = 32; w >= 1; w /= 2) {
    rb.push_back(std::max(n_huge, rows_above(deg, uint32_t(2 * w * S))));
    width.push_back(w);
  }
  rb.push_back(N);
  uint32_t max_blocks = 0;
  for (int x = 0; x < S; ++x) {
    uint32_t blocks = 0;
    int ns = 0;
    auto add = [&](int kind, int w, uint32_t b, uint32_t e, uint32_t per) {
      if (e <= b) return;
      P.sec[x][ns++] = prh::RankSection{kind, w, b, e, blocks, per};
      blocks += (e - b + per - 1) / per;
    };
    add(prh::kChunkedRows, 0, cfirst[x], cfirst[x + 1], 1);
    for (size_t b = 0; b + 1 < rb.size(); ++b) {
      if (rb[b + 1] <= rb[b]) continue;
      uint32_t k0, k1, e0, e1;
      PRH_CHECK(at(item, x, rb[b], k0));
      PRH_CHECK(at(item, x, rb[b + 1], k1));
      PRH_CHECK(at(eoff, x, rb[b], e0));
      PRH_CHECK(at(eoff, x, rb[b + 1], e1));
      if (k1 <= k0) continue;
      const int w = width[b];
      const uint64_t unit = w == 64 ? prh::kWavesPerBlock : uint64_t(prh::kWavesPerBlock) * (prh::kWave / w);
      const uint64_t nb = std::max<uint64_t>(1, (uint64_t(e1 - e0) + kEdgesPerBlock - 1) / kEdgesPerBlock);
      uint64_t per = (uint64_t(k1 - k0) + nb - 1) / nb;
      per = std::max<uint64_t>(unit, (per + unit - 1) / unit * unit);
      add(w == 64 ? prh::kWaveRows : prh::kGroupRows, w, B.item[x] + k0, B.item[x] + k1, uint32_t(per));
    }
    P.nsec[x] = ns;
    max_blocks = std::max(max_blocks, blocks);
  }
  P.blocks_per_slice = max_blocks;
  PRH_HIP(c, hipStreamSynchronize(s));
  g->sliced = P;
  g->sliced_slices = S;
  g->sliced_huge = t.huge_degree;
  g->sliced_block_edges = t.block_edges;
  g->sliced_valid = true;
  return PRH_SUCCESS;
}

void free_binned(prh_context *c, BinnedCache &bc) {
  for (auto &b : bc.bufs) dev_free(c, b.first, b.second);
  bc.bufs.clear();
  bc.valid = false;
}

// Propagation-blocked plan (prh::BinnedPlan) for blocks and bins of B = tuning.blocking layout rows:
//   1. a stable sort of the pull-CSR slots by source, then a stable sort of that order by key2 = bin(row) *
//      nblocks + block(source), give the slot order (tile, source, row), perm2[j] = CSR slot of slot j;
//   2. per slot: its local destination with the pair and run start flags; an exclusive scan of the pair starts
//      numbers the pairs in slot order, and each pair records its source, tile key and key1 = its source block;
//   3. a stable sort of the pairs by key1 gives the X order (block, bin, source), perm1[k] = pair of X entry k;
//      src16 in X order; runs (tiles) start where the pairs' tile key changes: per run, its first X entry minus
//      its first pair number, and per 64 slots the pair and run of the first slot (with the offsets of that run
//      and the next);
//   4. block starts in X order (lower bounds of the sorted key1) and bin starts from the row degrees cut the two
//      orders into items; split bins get partial rows.
// Builds the plan of the pull CSR (off, idx) with rows of degree rowdeg (host), N rows, M slots, into bc.
int build_binned(prh_context *c, uint32_t N, uint64_t M, const uint32_t *off, const uint32_t *idx,
                 const std::vector<uint32_t> &rowdeg, bool keep_slot, BinnedCache &bc) {
  const int B = c->tuning.blocking;
  if (bc.valid && bc.B == B && (bc.slot || !keep_slot)) return PRH_SUCCESS;
  hipStream_t s = c->stream;
  PRH_HIP(c, hipStreamSynchronize(s));
  free_binned(c, bc);
  const uint64_t nb = (uint64_t(N) + B - 1) / B;
  if (nb * nb > (uint64_t(1) << 32))
    return fail(c, PRH_ERROR_UNSUPPORTED, "blocking %d: %llu blocks per side exceed 32-bit tile keys", B,
                (unsigned long long)nb);
  int key2_bits = 0, key1_bits = 0, src_bits = 0;
  while ((uint64_t(1) << key2_bits) < nb * nb) ++key2_bits;
  while ((uint64_t(1) << key1_bits) < nb) ++key1_bits;
  while ((uint64_t(1) << src_bits) < N) ++src_bits;
  key2_bits = std::max(key2_bits, 1);
  key1_bits = std::max(key1_bits, 1);
  src_bits = std::max(src_bits, 1);
  struct Temps {
    prh_context *c;
    std::vector<std::pair<void *, size_t>> v;
    ~Temps() {
      for (auto &b : v) dev_free(c, b.first, b.second);
    }
  } tmp{c, {}};
  auto temp = [&](auto **p, size_t bytes) -> int {
    void *q = nullptr;
    PRH_CHECK(dev_alloc(c, &q, bytes ? bytes : 16));
    tmp.v.emplace_back(q, bytes ? bytes : 16);
    *p = static_cast<std::remove_pointer_t<decltype(p)>>(q);
    return PRH_SUCCESS;
  };
  auto keep = [&](auto **p, size_t bytes) -> int {
    void *q = nullptr;
    PRH_CHECK(dev_alloc(c, &q, bytes ? bytes : 16));
    bc.bufs.emplace_back(q, bytes ? bytes : 16);
    *p = static_cast<std::remove_pointer_t<decltype(p)>>(q);
    return PRH_SUCCESS;
  };
  prh::BinnedPlan P{};
  P.B = B;
  P.N = N;
  const uint64_t Mp = (M + 127) / 128 * 128;  // slots padded to whole pairs of 64-slot groups
  uint16_t *dst16, *src16 = nullptr;
  uint4 *grp;
  uint32_t *run_off = nullptr;
  double *X = nullptr;
  PRH_CHECK(keep(&dst16, Mp * 2));
  PRH_CHECK(keep(&grp, (M + 63) / 64 * sizeof(uint4)));
  std::vector<uint32_t> bstart(nb + 1, 0);
  uint32_t n_runs = 0;
  uint64_t Q = 0;
  if (M) {
    uint32_t *key, *row, *tk, *pa, *perm2, *starts;
    for (uint32_t **p : {&key, &row, &tk, &pa}) PRH_CHECK(temp(p, M * 4));
    if (keep_slot) PRH_CHECK(keep(&perm2, M * 4));
    else PRH_CHECK(temp(&perm2, M * 4));
    PRH_CHECK(temp(&starts, (nb + 1) * 4));
    size_t b_src = 0, b_key2 = 0, b_key1 = 0, b_scan = 0;
    PRH_HIP(c, prh::sort_pairs_u32(nullptr, b_src, idx, tk, nullptr, pa, M, src_bits, s));
    PRH_HIP(c, prh::sort_pairs_u32(nullptr, b_key2, tk, key, pa, perm2, M, key2_bits, s));
    PRH_HIP(c, prh::sort_pairs_u32(nullptr, b_key1, key, tk, nullptr, pa, M, key1_bits, s));
    PRH_HIP(c, prh::scan_u32(nullptr, b_scan, row, tk, M, s));
    const size_t work_bytes = std::max({b_src, b_key2, b_key1, b_scan});
    uint8_t *work;
    PRH_CHECK(temp(&work, work_bytes));
    size_t wb = work_bytes;
    // 1. slot order: by source (pa), then by tile, stably (perm2); key then holds the sorted tile keys.
    PRH_HIP(c, prh::sort_pairs_u32(work, wb, idx, tk, nullptr, pa, M, src_bits, s));
    PRH_HIP(c, prh::launch_bin_keys(off, idx, N, uint32_t(B), uint32_t(nb), key, row, s));
    PRH_HIP(c, prh::launch_gather_u32(key, pa, M, tk, s));
    wb = work_bytes;
    PRH_HIP(c, prh::sort_pairs_u32(work, wb, tk, key, pa, perm2, M, key2_bits, s));
    // 2. slot flags (pf in pa) and pair numbers (pid in tk).
    PRH_HIP(c, prh::launch_pair_slots(perm2, idx, row, key, M, uint32_t(B), dst16, pa, s));
    if (Mp > M) PRH_HIP(c, hipMemsetAsync(dst16 + M, 0, (Mp - M) * 2, s));
    wb = work_bytes;
    PRH_HIP(c, prh::scan_u32(work, wb, pa, tk, M, s));
    uint32_t last[2];
    PRH_HIP(c, hipMemcpyAsync(&last[0], tk + (M - 1), 4, hipMemcpyDeviceToHost, s));
    PRH_HIP(c, hipMemcpyAsync(&last[1], pa + (M - 1), 4, hipMemcpyDeviceToHost, s));
    PRH_HIP(c, hipStreamSynchronize(s));
    Q = uint64_t(last[0]) + last[1];
    uint32_t *pkey2, *key1, *k1s, *perm1, *inv, *rf, *rid;
    for (uint32_t **p : {&pkey2, &key1, &k1s, &perm1, &inv, &rf, &rid}) PRH_CHECK(temp(p, Q * 4));
    uint32_t *psrc = row;  // row is free once the flags are written
    PRH_HIP(c, prh::launch_pair_fill(pa, tk, perm2, idx, key, M, uint32_t(B), psrc, pkey2, key1, s));
    // 3. X order of the pairs, src16, runs and the group table.
    wb = work_bytes;
    PRH_HIP(c, prh::sort_pairs_u32(work, wb, key1, k1s, nullptr, perm1, Q, key1_bits, s));
    PRH_CHECK(keep(&src16, Q * 2));
    PRH_CHECK(keep(&X, Q * 8));
    PRH_HIP(c, prh::launch_pair_expand(perm1, psrc, Q, uint32_t(B), src16, inv, s));
    PRH_HIP(c, prh::launch_run_flags(pkey2, Q, rf, s));
    wb = work_bytes;
    PRH_HIP(c, prh::scan_u32(work, wb, rf, rid, Q, s));
    PRH_HIP(c, hipMemcpyAsync(&last[0], rid + (Q - 1), 4, hipMemcpyDeviceToHost, s));
    PRH_HIP(c, hipMemcpyAsync(&last[1], rf + (Q - 1), 4, hipMemcpyDeviceToHost, s));
    PRH_HIP(c, hipStreamSynchronize(s));
    n_runs = last[0] + last[1];
    PRH_CHECK(keep(&run_off, size_t(n_runs) * 4));
    PRH_HIP(c, prh::launch_pair_runs(rf, rid, inv, Q, run_off, s));
    PRH_HIP(c, prh::launch_group_table(pa, tk, rf, rid, run_off, n_runs, M, grp, s));
    // 4. block starts in X order.
    PRH_HIP(c, prh::launch_lower_bounds(k1s, Q, uint32_t(nb), starts, s));
    PRH_HIP(c, hipMemcpyAsync(bstart.data(), starts, (nb + 1) * 4, hipMemcpyDeviceToHost, s));
    PRH_HIP(c, hipStreamSynchronize(s));
    if (keep_slot) P.slot = perm2;
  } else {
    PRH_CHECK(keep(&src16, 2));
    PRH_CHECK(keep(&X, 16));
  }
  // Items: expand items of at most 32 per CU over all pairs, accumulate items of at most 2 per CU over all slots
  // (fewer split bins, hence fewer partial rows); both at least 16K entries. Measured on RMAT-24 (B = 8192) per
  // update: expand spans of 6.4K, 12.9K, 16K, 32K and 64K entries 763, 712, 709, 724 and 744 us; accumulate
  // spans of M/(k CUs) for k = 8, 4, 3, 2, 3/2, 1 and 1/2 733, 709, 711, 696, 710, 807 and 1195 us (k = 2: as many
  // items as the two resident workgroups of every CU).
  hipDeviceProp_t prop{};
  PRH_HIP(c, hipGetDeviceProperties(&prop, c->device));
  const uint64_t cus = uint64_t(std::max(prop.multiProcessorCount, 1));
  const uint64_t span_ex = std::max<uint64_t>(Q / (cus * 32), 16384);
  const uint64_t span = std::max<uint64_t>(M / (cus * 2), 16384);
  std::vector<uint4> ex, ac, multi;
  for (uint64_t b = 0; b < nb; ++b)
    for (uint64_t k = bstart[b]; k < bstart[b + 1]; k += span_ex)
      ex.push_back(make_uint4(uint32_t(b), uint32_t(k), uint32_t(std::min<uint64_t>(bstart[b + 1], k + span_ex)), 0));
  uint32_t rows = 0;
  uint64_t j = 0;
  for (uint64_t b = 0; b < nb; ++b) {
    const uint64_t r1 = std::min<uint64_t>(N, (b + 1) * uint64_t(B));
    uint64_t e = 0;
    for (uint64_t r = b * uint64_t(B); r < r1; ++r) e += rowdeg[r];
    const uint64_t n = std::max<uint64_t>(1, (e + span - 1) / span);
    if (n == 1) {
      ac.push_back(make_uint4(uint32_t(b), uint32_t(j), uint32_t(j + e), prh::kWholeBin));
    } else {
      // Finish units: a thread per vertex over 1024 vertices, or (more than 16 rows) 16 waves over 64 vertices.
      const uint32_t bin_rows = uint32_t(r1 - b * uint64_t(B));
      const bool wide = n > 16;
      const uint32_t step = wide ? 64u : uint32_t(prh::kBinBlock);
      for (uint32_t f = 0; f < bin_rows; f += step)
        multi.push_back(make_uint4(uint32_t(b), rows, uint32_t(n), f | (wide ? 0x80000000u : 0u)));
      for (uint64_t q = 0; q < n; ++q) {
        const uint64_t j0 = j + e * q / n, j1 = j + e * (q + 1) / n;
        ac.push_back(make_uint4(uint32_t(b), uint32_t(j0), uint32_t(j1), rows++));
      }
    }
    j += e;
  }
  if (j != M) return fail(c, PRH_ERROR_DEVICE, "binned layout: %llu slots of %llu", (unsigned long long)j,
                          (unsigned long long)M);
  uint4 *tables;
  PRH_CHECK(keep(&tables, (ex.size() + ac.size() + multi.size()) * sizeof(uint4)));
  if (!ex.empty()) PRH_HIP(c, hipMemcpy(tables, ex.data(), ex.size() * sizeof(uint4), hipMemcpyHostToDevice));
  PRH_HIP(c, hipMemcpy(tables + ex.size(), ac.data(), ac.size() * sizeof(uint4), hipMemcpyHostToDevice));
  if (!multi.empty())
    PRH_HIP(c, hipMemcpy(tables + ex.size() + ac.size(), multi.data(), multi.size() * sizeof(uint4),
                         hipMemcpyHostToDevice));
  double *partial;
  PRH_CHECK(keep(&partial, size_t(rows) * B * 8));
  P.src16 = src16;
  P.dst16 = dst16;
  P.X = X;
  P.run_off = run_off;
  P.grp = grp;
  P.n_runs = n_runs;
  P.Q = Q;
  P.ex_item = tables;
  P.ac_item = tables + ex.size();
  P.multi = tables + ex.size() + ac.size();
  P.partial = partial;
  P.n_ex = uint32_t(ex.size());
  P.n_ac = uint32_t(ac.size());
  P.n_multi = uint32_t(multi.size());
  PRH_HIP(c, hipStreamSynchronize(s));
  bc.plan = P;
  bc.B = B;
  bc.valid = true;
  bc.slot = keep_slot;
  return PRH_SUCCESS;
}

// The plan of the layout in-CSR (rows in in-degree order).
int build_binned_plan(prh_context *c, prh_graph *g, bool keep_slot = false) {
  return build_binned(c, g->N, g->M, g->l_in_off, g->l_in_src, g->sorted_deg, keep_slot, g->bin);
}

// The plan of the layout out-CSR (the transpose: rows are sources), for the weighted adjoint.
int build_binned_plan_t(prh_context *c, prh_graph *g) {
  if (g->bin_t.valid && g->bin_t.B == c->tuning.blocking) return PRH_SUCCESS;
  std::vector<uint32_t> outdeg(g->N);
  if (g->N) PRH_HIP(c, hipMemcpy(outdeg.data(), g->l_outdeg, size_t(g->N) * 4, hipMemcpyDeviceToHost));
  return build_binned(c, g->N, g->M, g->l_out_off, g->l_out_dst, outdeg, true, g->bin_t);
}

// Rank-kernel plan over the degree-ordered layout for the context's tuning (cached in the graph):
//   rows with in-degree > huge_degree: chunks of huge_degree edges, one workgroup each;
//   switch_degree < in-degree <= huge_degree: one wave per row;
//   in-degree <= switch_degree: bands of equal lanes-per-row W, W = pow2ceil(d)/2 in 1..32 (or the fixed
//   low_width), each band a contiguous row range.
int build_rank_plan(prh_context *c, prh_graph *g) {
  const prh_tuning &t = c->tuning;
  if (g->plan_valid && std::memcmp(&g->plan_tuning, &t, sizeof(t)) == 0) return PRH_SUCCESS;
  const uint32_t N = g->N;
  const std::vector<uint32_t> &deg = g->sorted_deg;
  const uint64_t kEdgesPerBlock = t.block_edges;
  PRH_HIP(c, hipStreamSynchronize(c->stream));  // the old tables may still be in use
  auto above = [&](uint32_t x) { return rows_above(deg, x); };
  auto edges = [&](uint32_t b, uint32_t e) {
    uint64_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += deg[i];
    return s;
  };
  prh::RankPlan p{};
  uint32_t blocks = 0;
  auto add = [&](int kind, int width, uint32_t b, uint32_t e, uint32_t per_block, uint32_t nblocks) {
    if (e <= b || p.nsec >= prh::kMaxSections) return;
    p.sec[p.nsec++] = prh::RankSection{kind, width, b, e, blocks, per_block};
    blocks += nblocks;
  };
  // Chunked rows and their tables.
  const uint32_t chunk = std::max<uint32_t>(t.huge_degree, 64);
  const uint32_t n_chunked = above(t.huge_degree);
  std::vector<uint32_t> item_row, item_beg, item_end, row_first(size_t(n_chunked) + 1);
  uint64_t off = 0;
  for (uint32_t r = 0; r < n_chunked; ++r) {
    row_first[r] = uint32_t(item_row.size());
    for (uint64_t e = 0; e < deg[r]; e += chunk) {
      item_row.push_back(r);
      item_beg.push_back(uint32_t(off + e));
      item_end.push_back(uint32_t(off + std::min<uint64_t>(deg[r], e + chunk)));
    }
    off += deg[r];
  }
  row_first[n_chunked] = uint32_t(item_row.size());
  const uint32_t n_items = uint32_t(item_row.size());
  // (Re)allocate and upload the chunk tables.
  const size_t tb = (3 * size_t(n_items) + n_chunked + 1) * 4;
  dev_free(c, g->chunk_tables, g->chunk_tables_bytes);
  dev_free(c, g->chunk_partial, g->chunk_partial_bytes);
  g->chunk_tables = nullptr;
  g->chunk_partial = nullptr;
  g->plan_valid = false;
  PRH_CHECK(dev_alloc(c, reinterpret_cast<void **>(&g->chunk_tables), tb));
  g->chunk_tables_bytes = tb;
  PRH_CHECK(dev_alloc(c, reinterpret_cast<void **>(&g->chunk_partial), std::max<size_t>(16, size_t(n_items) * 8)));
  g->chunk_partial_bytes = std::max<size_t>(16, size_t(n_items) * 8);
  uint32_t *tp = g->chunk_tables;
  if (n_items) {
    PRH_HIP(c, hipMemcpy(tp, item_row.data(), size_t(n_items) * 4, hipMemcpyHostToDevice));
    PRH_HIP(c, hipMemcpy(tp + n_items, item_beg.data(), size_t(n_items) * 4, hipMemcpyHostToDevice));
    PRH_HIP(c, hipMemcpy(tp + 2 * size_t(n_items), item_end.data(), size_t(n_items) * 4, hipMemcpyHostToDevice));
  }
  PRH_HIP(c, hipMemcpy(tp + 3 * size_t(n_items), row_first.data(), (size_t(n_chunked) + 1) * 4,
                       hipMemcpyHostToDevice));
  p.item_row = tp;
  p.item_beg = tp + n_items;
  p.item_end = tp + 2 * size_t(n_items);
  p.row_first = tp + 3 * size_t(n_items);
  p.partial = g->chunk_partial;
  p.n_chunked = n_chunked;
  add(prh::kChunkedRows, 0, 0, n_items, 1, n_items);
  // Wave rows.
  const uint32_t w_end = above(t.switch_degree);
  auto row_section = [&](int kind, int width, uint32_t b, uint32_t e) {
    if (e <= b) return;
    const uint32_t unit = kind == prh::kWaveRows ? prh::kWavesPerBlock : prh::kWavesPerBlock * (prh::kWave / width);
    const uint64_t nb = std::max<uint64_t>(1, (edges(b, e) + kEdgesPerBlock - 1) / kEdgesPerBlock);
    uint64_t per = (uint64_t(e - b) + nb - 1) / nb;
    per = std::max<uint64_t>(unit, (per + unit - 1) / unit * unit);
    add(kind, width, b, e, uint32_t(per), uint32_t((uint64_t(e - b) + per - 1) / per));
  };
  row_section(prh::kWaveRows, 64, n_chunked, std::max(n_chunked, w_end));
  // Group-row bands.
  const uint32_t lb = std::max(n_chunked, w_end);
  if (t.low_width) {
    row_section(prh::kGroupRows, t.low_width, lb, N);
  } else {
    uint32_t b = lb;
    for (int width = 32; width >= 1; width /= 2) {
      // rows with degree in (width, 2*width] get `width` lanes; width 1 also takes degrees 1 and 2
      const uint32_t e = width == 1 ? N : std::max(b, above(uint32_t(width)));
      row_section(prh::kGroupRows, width, b, e);
      b = e;
    }
  }
  p.blocks = blocks;
  g->plan = p;
  g->plan_tuning = t;
  g->plan_valid = true;
  g->n_chunked = n_chunked;
  return PRH_SUCCESS;
}

// Batch arrays staged in the context (host or device source), checked against N and mapped to layout ids.
struct DevBatch {
  const uint32_t *delu = nullptr, *delv = nullptr, *insu = nullptr;
  size_t ND = 0, NI = 0;
};

int stage_batch(prh_context *c, const prh_graph *g, const uint32_t *du, const uint32_t *dv, size_t nd,
                const uint32_t *iu, size_t ni, bool host, DevBatch &b) {
  b.ND = nd;
  b.NI = ni;
  if (nd + ni == 0) return PRH_SUCCESS;
  if ((nd && (!du || !dv)) || (ni && !iu)) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "null batch array");
  PRH_CHECK(ensure(c, c->batch, 2 * nd + ni));
  uint32_t *d = c->batch.p;
  const hipMemcpyKind kind = host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
  if (nd) {
    PRH_HIP(c, hipMemcpyAsync(d, du, nd * 4, kind, c->stream));
    PRH_HIP(c, hipMemcpyAsync(d + nd, dv, nd * 4, kind, c->stream));
  }
  if (ni) PRH_HIP(c, hipMemcpyAsync(d + 2 * nd, iu, ni * 4, kind, c->stream));
  const uint32_t N = g->N;
  PRH_CHECK(ensure(c, c->flags, 4));
  PRH_HIP(c, prh::launch_fill_i32(c->flags.p, 0, 1, c->stream));
  PRH_HIP(c, prh::launch_validate_ids(d, 2 * nd + ni, N, c->flags.p, 1, c->stream));
  PRH_HIP(c, hipMemcpyAsync(c->pinned, c->flags.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  PRH_HIP(c, hipStreamSynchronize(c->stream));
  if (*reinterpret_cast<int *>(c->pinned)) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "batch vertex id >= N (%u)", N);
  PRH_HIP(c, prh::launch_map_ids(d, 2 * nd + ni, g->inv, c->stream));
  b.delu = d;
  b.delv = d + nd;
  b.insu = d + 2 * nd;
  return PRH_SUCCESS;
}

// Out-degree schedule of the layout's out-CSR for the extension kernels, built once per graph (one readback of
// the part sizes, as the reference reads its partition point back, pagerankCuda.hxx:633), with the chunk table of
// the rows above kExtendChunk edges, so that a hub's edges are spread over many waves.
int out_schedule(prh_context *c, prh_graph *g, prh::ExtendArgs &E) {
  const uint32_t N = g->N;
  if (!g->out_valid) {
    hipStream_t s = c->stream;
    if (!g->out_order) PRH_CHECK(dev_alloc(c, reinterpret_cast<void **>(&g->out_order), std::max<size_t>(N, 1) * 4));
    PRH_CHECK(ensure(c, c->tiles, size_t(3) * prh::partition_tiles(N)));
    PRH_CHECK(ensure(c, c->counts, 8));
    PRH_HIP(c, prh::launch_partition(g->l_out_off, N, 64u, prh::kExtendChunk, g->out_order, c->tiles.p, c->counts.p,
                                     s));
    PRH_HIP(c, hipMemcpyAsync(c->pinned, c->counts.p, 12, hipMemcpyDeviceToHost, s));
    PRH_HIP(c, hipStreamSynchronize(s));
    const uint32_t *cnt = reinterpret_cast<const uint32_t *>(c->pinned);
    g->out_n[0] = cnt[0];
    g->out_n[1] = cnt[1];
    g->out_n[2] = cnt[2];
    const uint32_t nh = g->out_n[2];
    g->out_n_chunks = 0;
    if (nh) {
      uint32_t *ranges = nullptr;
      PRH_CHECK(dev_alloc(c, reinterpret_cast<void **>(&ranges), size_t(2) * nh * 4));
      std::vector<uint32_t> be(size_t(2) * nh);
      hipError_t e = prh::launch_gather_ranges(g->l_out_off, g->out_order + cnt[0] + cnt[1], nh, ranges, ranges + nh,
                                               s);
      if (e == hipSuccess) e = hipMemcpyAsync(be.data(), ranges, be.size() * 4, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      dev_free(c, ranges, size_t(2) * nh * 4);
      if (e != hipSuccess) return fail(c, PRH_ERROR_DEVICE, "out schedule: %s", hipGetErrorString(e));
      std::vector<uint32_t> cb, ce, cr;
      std::vector<uint32_t> rows(nh);
      PRH_HIP(c, hipMemcpy(rows.data(), g->out_order + cnt[0] + cnt[1], size_t(nh) * 4, hipMemcpyDeviceToHost));
      for (uint32_t h = 0; h < nh; ++h)
        for (uint64_t x = be[h]; x < be[nh + h]; x += prh::kExtendChunk) {
          cb.push_back(uint32_t(x));
          ce.push_back(uint32_t(std::min<uint64_t>(be[nh + h], x + prh::kExtendChunk)));
          cr.push_back(rows[h]);
        }
      const size_t nc = cb.size();
      PRH_CHECK(dev_alloc(c, reinterpret_cast<void **>(&g->out_chunks), std::max<size_t>(3 * nc, 1) * 4));
      PRH_HIP(c, hipMemcpy(g->out_chunks, cb.data(), nc * 4, hipMemcpyHostToDevice));
      PRH_HIP(c, hipMemcpy(g->out_chunks + nc, ce.data(), nc * 4, hipMemcpyHostToDevice));
      PRH_HIP(c, hipMemcpy(g->out_chunks + 2 * nc, cr.data(), nc * 4, hipMemcpyHostToDevice));
      g->out_n_chunks = uint32_t(nc);
    }
    g->out_valid = true;
  }
  E.sched = prh::Schedule{g->out_order, g->out_n[0], g->out_n[1], g->out_n[2]};
  if (g->out_n_chunks) {
    const uint32_t nc = g->out_n_chunks;
    E.chunk_beg = g->out_chunks;
    E.chunk_end = g->out_chunks + nc;
    E.chunk_row = g->out_chunks + 2 * nc;
    E.n_chunks = nc;
  }
  return PRH_SUCCESS;
}

// Marking phase of the dynamic methods (pagerankCuda.hxx:821 with the fm lambdas :926-928, :955-957, :984-986),
// on the layout's out-CSR with layout ids.
// xo: the out-degree schedule and chunk table (out_schedule), needed by the traversal levels only.
int mark_affected(prh_context *c, const prh_graph *g, int method, const DevBatch &b, const prh::ExtendArgs &xo) {
  const uint32_t N = g->N;
  hipStream_t s = c->stream;
  PRH_HIP(c, prh::launch_fill_u8(c->vaff.p, 0, N, s));
  PRH_HIP(c, prh::launch_fill_u8(c->naff.p, 0, N, s));
  if (b.ND + b.NI == 0) return PRH_SUCCESS;
  PRH_HIP(c, prh::launch_mark_partial(c->naff.p, c->vaff.p, b.delu, b.delv, b.ND, b.insu, b.NI, s));
  // The flagged vertices are the batch's sources: extend from the batch list.
  const size_t nb = b.ND + b.NI + 1;
  PRH_CHECK(ensure(c, c->ext_chunks, 2 * nb));
  size_t tmp_bytes = 0;
  PRH_HIP(c, prh::launch_extend_batch(c->vaff.p, c->naff.p, b.delu, b.ND, b.insu, b.NI, g->l_out_off, g->l_out_dst,
                                      c->ext_chunks.p, c->ext_chunks.p + nb, nullptr, tmp_bytes, s));
  PRH_CHECK(ensure(c, c->ext_tmp, std::max<size_t>(tmp_bytes, 16)));
  PRH_HIP(c, prh::launch_extend_batch(c->vaff.p, c->naff.p, b.delu, b.ND, b.insu, b.NI, g->l_out_off, g->l_out_dst,
                                      c->ext_chunks.p, c->ext_chunks.p + nb, c->ext_tmp.p, tmp_bytes, s));
  if (method != PRH_DYNAMIC_TRAVERSAL) return PRH_SUCCESS;
  // Traversal closure: breadth-first levels from every marked vertex until no vertex is newly marked.
  PRH_CHECK(ensure(c, c->flags, 4));
  uint8_t *front = c->naff.p, *fresh = c->buff.p;
  PRH_HIP(c, hipMemcpyAsync(front, c->vaff.p, N, hipMemcpyDeviceToDevice, s));
  for (;;) {
    PRH_HIP(c, prh::launch_fill_u8(fresh, 0, N, s));
    PRH_HIP(c, prh::launch_fill_i32(c->flags.p + 1, 0, 1, s));
    prh::ExtendArgs T = xo;
    T.vaff = c->vaff.p;
    T.naff = front;
    T.off = g->l_out_off;
    T.dst = g->l_out_dst;
    T.fresh = fresh;
    T.changed = c->flags.p + 1;
    T.direction = nullptr;
    PRH_HIP(c, prh::launch_extend(T, nullptr, nullptr, 0, 0.0, s));
    PRH_HIP(c, hipMemcpyAsync(c->pinned, c->flags.p + 1, sizeof(int), hipMemcpyDeviceToHost, s));
    PRH_HIP(c, hipStreamSynchronize(s));
    if (*reinterpret_cast<int *>(c->pinned) == 0) break;
    std::swap(front, fresh);
  }
  return PRH_SUCCESS;
}

int ensure_workspace(prh_context *c, uint32_t N, bool dynamic, int L) {
  for (int i = 0; i < 2; ++i) {
    PRH_CHECK(ensure(c, c->rank[i], N));
    PRH_CHECK(ensure(c, c->contrib[i], N));
  }
  if (dynamic) {
    PRH_CHECK(ensure(c, c->sched_out, N));
    PRH_CHECK(ensure(c, c->vaff, N));
    PRH_CHECK(ensure(c, c->naff, (size_t(N) + 3) / 4 * 4));  // whole 32-bit words (the batch extension's claims)
    PRH_CHECK(ensure(c, c->buff, N));
  }
  PRH_CHECK(ensure(c, c->loop, size_t(L) + 1));
  PRH_CHECK(ensure(c, c->flags, 4));
  return PRH_SUCCESS;
}

// Start ranks in layout order: rank[0][i] = q[perm[i]] (q staged through contrib[1], which the initialisation
// overwrites).
int load_initial_ranks(prh_context *c, const prh_graph *g, const double *q, bool host) {
  const double *src = q;
  if (host) {
    PRH_HIP(c, hipMemcpyAsync(c->contrib[1].p, q, size_t(g->N) * 8, hipMemcpyHostToDevice, c->stream));
    src = c->contrib[1].p;
  }
  PRH_HIP(c, prh::launch_gather_f64(src, g->perm, g->N, c->rank[0].p, c->stream));
  return PRH_SUCCESS;
}

int solve_impl(prh_context *c, const prh_graph *g, int method, const prh_options *o, const uint32_t *du,
               const uint32_t *dv, size_t nd, const uint32_t *iu, size_t ni, const double *q, double *ranks,
               prh_stats *st, bool host) {
  if (!c) return PRH_ERROR_INVALID_ARGUMENT;
  if (!g || !o || !valid_method(method)) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "null graph/options or bad method");
  if (g->ctx != c) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "graph belongs to another context");
  if (o->repeat < 1 || o->max_iterations < 0 || !(o->tolerance >= 0) || !(o->damping > 0 && o->damping < 1))
    return fail(c, PRH_ERROR_INVALID_ARGUMENT, "bad options (repeat >= 1, max_iterations >= 0, tolerance >= 0, 0 < damping < 1)");
  if (method != PRH_STATIC && !q) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "initial ranks required for this method");
  const uint32_t N = g->N;
  prh_stats stats{};
  if (N == 0) {  // the reference returns an empty result for an empty graph (pagerankCuda.hxx:869)
    if (st) *st = stats;
    return PRH_SUCCESS;
  }
  if (!ranks) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "null ranks");
  const bool dynamic = method >= PRH_DYNAMIC_TRAVERSAL;
  const bool frontier = method >= PRH_DYNAMIC_FRONTIER;
  if (dynamic && !g->out_off) return fail(c, PRH_ERROR_UNSUPPORTED, "dynamic method needs the graph's out-CSR");
  const int L = o->max_iterations;
  const int variant = method == PRH_PRUNE_DYNAMIC_FRONTIER ? prh::kPrune
                      : method == PRH_DYNAMIC_FRONTIER    ? prh::kFrontier
                      : method == PRH_DYNAMIC_TRAVERSAL   ? prh::kTraversal
                                                          : prh::kPlain;
  hipStream_t s = c->stream;
  PRH_HIP(c, hipSetDevice(c->device));
  PRH_CHECK(ensure_workspace(c, N, dynamic, L));
  prh_graph *gm = const_cast<prh_graph *>(g);
  const bool binned = c->tuning.blocking > 0;
  const bool sliced = !binned && variant == prh::kPlain && c->tuning.source_slices > 0;
  if (binned) PRH_CHECK(build_binned_plan(c, gm));
  else if (sliced) PRH_CHECK(build_sliced_plan(c, gm));
  else PRH_CHECK(build_rank_plan(c, gm));
  DevBatch b;
  if (dynamic) PRH_CHECK(stage_batch(c, g, du, dv, nd, iu, ni, host, b));
  if (q && !host) {  // device initial ranks: reject NaN/Inf (the reference would propagate them)
    PRH_HIP(c, prh::launch_fill_i32(c->flags.p, 0, 1, s));
    PRH_HIP(c, prh::launch_check_finite(q, N, c->flags.p, 1, s));
    PRH_HIP(c, hipMemcpyAsync(c->pinned, c->flags.p, sizeof(int), hipMemcpyDeviceToHost, s));
    PRH_HIP(c, hipStreamSynchronize(s));
    if (*reinterpret_cast<int *>(c->pinned)) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "initial ranks not finite");
  }
  if (q && host)
    for (uint32_t v = 0; v < N; ++v)
      if (!std::isfinite(q[v])) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "initial ranks not finite (vertex %u)", v);
  if (c->ktiming && c->kev.size() < size_t(2) * L) {
    size_t old = c->kev.size();
    c->kev.resize(size_t(2) * L);
    for (size_t i = old; i < c->kev.size(); ++i) PRH_HIP(c, hipEventCreate(&c->kev[i]));
  }
  double ti = 0, tm = 0, tc = 0;
  int iterations = 0;
  unsigned long long *err = c->loop.p;
  int *done = reinterpret_cast<int *>(c->loop.p + L);
  const int check = std::max(1, c->tuning.check_interval);
  const prh::RankPlan &plan = g->plan;
  // Extension direction (frontier variants): 0 push only, 1 chosen per iteration from the rank launch's costs,
  // 2 pull only. The pull form splits rows at in-degree 64 (a wave each) and 4 (8 lanes each).
  const int dir_mode = frontier ? c->tuning.extension_direction : 0;
  const bool dir_auto = dir_mode == 1 && !binned;
  // Per-workgroup extension costs: one slot per workgroup of the launches that finish vertices.
  const uint32_t stat_slots = binned ? g->bin.plan.n_ac + g->bin.plan.n_multi
                                     : plan.blocks + (plan.n_chunked ? prh::finish_chunked_grid(plan.n_chunked) : 0);
  if (dir_auto) PRH_CHECK(ensure(c, c->ext_stats, 2 * size_t(std::max<uint32_t>(stat_slots, 1))));
  int *dir = frontier ? c->flags.p + 2 : nullptr;
  const prh::PullArgs pull{c->vaff.p, c->naff.p, g->l_in_off, g->l_in_src, rows_above(g->sorted_deg, 64),
                           rows_above(g->sorted_deg, 4), N, dir};
  for (int rep = 0; rep < o->repeat; ++rep) {
    if (q) PRH_CHECK(load_initial_ranks(c, g, q, host));  // untimed, as the reference's q copy (pagerankCuda.hxx:810)
    PRH_HIP(c, hipEventRecord(c->ev[0], s));
    // Initialisation. The degree classes the reference partitions here (pagerankCuda.hxx:811-814) are the row
    // ranges of the layout, fixed when the graph was prepared; the extension kernels' out-degree split is per solve.
    // The out-degree schedule serves the traversal levels and the row plan's frontier extension; the initial
    // extension runs from the batch list and the blocked plan extends inside its update.
    prh::ExtendArgs xo{};
    if (dynamic && (method == PRH_DYNAMIC_TRAVERSAL || (frontier && !binned))) PRH_CHECK(out_schedule(c, gm, xo));
    PRH_HIP(c, prh::launch_init_ranks(c->rank[0].p, c->rank[1].p, c->contrib[0].p, c->contrib[1].p, g->l_outdeg, N,
                                      q != nullptr, s));
    PRH_HIP(c, hipEventRecord(c->ev[1], s));
    if (dynamic) PRH_CHECK(mark_affected(c, g, method, b, xo));
    PRH_HIP(c, hipEventRecord(c->ev[2], s));
    // Iterations.
    PRH_HIP(c, prh::launch_fill_u64(err, 0ull, size_t(L) + 1, s));
    if (frontier && dir_mode != 1) PRH_HIP(c, prh::launch_fill_i32(dir, dir_mode == 2 ? 1 : 0, 1, s));
    const double P = o->damping;
    const double C0 = (1 - P) / N;
    int n = -1, launched = 0;
    for (int l = 0; l < L; ++l) {
      prh::RankArgs A{c->rank[(l + 1) & 1].p, c->rank[l & 1].p, c->contrib[(l + 1) & 1].p, c->contrib[l & 1].p,
                      c->naff.p, c->vaff.p, g->l_in_off, g->l_in_src, g->l_outdeg, C0, P, o->frontier_tolerance,
                      o->prune_tolerance, dir_auto ? c->ext_stats.p : nullptr};
      if (c->ktiming) PRH_HIP(c, hipEventRecord(c->kev[2 * l], s));
      hipEvent_t mark = c->ktiming ? c->kev[2 * l + 1] : nullptr;
      if (binned)
        PRH_HIP(c, prh::launch_binned_update(variant, A, g->bin.plan, err, done, l, o->tolerance, frontier && l > 0, s,
                                             mark));
      else if (sliced) PRH_HIP(c, prh::launch_sliced_update(A, g->sliced, err, done, l, o->tolerance, s, mark));
      else PRH_HIP(c, prh::launch_rank_update(variant, A, plan, err, done, l, o->tolerance, s, mark));
      if (frontier && !binned) {  // the blocked plan extends inside the next iteration's update
        if (dir_auto) PRH_HIP(c, prh::launch_direction(c->ext_stats.p, stat_slots, dir, err, done, l, o->tolerance, s));
        prh::ExtendArgs E = xo;
        E.vaff = c->vaff.p;
        E.naff = c->naff.p;
        E.off = g->l_out_off;
        E.dst = g->l_out_dst;
        E.fresh = nullptr;
        E.changed = nullptr;
        E.direction = dir;
        PRH_HIP(c, prh::launch_extend(E, err, done, l, o->tolerance, s));
        if (dir_mode != 0) PRH_HIP(c, prh::launch_extend_pull(pull, err, done, l, o->tolerance, s));
      }
      launched = l + 1;
      if (launched % check == 0 || launched == L) {
        PRH_HIP(c, hipMemcpyAsync(c->pinned, err + l, 8, hipMemcpyDeviceToHost, s));
        PRH_HIP(c, hipMemcpyAsync(c->pinned + 1, done, sizeof(int), hipMemcpyDeviceToHost, s));
        PRH_HIP(c, hipStreamSynchronize(s));
        double el;
        std::memcpy(&el, c->pinned, 8);
        const int d = *reinterpret_cast<const int *>(c->pinned + 1);
        if (d > 0) { n = d; break; }
        if (el < o->tolerance) { n = launched; break; }
      }
    }
    if (n < 0) n = launched;
    PRH_HIP(c, hipEventRecord(c->ev[3], s));
    PRH_HIP(c, hipEventSynchronize(c->ev[3]));
    float t01 = 0, t12 = 0, t23 = 0;
    PRH_HIP(c, hipEventElapsedTime(&t01, c->ev[0], c->ev[1]));
    PRH_HIP(c, hipEventElapsedTime(&t12, c->ev[1], c->ev[2]));
    PRH_HIP(c, hipEventElapsedTime(&t23, c->ev[2], c->ev[3]));
    ti += t01;
    tm += t12;
    tc += t23;
    iterations = n;
    if (c->ktiming) {
      for (int l = 0; l < n; ++l) {
        float ms = 0;
        PRH_HIP(c, hipEventElapsedTime(&ms, c->kev[2 * l], c->kev[2 * l + 1]));
        c->k_ms += ms;
        ++c->k_launches;
      }
    }
  }
  // Final ranks are in the buffer written by the last iteration (pagerankCuda.hxx:688, :826); back to the
  // caller's vertex order (scatter through perm), as the reference scatters to span (:827).
  const double *res = c->rank[iterations & 1].p;
  if (host) {
    double *tmp = c->contrib[(iterations + 1) & 1].p;
    PRH_HIP(c, prh::launch_scatter_f64(res, g->perm, N, tmp, s));
    PRH_HIP(c, hipMemcpyAsync(ranks, tmp, size_t(N) * 8, hipMemcpyDeviceToHost, s));
  } else {
    PRH_HIP(c, prh::launch_scatter_f64(res, g->perm, N, ranks, s));
  }
  PRH_HIP(c, hipStreamSynchronize(s));
  stats.iterations = iterations;
  stats.initialization_time = float(ti / o->repeat);
  stats.marking_time = float(tm / o->repeat);
  stats.computation_time = float(tc / o->repeat);
  stats.time = float((ti + tm + tc) / o->repeat);
  if (st) *st = stats;
  return PRH_SUCCESS;
}

// Capture `body` (launches on the context stream only) into an executable graph.
template <class F>
int capture_graph(prh_context *c, F body, hipGraphExec_t *exec) {
  hipGraph_t graph = nullptr;
  PRH_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  const int st = body();
  const hipError_t e = hipStreamEndCapture(c->stream, &graph);
  if (st != PRH_SUCCESS || e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return st != PRH_SUCCESS ? st : fail(c, PRH_ERROR_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
  }
  const hipError_t ie = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ie != hipSuccess) return fail(c, PRH_ERROR_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(ie));
  return PRH_SUCCESS;
}

// Edge chunk of a heavy vertex for the batched solver.
constexpr uint32_t kPprChunk = 1024;

int ppr_impl(prh_context *c, const prh_graph *g, int k, const uint32_t *so, const uint32_t *seeds,
             const prh_options *o, double *ranks, int *iterations, prh_stats *st) {
  if (!c) return PRH_ERROR_INVALID_ARGUMENT;
  if (!g || !o || !so || !ranks || !iterations || k < 1 || k > prh::kPprCols)
    return fail(c, PRH_ERROR_INVALID_ARGUMENT, "null argument or k outside 1..64");
  if (g->ctx != c) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "graph belongs to another context");
  if (o->max_iterations < 0 || !(o->tolerance >= 0) || !(o->damping > 0 && o->damping < 1))
    return fail(c, PRH_ERROR_INVALID_ARGUMENT, "bad options");
  const uint32_t N = g->N;
  const int L = o->max_iterations;
  // Seed table: one row per distinct seed vertex, s_j[v] = (occurrences of v in seed set j) / |seed set j|.
  std::vector<int32_t> row(N, -1);
  std::vector<double> mass;
  int32_t rows = 0;
  for (int j = 0; j < k; ++j) {
    if (so[j + 1] <= so[j]) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "seed set %d is empty", j);
    const double share = 1.0 / double(so[j + 1] - so[j]);
    for (uint32_t i = so[j]; i < so[j + 1]; ++i) {
      const uint32_t v = seeds[i];
      if (v >= N) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "seed %u >= N", v);
      if (row[v] < 0) {
        row[v] = rows++;
        mass.resize(size_t(rows) * prh::kPprCols, 0.0);
      }
      mass[size_t(row[v]) * prh::kPprCols + j] += share;
    }
  }
  hipStream_t s = c->stream;
  PRH_HIP(c, hipSetDevice(c->device));
  // Light / heavy split of the in-degree at the chunk size, and the heavy chunks (built on the host: few).
  const uint32_t nt = prh::partition_tiles(N);
  PRH_CHECK(ensure(c, c->sched_in, N));
  PRH_CHECK(ensure(c, c->tiles, size_t(3) * nt));
  PRH_CHECK(ensure(c, c->counts, 8));
  PRH_HIP(c, prh::launch_partition(g->in_off, N, kPprChunk, 0xFFFFFFFFu, c->sched_in.p, c->tiles.p, c->counts.p, s));
  PRH_HIP(c, hipMemcpyAsync(c->pinned, c->counts.p, 8, hipMemcpyDeviceToHost, s));
  PRH_HIP(c, hipStreamSynchronize(s));
  const uint32_t n_light = reinterpret_cast<uint32_t *>(c->pinned)[0];
  const uint32_t n_heavy = reinterpret_cast<uint32_t *>(c->pinned)[1];
  std::vector<uint32_t> hv(n_heavy), hb(n_heavy), he(n_heavy);
  PRH_CHECK(ensure(c, c->scratch_u32, size_t(2) * n_heavy + 1));
  if (n_heavy) {
    PRH_HIP(c, prh::launch_gather_ranges(g->in_off, c->sched_in.p + n_light, n_heavy, c->scratch_u32.p,
                                         c->scratch_u32.p + n_heavy, s));
    PRH_HIP(c, hipMemcpyAsync(hv.data(), c->sched_in.p + n_light, size_t(n_heavy) * 4, hipMemcpyDeviceToHost, s));
    PRH_HIP(c, hipMemcpyAsync(hb.data(), c->scratch_u32.p, size_t(n_heavy) * 4, hipMemcpyDeviceToHost, s));
    PRH_HIP(c, hipMemcpyAsync(he.data(), c->scratch_u32.p + n_heavy, size_t(n_heavy) * 4, hipMemcpyDeviceToHost, s));
    PRH_HIP(c, hipStreamSynchronize(s));
  }
  std::vector<uint32_t> cb, ce, hf(size_t(n_heavy) + 1);
  for (uint32_t h = 0; h < n_heavy; ++h) {
    hf[h] = uint32_t(cb.size());
    for (uint32_t e = hb[h]; e < he[h]; e += kPprChunk) {
      cb.push_back(e);
      ce.push_back(std::min(he[h], e + kPprChunk));
    }
  }
  hf[n_heavy] = uint32_t(cb.size());
  const uint32_t n_chunks = uint32_t(cb.size());
  // Device buffers.
  const size_t NC = size_t(N) * prh::kPprCols;
  for (int i = 0; i < 2; ++i) {
    PRH_CHECK(ensure(c, c->ppr_rank[i], NC));
    PRH_CHECK(ensure(c, c->ppr_contrib[i], NC));
  }
  PRH_CHECK(ensure(c, c->ppr_out, size_t(N) * k));
  PRH_CHECK(ensure(c, c->ppr_partial, std::max<size_t>(1, size_t(n_chunks) * prh::kPprCols)));
  PRH_CHECK(ensure(c, c->ppr_mass, std::max<size_t>(1, mass.size())));
  PRH_CHECK(ensure(c, c->ppr_row, N));
  PRH_CHECK(ensure(c, c->ppr_tables, size_t(2) * n_chunks + size_t(2) * n_heavy + 1));
  PRH_CHECK(ensure(c, c->ppr_err, std::max<size_t>(1, size_t(L) * prh::kPprCols)));
  PRH_CHECK(ensure(c, c->ppr_state, 1));
  uint32_t *t = c->ppr_tables.p;
  PRH_HIP(c, hipMemcpyAsync(c->ppr_row.p, row.data(), size_t(N) * 4, hipMemcpyHostToDevice, s));
  if (!mass.empty()) PRH_HIP(c, hipMemcpyAsync(c->ppr_mass.p, mass.data(), mass.size() * 8, hipMemcpyHostToDevice, s));
  if (n_chunks) {
    PRH_HIP(c, hipMemcpyAsync(t, cb.data(), size_t(n_chunks) * 4, hipMemcpyHostToDevice, s));
    PRH_HIP(c, hipMemcpyAsync(t + n_chunks, ce.data(), size_t(n_chunks) * 4, hipMemcpyHostToDevice, s));
  }
  if (n_heavy) PRH_HIP(c, hipMemcpyAsync(t + 2 * n_chunks, hv.data(), size_t(n_heavy) * 4, hipMemcpyHostToDevice, s));
  PRH_HIP(c, hipMemcpyAsync(t + 2 * n_chunks + n_heavy, hf.data(), (size_t(n_heavy) + 1) * 4, hipMemcpyHostToDevice, s));
  prh::PprArgs A{};
  for (int i = 0; i < 2; ++i) {
    A.rank[i] = c->ppr_rank[i].p;
    A.contrib[i] = c->ppr_contrib[i].p;
  }
  A.off = g->in_off;
  A.src = g->in_src;
  A.deg = g->outdeg;
  A.seed_row = c->ppr_row.p;
  A.seed_mass = c->ppr_mass.p;
  A.light = c->sched_in.p;
  A.n_light = n_light;
  A.chunk_beg = t;
  A.chunk_end = t + n_chunks;
  A.n_chunks = n_chunks;
  A.heavy_v = t + 2 * n_chunks;
  A.heavy_first = t + 2 * n_chunks + n_heavy;
  A.n_heavy = n_heavy;
  A.partial = c->ppr_partial.p;
  A.state = c->ppr_state.p;
  A.err = c->ppr_err.p;
  A.P = o->damping;
  A.tol = o->tolerance;
  A.k = k;
  A.max_iter = L;
  A.N = N;
  // Initialization (timed), then iterations as replays of a captured run of check_interval iterations.
  PRH_HIP(c, hipEventRecord(c->ev[0], s));
  PRH_HIP(c, prh::launch_ppr_reset(A, s));
  PRH_HIP(c, prh::launch_ppr_init(A, s));
  PRH_HIP(c, hipEventRecord(c->ev[1], s));
  const int chunk = std::max(1, c->tuning.check_interval);
  hipGraphExec_t exec = nullptr;
  int status = L > 0 ? capture_graph(c, [&]() -> int {
    for (int i = 0; i < chunk; ++i) PRH_HIP(c, prh::launch_ppr_iteration(A, s));
    return PRH_SUCCESS;
  }, &exec) : PRH_SUCCESS;
  prh::PprState hs{};
  for (int launched = 0; status == PRH_SUCCESS && launched < L; launched += chunk) {
    hipError_t e = hipGraphLaunch(exec, s);
    if (e == hipSuccess) e = hipMemcpyAsync(c->pinned, c->ppr_state.p, sizeof(prh::PprState), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      status = fail(c, PRH_ERROR_DEVICE, "batched iterations: %s", hipGetErrorString(e));
      break;
    }
    std::memcpy(&hs, c->pinned, sizeof(hs));
    bool all = true;
    for (int j = 0; j < k; ++j) all = all && hs.done[j] != 0;
    if (all) break;
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  PRH_CHECK(status);
  PRH_HIP(c, hipEventRecord(c->ev[2], s));
  PRH_HIP(c, prh::launch_ppr_collect(A, c->ppr_out.p, s));
  PRH_HIP(c, hipMemcpyAsync(ranks, c->ppr_out.p, size_t(N) * k * 8, hipMemcpyDeviceToHost, s));
  PRH_HIP(c, hipMemcpyAsync(c->pinned, c->ppr_state.p, sizeof(prh::PprState), hipMemcpyDeviceToHost, s));
  PRH_HIP(c, hipStreamSynchronize(s));
  std::memcpy(&hs, c->pinned, sizeof(hs));
  int most = 0;
  for (int j = 0; j < k; ++j) {
    iterations[j] = hs.done[j] ? hs.done[j] : L;
    most = std::max(most, iterations[j]);
  }
  float t01 = 0, t12 = 0;
  PRH_HIP(c, hipEventElapsedTime(&t01, c->ev[0], c->ev[1]));
  PRH_HIP(c, hipEventElapsedTime(&t12, c->ev[1], c->ev[2]));
  if (st) *st = prh_stats{most, t01 + t12, t01, 0.0f, t12};
  return PRH_SUCCESS;
}

// The degree-ordered layout of a freshly uploaded graph (see layout.hip): perm = vertices by in-degree, descending,
// stable in id (counting sort on the host; degrees above kDegreeBuckets are few and sorted directly), then the
// relabelled in-CSR (rows sorted by new source id), out-degrees and, if present, out-CSR on the device.
constexpr uint32_t kDegreeBuckets = 1u << 20;

int build_layout(prh_context *c, prh_graph *g, const uint32_t *in_off, const uint32_t *outdeg, bool with_out) {
  const auto t0 = std::chrono::steady_clock::now();
  const uint32_t N = g->N;
  const uint64_t M = g->M;
  hipStream_t s = c->stream;
  std::vector<uint32_t> perm(N), big;
  std::vector<uint64_t> cnt(size_t(kDegreeBuckets) + 1, 0);
  for (uint32_t v = 0; v < N; ++v) {
    const uint32_t d = in_off[v + 1] - in_off[v];
    if (d > kDegreeBuckets) big.push_back(v);
    else ++cnt[d];
  }
  std::stable_sort(big.begin(), big.end(), [&](uint32_t a, uint32_t b) {
    return in_off[a + 1] - in_off[a] > in_off[b + 1] - in_off[b];
  });
  uint64_t pos = big.size();
  for (int64_t d = kDegreeBuckets; d >= 0; --d) {
    const uint64_t n = cnt[d];
    cnt[d] = pos;
    pos += n;
  }
  std::copy(big.begin(), big.end(), perm.begin());
  for (uint32_t v = 0; v < N; ++v) {
    const uint32_t d = in_off[v + 1] - in_off[v];
    if (d <= kDegreeBuckets) perm[cnt[d]++] = v;
  }
  std::vector<uint32_t>().swap(big);
  std::vector<uint64_t>().swap(cnt);
  g->sorted_deg.resize(N);
  std::vector<uint32_t> loff(size_t(N) + 1), oout(with_out ? size_t(N) + 1 : 0);
  uint64_t a = 0, b = 0;
  for (uint32_t i = 0; i < N; ++i) {
    const uint32_t v = perm[i];
    const uint32_t d = in_off[v + 1] - in_off[v];
    g->sorted_deg[i] = d;
    loff[i] = uint32_t(a);
    a += d;
    if (with_out) {
      oout[i] = uint32_t(b);
      b += outdeg[v];
    }
  }
  loff[N] = uint32_t(a);
  if (with_out) oout[N] = uint32_t(b);
  // Device side.
  auto alloc = [&](uint32_t **p, size_t n) { return dev_alloc(c, reinterpret_cast<void **>(p), n * 4); };
  PRH_CHECK(alloc(&g->perm, N));
  PRH_CHECK(alloc(&g->inv, N));
  PRH_CHECK(alloc(&g->l_in_off, size_t(N) + 1));
  PRH_CHECK(alloc(&g->l_in_src, M));
  PRH_CHECK(alloc(&g->l_outdeg, N));
  if (with_out) {
    PRH_CHECK(alloc(&g->l_out_off, size_t(N) + 1));
    PRH_CHECK(alloc(&g->l_out_dst, M));
  }
  PRH_HIP(c, hipMemcpyAsync(g->perm, perm.data(), size_t(N) * 4, hipMemcpyHostToDevice, s));
  PRH_HIP(c, hipMemcpyAsync(g->l_in_off, loff.data(), (size_t(N) + 1) * 4, hipMemcpyHostToDevice, s));
  if (with_out) PRH_HIP(c, hipMemcpyAsync(g->l_out_off, oout.data(), (size_t(N) + 1) * 4, hipMemcpyHostToDevice, s));
  PRH_HIP(c, prh::launch_invert(g->perm, N, g->inv, s));
  PRH_HIP(c, prh::launch_gather_u32(g->outdeg, g->perm, N, g->l_outdeg, s));
  if (with_out) PRH_HIP(c, prh::launch_relabel_rows(g->out_off, g->out_dst, g->perm, g->inv, g->l_out_off, N,
                                                    g->l_out_dst, s));
  // Relabelled in-rows into a scratch array, then sorted per row into place.
  uint32_t *tmp = nullptr;
  void *sort_tmp = nullptr;
  size_t sort_bytes = 0;
  int st = alloc(&tmp, M);
  if (st == PRH_SUCCESS) {
    hipError_t e = prh::launch_relabel_rows(g->in_off, g->in_src, g->perm, g->inv, g->l_in_off, N, tmp, s);
    if (e == hipSuccess) e = prh::sort_rows(nullptr, sort_bytes, tmp, g->l_in_src, M, N, g->l_in_off, s);
    if (e != hipSuccess) st = fail(c, PRH_ERROR_DEVICE, "layout: %s", hipGetErrorString(e));
  }
  if (st == PRH_SUCCESS) st = dev_alloc(c, &sort_tmp, sort_bytes);
  if (st == PRH_SUCCESS) {
    hipError_t e = prh::sort_rows(sort_tmp, sort_bytes, tmp, g->l_in_src, M, N, g->l_in_off, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) st = fail(c, PRH_ERROR_DEVICE, "layout row sort: %s", hipGetErrorString(e));
  }
  (void)hipStreamSynchronize(s);
  dev_free(c, sort_tmp, sort_bytes);
  dev_free(c, tmp, (M ? M : 4) * 4);
  PRH_CHECK(st);
  g->prep_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return PRH_SUCCESS;
}

// Up to L iterations of step(l, err, done) -- launches honouring the device stop control of iteration l -- with
// the device flag polled every check_interval iterations. *n receives the number of iterations performed.
template <class F>
int run_iterations(prh_context *c, int L, double tol, F step, int *n) {
  hipStream_t s = c->stream;
  PRH_CHECK(ensure(c, c->loop, size_t(L) + 1));
  unsigned long long *err = c->loop.p;
  int *done = reinterpret_cast<int *>(c->loop.p + L);
  PRH_HIP(c, prh::launch_fill_u64(err, 0ull, size_t(L) + 1, s));
  const int check = std::max(1, c->tuning.check_interval);
  int result = -1, launched = 0;
  for (int l = 0; l < L; ++l) {
    PRH_CHECK(step(l, err, done));
    launched = l + 1;
    if (launched % check == 0 || launched == L) {
      PRH_HIP(c, hipMemcpyAsync(c->pinned, err + l, 8, hipMemcpyDeviceToHost, s));
      PRH_HIP(c, hipMemcpyAsync(c->pinned + 1, done, sizeof(int), hipMemcpyDeviceToHost, s));
      PRH_HIP(c, hipStreamSynchronize(s));
      double el;
      std::memcpy(&el, c->pinned, 8);
      const int d = *reinterpret_cast<const int *>(c->pinned + 1);
      if (d > 0) { result = d; break; }
      if (el < tol) { result = launched; break; }
    }
  }
  *n = result < 0 ? launched : result;
  return PRH_SUCCESS;
}

// Device copy of a caller array: the pointer itself for device memory, else staged into `buf`.
template <class T>
int stage(prh_context *c, const T *src, size_t n, bool device, DevBuf<T> &buf, const T **out) {
  if (device) {
    *out = src;
    return PRH_SUCCESS;
  }
  PRH_CHECK(ensure(c, buf, std::max<size_t>(1, n)));
  if (n) PRH_HIP(c, hipMemcpyAsync(buf.p, src, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
  *out = buf.p;
  return PRH_SUCCESS;
}

// Shared set-up of the weighted solves: the raw weights staged, W (out-weight sums) in the caller's order and in
// layout order, the layout edge maps (once per graph), and the iteration buffers.
int weighted_common(prh_context *c, const prh_graph *g, const float *w, const prh_options *o, bool device,
                    const float **w_out) {
  if (!g || !o || !w) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "null argument");
  if (g->ctx != c) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "graph belongs to another context");
  if (!g->out_off || !g->l_out_off) return fail(c, PRH_ERROR_UNSUPPORTED, "weighted solves need the graph's out-CSR");
  if (o->max_iterations < 0 || !(o->tolerance >= 0) || !(o->damping > 0 && o->damping < 1))
    return fail(c, PRH_ERROR_INVALID_ARGUMENT, "bad options");
  PRH_HIP(c, hipSetDevice(c->device));
  const uint32_t N = g->N;
  const uint64_t M = g->M;
  hipStream_t hs = c->stream;
  if (!g->l_in_eid) {
    prh_graph *gm = const_cast<prh_graph *>(g);
    PRH_CHECK(dev_alloc(c, reinterpret_cast<void **>(&gm->l_in_eid), (M ? M : 4) * 4));
    PRH_CHECK(dev_alloc(c, reinterpret_cast<void **>(&gm->l_out_eid), (M ? M : 4) * 4));
    PRH_CHECK(ensure(c, c->flags, 4));
    PRH_HIP(c, prh::launch_fill_i32(c->flags.p, 0, 1, hs));
    PRH_HIP(c, prh::launch_layout_edge_map(g->l_in_off, g->l_in_src, g->perm, g->out_off, g->out_dst, N, true,
                                           gm->l_in_eid, c->flags.p, hs));
    PRH_HIP(c, prh::launch_layout_edge_map(g->l_out_off, g->l_out_dst, g->perm, g->out_off, g->out_dst, N, false,
                                           gm->l_out_eid, c->flags.p, hs));
    PRH_HIP(c, hipMemcpyAsync(c->pinned, c->flags.p, sizeof(int), hipMemcpyDeviceToHost, hs));
    PRH_HIP(c, hipStreamSynchronize(hs));
    if (*reinterpret_cast<int *>(c->pinned)) {
      dev_free(c, gm->l_in_eid, (M ? M : 4) * 4);
      dev_free(c, gm->l_out_eid, (M ? M : 4) * 4);
      gm->l_in_eid = gm->l_out_eid = nullptr;
      return fail(c, PRH_ERROR_INVALID_GRAPH, "in-CSR and out-CSR disagree or out-rows are not ascending");
    }
  }
  PRH_CHECK(stage(c, w, size_t(M), device, c->wt_wout, w_out));
  PRH_CHECK(ensure(c, c->wt_W, N));
  PRH_CHECK(ensure(c, c->wt_Wl, N));
  PRH_HIP(c, prh::launch_out_weight_sum(g->out_off, *w_out, N, c->wt_W.p, hs));
  PRH_HIP(c, prh::launch_gather_f64(c->wt_W.p, g->perm, N, c->wt_Wl.p, hs));
  PRH_CHECK(ensure(c, c->wt_wl, std::max<uint64_t>(1, M)));
  for (int i = 0; i < 2; ++i) {
    PRH_CHECK(ensure(c, c->wt_x[i], N));
    PRH_CHECK(ensure(c, c->wt_z[i], N));
  }
  return PRH_SUCCESS;
}

// The low rows of a weighted schedule (order[0, n_low), degree <= 64) sorted by degree, descending, and cut into
// bands of W lanes per row for degree in (W, 2W] (W = 32 .. 2; W = 1 up to degree 2), each band on its own
// blocks of about block_edges edges.
int weighted_bands(prh_context *c, const uint32_t *off, DevBuf<uint32_t> &order, prh::WeightedArgs &A) {
  hipStream_t hs = c->stream;
  const uint32_t n = A.sched.n_low;
  A.n_bands = 0;
  if (!n) return PRH_SUCCESS;
  size_t sort_bytes = 0;
  PRH_CHECK(ensure(c, c->wt_sort, size_t(3) * n + 65));
  uint32_t *key = c->wt_sort.p, *key_s = key + n, *val_s = key_s + n, *starts = val_s + n;
  PRH_HIP(c, prh::sort_pairs_u32(nullptr, sort_bytes, key, key_s, order.p, val_s, n, 7, hs));
  PRH_CHECK(ensure(c, c->wt_sort_tmp, sort_bytes / 4 + 1));
  PRH_HIP(c, prh::launch_degree_keys(off, order.p, n, key, hs));
  PRH_HIP(c, prh::sort_pairs_u32(c->wt_sort_tmp.p, sort_bytes, key, key_s, order.p, val_s, n, 7, hs));
  PRH_HIP(c, hipMemcpyAsync(order.p, val_s, size_t(n) * 4, hipMemcpyDeviceToDevice, hs));
  PRH_HIP(c, prh::launch_lower_bounds(key_s, n, 64, starts, hs));
  PRH_HIP(c, hipMemcpyAsync(c->pinned, starts, 65 * 4, hipMemcpyDeviceToHost, hs));
  PRH_HIP(c, hipStreamSynchronize(hs));
  const uint32_t *lb = reinterpret_cast<const uint32_t *>(c->pinned);
  auto above = [&](uint32_t t) { return lb[64 - t]; };  // rows with degree > t
  const uint64_t edges_per_block = std::max<uint32_t>(c->tuning.block_edges, 256);
  uint32_t blocks = 0, prev = 0;
  for (int W = 32; W >= 1; W /= 2) {
    const uint32_t end = W == 1 ? n : above(uint32_t(W));
    if (end > prev) {
      const uint64_t unit = uint64_t(prh::kWavesPerBlock) * (prh::kWave / W);
      const uint64_t per = std::max<uint64_t>(unit, (edges_per_block / (2 * W) + unit - 1) / unit * unit);
      A.band[A.n_bands++] = prh::RankSection{prh::kGroupRows, W, prev, end, blocks, uint32_t(per)};
      blocks += uint32_t((end - prev + per - 1) / per);
      prev = end;
    }
  }
  return PRH_SUCCESS;
}

// Schedule of a weighted pull CSR: rows stably partitioned by degree into [<= 64 | <= kWeightedChunk | heavier]
// (`order`), and the heavier rows cut into chunks of kWeightedChunk edges (tables in wt_chunks, sums in
// wt_partial). Fills the schedule and chunk fields of A.
int weighted_schedule(prh_context *c, const uint32_t *off, uint32_t N, DevBuf<uint32_t> &order, prh::WeightedArgs &A) {
  hipStream_t hs = c->stream;
  PRH_CHECK(ensure(c, order, N));
  PRH_CHECK(ensure(c, c->tiles, size_t(3) * prh::partition_tiles(N)));
  PRH_CHECK(ensure(c, c->counts, 8));
  PRH_HIP(c, prh::launch_partition(off, N, 64u, prh::kWeightedChunk, order.p, c->tiles.p, c->counts.p, hs));
  PRH_HIP(c, hipMemcpyAsync(c->pinned, c->counts.p, 12, hipMemcpyDeviceToHost, hs));
  PRH_HIP(c, hipStreamSynchronize(hs));
  const uint32_t *cnt = reinterpret_cast<const uint32_t *>(c->pinned);
  A.sched = prh::Schedule{order.p, cnt[0], cnt[1], cnt[2]};
  A.n_chunks = 0;
  A.chunk_beg = A.chunk_end = A.heavy_first = nullptr;
  A.partial = nullptr;
  PRH_CHECK(weighted_bands(c, off, order, A));
  const uint32_t nh = A.sched.n_huge;
  if (!nh) return PRH_SUCCESS;
  PRH_CHECK(ensure(c, c->wt_chunks, size_t(2) * nh));
  PRH_HIP(c, prh::launch_gather_ranges(off, order.p + A.sched.n_low + A.sched.n_high, nh, c->wt_chunks.p,
                                       c->wt_chunks.p + nh, hs));
  std::vector<uint32_t> be(size_t(2) * nh);
  PRH_HIP(c, hipMemcpyAsync(be.data(), c->wt_chunks.p, be.size() * 4, hipMemcpyDeviceToHost, hs));
  PRH_HIP(c, hipStreamSynchronize(hs));
  std::vector<uint32_t> cb, ce, first(size_t(nh) + 1);
  for (uint32_t h = 0; h < nh; ++h) {
    first[h] = uint32_t(cb.size());
    for (uint64_t e = be[h]; e < be[nh + h]; e += prh::kWeightedChunk) {
      cb.push_back(uint32_t(e));
      ce.push_back(uint32_t(std::min<uint64_t>(be[nh + h], e + prh::kWeightedChunk)));
    }
  }
  first[nh] = uint32_t(cb.size());
  const size_t nc = cb.size();
  PRH_CHECK(ensure(c, c->wt_chunks, 2 * nc + nh + 1));
  PRH_HIP(c, hipMemcpyAsync(c->wt_chunks.p, cb.data(), nc * 4, hipMemcpyHostToDevice, hs));
  PRH_HIP(c, hipMemcpyAsync(c->wt_chunks.p + nc, ce.data(), nc * 4, hipMemcpyHostToDevice, hs));
  PRH_HIP(c, hipMemcpyAsync(c->wt_chunks.p + 2 * nc, first.data(), (size_t(nh) + 1) * 4, hipMemcpyHostToDevice, hs));
  PRH_CHECK(ensure(c, c->wt_partial, nc));
  PRH_HIP(c, hipStreamSynchronize(hs));
  A.chunk_beg = c->wt_chunks.p;
  A.chunk_end = c->wt_chunks.p + nc;
  A.heavy_first = c->wt_chunks.p + 2 * nc;
  A.partial = c->wt_partial.p;
  A.n_chunks = uint32_t(nc);
  return PRH_SUCCESS;
}

// A caller-order vector of N doubles in layout order: dst[i] = src[perm[i]].
int to_layout(prh_context *c, const prh_graph *g, const double *src, DevBuf<double> &dst) {
  PRH_CHECK(ensure(c, dst, g->N));
  PRH_HIP(c, prh::launch_gather_f64(src, g->perm, g->N, dst.p, c->stream));
  return PRH_SUCCESS;
}

// A layout-order result back to the caller's order (through `tmp` when the destination is on the host).
int from_layout(prh_context *c, const double *src, const uint32_t *map, size_t n, double *dst, bool device,
                DevBuf<double> &tmp) {
  if (device) {
    PRH_HIP(c, prh::launch_scatter_f64(src, map, n, dst, c->stream));
    return PRH_SUCCESS;
  }
  PRH_CHECK(ensure(c, tmp, std::max<size_t>(1, n)));
  PRH_HIP(c, prh::launch_scatter_f64(src, map, n, tmp.p, c->stream));
  if (n) PRH_HIP(c, hipMemcpyAsync(dst, tmp.p, n * 8, hipMemcpyDeviceToHost, c->stream));
  return PRH_SUCCESS;
}

int weighted_forward_impl(prh_context *c, const prh_graph *g, const float *w, const double *s, const prh_options *o,
                          double *ranks, prh_stats *st, bool device) {
  if (!c) return PRH_ERROR_INVALID_ARGUMENT;
  if (!s || !ranks) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "null argument");
  const float *w_out = nullptr;
  PRH_CHECK(weighted_common(c, g, w, o, device, &w_out));
  const uint32_t N = g->N;
  const uint64_t M = g->M;
  hipStream_t hs = c->stream;
  const double *term = nullptr;
  PRH_CHECK(stage(c, s, N, device, c->wt_term, &term));
  PRH_CHECK(to_layout(c, g, term, c->wt_terml));
  if (c->tuning.blocking > 0) {  // through the propagation-blocked plan of the layout in-CSR
    prh_graph *gm = const_cast<prh_graph *>(g);
    PRH_CHECK(build_binned_plan(c, gm, true));
    PRH_HIP(c, prh::launch_binned_weights(w_out, g->l_in_eid, g->bin.plan.slot, M, c->wt_wl.p, hs));
    prh::BinnedPlan plan = g->bin.plan;
    plan.ws = c->wt_wl.p;
    PRH_HIP(c, hipEventRecord(c->ev[0], hs));
    PRH_HIP(c, prh::launch_weighted_init(c->wt_terml.p, c->wt_Wl.p, N, false, c->wt_x[0].p, c->wt_z[0].p, hs));
    PRH_HIP(c, hipEventRecord(c->ev[1], hs));
    int n = 0;
    PRH_CHECK(run_iterations(c, o->max_iterations, o->tolerance, [&](int l, unsigned long long *err, int *done) -> int {
      prh::RankArgs A{c->wt_x[(l + 1) & 1].p, c->wt_x[l & 1].p, c->wt_z[(l + 1) & 1].p, c->wt_z[l & 1].p, nullptr,
                      nullptr, g->l_in_off, g->l_in_src, nullptr, 0.0, o->damping, 0.0, 0.0, nullptr};
      A.term = c->wt_terml.p;
      A.W = c->wt_Wl.p;
      PRH_HIP(c, prh::launch_binned_update(prh::kWeightedForward, A, plan, err, done, l, o->tolerance, false, hs));
      return PRH_SUCCESS;
    }, &n));
    PRH_HIP(c, hipEventRecord(c->ev[2], hs));
    PRH_CHECK(from_layout(c, c->wt_x[n & 1].p, g->perm, N, ranks, device, c->wt_outN));
    PRH_HIP(c, hipStreamSynchronize(hs));
    float t01 = 0, t12 = 0;
    PRH_HIP(c, hipEventElapsedTime(&t01, c->ev[0], c->ev[1]));
    PRH_HIP(c, hipEventElapsedTime(&t12, c->ev[1], c->ev[2]));
    if (st) *st = prh_stats{n, t01 + t12, t01, 0.0f, t12};
    return PRH_SUCCESS;
  }
  PRH_HIP(c, prh::launch_permute_weights(w_out, g->l_in_eid, M, c->wt_wl.p, hs));
  prh::WeightedArgs A0{};
  PRH_CHECK(weighted_schedule(c, g->l_in_off, N, c->sched_in, A0));
  PRH_HIP(c, hipEventRecord(c->ev[0], hs));
  PRH_HIP(c, prh::launch_weighted_init(c->wt_terml.p, c->wt_Wl.p, N, false, c->wt_x[0].p, c->wt_z[0].p, hs));
  PRH_HIP(c, hipEventRecord(c->ev[1], hs));
  int n = 0;
  PRH_CHECK(run_iterations(c, o->max_iterations, o->tolerance, [&](int l, unsigned long long *err, int *done) -> int {
    prh::WeightedArgs A = A0;
    A.off = g->l_in_off;
    A.idx = g->l_in_src;
    A.w = c->wt_wl.p;
    A.z = c->wt_z[l & 1].p;
    A.x_new = c->wt_x[(l + 1) & 1].p;
    A.z_new = c->wt_z[(l + 1) & 1].p;
    A.x_old = c->wt_x[l & 1].p;
    A.term = c->wt_terml.p;
    A.W = c->wt_Wl.p;
    A.P = o->damping;
    PRH_HIP(c, prh::launch_weighted_step(false, A, err, done, l, o->tolerance, hs));
    return PRH_SUCCESS;
  }, &n));
  PRH_HIP(c, hipEventRecord(c->ev[2], hs));
  PRH_CHECK(from_layout(c, c->wt_x[n & 1].p, g->perm, N, ranks, device, c->wt_outN));
  PRH_HIP(c, hipStreamSynchronize(hs));
  float t01 = 0, t12 = 0;
  PRH_HIP(c, hipEventElapsedTime(&t01, c->ev[0], c->ev[1]));
  PRH_HIP(c, hipEventElapsedTime(&t12, c->ev[1], c->ev[2]));
  if (st) *st = prh_stats{n, t01 + t12, t01, 0.0f, t12};
  return PRH_SUCCESS;
}

int weighted_adjoint_impl(prh_context *c, const prh_graph *g, const float *w, const double *r, const double *gr,
                          const prh_options *o, double *lam, double *grad_w, double *grad_s, prh_stats *st,
                          bool device) {
  if (!c) return PRH_ERROR_INVALID_ARGUMENT;
  if (!r || !gr || !grad_w || !grad_s) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "null argument");
  const float *w_out = nullptr;
  PRH_CHECK(weighted_common(c, g, w, o, device, &w_out));
  const uint32_t N = g->N;
  const uint64_t M = g->M;
  hipStream_t hs = c->stream;
  const double *term = nullptr, *rr = nullptr;
  PRH_CHECK(stage(c, gr, N, device, c->wt_term, &term));
  PRH_CHECK(stage(c, r, N, device, c->wt_r, &rr));
  PRH_CHECK(to_layout(c, g, term, c->wt_terml));
  PRH_CHECK(to_layout(c, g, rr, c->wt_rl));
  const bool blocked = c->tuning.blocking > 0;  // through the propagation-blocked plan of the layout out-CSR
  prh::BinnedPlan plan{};
  prh::WeightedArgs A0{};
  if (blocked) {
    prh_graph *gm = const_cast<prh_graph *>(g);
    PRH_CHECK(build_binned_plan_t(c, gm));
    PRH_HIP(c, prh::launch_binned_weights(w_out, g->l_out_eid, g->bin_t.plan.slot, M, c->wt_wl.p, hs));
    plan = g->bin_t.plan;
    plan.ws = c->wt_wl.p;
  } else {
    PRH_HIP(c, prh::launch_permute_weights(w_out, g->l_out_eid, M, c->wt_wl.p, hs));
    PRH_CHECK(weighted_schedule(c, g->l_out_off, N, c->sched_out, A0));
  }
  PRH_HIP(c, hipEventRecord(c->ev[0], hs));
  PRH_HIP(c, prh::launch_weighted_init(c->wt_terml.p, c->wt_Wl.p, N, true, c->wt_x[0].p, c->wt_z[0].p, hs));
  PRH_HIP(c, hipEventRecord(c->ev[1], hs));
  int n = 0;
  PRH_CHECK(run_iterations(c, o->max_iterations, o->tolerance, [&](int l, unsigned long long *err, int *done) -> int {
    if (blocked) {  // lam_new = g + (P / W) * sum over out-edges of w * lam_old, gathered straight from lam_old
      double *x_new = c->wt_x[(l + 1) & 1].p, *x_old = c->wt_x[l & 1].p;
      prh::RankArgs A{x_new, x_old, x_new, x_old, nullptr, nullptr, g->l_out_off, g->l_out_dst, nullptr,
                      0.0, o->damping, 0.0, 0.0, nullptr};
      A.term = c->wt_terml.p;
      A.W = c->wt_Wl.p;
      PRH_HIP(c, prh::launch_binned_update(prh::kWeightedAdjoint, A, plan, err, done, l, o->tolerance, false, hs));
      return PRH_SUCCESS;
    }
    prh::WeightedArgs A = A0;
    A.off = g->l_out_off;
    A.idx = g->l_out_dst;
    A.w = c->wt_wl.p;
    A.z = c->wt_x[l & 1].p;
    A.x_new = c->wt_x[(l + 1) & 1].p;
    A.z_new = c->wt_x[(l + 1) & 1].p;
    A.x_old = c->wt_x[l & 1].p;
    A.term = c->wt_terml.p;
    A.W = c->wt_Wl.p;
    A.P = o->damping;
    PRH_HIP(c, prh::launch_weighted_step(true, A, err, done, l, o->tolerance, hs));
    return PRH_SUCCESS;
  }, &n));
  const double *lam_l = c->wt_x[n & 1].p;
  PRH_CHECK(ensure(c, c->wt_gwl, std::max<uint64_t>(1, M)));
  PRH_CHECK(ensure(c, c->wt_gsl, N));
  PRH_HIP(c, prh::launch_weighted_grad(g->l_out_off, g->l_out_dst, c->wt_Wl.p, c->wt_rl.p, lam_l, c->wt_terml.p,
                                       o->damping, N, c->wt_gwl.p, c->wt_gsl.p, hs));
  PRH_HIP(c, hipEventRecord(c->ev[2], hs));
  if (lam) PRH_CHECK(from_layout(c, lam_l, g->perm, N, lam, device, c->wt_z[0]));
  PRH_CHECK(from_layout(c, c->wt_gwl.p, g->l_out_eid, M, grad_w, device, c->wt_outM));
  PRH_CHECK(from_layout(c, c->wt_gsl.p, g->perm, N, grad_s, device, c->wt_outN));
  PRH_HIP(c, hipStreamSynchronize(hs));
  float t01 = 0, t12 = 0;
  PRH_HIP(c, hipEventElapsedTime(&t01, c->ev[0], c->ev[1]));
  PRH_HIP(c, hipEventElapsedTime(&t12, c->ev[1], c->ev[2]));
  if (st) *st = prh_stats{n, t01 + t12, t01, 0.0f, t12};
  return PRH_SUCCESS;
}

}  // namespace

extern "C" {

int prh_weighted_pagerank(prh_context *c, const prh_graph *g, const float *w, const double *s, const prh_options *o,
                          double *ranks, prh_stats *st, int device_pointers) {
  if (c && g && g->N == 0) {
    if (st) *st = prh_stats{};
    return PRH_SUCCESS;
  }
  return weighted_forward_impl(c, g, w, s, o, ranks, st, device_pointers != 0);
}

int prh_weighted_adjoint(prh_context *c, const prh_graph *g, const float *w, const double *r, const double *gr,
                         const prh_options *o, double *lam, double *grad_w, double *grad_s, prh_stats *st,
                         int device_pointers) {
  if (c && g && g->N == 0) {
    if (st) *st = prh_stats{};
    return PRH_SUCCESS;
  }
  return weighted_adjoint_impl(c, g, w, r, gr, o, lam, grad_w, grad_s, st, device_pointers != 0);
}

int prh_ppr_batched(prh_context *c, const prh_graph *g, int k, const uint32_t *so, const uint32_t *seeds,
                    const prh_options *o, double *ranks, int *iterations, prh_stats *st) {
  if (c && g && g->N == 0) {
    if (st) *st = prh_stats{};
    if (iterations)
      for (int j = 0; j < k && j < prh::kPprCols; ++j) iterations[j] = 0;
    return PRH_SUCCESS;
  }
  return ppr_impl(c, g, k, so, seeds, o, ranks, iterations, st);
}

void prh_default_options(prh_options *o) {
  if (!o) return;
  o->repeat = 1;
  o->tolerance = 1e-10;
  o->frontier_tolerance = 1e-6;
  o->prune_tolerance = 1e-6;
  o->damping = 0.85;
  o->max_iterations = 500;
}

void prh_default_tuning(prh_tuning *t) {
  if (!t) return;
  t->switch_degree = 64;
  t->huge_degree = 16384;
  t->low_width = 0;
  t->check_interval = 8;
  t->source_slices = 0;
  t->block_edges = 2048;
  t->extension_direction = 1;
  t->blocking = 0;
}

int prh_context_create(int device, prh_context **out) {
  if (!out) return PRH_ERROR_INVALID_ARGUMENT;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
    (void)hipGetLastError();
    return PRH_ERROR_DEVICE;
  }
  prh_context *c = new prh_context();
  c->device = device;
  prh_default_tuning(&c->tuning);
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void **>(&c->pinned), kPinnedBytes, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    prh_context_destroy(c);
    return PRH_ERROR_DEVICE;
  }
  for (auto &e : c->ev)
    if (hipEventCreate(&e) != hipSuccess) {
      prh_context_destroy(c);
      return PRH_ERROR_DEVICE;
    }
  *out = c;
  return PRH_SUCCESS;
}

void prh_context_destroy(prh_context *c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (int i = 0; i < 2; ++i) {
    release(c, c->rank[i]);
    release(c, c->contrib[i]);
  }
  release(c, c->vaff);
  release(c, c->naff);
  release(c, c->buff);
  release(c, c->sched_in);
  release(c, c->sched_out);
  release(c, c->tiles);
  release(c, c->counts);
  release(c, c->ext_stats);
  release(c, c->ext_chunks);
  release(c, c->ext_tmp);
  release(c, c->batch);
  release(c, c->scratch_u32);
  release(c, c->loop);
  release(c, c->flags);
  for (int i = 0; i < 2; ++i) {
    release(c, c->ppr_rank[i]);
    release(c, c->ppr_contrib[i]);
  }
  release(c, c->ppr_partial);
  release(c, c->ppr_mass);
  release(c, c->ppr_out);
  release(c, c->ppr_row);
  release(c, c->ppr_tables);
  release(c, c->ppr_err);
  release(c, c->ppr_state);
  for (int i = 0; i < 2; ++i) {
    release(c, c->wt_x[i]);
    release(c, c->wt_z[i]);
  }
  for (DevBuf<double> *b : {&c->wt_W, &c->wt_Wl, &c->wt_term, &c->wt_terml, &c->wt_r, &c->wt_rl, &c->wt_gwl,
                            &c->wt_gsl, &c->wt_outM, &c->wt_outN})
    release(c, *b);
  release(c, c->wt_wout);
  release(c, c->wt_wl);
  release(c, c->wt_chunks);
  release(c, c->wt_sort);
  release(c, c->wt_sort_tmp);
  release(c, c->wt_partial);
  for (auto &e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &e : c->kev)
    if (e) (void)hipEventDestroy(e);
  if (c->pinned) (void)hipHostFree(c->pinned);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char *prh_context_last_error(const prh_context *c) { return c ? c->error.c_str() : "null context"; }

int prh_context_set_memory_limit(prh_context *c, size_t bytes) {
  if (!c) return PRH_ERROR_INVALID_ARGUMENT;
  c->limit = bytes;
  return PRH_SUCCESS;
}

size_t prh_context_memory_used(const prh_context *c) { return c ? c->used : 0; }

int prh_context_set_tuning(prh_context *c, const prh_tuning *t) {
  if (!c || !t) return PRH_ERROR_INVALID_ARGUMENT;
  const int w = t->low_width;
  const int sl = t->source_slices;
  if (!(w == 0 || w == 1 || w == 2 || w == 4 || w == 8 || w == 16 || w == 32) || t->check_interval < 1 ||
      t->huge_degree < t->switch_degree || !(sl == 0 || sl == 1 || sl == 2 || sl == 4 || sl == 8) ||
      t->block_edges < 256 || t->extension_direction < 0 || t->extension_direction > 2 ||
      !(t->blocking == 0 || t->blocking == 1024 || t->blocking == 2048 || t->blocking == 4096 ||
        t->blocking == 8192 || t->blocking == 16384))
    return fail(c, PRH_ERROR_INVALID_ARGUMENT, "bad tuning");
  c->tuning = *t;
  return PRH_SUCCESS;
}

int prh_context_get_tuning(const prh_context *c, prh_tuning *t) {
  if (!c || !t) return PRH_ERROR_INVALID_ARGUMENT;
  *t = c->tuning;
  return PRH_SUCCESS;
}

void *prh_context_stream(prh_context *c) { return c ? reinterpret_cast<void *>(c->stream) : nullptr; }

int prh_graph_create(prh_context *c, uint32_t N, uint64_t M, const uint32_t *in_off, const uint32_t *in_src,
                     const uint32_t *outdeg, const uint32_t *out_off, const uint32_t *out_dst, prh_graph **out) {
  if (!c || !out) return PRH_ERROR_INVALID_ARGUMENT;
  *out = nullptr;
  if (M >= (uint64_t(1) << 32)) return fail(c, PRH_ERROR_INVALID_GRAPH, "M = %llu does not fit uint32 offsets", (unsigned long long)M);
  if (N > 0 && (!in_off || (M && !in_src) || !outdeg))
    return fail(c, PRH_ERROR_INVALID_ARGUMENT, "null in-CSR array");
  if ((out_off == nullptr) != (out_dst == nullptr) || (out_off && M && !out_dst))
    return fail(c, PRH_ERROR_INVALID_ARGUMENT, "out-CSR needs both offsets and targets");
  PRH_HIP(c, hipSetDevice(c->device));
  prh_graph *g = new prh_graph();
  g->ctx = c;
  g->N = N;
  g->M = M;
  if (N == 0) {
    *out = g;
    return PRH_SUCCESS;
  }
  auto up = [&](uint32_t **dst, const uint32_t *src, size_t n) -> int {
    PRH_CHECK(dev_alloc(c, reinterpret_cast<void **>(dst), n * 4));
    PRH_HIP(c, hipMemcpyAsync(*dst, src, n * 4, hipMemcpyHostToDevice, c->stream));
    return PRH_SUCCESS;
  };
  int st = up(&g->in_off, in_off, size_t(N) + 1);
  if (st == PRH_SUCCESS) st = up(&g->in_src, in_src, M);
  if (st == PRH_SUCCESS) st = up(&g->outdeg, outdeg, N);
  if (st == PRH_SUCCESS && out_off) st = up(&g->out_off, out_off, size_t(N) + 1);
  if (st == PRH_SUCCESS && out_off) st = up(&g->out_dst, out_dst, M);
  if (st == PRH_SUCCESS) st = ensure(c, c->flags, 4);
  if (st == PRH_SUCCESS) {
    hipError_t e = prh::launch_fill_i32(c->flags.p, 0, 1, c->stream);
    if (e == hipSuccess)
      e = prh::launch_validate_graph(g->in_off, g->in_src, g->outdeg, g->out_off, g->out_dst, N, M, c->flags.p,
                                     c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(c->pinned, c->flags.p, sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) st = fail(c, PRH_ERROR_DEVICE, "graph validation: %s", hipGetErrorString(e));
    else {
      const int bits = *reinterpret_cast<int *>(c->pinned);
      if (bits)
        st = fail(c, PRH_ERROR_INVALID_GRAPH, "invalid graph (bits 0x%x:%s%s%s%s%s%s)", bits,
                  bits & 1 ? " in-offsets" : "", bits & 2 ? " dead-end" : "", bits & 4 ? " out-degree-mismatch" : "",
                  bits & 8 ? " in-id-range" : "", bits & 16 ? " out-id-range" : "", bits & 32 ? " out-offsets" : "");
    }
  }
  if (st == PRH_SUCCESS) st = build_layout(c, g, in_off, outdeg, out_off != nullptr);
  if (st != PRH_SUCCESS) {
    prh_graph_destroy(g);
    return st;
  }
  *out = g;
  return PRH_SUCCESS;
}

void prh_graph_destroy(prh_graph *g) {
  if (!g) return;
  prh_context *c = g->ctx;
  if (c && c->stream) (void)hipStreamSynchronize(c->stream);
  const size_t N = g->N, M = g->M;
  dev_free(c, g->in_off, (N + 1) * 4);
  dev_free(c, g->in_src, (M ? M : 4) * 4);
  dev_free(c, g->outdeg, N * 4);
  dev_free(c, g->out_off, (N + 1) * 4);
  dev_free(c, g->out_dst, (M ? M : 4) * 4);
  dev_free(c, g->l_in_eid, (M ? M : 4) * 4);
  dev_free(c, g->l_out_eid, (M ? M : 4) * 4);
  dev_free(c, g->perm, N * 4);
  dev_free(c, g->inv, N * 4);
  dev_free(c, g->l_in_off, (N + 1) * 4);
  dev_free(c, g->l_in_src, (M ? M : 4) * 4);
  dev_free(c, g->l_outdeg, N * 4);
  dev_free(c, g->l_out_off, (N + 1) * 4);
  dev_free(c, g->l_out_dst, (M ? M : 4) * 4);
  dev_free(c, g->chunk_tables, g->chunk_tables_bytes);
  dev_free(c, g->chunk_partial, g->chunk_partial_bytes);
  dev_free(c, g->out_order, std::max<size_t>(N, 1) * 4);
  dev_free(c, g->out_chunks, std::max<size_t>(3 * size_t(g->out_n_chunks), 1) * 4);
  free_sliced(c, g);
  free_binned(c, g->bin);
  free_binned(c, g->bin_t);
  delete g;
}

uint32_t prh_graph_order(const prh_graph *g) { return g ? g->N : 0; }
uint64_t prh_graph_size(const prh_graph *g) { return g ? g->M : 0; }

int prh_solve(prh_context *c, const prh_graph *g, prh_method method, int partition, const prh_options *o,
              const uint32_t *du, const uint32_t *dv, size_t nd, const uint32_t *iu, size_t ni, const double *q,
              double *ranks, prh_stats *st) {
  (void)partition;
  return solve_impl(c, g, method, o, du, dv, nd, iu, ni, q, ranks, st, true);
}

int prh_solve_device(prh_context *c, const prh_graph *g, prh_method method, int partition, const prh_options *o,
                     const uint32_t *du, const uint32_t *dv, size_t nd, const uint32_t *iu, size_t ni,
                     const double *q, double *ranks, prh_stats *st) {
  (void)partition;
  return solve_impl(c, g, method, o, du, dv, nd, iu, ni, q, ranks, st, false);
}

// ---- autotuner ----------------------------------------------------------------------------------------------

}  // extern "C"

namespace {

// Key of a (device, graph) pair: the device's architecture and compute-unit count, N, M and an FNV-1a hash of
// the graph's in-degree sequence in layout order (which fixes every launch plan).
std::string tuning_key(prh_context *c, const prh_graph *g) {
  hipDeviceProp_t prop{};
  if (hipGetDeviceProperties(&prop, c->device) != hipSuccess) (void)hipGetLastError();
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](uint64_t x) {
    for (int k = 0; k < 8; ++k) {
      h ^= (x >> (8 * k)) & 0xFF;
      h *= 1099511628211ull;
    }
  };
  mix(g->N);
  mix(g->M);
  for (uint32_t d : g->sorted_deg) mix(d);
  char buf[160];
  std::string arch(prop.gcnArchName);
  arch = arch.substr(0, arch.find(':'));
  std::snprintf(buf, sizeof(buf), "%s-cu%d-n%u-m%llu-%016llx", arch.c_str(), prop.multiProcessorCount, g->N,
                (unsigned long long)g->M, (unsigned long long)h);
  return buf;
}

constexpr int kTuningCount = 8;
const char *const kTuningFields[kTuningCount] = {"switch_degree", "huge_degree", "low_width", "check_interval",
                                                 "source_slices", "block_edges", "extension_direction", "blocking"};

void tuning_to_values(const prh_tuning &t, long long v[kTuningCount]) {
  v[0] = t.switch_degree;
  v[1] = t.huge_degree;
  v[2] = t.low_width;
  v[3] = t.check_interval;
  v[4] = t.source_slices;
  v[5] = t.block_edges;
  v[6] = t.extension_direction;
  v[7] = t.blocking;
}

void values_to_tuning(const long long v[kTuningCount], prh_tuning &t) {
  t.switch_degree = uint32_t(v[0]);
  t.huge_degree = uint32_t(v[1]);
  t.low_width = int(v[2]);
  t.check_interval = int(v[3]);
  t.source_slices = int(v[4]);
  t.block_edges = uint32_t(v[5]);
  t.extension_direction = int(v[6]);
  t.blocking = int(v[7]);
}

// The cache is a JSON object with one entry per line:
//   {
//   "<key>": {"switch_degree": 128, ..., "ms_per_iteration": 1.92},
//   ...
//   }
// Lines that do not parse as an entry are ignored (a hand-edited file degrades to a cache miss).
struct CacheEntry {
  std::string key;
  long long v[kTuningCount];
  double ms;
};

std::vector<CacheEntry> read_cache(const char *path) {
  std::vector<CacheEntry> out;
  FILE *f = std::fopen(path, "r");
  if (!f) return out;
  char line[1024];
  while (std::fgets(line, sizeof(line), f)) {
    const char *q0 = std::strchr(line, '"');
    const char *q1 = q0 ? std::strchr(q0 + 1, '"') : nullptr;
    if (!q1) continue;
    CacheEntry e;
    e.key.assign(q0 + 1, q1);
    bool ok = true;
    for (int k = 0; k < kTuningCount && ok; ++k) {
      const std::string pat = std::string("\"") + kTuningFields[k] + "\":";
      const char *p = std::strstr(q1, pat.c_str());
      ok = p && std::sscanf(p + pat.size(), "%lld", &e.v[k]) == 1;
    }
    const char *pm = std::strstr(q1, "\"ms_per_iteration\":");
    if (!ok || !pm || std::sscanf(pm + 19, "%lf", &e.ms) != 1) continue;
    out.push_back(e);
  }
  std::fclose(f);
  return out;
}

bool write_cache(const char *path, const std::vector<CacheEntry> &entries) {
  const std::string tmp = std::string(path) + ".tmp";
  FILE *f = std::fopen(tmp.c_str(), "w");
  if (!f) return false;
  std::fprintf(f, "{\n");
  for (size_t i = 0; i < entries.size(); ++i) {
    const CacheEntry &e = entries[i];
    std::fprintf(f, "\"%s\": {", e.key.c_str());
    for (int k = 0; k < kTuningCount; ++k) std::fprintf(f, "\"%s\": %lld, ", kTuningFields[k], e.v[k]);
    std::fprintf(f, "\"ms_per_iteration\": %.6f}%s\n", e.ms, i + 1 < entries.size() ? "," : "");
  }
  std::fprintf(f, "}\n");
  const bool ok = std::fclose(f) == 0;
  return ok && std::rename(tmp.c_str(), path) == 0;
}

}  // namespace

extern "C" {

int prh_autotune(prh_context *c, const prh_graph *g, const char *cache_path, int iterations, prh_tuning *chosen,
                 double *ms_per_iteration, int *from_cache) {
  if (!c) return PRH_ERROR_INVALID_ARGUMENT;
  if (!g || g->ctx != c || iterations < 1) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "bad graph or iterations");
  if (from_cache) *from_cache = 0;
  PRH_HIP(c, hipSetDevice(c->device));
  const prh_tuning base = c->tuning;
  if (g->N == 0) {
    if (chosen) *chosen = base;
    if (ms_per_iteration) *ms_per_iteration = 0;
    return PRH_SUCCESS;
  }
  const std::string key = tuning_key(c, g);
  std::vector<CacheEntry> cache;
  if (cache_path) {
    cache = read_cache(cache_path);
    for (const CacheEntry &e : cache)
      if (e.key == key) {
        prh_tuning t = base;
        values_to_tuning(e.v, t);
        PRH_CHECK(prh_context_set_tuning(c, &t));
        if (chosen) *chosen = t;
        if (ms_per_iteration) *ms_per_iteration = e.ms;
        if (from_cache) *from_cache = 1;
        return PRH_SUCCESS;
      }
  }
  // Candidates: the row plan over switch degree x chunk threshold x block size, the sliced plan over slice
  // count x block size, and the propagation-blocked plan over its block width (graphs of 2^20 edges and more,
  // where its streaming passes pay). Each is timed on `iterations` static iterations (tolerance 0) after one
  // warm-up solve.
  std::vector<prh_tuning> cand;
  for (uint32_t sw : {64u, 128u, 256u})
    for (uint32_t hu : {4096u, 16384u})
      for (uint32_t be : {2048u, 4096u}) {
        prh_tuning t = base;
        t.switch_degree = sw;
        t.huge_degree = hu;
        t.low_width = 0;
        t.source_slices = 0;
        t.block_edges = be;
        t.blocking = 0;
        cand.push_back(t);
      }
  for (int sl : {1, 8})
    for (uint32_t be : {2048u, 4096u}) {
      prh_tuning t = base;
      t.switch_degree = 128;
      t.huge_degree = 16384;
      t.low_width = 0;
      t.source_slices = sl;
      t.block_edges = be;
      t.blocking = 0;
      cand.push_back(t);
    }
  if (g->M >= (uint64_t(1) << 20))
    for (int bw : {2048, 4096, 8192, 16384}) {
      const uint64_t nb = (uint64_t(g->N) + bw - 1) / bw;
      if (nb * nb > (uint64_t(1) << 32)) continue;
      prh_tuning t = base;
      t.switch_degree = 64;
      t.huge_degree = 16384;
      t.low_width = 0;
      t.source_slices = 0;
      t.block_edges = 2048;
      t.blocking = bw;
      cand.push_back(t);
    }
  double *ranks = nullptr;
  PRH_CHECK(dev_alloc(c, reinterpret_cast<void **>(&ranks), size_t(g->N) * 8));
  prh_options o;
  prh_default_options(&o);
  o.tolerance = 0.0;
  int best = -1;
  double best_ms = 0.0;
  int status = PRH_SUCCESS;
  for (size_t i = 0; i < cand.size() && status == PRH_SUCCESS; ++i) {
    status = prh_context_set_tuning(c, &cand[i]);
    prh_stats st{};
    o.max_iterations = 2;
    if (status == PRH_SUCCESS) status = solve_impl(c, g, PRH_STATIC, &o, nullptr, nullptr, 0, nullptr, 0, nullptr,
                                                   ranks, &st, false);
    o.max_iterations = iterations;
    if (status == PRH_SUCCESS) status = solve_impl(c, g, PRH_STATIC, &o, nullptr, nullptr, 0, nullptr, 0, nullptr,
                                                   ranks, &st, false);
    if (status != PRH_SUCCESS) break;
    const double ms = double(st.computation_time) / std::max(st.iterations, 1);
    if (best < 0 || ms < best_ms) {
      best = int(i);
      best_ms = ms;
    }
  }
  dev_free(c, ranks, size_t(g->N) * 8);
  if (status != PRH_SUCCESS) {
    c->tuning = base;
    return status;
  }
  PRH_CHECK(prh_context_set_tuning(c, &cand[best]));
  if (chosen) *chosen = cand[best];
  if (ms_per_iteration) *ms_per_iteration = best_ms;
  if (cache_path) {
    CacheEntry e;
    e.key = key;
    tuning_to_values(cand[best], e.v);
    e.ms = best_ms;
    bool replaced = false;
    for (CacheEntry &old : cache)
      if (old.key == key) {
        old = e;
        replaced = true;
      }
    if (!replaced) cache.push_back(e);
    if (!write_cache(cache_path, cache)) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "cannot write %s", cache_path);
  }
  return PRH_SUCCESS;
}

int prh_partition(prh_context *c, uint32_t N, const uint32_t *off, uint32_t d, uint32_t *order, uint32_t *np) {
  if (!c) return PRH_ERROR_INVALID_ARGUMENT;
  if (N == 0) {
    if (np) *np = 0;
    return PRH_SUCCESS;
  }
  if (!off || !order || !np) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "null argument");
  PRH_HIP(c, hipSetDevice(c->device));
  hipStream_t s = c->stream;
  PRH_CHECK(ensure(c, c->scratch_u32, size_t(N) + 1));
  PRH_CHECK(ensure(c, c->sched_in, N));
  PRH_CHECK(ensure(c, c->tiles, size_t(3) * prh::partition_tiles(N)));
  PRH_CHECK(ensure(c, c->counts, 8));
  PRH_HIP(c, hipMemcpyAsync(c->scratch_u32.p, off, (size_t(N) + 1) * 4, hipMemcpyHostToDevice, s));
  PRH_HIP(c, prh::launch_partition(c->scratch_u32.p, N, d, 0xFFFFFFFFu, c->sched_in.p, c->tiles.p, c->counts.p, s));
  PRH_HIP(c, hipMemcpyAsync(order, c->sched_in.p, size_t(N) * 4, hipMemcpyDeviceToHost, s));
  PRH_HIP(c, hipMemcpyAsync(c->pinned, c->counts.p, 4, hipMemcpyDeviceToHost, s));
  PRH_HIP(c, hipStreamSynchronize(s));
  *np = *reinterpret_cast<uint32_t *>(c->pinned);
  return PRH_SUCCESS;
}

int prh_mark_affected(prh_context *c, const prh_graph *g, prh_method method, const uint32_t *du, const uint32_t *dv,
                      size_t nd, const uint32_t *iu, size_t ni, uint8_t *affected) {
  if (!c) return PRH_ERROR_INVALID_ARGUMENT;
  if (!g || !affected || method < PRH_DYNAMIC_TRAVERSAL || method > PRH_PRUNE_DYNAMIC_FRONTIER)
    return fail(c, PRH_ERROR_INVALID_ARGUMENT, "null argument or non-dynamic method");
  if (!g->out_off) return fail(c, PRH_ERROR_UNSUPPORTED, "marking needs the graph's out-CSR");
  if (g->N == 0) return PRH_SUCCESS;
  PRH_HIP(c, hipSetDevice(c->device));
  PRH_CHECK(ensure_workspace(c, g->N, true, 1));
  DevBatch b;
  PRH_CHECK(stage_batch(c, g, du, dv, nd, iu, ni, true, b));
  prh::ExtendArgs xo{};
  if (method == PRH_DYNAMIC_TRAVERSAL) PRH_CHECK(out_schedule(c, const_cast<prh_graph *>(g), xo));
  PRH_CHECK(mark_affected(c, g, method, b, xo));
  // back to the caller's vertex order (naff is free after marking)
  PRH_HIP(c, prh::launch_scatter_u8(c->vaff.p, g->perm, g->N, c->naff.p, c->stream));
  PRH_HIP(c, hipMemcpyAsync(affected, c->naff.p, g->N, hipMemcpyDeviceToHost, c->stream));
  PRH_HIP(c, hipStreamSynchronize(c->stream));
  return PRH_SUCCESS;
}

int prh_kernel_timing_enable(prh_context *c, int enable) {
  if (!c) return PRH_ERROR_INVALID_ARGUMENT;
  c->ktiming = enable != 0;
  return PRH_SUCCESS;
}

int prh_kernel_timing_read(prh_context *c, uint64_t *launches, double *total_ms) {
  if (!c) return PRH_ERROR_INVALID_ARGUMENT;
  if (launches) *launches = c->k_launches;
  if (total_ms) *total_ms = c->k_ms;
  c->k_launches = 0;
  c->k_ms = 0.0;
  return PRH_SUCCESS;
}

}  // extern "C"

// ---- streaming dynamic session -------------------------------------------------------------------------------
//
// A device-resident copy of a graph's degree-ordered layout that batches update in place (stream.hip), with
// the ranks kept between batches and each batch's work replayed from captured HIP graphs:
//   apply graph (one per buffer parity): batch -> keys -> sort -> tidy -> new in- and out-CSR (into the other
//       buffers) -> out-degrees -> chunk tables -> solver initialisation and, for DF / DF-P, the initial marking;
//   iteration graph (one per parity): kChunk iterations of the rank update (and, for DF / DF-P, the direction
//       choice and both extension forms) followed each by the device-side convergence step; replayed until the
//       device reports the loop done.
// Only the batch upload, one flag read after the apply and one read per iteration chunk touch the host.

struct prh_stream {
  prh_context *ctx = nullptr;
  const prh_graph *g = nullptr;
  int method = 0, variant = 0;
  prh_options opt{};
  uint32_t N = 0;
  uint64_t M = 0;          // current edge count
  uint64_t cap_edges = 0;  // capacity of the edge arrays
  uint32_t cap_batch = 0;  // capacity per batch side (deletions, insertions)
  uint32_t *in_off[2] = {}, *in_src[2] = {}, *out_off[2] = {}, *out_dst[2] = {};
  int cur = 0;             // buffers holding the current graph
  uint32_t *outdeg = nullptr, *new_deg = nullptr, *sched_order = nullptr;
  uint32_t sched_low = 0, sched_high = 0;
  int *delta = nullptr;
  uint8_t *changed = nullptr;
  uint32_t *h_batch = nullptr;  // pinned: del_u, del_v, ins_u, ins_v [cap_batch each], nd, ni
  uint32_t *d_batch = nullptr;
  unsigned long long *keys[4] = {};     // raw keys: in-CSR del, in-CSR ins, out-CSR del, out-CSR ins
  unsigned long long *compact[4] = {};  // sorted, tidied
  unsigned long long *sorted = nullptr;
  uint32_t *valid = nullptr, *pos = nullptr, *counts = nullptr;  // counts[4]: sizes of compact[]
  void *sort_tmp = nullptr, *scan_tmp = nullptr;
  size_t sort_bytes = 0, scan_bytes = 0;
  double *rank[2] = {}, *contrib[2] = {}, *stats = nullptr;
  uint8_t *vaff = nullptr, *naff = nullptr;
  int *ctl = nullptr;  // [0] done, [1] iteration index (adjacent, as LoopCtl expects), [2] direction, [3] flags
  unsigned long long *err = nullptr;
  prh::RankPlan plan{};
  uint32_t n_items = 0, chunk = 0, stat_slots = 0, pull_wave = 0, pull_group = 0;
  hipGraphExec_t apply[2] = {}, iter[2] = {};
  std::vector<std::pair<void *, size_t>> bufs;
  bool broken = false;
};

namespace {

constexpr int kStreamChunk = 8;  // iterations per replay of the iteration graph (even: ping-pong buffers realign)

int stream_alloc(prh_stream *S, void **p, size_t bytes) {
  PRH_CHECK(dev_alloc(S->ctx, p, bytes));
  S->bufs.emplace_back(*p, bytes ? bytes : 16);
  return PRH_SUCCESS;
}

template <class T>
int stream_alloc(prh_stream *S, T **p, size_t n) {
  void *v = nullptr;
  PRH_CHECK(stream_alloc(S, &v, n * sizeof(T)));
  *p = static_cast<T *>(v);
  return PRH_SUCCESS;
}

void stream_free(prh_stream *S) {
  for (int k = 0; k < 2; ++k) {
    if (S->apply[k]) (void)hipGraphExecDestroy(S->apply[k]);
    if (S->iter[k]) (void)hipGraphExecDestroy(S->iter[k]);
  }
  for (auto &b : S->bufs) dev_free(S->ctx, b.first, b.second);
  S->bufs.clear();
  if (S->h_batch) (void)hipHostFree(S->h_batch);
  S->h_batch = nullptr;
}

// One CSR's tidy-and-apply: keys k_del / k_ins (raw) of the current rows (off, idx) into new rows (noff, nidx).
int stream_apply_csr(prh_stream *S, hipStream_t s, int kd, int ki, const uint32_t *off, const uint32_t *idx,
                     uint32_t *noff, uint32_t *nidx, int empty_bit) {
  prh_context *c = S->ctx;
  const uint32_t cap = S->cap_batch, N = S->N;
  for (int k : {kd, ki}) {
    PRH_HIP(c, prh::sort_keys_u64(S->sort_tmp, S->sort_bytes, S->keys[k], S->sorted, cap, s));
    PRH_HIP(c, prh::launch_batch_valid(S->sorted, cap, off, idx, k == kd, S->valid, s));
    PRH_HIP(c, prh::scan_u32(S->scan_tmp, S->scan_bytes, S->valid, S->pos, cap, s));
    PRH_HIP(c, prh::launch_batch_compact(S->sorted, S->valid, S->pos, cap, S->compact[k], S->counts + k, s));
  }
  PRH_HIP(c, prh::launch_fill_i32(S->delta, 0, N, s));
  PRH_HIP(c, prh::launch_fill_u8(S->changed, 0, N, s));
  PRH_HIP(c, prh::launch_row_delta(S->compact[kd], S->counts + kd, S->compact[ki], S->counts + ki, 2 * cap, S->delta,
                                   S->changed, s));
  PRH_HIP(c, prh::launch_new_degrees(off, S->delta, N, S->new_deg, S->ctl + 3, empty_bit, s));
  PRH_HIP(c, prh::scan_u32(S->scan_tmp, S->scan_bytes, S->new_deg, noff, size_t(N) + 1, s));
  PRH_HIP(c, prh::launch_rewrite(off, idx, noff, S->changed, S->compact[kd], S->counts + kd, S->compact[ki],
                                 S->counts + ki, N, S->cap_edges, nidx, s));
  PRH_HIP(c, prh::launch_insert(off, idx, noff, S->compact[kd], S->counts + kd, S->compact[ki], S->counts + ki, cap,
                                nidx, s));
  return PRH_SUCCESS;
}

// Work of one batch from buffers p to buffers 1 - p, up to the first iteration (captured once per parity).
int stream_apply_body(prh_stream *S, int p) {
  prh_context *c = S->ctx;
  hipStream_t s = c->stream;
  const uint32_t N = S->N, cap = S->cap_batch;
  const int q = 1 - p;
  const uint32_t *b = S->d_batch;
  PRH_HIP(c, prh::launch_fill_i32(S->ctl, 0, 4, s));
  PRH_HIP(c, prh::launch_batch_keys(b, b + cap, b + 4 * size_t(cap), cap, S->g->inv, N, S->keys[0], S->keys[2],
                                    S->ctl + 3, 1, s));
  PRH_HIP(c, prh::launch_batch_keys(b + 2 * size_t(cap), b + 3 * size_t(cap), b + 4 * size_t(cap) + 1, cap,
                                    S->g->inv, N, S->keys[1], S->keys[3], S->ctl + 3, 1, s));
  PRH_CHECK(stream_apply_csr(S, s, 0, 1, S->in_off[p], S->in_src[p], S->in_off[q], S->in_src[q], 0));
  PRH_CHECK(stream_apply_csr(S, s, 2, 3, S->out_off[p], S->out_dst[p], S->out_off[q], S->out_dst[q], 2));
  PRH_HIP(c, prh::launch_degrees(S->out_off[q], N, S->outdeg, s));
  PRH_HIP(c, prh::launch_refresh_chunks(S->plan.item_row, S->plan.row_first, S->n_items, S->in_off[q], S->chunk,
                                        const_cast<uint32_t *>(S->plan.item_beg),
                                        const_cast<uint32_t *>(S->plan.item_end), s));
  // Solver initialisation (pagerankCuda.hxx:811-821): ranks from the previous batch's result (or 1/N for Static),
  // contributions with the new out-degrees, loop state cleared, and the DF initial marking on the new graph.
  PRH_HIP(c, prh::launch_fill_u64(S->err, 0ull, size_t(S->opt.max_iterations) + 1, s));
  PRH_HIP(c, prh::launch_init_ranks(S->rank[0], S->rank[1], S->contrib[0], S->contrib[1], S->outdeg, N,
                                    S->method != PRH_STATIC, s));
  if (S->variant >= prh::kFrontier) {
    PRH_HIP(c, prh::launch_fill_u8(S->vaff, 0, N, s));
    PRH_HIP(c, prh::launch_fill_u8(S->naff, 0, N, s));
    PRH_HIP(c, prh::launch_mark_batch(S->naff, S->vaff, S->compact[2], S->counts + 2, S->compact[3], S->counts + 3,
                                      2 * cap, s));
    const prh::Schedule os{S->sched_order, S->sched_low, S->sched_high, 0};
    prh::ExtendArgs E{S->vaff, S->naff, S->out_off[q], S->out_dst[q], os, nullptr, nullptr, nullptr};
    PRH_HIP(c, prh::launch_extend(E, nullptr, nullptr, 0, 0.0, s));
    const int mode = c->tuning.extension_direction;
    if (mode != 1) PRH_HIP(c, prh::launch_fill_i32(S->ctl + 2, mode == 2 ? 1 : 0, 1, s));
  }
  return PRH_SUCCESS;
}

// kStreamChunk iterations on buffers q (captured once per parity); the iteration index comes from ctl[1].
int stream_iter_body(prh_stream *S, int q) {
  prh_context *c = S->ctx;
  hipStream_t s = c->stream;
  const prh_options &o = S->opt;
  const double P = o.damping, C0 = (1 - P) / S->N;
  const bool frontier = S->variant >= prh::kFrontier;
  const int mode = c->tuning.extension_direction;
  unsigned long long *err = S->err;
  int *done = S->ctl, *dir = S->ctl + 2;
  const prh::Schedule os{S->sched_order, S->sched_low, S->sched_high, 0};
  const prh::PullArgs pull{S->vaff, S->naff, S->in_off[q], S->in_src[q], S->pull_wave, S->pull_group, S->N, dir};
  for (int i = 0; i < kStreamChunk; ++i) {
    prh::RankArgs A{S->rank[(i + 1) & 1], S->rank[i & 1], S->contrib[(i + 1) & 1], S->contrib[i & 1], S->naff,
                    S->vaff, S->in_off[q], S->in_src[q], S->outdeg, C0, P, o.frontier_tolerance, o.prune_tolerance,
                    frontier && mode == 1 ? S->stats : nullptr};
    PRH_HIP(c, prh::launch_rank_update(S->variant, A, S->plan, err, done, -1, o.tolerance, s));
    if (frontier) {
      if (mode == 1) PRH_HIP(c, prh::launch_direction(S->stats, S->stat_slots, dir, err, done, -1, o.tolerance, s));
      prh::ExtendArgs E{S->vaff, S->naff, S->out_off[q], S->out_dst[q], os, nullptr, nullptr, dir};
      PRH_HIP(c, prh::launch_extend(E, err, done, -1, o.tolerance, s));
      if (mode != 0) PRH_HIP(c, prh::launch_extend_pull(pull, err, done, -1, o.tolerance, s));
    }
    PRH_HIP(c, prh::launch_advance(S->ctl + 1, done, err, o.tolerance, o.max_iterations, s));
  }
  return PRH_SUCCESS;
}

}  // namespace

extern "C" {

int prh_stream_create(prh_context *c, const prh_graph *g, prh_method method, const prh_options *options,
                      uint32_t max_batch, uint64_t max_edges, const double *initial_ranks, prh_stream **out) {
  if (!c || !out) return PRH_ERROR_INVALID_ARGUMENT;
  *out = nullptr;
  if (!g || !options || g->ctx != c) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "null graph/options or foreign graph");
  if (method == PRH_DYNAMIC_TRAVERSAL)
    return fail(c, PRH_ERROR_UNSUPPORTED, "the traversal closure needs host-side level checks; not streamed");
  if (!valid_method(method)) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "bad method");
  if (!g->l_out_off) return fail(c, PRH_ERROR_UNSUPPORTED, "the graph was uploaded without its out-CSR");
  if (g->N == 0 || max_batch == 0) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "empty graph or zero batch capacity");
  const prh_options &o = *options;
  if (o.max_iterations < 1 || !(o.tolerance >= 0) || !(o.damping > 0 && o.damping < 1))
    return fail(c, PRH_ERROR_INVALID_ARGUMENT, "bad options");
  const uint64_t cap_edges = max_edges ? max_edges : g->M + 16ull * max_batch;
  if (cap_edges < g->M || cap_edges >= (uint64_t(1) << 32))
    return fail(c, PRH_ERROR_INVALID_ARGUMENT, "edge capacity must hold the graph and stay below 2^32");
  PRH_HIP(c, hipSetDevice(c->device));
  PRH_CHECK(build_rank_plan(c, const_cast<prh_graph *>(g)));
  auto *S = new prh_stream;
  S->ctx = c;
  S->g = g;
  S->method = method;
  S->variant = method == PRH_PRUNE_DYNAMIC_FRONTIER ? prh::kPrune
               : method == PRH_DYNAMIC_FRONTIER      ? prh::kFrontier
                                                     : prh::kPlain;
  S->opt = o;
  S->opt.repeat = 1;
  S->N = g->N;
  S->M = g->M;
  S->cap_edges = cap_edges;
  S->cap_batch = max_batch;
  const uint32_t N = g->N;
  const size_t N1 = size_t(N) + 1;
  hipStream_t s = c->stream;
  int st = PRH_SUCCESS;
  auto step = [&](int r) {
    if (st == PRH_SUCCESS) st = r;
  };
  for (int k = 0; k < 2; ++k) {
    step(stream_alloc(S, &S->in_off[k], N1));
    step(stream_alloc(S, &S->in_src[k], cap_edges));
    step(stream_alloc(S, &S->out_off[k], N1));
    step(stream_alloc(S, &S->out_dst[k], cap_edges));
    step(stream_alloc(S, &S->rank[k], N));
    step(stream_alloc(S, &S->contrib[k], N));
  }
  step(stream_alloc(S, &S->outdeg, N));
  step(stream_alloc(S, &S->new_deg, N1));
  step(stream_alloc(S, &S->sched_order, N));
  step(stream_alloc(S, &S->delta, N));
  step(stream_alloc(S, &S->changed, N));
  step(stream_alloc(S, &S->d_batch, 4 * size_t(max_batch) + 2));
  for (int k = 0; k < 4; ++k) {
    step(stream_alloc(S, &S->keys[k], max_batch));
    step(stream_alloc(S, &S->compact[k], max_batch));
  }
  step(stream_alloc(S, &S->sorted, max_batch));
  step(stream_alloc(S, &S->valid, max_batch));
  step(stream_alloc(S, &S->pos, max_batch));
  step(stream_alloc(S, &S->counts, 4));
  step(stream_alloc(S, &S->vaff, N));
  step(stream_alloc(S, &S->naff, N));
  step(stream_alloc(S, &S->ctl, 4));
  step(stream_alloc(S, &S->err, size_t(o.max_iterations) + 1));
  if (st == PRH_SUCCESS && hipHostMalloc(reinterpret_cast<void **>(&S->h_batch), (4 * size_t(max_batch) + 2) * 4) !=
                               hipSuccess) {
    (void)hipGetLastError();
    S->h_batch = nullptr;
    st = fail(c, PRH_ERROR_OUT_OF_MEMORY, "pinned batch staging");
  }
  // Temporary storage of the sort and of the scans (the largest scan is the N + 1 offsets).
  if (st == PRH_SUCCESS) {
    hipError_t e = prh::sort_keys_u64(nullptr, S->sort_bytes, S->keys[0], S->sorted, max_batch, s);
    size_t b1 = 0, b2 = 0;
    if (e == hipSuccess) e = prh::scan_u32(nullptr, b1, S->valid, S->pos, max_batch, s);
    if (e == hipSuccess) e = prh::scan_u32(nullptr, b2, S->new_deg, S->in_off[0], N1, s);
    S->scan_bytes = std::max(b1, b2);
    if (e != hipSuccess) st = fail(c, PRH_ERROR_DEVICE, "stream temporaries: %s", hipGetErrorString(e));
  }
  step(stream_alloc(S, &S->sort_tmp, S->sort_bytes));
  step(stream_alloc(S, &S->scan_tmp, S->scan_bytes));
  // Rank plan: the graph's sections, with the session's own chunk tables (refreshed after every batch).
  if (st == PRH_SUCCESS) {
    S->plan = g->plan;
    S->chunk = std::max<uint32_t>(c->tuning.huge_degree, 64);
    const uint32_t nc = g->plan.n_chunked;
    S->n_items = nc ? uint32_t((g->chunk_tables_bytes / 4 - nc - 1) / 3) : 0;
    uint32_t *t = nullptr;
    step(stream_alloc(S, &t, 3 * size_t(S->n_items) + nc + 1));
    double *part = nullptr;
    step(stream_alloc(S, &part, std::max<uint32_t>(S->n_items, 1)));
    if (st == PRH_SUCCESS) {
      PRH_HIP(c, hipMemcpyAsync(t, g->plan.item_row, size_t(S->n_items) * 4, hipMemcpyDeviceToDevice, s));
      PRH_HIP(c, hipMemcpyAsync(t + 3 * size_t(S->n_items), g->plan.row_first, (size_t(nc) + 1) * 4,
                                hipMemcpyDeviceToDevice, s));
      S->plan.item_row = t;
      S->plan.item_beg = t + S->n_items;
      S->plan.item_end = t + 2 * size_t(S->n_items);
      S->plan.row_first = t + 3 * size_t(S->n_items);
      S->plan.partial = part;
    }
    S->stat_slots = S->plan.blocks + (nc ? prh::finish_chunked_grid(nc) : 0);
    step(stream_alloc(S, &S->stats, 2 * size_t(std::max<uint32_t>(S->stat_slots, 1))));
    S->pull_wave = rows_above(g->sorted_deg, 64);
    S->pull_group = rows_above(g->sorted_deg, 4);
  }
  // Current graph: the layout's in-CSR as is; its out-CSR with every row sorted (the batch tidy searches it).
  if (st == PRH_SUCCESS) {
    PRH_HIP(c, hipMemcpyAsync(S->in_off[0], g->l_in_off, N1 * 4, hipMemcpyDeviceToDevice, s));
    PRH_HIP(c, hipMemcpyAsync(S->in_src[0], g->l_in_src, g->M * 4, hipMemcpyDeviceToDevice, s));
    PRH_HIP(c, hipMemcpyAsync(S->out_off[0], g->l_out_off, N1 * 4, hipMemcpyDeviceToDevice, s));
    PRH_HIP(c, hipMemcpyAsync(S->outdeg, g->l_outdeg, size_t(N) * 4, hipMemcpyDeviceToDevice, s));
    size_t sb = 0;
    hipError_t e = prh::sort_rows(nullptr, sb, g->l_out_dst, S->out_dst[0], g->M, N, g->l_out_off, s);
    void *tmp = nullptr;
    if (e == hipSuccess) st = dev_alloc(c, &tmp, sb);
    if (st == PRH_SUCCESS && e == hipSuccess) e = prh::sort_rows(tmp, sb, g->l_out_dst, S->out_dst[0], g->M, N,
                                                                 g->l_out_off, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    dev_free(c, tmp, sb);
    if (e != hipSuccess && st == PRH_SUCCESS) st = fail(c, PRH_ERROR_DEVICE, "stream setup: %s", hipGetErrorString(e));
  }
  // Out-degree schedule of the extension (fixed for the session; rows that outgrow the low class stay correct).
  if (st == PRH_SUCCESS && S->variant >= prh::kFrontier) {
    prh::ExtendArgs xo{};
    st = out_schedule(c, const_cast<prh_graph *>(g), xo);
    if (st == PRH_SUCCESS) {  // the session's kernels take the huge rows as high ones (a wave each)
      PRH_HIP(c, hipMemcpyAsync(S->sched_order, xo.sched.order, size_t(N) * 4, hipMemcpyDeviceToDevice, s));
      S->sched_low = xo.sched.n_low;
      S->sched_high = xo.sched.n_high + xo.sched.n_huge;
    }
  }
  // Initial ranks (layout order): the caller's, or a static solve of the base graph.
  if (st == PRH_SUCCESS) {
    double *tmp = nullptr;
    st = dev_alloc(c, reinterpret_cast<void **>(&tmp), size_t(N) * 8);
    if (st == PRH_SUCCESS) {
      if (initial_ranks) {
        for (uint32_t v = 0; v < N && st == PRH_SUCCESS; ++v)
          if (!std::isfinite(initial_ranks[v])) st = fail(c, PRH_ERROR_INVALID_ARGUMENT, "initial ranks not finite");
        if (st == PRH_SUCCESS && hipMemcpy(tmp, initial_ranks, size_t(N) * 8, hipMemcpyHostToDevice) != hipSuccess)
          st = fail(c, PRH_ERROR_DEVICE, "initial ranks upload");
      } else {
        prh_stats stt{};
        st = solve_impl(c, g, PRH_STATIC, &o, nullptr, nullptr, 0, nullptr, 0, nullptr, tmp, &stt, false);
      }
      if (st == PRH_SUCCESS) {
        hipError_t e = prh::launch_gather_f64(tmp, g->perm, N, S->rank[0], s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = fail(c, PRH_ERROR_DEVICE, "initial ranks: %s", hipGetErrorString(e));
      }
      dev_free(c, tmp, size_t(N) * 8);
    }
  }
  if (st != PRH_SUCCESS) {
    stream_free(S);
    delete S;
    return st;
  }
  *out = S;
  return PRH_SUCCESS;
}

int prh_stream_update(prh_stream *S, const uint32_t *del_u, const uint32_t *del_v, size_t nd, const uint32_t *ins_u,
                      const uint32_t *ins_v, size_t ni, double *ranks, prh_stats *stats) {
  if (!S) return PRH_ERROR_INVALID_ARGUMENT;
  prh_context *c = S->ctx;
  if (S->broken) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "stream is unusable after an earlier failed update");
  if (nd > S->cap_batch || ni > S->cap_batch)
    return fail(c, PRH_ERROR_INVALID_ARGUMENT, "batch of %zu deletions / %zu insertions exceeds the capacity %u", nd,
                ni, S->cap_batch);
  if ((nd && (!del_u || !del_v)) || (ni && (!ins_u || !ins_v)))
    return fail(c, PRH_ERROR_INVALID_ARGUMENT, "null batch array");
  if (S->M + ni > S->cap_edges)
    return fail(c, PRH_ERROR_OUT_OF_MEMORY, "the graph could outgrow the stream's edge capacity %llu",
                (unsigned long long)S->cap_edges);
  PRH_HIP(c, hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const uint32_t cap = S->cap_batch;
  uint32_t *h = S->h_batch;
  if (nd) {
    std::memcpy(h, del_u, nd * 4);
    std::memcpy(h + cap, del_v, nd * 4);
  }
  if (ni) {
    std::memcpy(h + 2 * size_t(cap), ins_u, ni * 4);
    std::memcpy(h + 3 * size_t(cap), ins_v, ni * 4);
  }
  h[4 * size_t(cap)] = uint32_t(nd);
  h[4 * size_t(cap) + 1] = uint32_t(ni);
  PRH_HIP(c, hipEventRecord(c->ev[0], s));
  PRH_HIP(c, hipMemcpyAsync(S->d_batch, h, (4 * size_t(cap) + 2) * 4, hipMemcpyHostToDevice, s));
  const int p = S->cur, q = 1 - p;
  if (!S->apply[p]) PRH_CHECK(capture_graph(c, [&]() { return stream_apply_body(S, p); }, &S->apply[p]));
  PRH_HIP(c, hipGraphLaunch(S->apply[p], s));
  PRH_HIP(c, hipEventRecord(c->ev[1], s));
  PRH_HIP(c, hipMemcpyAsync(c->pinned, S->ctl + 3, 4, hipMemcpyDeviceToHost, s));
  PRH_HIP(c, hipMemcpyAsync(c->pinned + 1, S->in_off[q] + S->N, 4, hipMemcpyDeviceToHost, s));
  PRH_HIP(c, hipStreamSynchronize(s));
  const int flags = *reinterpret_cast<const int *>(c->pinned);
  if (flags) {
    S->broken = true;
    return fail(c, PRH_ERROR_INVALID_ARGUMENT, flags & 1 ? "batch vertex id out of range"
                                                          : "the batch leaves a vertex without out-edges (dead end)");
  }
  S->M = *reinterpret_cast<const uint32_t *>(c->pinned + 1);
  S->cur = q;
  if (!S->iter[q]) PRH_CHECK(capture_graph(c, [&]() { return stream_iter_body(S, q); }, &S->iter[q]));
  int n = 0;
  for (int launched = 0; launched < S->opt.max_iterations; launched += kStreamChunk) {
    PRH_HIP(c, hipGraphLaunch(S->iter[q], s));
    PRH_HIP(c, hipMemcpyAsync(c->pinned, S->ctl, 4, hipMemcpyDeviceToHost, s));
    PRH_HIP(c, hipStreamSynchronize(s));
    n = *reinterpret_cast<const int *>(c->pinned);
    if (n) break;
  }
  if (n & 1) {  // the result is in buffer 1: bring it to buffer 0 for the next batch
    PRH_HIP(c, prh::launch_copy_f64(S->rank[0], S->rank[1], S->N, s));
    PRH_HIP(c, prh::launch_copy_f64(S->contrib[0], S->contrib[1], S->N, s));
  }
  PRH_HIP(c, hipEventRecord(c->ev[2], s));
  PRH_HIP(c, hipEventSynchronize(c->ev[2]));
  float t01 = 0, t12 = 0;
  PRH_HIP(c, hipEventElapsedTime(&t01, c->ev[0], c->ev[1]));
  PRH_HIP(c, hipEventElapsedTime(&t12, c->ev[1], c->ev[2]));
  if (stats) {
    stats->iterations = n;
    stats->initialization_time = t01;  // batch upload, application and marking
    stats->marking_time = 0;
    stats->computation_time = t12;
    stats->time = t01 + t12;
  }
  if (ranks) PRH_CHECK(prh_stream_ranks(S, ranks));
  return PRH_SUCCESS;
}

int prh_stream_ranks(prh_stream *S, double *ranks) {
  if (!S || !ranks) return S ? fail(S->ctx, PRH_ERROR_INVALID_ARGUMENT, "null ranks") : PRH_ERROR_INVALID_ARGUMENT;
  prh_context *c = S->ctx;
  PRH_CHECK(ensure(c, c->contrib[0], S->N));
  PRH_HIP(c, prh::launch_scatter_f64(S->rank[0], S->g->perm, S->N, c->contrib[0].p, c->stream));
  PRH_HIP(c, hipMemcpyAsync(ranks, c->contrib[0].p, size_t(S->N) * 8, hipMemcpyDeviceToHost, c->stream));
  PRH_HIP(c, hipStreamSynchronize(c->stream));
  return PRH_SUCCESS;
}

uint64_t prh_stream_size(const prh_stream *S) { return S ? S->M : 0; }

void prh_stream_destroy(prh_stream *S) {
  if (!S) return;
  (void)hipSetDevice(S->ctx->device);
  (void)hipStreamSynchronize(S->ctx->stream);
  stream_free(S);
  delete S;
}

}  // extern "C"
