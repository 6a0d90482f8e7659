// plans.hip -- the degree-ordered layout of an uploaded graph and the launch plans built from it: the row
// plan of the rank kernel, the source-sliced plan, the propagation-blocked plan, and the out-degree schedule of
// the extension kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "capi_internal.hpp"

namespace prhc {

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
  PRH_CHECK(ensure_in_csr(c, g));
  const prh_tuning &t = c->tuning;
  const int S = t.source_slices;
  if (g->sliced_valid && g->sliced_slices == S && g->sliced_huge == t.huge_degree &&
      g->sliced_block_edges == t.block_edges)
    return PRH_SUCCESS;
  PrepTimer timer(g->sliced_prep);
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
  const uint32_t N = bc.plan.N;
  dev_free(c, bc.base_off, bc.base_off ? (size_t(N) + 1) * 4 : 0);
  dev_free(c, bc.base_src, bc.base_src ? std::max<size_t>(bc.base_cap, 4) * 4 : 0);
  dev_free(c, bc.slot_of, bc.slot_of ? std::max<size_t>(bc.slot_of_n, 4) * 4 : 0);
  free_overlay(c, bc);
  bc.base_off = bc.base_src = bc.slot_of = nullptr;
  bc.base_M = bc.base_cap = bc.slot_of_n = 0;
}

void free_overlay(prh_context *c, BinnedCache &bc) {
  const size_t nb = (size_t(bc.plan.N) + std::max(bc.B, 1) - 1) / std::max(bc.B, 1);
  dev_free(c, bc.ov_key, bc.ov_key ? bc.ov_cap * 8 : 0);
  dev_free(c, bc.ovX, bc.ovX ? bc.ov_cap * 8 : 0);
  dev_free(c, bc.ov_off, bc.ov_off ? (nb + 1) * 4 : 0);
  bc.ov_key = nullptr;
  bc.ovX = nullptr;
  bc.ov_off = nullptr;
  bc.ov_cap = 0;
  bc.plan.ov_n = 0;
  bc.plan.ov_key = nullptr;
  bc.plan.ov_off = nullptr;
  bc.plan.ovX = nullptr;
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
  PrepTimer timer(bc.prep);
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
  // Items: expand items of 16K pairs whatever the graph's size, accumulate items of at most 2 per CU over all slots
  // (fewer split bins, hence fewer partial rows), at least 16K slots. Measured per update (B = 8192): on RMAT-24,
  // expand spans of 6.4K, 8K, 12K, 12.9K, 16K, 32K and 64K entries 763, 697, 684, 712, 675-709, 724 and 744 us;
  // on RMAT-26 with edge factor 30 (0.78G pairs), 8K, 12K, 16K, 24K, 32K, 64K and 95K (Q / (32 CUs)) 4418-4450,
  // 4348-4350, 4257-4366, 4414, 4315, 4364 and 4513-4574 us. Accumulate spans of M/(k CUs) for k = 8, 4, 3, 2, 3/2,
  // 1 and 1/2 on RMAT-24: 733, 709, 700-711, 675-696, 710, 807 and 1195 us (k = 2: as many items as the two resident
  // workgroups of every CU); on RMAT-26, k = 1, 2, 3 and 4: 5580, 4257-4366, 4246 and 4316 us.
  hipDeviceProp_t prop{};
  PRH_HIP(c, hipGetDeviceProperties(&prop, c->device));
  const uint64_t cus = uint64_t(std::max(prop.multiProcessorCount, 1));
  const uint64_t span_ex = 16384;
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
int build_binned_plan(prh_context *c, prh_graph *g, bool keep_slot) {
  // A graph with its out-CSR can be updated in place, which patches the plan through its slot map and the map's
  // inverse (in-CSR edge -> plan slot). Both are built with the plan, as part of its one-time preparation.
  const bool updatable = g->l_out_off != nullptr;
  BinnedCache &bc = g->bin;
  if (!(bc.valid && bc.B == c->tuning.blocking && (bc.slot || !(keep_slot || updatable)))) {
    // A new plan is built from the current in-CSR; the old plan's base, and the net change kept against it, go.
    PRH_CHECK(ensure_in_csr(c, g));
    g->net_del_n = g->net_ins_n = 0;
  }
  PRH_CHECK(build_binned(c, g->N, g->M, g->l_in_off, g->l_in_src, g->sorted_deg, keep_slot || updatable, bc));
  if (!updatable || bc.B > 8192 || !bc.plan.slot || bc.slot_of || bc.base_off) return PRH_SUCCESS;
  prhc::Prep inv;
  {
    PrepTimer timer(inv);
    PRH_CHECK(dev_alloc(c, reinterpret_cast<void **>(&bc.slot_of), std::max<size_t>(g->M, 4) * 4));
    bc.slot_of_n = g->M;
    if (g->M) PRH_HIP(c, prh::launch_invert(bc.plan.slot, uint32_t(g->M), bc.slot_of, c->stream));
    PRH_HIP(c, hipStreamSynchronize(c->stream));
  }
  bc.prep.ms = (bc.prep.reported ? 0.0f : bc.prep.ms) + inv.ms;
  bc.prep.reported = false;
  return PRH_SUCCESS;
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
  PRH_CHECK(ensure_in_csr(c, g));
  const prh_tuning &t = c->tuning;
  if (g->plan_valid && std::memcmp(&g->plan_tuning, &t, sizeof(t)) == 0) return PRH_SUCCESS;
  PrepTimer timer(g->plan_prep);
  const uint32_t N = g->N;
  const std::vector<uint32_t> &deg = g->sorted_deg;
  const uint64_t kEdgesPerBlock = t.block_edges;
  PRH_HIP(c, hipStreamSynchronize(c->stream));  // the old tables may still be in use
  auto above = [&](uint32_t x) { return layout_rows_above(g, x); };
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
  } else if (g->layout == 1) {  // rows in the caller's order: one width for all, half the largest in-degree
    const uint32_t dmax = N ? *std::max_element(deg.begin(), deg.end()) : 1;
    int width = 1;
    while (width < 32 && uint32_t(2 * width) < dmax) width *= 2;
    row_section(prh::kGroupRows, width, 0, N);
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

// Banded plan: the window reaches H = the largest |source - destination| of the layout in-CSR (rows in the caller's
// order) on each side of a chunk, capped at what the larger ring holds (edges beyond it are gathered directly). The
// smaller ring (two workgroups per CU) serves H <= 4544; each workgroup sweeps about N / wgs rows.
int build_window_plan(prh_context *c, prh_graph *g) {
  PRH_CHECK(ensure_in_csr(c, g));
  if (g->win_valid) return PRH_SUCCESS;
  if (g->layout != 1) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "the banded plan needs rows in the caller's order");
  if (g->N > (1u << 31)) return fail(c, PRH_ERROR_UNSUPPORTED, "the banded plan takes at most 2^31 vertices");
  PrepTimer timer(g->win_prep);
  hipStream_t s = c->stream;
  uint32_t *spread = nullptr;
  PRH_CHECK(dev_alloc(c, reinterpret_cast<void **>(&spread), 4));
  uint32_t H = 0;
  hipError_t e = prh::launch_max_spread(g->l_in_off, g->l_in_src, g->N, spread, s);
  if (e == hipSuccess) e = hipMemcpyAsync(&H, spread, 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  dev_free(c, spread, 4);
  if (e != hipSuccess) return fail(c, PRH_ERROR_DEVICE, "banded plan: %s", hipGetErrorString(e));
  const uint32_t R = uint32_t(prh::kBinBlock);
  const uint32_t h_small = (prh::kWindowRingSmall - R) / 2, h_large = (prh::kWindowRingLarge - R) / 2;
  hipDeviceProp_t prop{};
  PRH_HIP(c, hipGetDeviceProperties(&prop, c->device));
  const uint32_t cus = uint32_t(std::max(prop.multiProcessorCount, 1));
  prh::WindowPlan W{};
  W.N = g->N;
  W.S = H <= h_small ? prh::kWindowRingSmall : prh::kWindowRingLarge;
  W.H = std::min(H, W.S == prh::kWindowRingSmall ? h_small : h_large);
  const uint32_t chunks = (g->N + R - 1) / R;
  W.wgs = std::max(1u, std::min((W.S == prh::kWindowRingSmall ? 2u : 1u) * cus, chunks));
  g->win = W;
  g->win_valid = true;
  return PRH_SUCCESS;
}

// The layout in the caller's order (mode 1: row i is vertex i; the layout arrays are copies of the uploaded ones) or
// by in-degree (mode 0, from the uploaded arrays read back to the host). Plans built on the old layout are dropped.
int relayout(prh_context *c, prh_graph *g, int mode) {
  if (g->layout == mode) return PRH_SUCCESS;
  if (g->updated) return fail(c, PRH_ERROR_UNSUPPORTED, "a graph updated in place keeps its layout");
  hipStream_t s = c->stream;
  PRH_HIP(c, hipStreamSynchronize(s));
  const uint32_t N = g->N;
  const uint64_t M = g->M;
  const size_t N1 = size_t(N) + 1;
  const bool with_out = g->out_off != nullptr;
  // Everything built on the old layout.
  g->plan_valid = false;
  g->win_valid = false;
  free_sliced(c, g);
  free_binned(c, g->bin);
  free_binned(c, g->bin_t);
  if (g->out_valid) {
    dev_free(c, g->out_chunks, std::max<size_t>(3 * size_t(g->out_n_chunks), 1) * 4);
    g->out_chunks = nullptr;
    g->out_n_chunks = 0;
    g->out_valid = false;
  }
  dev_free(c, g->l_in_eid, (M ? M : 4) * 4);
  dev_free(c, g->l_out_eid, (M ? M : 4) * 4);
  g->l_in_eid = g->l_out_eid = nullptr;
  dev_free(c, g->perm, size_t(N) * 4);
  dev_free(c, g->inv, size_t(N) * 4);
  dev_free(c, g->l_in_off, N1 * 4);
  dev_free(c, g->l_in_src, std::max<size_t>(g->l_in_cap, 4) * 4);
  dev_free(c, g->l_outdeg, size_t(N) * 4);
  dev_free(c, g->l_out_off, N1 * 4);
  dev_free(c, g->l_out_dst, std::max<size_t>(g->l_cap, 4) * 4);
  g->perm = g->inv = g->l_in_off = g->l_in_src = g->l_outdeg = g->l_out_off = g->l_out_dst = nullptr;
  g->out_sorted = false;
  std::vector<uint32_t> in_off(N1), outdeg(N);
  PRH_HIP(c, hipMemcpy(in_off.data(), g->in_off, N1 * 4, hipMemcpyDeviceToHost));
  PRH_HIP(c, hipMemcpy(outdeg.data(), g->outdeg, size_t(N) * 4, hipMemcpyDeviceToHost));
  if (mode == 0) return build_layout(c, g, in_off.data(), outdeg.data(), with_out);
  PrepTimer timer(g->layout_prep);
  auto copy = [&](uint32_t **dst, const uint32_t *src, size_t n) -> int {
    PRH_CHECK(dev_alloc(c, reinterpret_cast<void **>(dst), std::max<size_t>(n, 4) * 4));
    if (n) PRH_HIP(c, hipMemcpyAsync(*dst, src, n * 4, hipMemcpyDeviceToDevice, s));
    return PRH_SUCCESS;
  };
  g->l_cap = g->l_in_cap = std::max<uint64_t>(M, 4);
  PRH_CHECK(copy(&g->l_in_off, g->in_off, N1));
  PRH_CHECK(copy(&g->l_in_src, g->in_src, M));
  PRH_CHECK(copy(&g->l_outdeg, g->outdeg, N));
  if (with_out) {
    PRH_CHECK(copy(&g->l_out_off, g->out_off, N1));
    PRH_CHECK(copy(&g->l_out_dst, g->out_dst, M));
  }
  PRH_CHECK(dev_alloc(c, reinterpret_cast<void **>(&g->perm), std::max<size_t>(N, 1) * 4));
  PRH_CHECK(dev_alloc(c, reinterpret_cast<void **>(&g->inv), std::max<size_t>(N, 1) * 4));
  PRH_HIP(c, prh::launch_iota_u32(g->perm, N, s));
  PRH_HIP(c, prh::launch_iota_u32(g->inv, N, s));
  g->sorted_deg.resize(N);
  for (uint32_t v = 0; v < N; ++v) g->sorted_deg[v] = in_off[v + 1] - in_off[v];
  PRH_HIP(c, hipStreamSynchronize(s));
  g->layout = 1;
  return PRH_SUCCESS;
}

// Out-degree schedule of the layout's out-CSR for the extension kernels, built once per graph (one readback of
// the part sizes, as the reference reads its partition point back, pagerankCuda.hxx:633), with the chunk table of
// the rows above kExtendChunk edges, so that a hub's edges are spread over many waves.
int out_schedule(prh_context *c, prh_graph *g, prh::ExtendArgs &E) {
  const uint32_t N = g->N;
  if (!g->out_valid) {
    PrepTimer timer(g->out_prep);
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

// The degree-ordered layout of a freshly uploaded graph (see layout.hip): perm = vertices by in-degree, descending,
// stable in id (counting sort on the host; degrees above kDegreeBuckets are few and sorted directly), then the
// relabelled in-CSR (rows sorted by new source id), out-degrees and, if present, out-CSR on the device.
constexpr uint32_t kDegreeBuckets = 1u << 20;

int build_layout(prh_context *c, prh_graph *g, const uint32_t *in_off, const uint32_t *outdeg, bool with_out) {
  PrepTimer timer(g->layout_prep);
  g->layout = 0;
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
  g->l_cap = g->l_in_cap = std::max<uint64_t>(M, 4);
  PRH_CHECK(alloc(&g->l_in_src, g->l_in_cap));
  PRH_CHECK(alloc(&g->l_outdeg, N));
  if (with_out) {
    PRH_CHECK(alloc(&g->l_out_off, size_t(N) + 1));
    PRH_CHECK(alloc(&g->l_out_dst, g->l_cap));
  }
  PRH_HIP(c, hipMemcpyAsync(g->perm, perm.data(), size_t(N) * 4, hipMemcpyHostToDevice, s));
  PRH_HIP(c, hipMemcpyAsync(g->l_in_off, loff.data(), (size_t(N) + 1) * 4, hipMemcpyHostToDevice, s));
  if (with_out) PRH_HIP(c, hipMemcpyAsync(g->l_out_off, oout.data(), (size_t(N) + 1) * 4, hipMemcpyHostToDevice, s));
  PRH_HIP(c, prh::launch_invert(g->perm, N, g->inv, s));
  PRH_HIP(c, prh::launch_gather_u32(g->outdeg, g->perm, N, g->l_outdeg, s));
  // Relabelled in-rows (and out-rows) into a scratch array, then sorted per row into place: in-rows by new source
  // id for the rank kernels' gathers, out-rows by new target id for the batch tidy's binary searches.
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
    if (e == hipSuccess && with_out)
      e = prh::launch_relabel_rows(g->out_off, g->out_dst, g->perm, g->inv, g->l_out_off, N, tmp, s);
    if (e == hipSuccess && with_out) e = prh::sort_rows(sort_tmp, sort_bytes, tmp, g->l_out_dst, M, N, g->l_out_off, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) st = fail(c, PRH_ERROR_DEVICE, "layout row sort: %s", hipGetErrorString(e));
  }
  (void)hipStreamSynchronize(s);
  dev_free(c, sort_tmp, sort_bytes);
  dev_free(c, tmp, (M ? M : 4) * 4);
  PRH_CHECK(st);
  g->out_sorted = with_out;
  return PRH_SUCCESS;
}

}  // namespace prhc
