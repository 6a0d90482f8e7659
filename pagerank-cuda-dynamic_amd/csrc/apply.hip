// apply.hip -- prh_graph_update: an edge batch applied to an uploaded graph on the device, and the graph's
// propagation-blocked plan patched in place instead of rebuilt.
//
// The reference builds every updated graph y on the host (tidyBatchUpdateU, applyBatchUpdateOmpU,
// transposeWithDegreeOmp; batch.hxx:201-247, transpose.hxx:109-134, main.cu:80-86) and uploads it for each solve,
// whose partition it then times (pagerankCuda.hxx:811-814). Uploading y here would mean a new degree-ordered layout
// and a new blocked plan per batch: three radix sorts over the edges, about as long as a static solve. Instead:
//   tidy        the batch's keys sorted per CSR and kept as tidyBatchUpdateU keeps them (deletions of existing edges,
//               insertions of absent ones, each once), with the session's kernels (stream.hip);
//   apply       the layout's CSRs rewritten into new arrays (rows stay ascending), the out-degrees recounted. With
//               the blocked plan patched (below) the in-CSR waits: its iterations read none, so the in-CSR is kept
//               as the plan's base and the net change since (deletions and insertions, composed per batch on the
//               host) and built by ensure_in_csr when a reader needs it (the row, banded and sliced plans, a new
//               blocked plan, a read-back, a session);
//   patch       the blocked plan keeps its slots: a deleted edge that is a slot gets the dead flag, an inserted one
//               that is not becomes an entry of the plan's overlay (sorted by destination), which accumulate adds
//               per bin. The vertex order (the layout's permutation) stays that of the upload.
// Everything else that depends on the edges (the row plan's chunk tables, the sliced plan, the extension schedule,
// the adjoint's plan) is rebuilt on its next use. The update's time is reported by the next solve, in its
// initialization and preparation times.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "capi_internal.hpp"

namespace {

using prhc::fail;
using prhc::Temps;

// Vertices with no out-edge. A wave counts its lanes' zeros with a ballot per step; lane 0 takes part in every step
// any lane of its wave takes part in (its vertex is the wave's lowest), so its count is the wave's.
__global__ __launch_bounds__(256) void count_zero_kernel(const uint32_t *deg, uint32_t N, uint32_t *out) {
  uint32_t z = 0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x)
    z += uint32_t(__popcll(__ballot(deg[v] == 0)));
  if (z && (threadIdx.x & 63) == 0) atomicAdd(out, z);
}

// The tidied keys of one CSR side: keys sorted, first of each key whose edge exists (deletions) or is absent
// (insertions), compacted into out with their count in *count.
int tidy(prh_context *c, const unsigned long long *keys, uint32_t cap, const uint32_t *off, const uint32_t *idx,
         bool deletions, unsigned long long *sorted, uint32_t *valid, uint32_t *pos, void *sort_tmp, size_t sort_bytes,
         void *scan_tmp, size_t scan_bytes, unsigned long long *out, uint32_t *count) {
  hipStream_t s = c->stream;
  if (!cap) return PRH_SUCCESS;
  PRH_HIP(c, prh::sort_keys_u64(sort_tmp, sort_bytes, keys, sorted, cap, s));
  PRH_HIP(c, prh::launch_batch_valid(sorted, cap, off, idx, deletions, valid, s));
  PRH_HIP(c, prh::scan_u32(scan_tmp, scan_bytes, valid, pos, cap, s));
  PRH_HIP(c, prh::launch_batch_compact(sorted, valid, pos, cap, out, count, s));
  return PRH_SUCCESS;
}

}  // namespace

namespace prhc {

// The plan's in-place patch for one batch (tidied in-CSR keys dkeys / ikeys, counts on the device, nd / ni on the
// host). On the first update the CSR the plan was built from becomes its base (ownership moves from the graph).
int patch_plan(prh_context *c, prh_graph *g, const unsigned long long *dkeys, const uint32_t *dcount, uint32_t nd,
               const unsigned long long *ikeys, const uint32_t *icount, uint32_t ni) {
  BinnedCache &bc = g->bin;
  hipStream_t s = c->stream;
  const uint32_t N = g->N;
  const uint32_t B = uint32_t(bc.B);
  const uint32_t nb = (N + B - 1) / B;
  if (!bc.slot_of) {  // slot map inverse: base in-CSR slot -> plan slot (built with the plan when it has room)
    PRH_CHECK(dev_alloc(c, reinterpret_cast<void **>(&bc.slot_of), std::max<size_t>(bc.base_M, 4) * 4));
    bc.slot_of_n = bc.base_M;
    if (bc.base_M) PRH_HIP(c, prh::launch_invert(bc.plan.slot, uint32_t(bc.base_M), bc.slot_of, s));
  }
  Temps tmp(c);
  uint8_t *gone = nullptr;
  unsigned long long *fresh = nullptr, *cand = nullptr, *cand_sorted = nullptr;
  uint32_t *words = nullptr;  // [0] the current overlay's size when it has none yet, [1] the next one's
  const uint32_t ov_cap = uint32_t(bc.ov_cap);
  PRH_CHECK(tmp.get(&gone, std::max<uint32_t>(ov_cap, 1)));
  PRH_CHECK(tmp.get(&fresh, std::max<uint32_t>(ni, 1)));
  PRH_CHECK(tmp.get(&words, 2));
  PRH_HIP(c, prh::launch_fill_u8(gone, 0, std::max<uint32_t>(ov_cap, 1), s));
  PRH_HIP(c, prh::launch_fill_u32(words, 0, 2, s));
  const uint32_t *ov_count = bc.ov_off ? bc.ov_off + nb : words;
  uint16_t *dst16 = const_cast<uint16_t *>(bc.plan.dst16);
  PRH_HIP(c, prh::launch_patch_slots(dkeys, dcount, nd, true, bc.base_off, bc.base_src, bc.slot_of, dst16,
                                     bc.ov_key, ov_count, gone, nullptr, s));
  PRH_HIP(c, prh::launch_patch_slots(ikeys, icount, ni, false, bc.base_off, bc.base_src, bc.slot_of, dst16,
                                     bc.ov_key, ov_count, gone, fresh, s));
  const uint32_t total = ov_cap + ni;
  if (total == 0) return PRH_SUCCESS;
  PRH_CHECK(tmp.get(&cand, total));
  PRH_CHECK(tmp.get(&cand_sorted, total));
  PRH_HIP(c, prh::launch_overlay_candidates(bc.ov_key, ov_count, ov_cap, gone, fresh, ni, cand, words + 1, s));
  size_t sort_bytes = 0;
  PRH_HIP(c, prh::sort_keys_u64(nullptr, sort_bytes, cand, cand_sorted, total, s));
  uint8_t *sort_tmp = nullptr;
  PRH_CHECK(tmp.get(&sort_tmp, sort_bytes));
  PRH_HIP(c, prh::sort_keys_u64(sort_tmp, sort_bytes, cand, cand_sorted, total, s));
  PRH_HIP(c, hipMemcpyAsync(c->pinned, words + 1, 4, hipMemcpyDeviceToHost, s));
  PRH_HIP(c, hipStreamSynchronize(s));
  const uint32_t n = *reinterpret_cast<const uint32_t *>(c->pinned);
  if (n > bc.ov_cap || !bc.ov_off) {
    const size_t cap = std::max<size_t>(size_t(n) * 2, 4096);
    free_overlay(c, bc);
    PRH_CHECK(dev_alloc(c, reinterpret_cast<void **>(&bc.ov_key), cap * 8));
    PRH_CHECK(dev_alloc(c, reinterpret_cast<void **>(&bc.ovX), cap * 8));
    PRH_CHECK(dev_alloc(c, reinterpret_cast<void **>(&bc.ov_off), (size_t(nb) + 1) * 4));
    bc.ov_cap = cap;
  }
  PRH_HIP(c, prh::launch_copy_u64(bc.ov_key, cand_sorted, n, s));
  PRH_HIP(c, prh::launch_overlay_bounds(bc.ov_key, words + 1, uint32_t(bc.ov_cap), B, nb, bc.ov_off, nullptr, 0, s));
  PRH_HIP(c, hipStreamSynchronize(s));
  bc.plan.ov_n = n ? uint32_t(bc.ov_cap) : 0;
  bc.plan.ov_key = bc.ov_key;
  bc.plan.ov_off = bc.ov_off;
  bc.plan.ovX = bc.ovX;
  return PRH_SUCCESS;
}

// A layout-order CSR on the device (rows = layout rows, ids = layout ids) as the caller's CSR on the host: row v is
// layout row inv[v] with every id mapped through perm, ascending (relabelled and sorted per row on the device, then
// copied back). off receives N + 1 offsets, idx M ids.
int csr_to_caller(prh_context *c, const uint32_t *perm_d, const uint32_t *inv_d, uint32_t N, const uint32_t *loff_d,
                  const uint32_t *lidx_d, uint64_t M, uint32_t *off, uint32_t *idx) {
  hipStream_t s = c->stream;
  Temps tmp(c);
  uint32_t *ldeg = nullptr, *cdeg = nullptr, *coff = nullptr, *rows = nullptr, *sorted = nullptr;
  PRH_CHECK(tmp.get(&ldeg, N));
  PRH_CHECK(tmp.get(&cdeg, size_t(N) + 1));
  PRH_CHECK(tmp.get(&coff, size_t(N) + 1));
  PRH_CHECK(tmp.get(&rows, M));
  PRH_CHECK(tmp.get(&sorted, M));
  size_t scan_bytes = 0, sort_bytes = 0;
  PRH_HIP(c, prh::scan_u32(nullptr, scan_bytes, cdeg, coff, size_t(N) + 1, s));
  PRH_HIP(c, prh::sort_rows(nullptr, sort_bytes, rows, sorted, M, N, coff, s));
  uint8_t *work = nullptr;
  PRH_CHECK(tmp.get(&work, std::max(scan_bytes, sort_bytes)));
  PRH_HIP(c, prh::launch_degrees(loff_d, N, ldeg, s));
  PRH_HIP(c, prh::launch_gather_u32(ldeg, inv_d, N, cdeg, s));
  PRH_HIP(c, prh::launch_fill_u32(cdeg + N, 0, 1, s));
  PRH_HIP(c, prh::scan_u32(work, scan_bytes, cdeg, coff, size_t(N) + 1, s));
  PRH_HIP(c, prh::launch_relabel_rows(loff_d, lidx_d, inv_d, perm_d, coff, N, rows, s));
  PRH_HIP(c, prh::sort_rows(work, sort_bytes, rows, sorted, M, N, coff, s));
  PRH_HIP(c, hipMemcpyAsync(off, coff, (size_t(N) + 1) * 4, hipMemcpyDeviceToHost, s));
  if (M) PRH_HIP(c, hipMemcpyAsync(idx, sorted, M * 4, hipMemcpyDeviceToHost, s));
  PRH_HIP(c, hipStreamSynchronize(s));
  if (off[N] != M)
    return fail(c, PRH_ERROR_DEVICE, "CSR read-back: %u edges in the offsets, %llu expected", off[N],
                (unsigned long long)M);
  return PRH_SUCCESS;
}

// One CSR side of an update: the rows of (off, idx), M edges, less the tidied deletions dk plus the tidied insertions
// ik (sorted keys row << 32 | column; their counts on the device at dcnt / icnt, at most nd / ni), written to (noff,
// nidx) of capacity ncap. Rows keep their columns ascending.
int apply_side(prh_context *c, Temps &tmp, uint32_t N, uint64_t M, const uint32_t *off, const uint32_t *idx,
               const unsigned long long *dk, const uint32_t *dcnt, uint32_t nd, const unsigned long long *ik,
               const uint32_t *icnt, uint32_t ni, uint32_t *noff, uint32_t *nidx, uint64_t ncap) {
  hipStream_t s = c->stream;
  const size_t N1 = size_t(N) + 1;
  PRH_CHECK(ensure(c, c->upd_delta, N));
  PRH_CHECK(ensure(c, c->upd_changed, N));
  PRH_CHECK(ensure(c, c->upd_deg, N1));
  PRH_CHECK(ensure(c, c->upd_ranges, N));
  int *delta = c->upd_delta.p;
  uint8_t *changed = c->upd_changed.p;
  uint32_t *new_deg = c->upd_deg.p;
  uint4 *ranges = c->upd_ranges.p;
  uint32_t *chpos = reinterpret_cast<uint32_t *>(delta);  // positions of the changed rows, once the degrees are spent
  size_t scan_bytes = 0, b1 = 0;
  PRH_HIP(c, prh::scan_u32(nullptr, scan_bytes, new_deg, noff, N1, s));
  PRH_HIP(c, prh::scan_u8(nullptr, b1, changed, chpos, N, s));
  scan_bytes = std::max(scan_bytes, b1);
  // rewrite units: one per changed row and a closing one (a side has at most nd + ni changed rows)
  const uint32_t max_units = nd + ni + 1;
  uint8_t *scan_tmp = nullptr;
  uint32_t *units = nullptr;
  int *flags = nullptr;
  PRH_CHECK(tmp.get(&scan_tmp, scan_bytes));
  PRH_CHECK(tmp.get(&units, 4 * size_t(max_units) + 1));
  PRH_CHECK(tmp.get(&flags, 1));
  PRH_HIP(c, prh::launch_fill_i32(delta, 0, N, s));
  PRH_HIP(c, prh::launch_fill_u8(changed, 0, N, s));
  PRH_HIP(c, prh::launch_row_ranges(dk, dcnt, ik, icnt, nd + ni, delta, changed, ranges, s));
  PRH_HIP(c, prh::launch_new_degrees(off, delta, N, new_deg, flags, 0, s));
  PRH_HIP(c, prh::scan_u32(scan_tmp, scan_bytes, new_deg, noff, N1, s));
  PRH_HIP(c, prh::scan_u8(scan_tmp, scan_bytes, changed, chpos, N, s));
  PRH_HIP(c, prh::launch_rewrite(off, idx, noff, changed, chpos, units, ranges, dk, ik, N, max_units, M, ncap, nidx,
                                 s));
  PRH_HIP(c, prh::launch_insert(off, idx, noff, ranges, dk, ik, icnt, ni, nidx, s));
  return PRH_SUCCESS;
}

// The net change of the in-CSR since the plan's base after one more tidied batch (d: nd edges it deletes, all
// present before it; i: ni edges it inserts, all absent before it; ascending keys on the device). The net change D
// (base edges now absent) and d are disjoint, as are I (edges now present that the base lacks) and i; a key in both
// D + d and I + i is an edge the batch restores (in D and i) or removes again (in I and d), and it drops out. So
// D' = (D + d) - (I + i) and I' = (I + i) - (D + d): two merges, a search of each side in the other and two
// compactions, on the device; the new counts land in g->net_counts (read back by the caller).
int compose_net(prh_context *c, prh_graph *g, Temps &tmp, const unsigned long long *d, uint32_t nd,
                const unsigned long long *i, uint32_t ni) {
  hipStream_t s = c->stream;
  const uint32_t nx = uint32_t(g->net_del_n) + nd, ny = uint32_t(g->net_ins_n) + ni;
  unsigned long long *x = nullptr, *y = nullptr;
  uint32_t *vx = nullptr, *vy = nullptr, *px = nullptr, *py = nullptr;
  PRH_CHECK(tmp.get(&x, std::max<uint32_t>(nx, 1)));
  PRH_CHECK(tmp.get(&y, std::max<uint32_t>(ny, 1)));
  PRH_CHECK(tmp.get(&vx, std::max<uint32_t>(nx, 1)));
  PRH_CHECK(tmp.get(&vy, std::max<uint32_t>(ny, 1)));
  PRH_CHECK(tmp.get(&px, std::max<uint32_t>(nx, 1)));
  PRH_CHECK(tmp.get(&py, std::max<uint32_t>(ny, 1)));
  auto merged = [&](const unsigned long long *a, uint32_t na, const unsigned long long *b, uint32_t nb,
                    unsigned long long *out) -> int {
    if (na && nb) {
      size_t bytes = 0;
      PRH_HIP(c, prh::merge_keys_u64(nullptr, bytes, a, na, b, nb, out, s));
      uint8_t *work = nullptr;
      PRH_CHECK(tmp.get(&work, bytes));
      PRH_HIP(c, prh::merge_keys_u64(work, bytes, a, na, b, nb, out, s));
    } else if (na || nb) {
      PRH_HIP(c, hipMemcpyAsync(out, na ? a : b, size_t(na ? na : nb) * 8, hipMemcpyDeviceToDevice, s));
    }
    return PRH_SUCCESS;
  };
  PRH_CHECK(merged(g->net_del, uint32_t(g->net_del_n), d, nd, x));
  PRH_CHECK(merged(g->net_ins, uint32_t(g->net_ins_n), i, ni, y));
  // Room for the new net change (the old arrays are read by the merges above: wait for them before a free).
  auto room = [&](unsigned long long *&arr, uint64_t &cap, uint32_t n) -> int {
    if (cap >= n) return PRH_SUCCESS;
    const uint64_t want = std::max<uint64_t>(uint64_t(n) + n / 2, 4096);
    unsigned long long *fresh = nullptr;
    PRH_CHECK(dev_alloc(c, reinterpret_cast<void **>(&fresh), want * 8));
    PRH_HIP(c, hipStreamSynchronize(s));
    dev_free(c, arr, cap * 8);
    arr = fresh;
    cap = want;
    return PRH_SUCCESS;
  };
  PRH_CHECK(room(g->net_del, g->net_del_cap, nx));
  PRH_CHECK(room(g->net_ins, g->net_ins_cap, ny));
  if (!g->net_counts) PRH_CHECK(dev_alloc(c, reinterpret_cast<void **>(&g->net_counts), 8));
  PRH_HIP(c, prh::launch_fill_u32(g->net_counts, 0, 2, s));
  size_t scan_bytes = 0;
  PRH_HIP(c, prh::scan_u32(nullptr, scan_bytes, vx, px, std::max(nx, ny), s));
  uint8_t *scan_tmp = nullptr;
  PRH_CHECK(tmp.get(&scan_tmp, scan_bytes));
  if (nx) {
    PRH_HIP(c, prh::launch_absent_from(x, nx, y, ny, vx, s));
    PRH_HIP(c, prh::scan_u32(scan_tmp, scan_bytes, vx, px, nx, s));
    PRH_HIP(c, prh::launch_batch_compact(x, vx, px, nx, g->net_del, g->net_counts + 0, s));
  }
  if (ny) {
    PRH_HIP(c, prh::launch_absent_from(y, ny, x, nx, vy, s));
    PRH_HIP(c, prh::scan_u32(scan_tmp, scan_bytes, vy, py, ny, s));
    PRH_HIP(c, prh::launch_batch_compact(y, vy, py, ny, g->net_ins, g->net_counts + 1, s));
  }
  return PRH_SUCCESS;
}

// New arrays for a CSR side of M2 edges: the side's spare arrays when they hold M2 edges, else new ones with room for
// growth (1/256 of the edges), which later updates reuse.
int side_arrays(prh_context *c, uint32_t N, uint64_t M2, uint32_t *&spare_off, uint32_t *&spare_idx,
                uint64_t &spare_cap, uint32_t **noff, uint32_t **nidx, uint64_t *ncap) {
  const size_t N1 = size_t(N) + 1;
  if (spare_idx && spare_cap >= M2) {
    *noff = spare_off;
    *nidx = spare_idx;
    *ncap = spare_cap;
    spare_off = spare_idx = nullptr;
    spare_cap = 0;
    return PRH_SUCCESS;
  }
  dev_free(c, spare_off, N1 * 4);
  dev_free(c, spare_idx, std::max<size_t>(spare_cap, 4) * 4);
  spare_off = spare_idx = nullptr;
  spare_cap = 0;
  *ncap = std::max<uint64_t>(M2 + M2 / 256, 4);
  PRH_CHECK(dev_alloc(c, reinterpret_cast<void **>(noff), N1 * 4));
  return dev_alloc(c, reinterpret_cast<void **>(nidx), *ncap * 4);
}

void free_side(prh_context *c, uint32_t N, uint32_t *off, uint32_t *idx, uint64_t cap) {
  dev_free(c, off, (size_t(N) + 1) * 4);
  dev_free(c, idx, std::max<size_t>(cap, 4) * 4);
}

int ensure_in_csr(prh_context *c, prh_graph *g) {
  if (!g->in_stale) return PRH_SUCCESS;
  const BinnedCache &bc = g->bin;
  if (!bc.base_off) return fail(c, PRH_ERROR_DEVICE, "internal: an unbuilt in-CSR without its base");
  PRH_HIP(c, hipSetDevice(c->device));
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t s = c->stream;
  const uint32_t N = g->N;
  const uint32_t nd = uint32_t(g->net_del_n), ni = uint32_t(g->net_ins_n);
  uint32_t *noff = nullptr, *nidx = nullptr;
  uint64_t ncap = 0;
  int st = side_arrays(c, N, g->M, g->spare_in_off, g->spare_in_src, g->spare_in_cap, &noff, &nidx, &ncap);
  {
    Temps tmp(c);
    unsigned long long *none = nullptr;
    uint32_t *cnt = nullptr;
    if (st == PRH_SUCCESS) st = tmp.get(&none, 1);
    if (st == PRH_SUCCESS) st = tmp.get(&cnt, 2);
    const uint32_t counts[2] = {nd, ni};
    if (st == PRH_SUCCESS && hipMemcpyAsync(cnt, counts, 8, hipMemcpyHostToDevice, s) != hipSuccess)
      st = fail(c, PRH_ERROR_DEVICE, "in-CSR build: upload");
    const unsigned long long *dk = nd ? g->net_del : none, *ik = ni ? g->net_ins : none;
    if (st == PRH_SUCCESS)
      st = apply_side(c, tmp, N, bc.base_M, bc.base_off, bc.base_src, dk, cnt, nd, ik, cnt + 1, ni, noff, nidx, ncap);
    if (hipStreamSynchronize(s) != hipSuccess && st == PRH_SUCCESS) st = fail(c, PRH_ERROR_DEVICE, "in-CSR build: sync");
    uint32_t edges = 0;
    if (st == PRH_SUCCESS && hipMemcpy(&edges, noff + N, 4, hipMemcpyDeviceToHost) != hipSuccess)
      st = fail(c, PRH_ERROR_DEVICE, "in-CSR build: readback");
    if (st == PRH_SUCCESS && edges != g->M)
      st = fail(c, PRH_ERROR_DEVICE, "internal: the in-CSR build gave %u edges, %llu expected", edges,
                (unsigned long long)g->M);
  }
  if (st != PRH_SUCCESS) {
    free_side(c, N, noff, nidx, ncap);
    return st;
  }
  g->l_in_off = noff;
  g->l_in_src = nidx;
  g->l_in_cap = ncap;
  g->in_stale = false;
  const float ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
  g->update_prep.ms = (g->update_prep.reported ? 0.0f : g->update_prep.ms) + ms;
  g->update_prep.reported = false;
  return PRH_SUCCESS;
}

}  // namespace prhc

using namespace prhc;

extern "C" {

int prh_graph_csr(prh_context *c, const prh_graph *g, int which, uint32_t *offsets, uint32_t *ids) {
  if (!c) return PRH_ERROR_INVALID_ARGUMENT;
  if (!g || g->ctx != c || !offsets || (g->M && !ids) || (which != 0 && which != 1))
    return fail(c, PRH_ERROR_INVALID_ARGUMENT, "null or foreign graph, null array or bad CSR selector");
  if (g->N == 0) {
    offsets[0] = 0;
    return PRH_SUCCESS;
  }
  if (which == 1 && !g->l_out_off) return fail(c, PRH_ERROR_UNSUPPORTED, "the graph was uploaded without its out-CSR");
  PRH_HIP(c, hipSetDevice(c->device));
  if (which == 0) PRH_CHECK(ensure_in_csr(c, const_cast<prh_graph *>(g)));
  PRH_HIP(c, hipStreamSynchronize(c->stream));
  return csr_to_caller(c, g->perm, g->inv, g->N, which ? g->l_out_off : g->l_in_off,
                       which ? g->l_out_dst : g->l_in_src, g->M, offsets, ids);
}

int prh_graph_update(prh_context *c, prh_graph *g, const uint32_t *del_u, const uint32_t *del_v, size_t nd,
                     const uint32_t *ins_u, const uint32_t *ins_v, size_t ni, float *update_ms) {
  if (!c) return PRH_ERROR_INVALID_ARGUMENT;
  if (update_ms) *update_ms = 0.0f;
  if (!g || g->ctx != c) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "null or foreign graph");
  if ((nd && (!del_u || !del_v)) || (ni && (!ins_u || !ins_v))) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "null batch array");
  if (!g->l_out_off) return fail(c, PRH_ERROR_UNSUPPORTED, "an update needs the graph's out-CSR (upload it with one)");
  if (nd >= (size_t(1) << 31) || ni >= (size_t(1) << 31)) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "batch too large");
  const uint32_t N = g->N;
  if (N == 0 || nd + ni == 0) return PRH_SUCCESS;
  PRH_HIP(c, hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const auto t0 = std::chrono::steady_clock::now();
  const uint64_t M = g->M;
  const size_t N1 = size_t(N) + 1;
  // The blocked plan (blocks of at most 8192 vertices) is patched. Its iterations read no in-CSR, so the update then
  // leaves the in-CSR unbuilt: it stays the plan's base with the net change since, built when something reads it
  // (ensure_in_csr). A plan with wider blocks has no room for the dead flag: it is rebuilt on its next use.
  BinnedCache &bc = g->bin;
  const bool patch = bc.valid && bc.B <= 8192 && bc.plan.slot;
  if (!patch) PRH_CHECK(ensure_in_csr(c, g));
  Temps tmp(c);
  // The layout's out-rows ascending (the tidy searches them), once per graph.
  if (!g->out_sorted) {
    uint32_t *sorted = nullptr;
    PRH_CHECK(dev_alloc(c, reinterpret_cast<void **>(&sorted), std::max<size_t>(g->l_cap, 4) * 4));
    size_t sb = 0;
    hipError_t e = prh::sort_rows(nullptr, sb, g->l_out_dst, sorted, M, N, g->l_out_off, s);
    uint8_t *work = nullptr;
    int st = e == hipSuccess ? tmp.get(&work, sb) : PRH_ERROR_DEVICE;
    if (st == PRH_SUCCESS) e = prh::sort_rows(work, sb, g->l_out_dst, sorted, M, N, g->l_out_off, s);
    if (st == PRH_SUCCESS && e == hipSuccess) e = hipStreamSynchronize(s);
    if (st != PRH_SUCCESS || e != hipSuccess) {
      dev_free(c, sorted, std::max<size_t>(g->l_cap, 4) * 4);
      return st != PRH_SUCCESS ? st : fail(c, PRH_ERROR_DEVICE, "out-row sort: %s", hipGetErrorString(e));
    }
    dev_free(c, g->l_out_dst, std::max<size_t>(g->l_cap, 4) * 4);
    g->l_out_dst = sorted;
    g->out_sorted = true;
  }
  // Batch upload: deletions (u, v), insertions (u, v), then the two sizes.
  const uint32_t cd = uint32_t(nd), ci = uint32_t(ni);
  uint32_t *d_batch = nullptr;
  PRH_CHECK(tmp.get(&d_batch, 2 * nd + 2 * ni + 2));
  if (nd) {
    PRH_HIP(c, hipMemcpyAsync(d_batch, del_u, nd * 4, hipMemcpyHostToDevice, s));
    PRH_HIP(c, hipMemcpyAsync(d_batch + nd, del_v, nd * 4, hipMemcpyHostToDevice, s));
  }
  if (ni) {
    PRH_HIP(c, hipMemcpyAsync(d_batch + 2 * nd, ins_u, ni * 4, hipMemcpyHostToDevice, s));
    PRH_HIP(c, hipMemcpyAsync(d_batch + 2 * nd + ni, ins_v, ni * 4, hipMemcpyHostToDevice, s));
  }
  uint32_t *sizes = d_batch + 2 * nd + 2 * ni;
  reinterpret_cast<uint32_t *>(c->pinned)[0] = cd;
  reinterpret_cast<uint32_t *>(c->pinned)[1] = ci;
  PRH_HIP(c, hipMemcpyAsync(sizes, c->pinned, 8, hipMemcpyHostToDevice, s));
  // Keys per side and CSR (0: in-CSR deletions, 1: in-CSR insertions, 2: out-CSR deletions, 3: out-CSR insertions),
  // tidied into compact[k] with their counts in counts[k].
  const uint32_t cap = std::max(cd, ci);
  unsigned long long *keys[4], *compact[4], *sorted = nullptr;
  for (int k = 0; k < 4; ++k) {
    PRH_CHECK(tmp.get(&keys[k], std::max<uint32_t>(k & 1 ? ci : cd, 1)));
    PRH_CHECK(tmp.get(&compact[k], std::max<uint32_t>(k & 1 ? ci : cd, 1)));
  }
  uint32_t *valid = nullptr, *pos = nullptr, *counts = nullptr;
  int *flags = nullptr;
  PRH_CHECK(tmp.get(&sorted, cap));
  PRH_CHECK(tmp.get(&valid, cap));
  PRH_CHECK(tmp.get(&pos, cap));
  PRH_CHECK(tmp.get(&counts, 4));
  PRH_CHECK(tmp.get(&flags, 1));
  PRH_HIP(c, prh::launch_fill_u32(counts, 0, 4, s));
  PRH_HIP(c, prh::launch_fill_i32(flags, 0, 1, s));
  size_t sort_bytes = 0, scan_bytes = 0;
  PRH_HIP(c, prh::sort_keys_u64(nullptr, sort_bytes, keys[0], sorted, cap, s));
  PRH_HIP(c, prh::scan_u32(nullptr, scan_bytes, valid, pos, cap, s));
  uint8_t *sort_tmp = nullptr, *scan_tmp = nullptr;
  PRH_CHECK(tmp.get(&sort_tmp, sort_bytes));
  PRH_CHECK(tmp.get(&scan_tmp, scan_bytes));
  if (cd) PRH_HIP(c, prh::launch_batch_keys(d_batch, d_batch + nd, sizes, cd, g->inv, N, keys[0], keys[2], flags, 1, s));
  if (ci)
    PRH_HIP(c, prh::launch_batch_keys(d_batch + 2 * nd, d_batch + 2 * nd + ni, sizes + 1, ci, g->inv, N, keys[1],
                                      keys[3], flags, 1, s));
  PRH_CHECK(tidy(c, keys[2], cd, g->l_out_off, g->l_out_dst, true, sorted, valid, pos, sort_tmp, sort_bytes,
                 scan_tmp, scan_bytes, compact[2], counts + 2));
  PRH_CHECK(tidy(c, keys[3], ci, g->l_out_off, g->l_out_dst, false, sorted, valid, pos, sort_tmp, sort_bytes,
                 scan_tmp, scan_bytes, compact[3], counts + 3));
  if (patch) {
    // The in-CSR keys of the tidied batch: its out-CSR keys transposed and sorted.
    for (int k = 0; k < 2; ++k) {
      const uint32_t n = k ? ci : cd;
      if (!n) continue;
      PRH_HIP(c, prh::launch_transpose_keys(compact[2 + k], counts + 2 + k, n, keys[k], s));
      PRH_HIP(c, prh::sort_keys_u64(sort_tmp, sort_bytes, keys[k], compact[k], n, s));
      PRH_HIP(c, hipMemcpyAsync(counts + k, counts + 2 + k, 4, hipMemcpyDeviceToDevice, s));
    }
  } else {
    PRH_CHECK(tidy(c, keys[0], cd, g->l_in_off, g->l_in_src, true, sorted, valid, pos, sort_tmp, sort_bytes,
                   scan_tmp, scan_bytes, compact[0], counts + 0));
    PRH_CHECK(tidy(c, keys[1], ci, g->l_in_off, g->l_in_src, false, sorted, valid, pos, sort_tmp, sort_bytes,
                   scan_tmp, scan_bytes, compact[1], counts + 1));
  }
  PRH_HIP(c, hipMemcpyAsync(c->pinned, counts, 16, hipMemcpyDeviceToHost, s));
  PRH_HIP(c, hipMemcpyAsync(c->pinned + 2, flags, 4, hipMemcpyDeviceToHost, s));
  PRH_HIP(c, hipStreamSynchronize(s));
  uint32_t cnt[4];
  std::memcpy(cnt, c->pinned, 16);
  if (*reinterpret_cast<const int *>(c->pinned + 2))
    return fail(c, PRH_ERROR_INVALID_ARGUMENT, "batch vertex id >= N (%u); the graph is unchanged", N);
  if (cnt[0] != cnt[2] || cnt[1] != cnt[3])
    return fail(c, PRH_ERROR_INVALID_GRAPH, "in-CSR and out-CSR disagree on the batch (%u/%u deletions, %u/%u "
                                            "insertions); the graph is unchanged", cnt[0], cnt[2], cnt[1], cnt[3]);
  const uint64_t M2 = M - cnt[0] + cnt[1];
  if (M2 >= (uint64_t(1) << 32)) return fail(c, PRH_ERROR_UNSUPPORTED, "the update would exceed 2^32 - 1 edges");
  // New CSRs, in the spare arrays an earlier update left or in new ones. Nothing of the graph changes before this
  // point, so an error above leaves it as it was.
  uint32_t *n_in_off = nullptr, *n_in_src = nullptr, *n_out_off = nullptr, *n_out_dst = nullptr;
  uint64_t n_in_cap = 0, n_out_cap = 0;
  int st = side_arrays(c, N, M2, g->spare_out_off, g->spare_out_dst, g->spare_out_cap, &n_out_off, &n_out_dst,
                       &n_out_cap);
  if (st == PRH_SUCCESS && !patch)
    st = side_arrays(c, N, M2, g->spare_in_off, g->spare_in_src, g->spare_in_cap, &n_in_off, &n_in_src, &n_in_cap);
  if (st == PRH_SUCCESS && !patch)
    st = apply_side(c, tmp, N, M, g->l_in_off, g->l_in_src, compact[0], counts + 0, cd, compact[1], counts + 1, ci,
                    n_in_off, n_in_src, n_in_cap);
  if (st == PRH_SUCCESS)
    st = apply_side(c, tmp, N, M, g->l_out_off, g->l_out_dst, compact[2], counts + 2, cd, compact[3], counts + 3, ci,
                    n_out_off, n_out_dst, n_out_cap);
  // Host copies of the tidied in-CSR keys: the layout's in-degrees kept on the host and the in-CSR's net change
  // follow them.
  std::vector<unsigned long long> hd(cnt[0]), hi(cnt[1]);
  if (st == PRH_SUCCESS && cnt[0] &&
      hipMemcpyAsync(hd.data(), compact[0], size_t(cnt[0]) * 8, hipMemcpyDeviceToHost, s) != hipSuccess)
    st = fail(c, PRH_ERROR_DEVICE, "batch readback");
  if (st == PRH_SUCCESS && cnt[1] &&
      hipMemcpyAsync(hi.data(), compact[1], size_t(cnt[1]) * 8, hipMemcpyDeviceToHost, s) != hipSuccess)
    st = fail(c, PRH_ERROR_DEVICE, "batch readback");
  if (st == PRH_SUCCESS && hipStreamSynchronize(s) != hipSuccess) st = fail(c, PRH_ERROR_DEVICE, "update: sync");
  if (st != PRH_SUCCESS) {
    (void)hipStreamSynchronize(s);
    free_side(c, N, n_in_off, n_in_src, n_in_cap);
    free_side(c, N, n_out_off, n_out_dst, n_out_cap);
    return st;
  }
  // Swap in the new CSRs; the old ones become the spare arrays of the next update. With the patched plan, the
  // in-CSR the plan was built from becomes its base on the first update (ownership moves from the graph), and an
  // in-CSR built since the last update becomes the spare of the next build.
  if (patch) {
    if (!bc.base_off) {
      bc.base_off = g->l_in_off;
      bc.base_src = g->l_in_src;
      bc.base_M = M;
      bc.base_cap = g->l_in_cap;
      g->net_del_n = g->net_ins_n = 0;
    } else if (g->l_in_off) {
      free_side(c, N, g->spare_in_off, g->spare_in_src, g->spare_in_cap);
      g->spare_in_off = g->l_in_off;
      g->spare_in_src = g->l_in_src;
      g->spare_in_cap = g->l_in_cap;
    }
    g->l_in_off = g->l_in_src = nullptr;
    g->l_in_cap = 0;
    g->in_stale = true;
  } else {
    g->spare_in_off = g->l_in_off;
    g->spare_in_src = g->l_in_src;
    g->spare_in_cap = g->l_in_cap;
    g->l_in_off = n_in_off;
    g->l_in_src = n_in_src;
    g->l_in_cap = n_in_cap;
    if (bc.valid) free_binned(c, bc);
  }
  g->spare_out_off = g->l_out_off;
  g->spare_out_dst = g->l_out_dst;
  g->spare_out_cap = g->l_cap;
  g->l_out_off = n_out_off;
  g->l_out_dst = n_out_dst;
  g->l_cap = n_out_cap;
  g->M = M2;
  PRH_HIP(c, prh::launch_degrees(g->l_out_off, N, g->l_outdeg, s));
  uint32_t *zeros = nullptr;
  PRH_CHECK(tmp.get(&zeros, 1));
  PRH_HIP(c, prh::launch_fill_u32(zeros, 0, 1, s));
  count_zero_kernel<<<std::min<uint32_t>(8192, (N + 255) / 256), 256, 0, s>>>(g->l_outdeg, N, zeros);
  PRH_HIP(c, hipGetLastError());
  // patch_plan reads its own words back through the pinned area: read the dead-end count after it
  if (patch) {
    PRH_CHECK(patch_plan(c, g, compact[0], counts + 0, cnt[0], compact[1], counts + 1, cnt[1]));
    PRH_CHECK(compose_net(c, g, tmp, compact[0], cnt[0], compact[1], cnt[1]));
    PRH_HIP(c, hipMemcpyAsync(c->pinned + 4, g->net_counts, 8, hipMemcpyDeviceToHost, s));
  }
  PRH_HIP(c, hipMemcpyAsync(c->pinned, zeros, 4, hipMemcpyDeviceToHost, s));
  PRH_HIP(c, hipStreamSynchronize(s));
  g->dead_ends = *reinterpret_cast<const uint32_t *>(c->pinned);
  if (patch) {
    uint32_t net[2];
    std::memcpy(net, c->pinned + 4, 8);
    g->net_del_n = net[0];
    g->net_ins_n = net[1];
  }
  for (unsigned long long k : hd) --g->sorted_deg[uint32_t(k >> 32)];
  for (unsigned long long k : hi) ++g->sorted_deg[uint32_t(k >> 32)];
  // What depends on the edges is rebuilt on its next use; the caller-order arrays are dropped.
  g->plan_valid = false;
  free_sliced(c, g);
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
  if (!g->updated) {
    dev_free(c, g->in_off, N1 * 4);
    dev_free(c, g->in_src, (M ? M : 4) * 4);
    dev_free(c, g->outdeg, size_t(N) * 4);
    dev_free(c, g->out_off, N1 * 4);
    dev_free(c, g->out_dst, (M ? M : 4) * 4);
    g->in_off = g->in_src = g->outdeg = g->out_off = g->out_dst = nullptr;
    g->updated = true;
  }
  const float ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
  g->update_prep.ms = (g->update_prep.reported ? 0.0f : g->update_prep.ms) + ms;
  g->update_prep.reported = false;
  if (update_ms) *update_ms = ms;
  return PRH_SUCCESS;
}

}  // extern "C"
