// session.hip -- the streaming dynamic session (prh_stream_*).
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

using namespace prhc;


// ---- streaming dynamic session -------------------------------------------------------------------------------
//
// A device-resident copy of a graph's degree-ordered layout that batches update in place (stream.hip), with
// the ranks kept between batches and each batch's work replayed from captured HIP graphs:
//   apply graph (one per buffer parity): batch -> keys -> sort -> tidy -> new in- and out-CSR (into the other
//       buffers) -> out-degrees -> the update plan's patch (blocked: dead slots and the overlay; rows: the chunk
//       tables) -> solver initialisation and, for DT / DF / DF-P, the initial marking;
//   closure graph (DT, one per parity): levels of the traversal closure, each run only while the level before it
//       marked a vertex; replayed until the device reports the closure done;
//   iteration graph (one per parity): kChunk iterations of the rank update (the blocked plan extends the frontier
//       inside it; the row plan adds the direction choice and both extension forms) followed each by the
//       device-side convergence step; replayed until the device reports the loop done.
// The rank update runs through the session's own propagation-blocked plan when the context's tuning asks for
// blocks of at most 8192 vertices (the plan of the base graph, patched per batch as prh_graph_update patches a
// graph's), else through the row plan.
// Only the batch upload, one flag read after the apply and one read per closure or iteration chunk touch the host.

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
  uint32_t *perm = nullptr, *inv = nullptr;  // the vertex order (layout row -> vertex, vertex -> layout row)
  uint32_t sched_low = 0, sched_high = 0;
  int *delta = nullptr;
  uint8_t *changed = nullptr;
  uint4 *ranges = nullptr;  // per changed row: its deletions and insertions in the compacted keys
  uint32_t *units = nullptr;  // the rewrite's runs of rows: 4 (2 cap_batch + 1) + 1 words
  uint32_t *h_batch = nullptr;  // pinned: del_u, del_v, ins_u, ins_v [cap_batch each], nd, ni
  uint32_t *d_batch = nullptr;
  unsigned long long *keys[4] = {};     // raw keys: in-CSR del, in-CSR ins, out-CSR del, out-CSR ins
  unsigned long long *compact[4] = {};  // sorted, tidied
  unsigned long long *sorted = nullptr;
  uint32_t *valid = nullptr, *pos = nullptr, *counts = nullptr;  // counts[4]: sizes of compact[]
  uint32_t *ext_cnt = nullptr, *ext_start = nullptr;  // the initial extension's chunks per batch entry [2 cap + 1]
  void *sort_tmp = nullptr, *scan_tmp = nullptr;
  size_t sort_bytes = 0, scan_bytes = 0;
  double *rank[2] = {}, *contrib[2] = {}, *stats = nullptr;
  uint8_t *vaff = nullptr, *naff = nullptr;
  int *ctl = nullptr;  // [0] done, [1] iteration index (adjacent, as LoopCtl expects), [2] direction, [3] flags
  unsigned long long *err = nullptr;
  prh::RankPlan plan{};
  uint32_t n_items = 0, chunk = 0, stat_slots = 0, pull_wave = 0, pull_group = 0;
  hipGraphExec_t apply[2] = {}, iter[2] = {}, closure[2] = {};
  std::vector<std::pair<void *, size_t>> bufs;
  bool broken = false;
  // Propagation-blocked update (blocked): the session's plan of the base graph, its own copy of the slot words (the
  // dead flags), the base in-CSR with the slot map's inverse, and the overlay at a fixed capacity.
  bool blocked = false;
  BinnedCache bc;
  prh::BinnedPlan bplan{};
  uint32_t *base_off = nullptr, *base_src = nullptr, *slot_of = nullptr;
  unsigned long long *ov_key = nullptr, *ov_cand = nullptr, *ov_sorted = nullptr, *fresh = nullptr;
  uint32_t *ov_off = nullptr, *ov_words = nullptr;  // ov_words[0]: the next overlay's size
  uint8_t *ov_gone = nullptr;
  double *ovX = nullptr;
  void *ov_sort_tmp = nullptr;
  size_t ov_sort_bytes = 0;
  uint32_t ov_cap = 0, nb = 0;
  // Traversal closure (DT): two frontier buffers and the level words: [0] level, [1] stop, [2], [3] changed flags of
  // even and odd levels.
  uint8_t *front[2] = {};
  int *lvl = nullptr;
};

namespace {

constexpr int kStreamChunk = 8;  // iterations per replay of the iteration graph (even: ping-pong buffers realign)
constexpr int kClosureLevels = 8;  // traversal levels per replay of the closure graph (even: its buffers realign)

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
    if (S->closure[k]) (void)hipGraphExecDestroy(S->closure[k]);
  }
  free_binned(S->ctx, S->bc);
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
  PRH_HIP(c, prh::launch_row_ranges(S->compact[kd], S->counts + kd, S->compact[ki], S->counts + ki, 2 * cap,
                                    S->delta, S->changed, S->ranges, s));
  PRH_HIP(c, prh::launch_new_degrees(off, S->delta, N, S->new_deg, S->ctl + 3, empty_bit, s));
  PRH_HIP(c, prh::scan_u32(S->scan_tmp, S->scan_bytes, S->new_deg, noff, size_t(N) + 1, s));
  uint32_t *chpos = reinterpret_cast<uint32_t *>(S->delta);  // the degree changes are spent
  PRH_HIP(c, prh::scan_u8(S->scan_tmp, S->scan_bytes, S->changed, chpos, N, s));
  PRH_HIP(c, prh::launch_rewrite(off, idx, noff, S->changed, chpos, S->units, S->ranges, S->compact[kd],
                                 S->compact[ki], N, 2 * cap + 1, S->cap_edges, S->cap_edges, nidx, s));
  PRH_HIP(c, prh::launch_insert(off, idx, noff, S->ranges, S->compact[kd], S->compact[ki], S->counts + ki, cap, nidx,
                                s));
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
  PRH_HIP(c, prh::launch_batch_keys(b, b + cap, b + 4 * size_t(cap), cap, S->inv, N, S->keys[0], S->keys[2],
                                    S->ctl + 3, 1, s));
  PRH_HIP(c, prh::launch_batch_keys(b + 2 * size_t(cap), b + 3 * size_t(cap), b + 4 * size_t(cap) + 1, cap,
                                    S->inv, N, S->keys[1], S->keys[3], S->ctl + 3, 1, s));
  PRH_CHECK(stream_apply_csr(S, s, 0, 1, S->in_off[p], S->in_src[p], S->in_off[q], S->in_src[q], 0));
  PRH_CHECK(stream_apply_csr(S, s, 2, 3, S->out_off[p], S->out_dst[p], S->out_off[q], S->out_dst[q], 2));
  PRH_HIP(c, prh::launch_degrees(S->out_off[q], N, S->outdeg, s));
  if (S->blocked) {
    // The plan's patch (see prh_graph_update): dead flags, then the next overlay from the live entries and the
    // fresh insertions, sorted at the fixed capacity; an overlay past it sets flag bit 4.
    const uint32_t nbin = S->nb;
    uint16_t *dst16 = const_cast<uint16_t *>(S->bplan.dst16);
    PRH_HIP(c, prh::launch_fill_u8(S->ov_gone, 0, S->ov_cap, s));
    PRH_HIP(c, prh::launch_fill_u32(S->ov_words, 0, 1, s));
    PRH_HIP(c, prh::launch_patch_slots(S->compact[0], S->counts + 0, cap, true, S->base_off, S->base_src, S->slot_of,
                                       dst16, S->ov_key, S->ov_off + nbin, S->ov_gone, nullptr, s));
    PRH_HIP(c, prh::launch_patch_slots(S->compact[1], S->counts + 1, cap, false, S->base_off, S->base_src,
                                       S->slot_of, dst16, S->ov_key, S->ov_off + nbin, S->ov_gone, S->fresh, s));
    PRH_HIP(c, prh::launch_overlay_candidates(S->ov_key, S->ov_off + nbin, S->ov_cap, S->ov_gone, S->fresh, cap,
                                              S->ov_cand, S->ov_words, s));
    PRH_HIP(c, prh::sort_keys_u64(S->ov_sort_tmp, S->ov_sort_bytes, S->ov_cand, S->ov_sorted, S->ov_cap + cap, s));
    PRH_HIP(c, prh::launch_copy_u64(S->ov_key, S->ov_sorted, S->ov_cap, s));
    PRH_HIP(c, prh::launch_overlay_bounds(S->ov_key, S->ov_words, S->ov_cap, uint32_t(S->bc.B), nbin, S->ov_off,
                                          S->ctl + 3, 4, s));
  } else {
    PRH_HIP(c, prh::launch_refresh_chunks(S->plan.item_row, S->plan.row_first, S->n_items, S->in_off[q], S->chunk,
                                          const_cast<uint32_t *>(S->plan.item_beg),
                                          const_cast<uint32_t *>(S->plan.item_end), s));
  }
  // Solver initialisation (pagerankCuda.hxx:811-821): ranks from the previous batch's result (or 1/N for Static),
  // contributions with the new out-degrees, loop state cleared, and the DF initial marking on the new graph.
  PRH_HIP(c, prh::launch_fill_u64(S->err, 0ull, size_t(S->opt.max_iterations) + 1, s));
  PRH_HIP(c, prh::launch_init_ranks(S->rank[0], S->rank[1], S->contrib[0], S->contrib[1], S->outdeg, N,
                                    S->method != PRH_STATIC, s));
  if (S->variant >= prh::kTraversal) {
    PRH_HIP(c, prh::launch_fill_u8(S->vaff, 0, N, s));
    PRH_HIP(c, prh::launch_fill_u8(S->naff, 0, N, s));
    PRH_HIP(c, prh::launch_mark_batch(S->naff, S->vaff, S->compact[2], S->counts + 2, S->compact[3], S->counts + 3,
                                      2 * cap, s));
    // The initial extension from the batch's sources (each once, its out-edges in chunks over all waves), not a scan
    // of every vertex's flag: on RMAT-22 with batches of 1e-4 |E| that scan took 0.6 ms of a 5.5 ms batch.
    PRH_HIP(c, prh::launch_extend_batch_keys(S->vaff, S->naff, S->compact[2], S->counts + 2, S->compact[3],
                                             S->counts + 3, cap, S->out_off[q], S->out_dst[q], S->ext_cnt,
                                             S->ext_start, S->scan_tmp, S->scan_bytes, s));
    if (S->variant == prh::kTraversal) {  // the closure starts from every vertex marked so far
      PRH_HIP(c, prh::launch_copy_u8(S->front[0], S->vaff, N, s));
      PRH_HIP(c, prh::launch_fill_i32(S->lvl, 0, 4, s));
    } else if (!S->blocked) {
      const int mode = c->tuning.extension_direction;
      if (mode != 1) PRH_HIP(c, prh::launch_fill_i32(S->ctl + 2, mode == 2 ? 1 : 0, 1, s));
    }
  }
  return PRH_SUCCESS;
}

// kClosureLevels levels of the traversal closure on buffers q (captured once per parity), after
// pagerankAffectedTraversalCuW (pagerankCuda.hxx:438-456): level l extends from the vertices level l - 1 marked
// (front[l & 1]; level 0 from every vertex the initial marking set), marks their out-neighbours affected, and
// records the newly marked ones in front[(l + 1) & 1]. A level runs only while the one before it marked a vertex;
// the first level that finds nothing new sets lvl[1] (stop). Levels replay from an even index, so level l's
// buffers are those of capture slot l mod 2.
int stream_closure_body(prh_stream *S, int q) {
  prh_context *c = S->ctx;
  hipStream_t s = c->stream;
  const prh::Schedule os{S->sched_order, S->sched_low, S->sched_high, 0};
  for (int k = 0; k < kClosureLevels; ++k) {
    PRH_HIP(c, prh::launch_closure_step(S->lvl, k & 1, s));
    PRH_HIP(c, prh::launch_fill_u8(S->front[(k + 1) & 1], 0, S->N, s));
    prh::ExtendArgs E{S->vaff, S->front[k & 1], S->out_off[q], S->out_dst[q], os, S->front[(k + 1) & 1],
                      S->lvl + 2 + (k & 1), S->lvl + 1};
    PRH_HIP(c, prh::launch_extend(E, nullptr, nullptr, 0, 0.0, s));
    PRH_HIP(c, prh::launch_closure_advance(S->lvl, s));
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
                    frontier && mode == 1 && !S->blocked ? S->stats : nullptr};
    if (S->blocked) {  // the frontier's extension happens inside the update (from the second iteration on)
      PRH_HIP(c, prh::launch_binned_update(S->variant, A, S->bplan, err, done, -1, o.tolerance, frontier, s));
      PRH_HIP(c, prh::launch_advance(S->ctl + 1, done, err, o.tolerance, o.max_iterations, s));
      continue;
    }
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
  if (!valid_method(method)) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "bad method");
  if (!g->l_out_off) return fail(c, PRH_ERROR_UNSUPPORTED, "the graph was uploaded without its out-CSR");
  if (g->dead_ends) return fail(c, PRH_ERROR_INVALID_GRAPH, "the streamed graph has %u dead ends", g->dead_ends);
  if (g->N == 0 || max_batch == 0) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "empty graph or zero batch capacity");
  const prh_options &o = *options;
  if (o.max_iterations < 1 || !(o.tolerance >= 0) || !(o.damping > 0 && o.damping < 1))
    return fail(c, PRH_ERROR_INVALID_ARGUMENT, "bad options");
  const uint64_t cap_edges = max_edges ? max_edges : g->M + 16ull * max_batch;
  if (cap_edges < g->M || cap_edges >= (uint64_t(1) << 32))
    return fail(c, PRH_ERROR_INVALID_ARGUMENT, "edge capacity must hold the graph and stay below 2^32");
  PRH_HIP(c, hipSetDevice(c->device));
  if (c->tuning.layout != g->layout && !g->updated) PRH_CHECK(relayout(c, const_cast<prh_graph *>(g), c->tuning.layout));
  PRH_CHECK(build_rank_plan(c, const_cast<prh_graph *>(g)));  // builds the in-CSR an update may have left unbuilt
  auto *S = new prh_stream;
  S->ctx = c;
  S->g = g;
  S->method = method;
  S->variant = method == PRH_PRUNE_DYNAMIC_FRONTIER ? prh::kPrune
               : method == PRH_DYNAMIC_FRONTIER      ? prh::kFrontier
               : method == PRH_DYNAMIC_TRAVERSAL     ? prh::kTraversal
                                                     : prh::kPlain;
  S->blocked = c->tuning.blocking > 0 && c->tuning.blocking <= 8192;
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
  step(stream_alloc(S, &S->perm, N));
  step(stream_alloc(S, &S->inv, N));
  if (st == PRH_SUCCESS) {  // the session's own vertex order: the graph may change layout later
    PRH_HIP(c, hipMemcpyAsync(S->perm, g->perm, size_t(N) * 4, hipMemcpyDeviceToDevice, s));
    PRH_HIP(c, hipMemcpyAsync(S->inv, g->inv, size_t(N) * 4, hipMemcpyDeviceToDevice, s));
  }
  step(stream_alloc(S, &S->new_deg, N1));
  step(stream_alloc(S, &S->sched_order, N));
  step(stream_alloc(S, &S->delta, N));
  step(stream_alloc(S, &S->changed, N));
  step(stream_alloc(S, &S->ranges, N));
  step(stream_alloc(S, &S->units, 4 * (2 * size_t(max_batch) + 1) + 1));
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
  step(stream_alloc(S, &S->naff, (size_t(N) + 3) / 4 * 4));  // whole 32-bit words (the batch extension's claims)
  step(stream_alloc(S, &S->ext_cnt, 2 * size_t(max_batch) + 1));
  step(stream_alloc(S, &S->ext_start, 2 * size_t(max_batch) + 1));
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
    if (e == hipSuccess) e = prh::scan_u8(nullptr, b1, S->changed, S->pos, N, s);
    S->scan_bytes = std::max(S->scan_bytes, b1);
    if (e == hipSuccess)
      e = prh::launch_extend_batch_keys(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, max_batch, nullptr,
                                        nullptr, S->ext_cnt, S->ext_start, nullptr, b1, s);
    S->scan_bytes = std::max(S->scan_bytes, b1);
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
    S->pull_wave = layout_rows_above(g, 64);
    S->pull_group = layout_rows_above(g, 4);
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
  // The session's propagation-blocked plan: built from the base graph's layout in-CSR, with its own slot words (the
  // dead flags), the base in-CSR and the slot map's inverse, and an overlay of fixed capacity: twice the room the
  // edge capacity leaves above the base graph, plus two batches (every edge beyond the base that the graph can hold,
  // with room for base edges deleted and inserted edges kept).
  if (st == PRH_SUCCESS && S->blocked) {
    const uint64_t M = g->M;
    st = build_binned(c, N, M, g->l_in_off, g->l_in_src, g->sorted_deg, true, S->bc);
    if (st == PRH_SUCCESS) {
      const uint32_t B = uint32_t(S->bc.B);
      S->nb = (N + B - 1) / B;
      const uint64_t room = 2 * (cap_edges - M) + 2ull * max_batch;
      if (room >= (uint64_t(1) << 31)) st = fail(c, PRH_ERROR_INVALID_ARGUMENT, "edge capacity too far above the graph");
      S->ov_cap = uint32_t(room);
    }
    const uint64_t Mp = (M + 127) / 128 * 128;
    uint16_t *dst16 = nullptr;
    step(stream_alloc(S, &dst16, Mp));
    step(stream_alloc(S, &S->base_off, N1));
    step(stream_alloc(S, &S->base_src, std::max<uint64_t>(M, 1)));
    step(stream_alloc(S, &S->slot_of, std::max<uint64_t>(M, 1)));
    step(stream_alloc(S, &S->ov_key, S->ov_cap));
    step(stream_alloc(S, &S->ov_cand, size_t(S->ov_cap) + max_batch));
    step(stream_alloc(S, &S->ov_sorted, size_t(S->ov_cap) + max_batch));
    step(stream_alloc(S, &S->fresh, max_batch));
    step(stream_alloc(S, &S->ov_off, size_t(S->nb) + 1));
    step(stream_alloc(S, &S->ov_words, 1));
    step(stream_alloc(S, &S->ov_gone, S->ov_cap));
    step(stream_alloc(S, &S->ovX, S->ov_cap));
    if (st == PRH_SUCCESS) {
      hipError_t e = prh::sort_keys_u64(nullptr, S->ov_sort_bytes, S->ov_cand, S->ov_sorted, S->ov_cap + max_batch, s);
      if (e != hipSuccess) st = fail(c, PRH_ERROR_DEVICE, "overlay sort: %s", hipGetErrorString(e));
    }
    step(stream_alloc(S, &S->ov_sort_tmp, S->ov_sort_bytes));
    if (st == PRH_SUCCESS) {
      PRH_HIP(c, hipMemcpyAsync(dst16, S->bc.plan.dst16, Mp * 2, hipMemcpyDeviceToDevice, s));
      PRH_HIP(c, hipMemcpyAsync(S->base_off, g->l_in_off, N1 * 4, hipMemcpyDeviceToDevice, s));
      if (M) PRH_HIP(c, hipMemcpyAsync(S->base_src, g->l_in_src, M * 4, hipMemcpyDeviceToDevice, s));
      if (M) PRH_HIP(c, prh::launch_invert(S->bc.plan.slot, uint32_t(M), S->slot_of, s));
      PRH_HIP(c, prh::launch_fill_u32(S->ov_off, 0, size_t(S->nb) + 1, s));
      S->bplan = S->bc.plan;
      S->bplan.dst16 = dst16;
      S->bplan.ov_n = S->ov_cap;
      S->bplan.ov_key = S->ov_key;
      S->bplan.ov_off = S->ov_off;
      S->bplan.ovX = S->ovX;
    }
  }
  if (S->variant == prh::kTraversal) {
    step(stream_alloc(S, &S->front[0], N));
    step(stream_alloc(S, &S->front[1], N));
    step(stream_alloc(S, &S->lvl, 4));
  }
  // Out-degree schedule of the extension (fixed for the session; rows that outgrow the low class stay correct).
  if (st == PRH_SUCCESS && S->variant >= prh::kTraversal) {
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
        st = solve_impl(c, g, PRH_STATIC, &o, nullptr, nullptr, 0, nullptr, 0, nullptr, tmp, &stt, false, false);
      }
      if (st == PRH_SUCCESS) {
        hipError_t e = prh::launch_gather_f64(tmp, S->perm, N, S->rank[0], s);
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
  if (c->capture) {
    if (!S->apply[p]) PRH_CHECK(capture_graph(c, [&]() { return stream_apply_body(S, p); }, &S->apply[p]));
    PRH_HIP(c, hipGraphLaunch(S->apply[p], s));
  } else {
    PRH_CHECK(stream_apply_body(S, p));
  }
  PRH_HIP(c, hipEventRecord(c->ev[1], s));
  PRH_HIP(c, hipMemcpyAsync(c->pinned, S->ctl + 3, 4, hipMemcpyDeviceToHost, s));
  PRH_HIP(c, hipMemcpyAsync(c->pinned + 1, S->in_off[q] + S->N, 4, hipMemcpyDeviceToHost, s));
  PRH_HIP(c, hipStreamSynchronize(s));
  const int flags = *reinterpret_cast<const int *>(c->pinned);
  if (flags) {
    S->broken = true;
    return fail(c, PRH_ERROR_INVALID_ARGUMENT, flags & 1   ? "batch vertex id out of range"
                                               : flags & 2 ? "the batch leaves a vertex without out-edges (dead end)"
                                                           : "the update plan's overlay outgrew its capacity");
  }
  S->M = *reinterpret_cast<const uint32_t *>(c->pinned + 1);
  S->cur = q;
  if (S->variant == prh::kTraversal) {  // the closure's levels, replayed until one marks nothing new
    if (c->capture && !S->closure[q])
      PRH_CHECK(capture_graph(c, [&]() { return stream_closure_body(S, q); }, &S->closure[q]));
    int stop = 0;
    for (uint64_t lv = 0; !stop && lv <= uint64_t(S->N) + kClosureLevels; lv += kClosureLevels) {
      if (c->capture) PRH_HIP(c, hipGraphLaunch(S->closure[q], s));
      else PRH_CHECK(stream_closure_body(S, q));
      PRH_HIP(c, hipMemcpyAsync(c->pinned, S->lvl + 1, 4, hipMemcpyDeviceToHost, s));
      PRH_HIP(c, hipStreamSynchronize(s));
      stop = *reinterpret_cast<const int *>(c->pinned);
    }
  }
  PRH_HIP(c, hipEventRecord(c->ev[3], s));
  if (c->capture && !S->iter[q])
    PRH_CHECK(capture_graph(c, [&]() { return stream_iter_body(S, q); }, &S->iter[q]));
  int n = 0;
  for (int launched = 0; launched < S->opt.max_iterations; launched += kStreamChunk) {
    if (c->capture) PRH_HIP(c, hipGraphLaunch(S->iter[q], s));
    else PRH_CHECK(stream_iter_body(S, q));
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
  float t01 = 0, t13 = 0, t32 = 0;
  PRH_HIP(c, hipEventElapsedTime(&t01, c->ev[0], c->ev[1]));
  PRH_HIP(c, hipEventElapsedTime(&t13, c->ev[1], c->ev[3]));
  PRH_HIP(c, hipEventElapsedTime(&t32, c->ev[3], c->ev[2]));
  if (stats) {
    stats->iterations = n;
    stats->initialization_time = t01;  // batch upload, application, plan patch and the initial marking
    stats->marking_time = t13;         // the flag read and, for DT, the traversal closure
    stats->computation_time = t32;
    stats->time = t01 + t13 + t32;
    stats->preparation_time = 0;
  }
  if (ranks) PRH_CHECK(prh_stream_ranks(S, ranks));
  return PRH_SUCCESS;
}

int prh_stream_ranks(prh_stream *S, double *ranks) {
  if (!S || !ranks) return S ? fail(S->ctx, PRH_ERROR_INVALID_ARGUMENT, "null ranks") : PRH_ERROR_INVALID_ARGUMENT;
  prh_context *c = S->ctx;
  PRH_CHECK(ensure(c, c->contrib[0], S->N));
  PRH_HIP(c, prh::launch_scatter_f64(S->rank[0], S->perm, S->N, c->contrib[0].p, c->stream));
  PRH_HIP(c, hipMemcpyAsync(ranks, c->contrib[0].p, size_t(S->N) * 8, hipMemcpyDeviceToHost, c->stream));
  PRH_HIP(c, hipStreamSynchronize(c->stream));
  return PRH_SUCCESS;
}

uint64_t prh_stream_size(const prh_stream *S) { return S ? S->M : 0; }

int prh_stream_csr(prh_stream *S, int which, uint32_t *offsets, uint32_t *ids) {
  if (!S) return PRH_ERROR_INVALID_ARGUMENT;
  prh_context *c = S->ctx;
  if (!offsets || (S->M && !ids) || (which != 0 && which != 1))
    return fail(c, PRH_ERROR_INVALID_ARGUMENT, "null array or bad CSR selector");
  PRH_HIP(c, hipSetDevice(c->device));
  PRH_HIP(c, hipStreamSynchronize(c->stream));
  const int p = S->cur;
  return csr_to_caller(c, S->perm, S->inv, S->N, which ? S->out_off[p] : S->in_off[p],
                       which ? S->out_dst[p] : S->in_src[p], S->M, offsets, ids);
}

void prh_stream_destroy(prh_stream *S) {
  if (!S) return;
  (void)hipSetDevice(S->ctx->device);
  (void)hipStreamSynchronize(S->ctx->stream);
  stream_free(S);
  delete S;
}

}  // extern "C"
