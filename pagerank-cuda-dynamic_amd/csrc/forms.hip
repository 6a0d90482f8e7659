// forms.hip -- the two forms beside the plain solvers (no reference counterpart): batched personalized PageRank
// (k <= 64 teleport columns in one multi-column pull) and the weighted forward / adjoint solves.
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
namespace {

// Edge chunk of a heavy vertex for the batched solver.
constexpr uint32_t kPprChunk = 1024;

int ppr_impl(prh_context *c, const prh_graph *g, int k, const uint32_t *so, const uint32_t *seeds,
             const prh_options *o, double *ranks, int *iterations, prh_stats *st) {
  if (!c) return PRH_ERROR_INVALID_ARGUMENT;
  if (!g || !o || !so || !ranks || !iterations || k < 1 || k > prh::kPprCols)
    return fail(c, PRH_ERROR_INVALID_ARGUMENT, "null argument or k outside 1..64");
  if (g->ctx != c) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "graph belongs to another context");
  if (g->updated)
    return fail(c, PRH_ERROR_UNSUPPORTED, "the batched solver needs a graph as uploaded (this one was updated in place)");
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
  auto iterations_chunk = [&]() -> int {
    for (int i = 0; i < chunk; ++i) PRH_HIP(c, prh::launch_ppr_iteration(A, s));
    return PRH_SUCCESS;
  };
  hipGraphExec_t exec = nullptr;
  int status = L > 0 && c->capture ? capture_graph(c, iterations_chunk, &exec) : PRH_SUCCESS;
  prh::PprState hs{};
  for (int launched = 0; status == PRH_SUCCESS && launched < L; launched += chunk) {
    hipError_t e = hipSuccess;
    if (exec) e = hipGraphLaunch(exec, s);
    else if ((status = iterations_chunk()) != PRH_SUCCESS) break;
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
  PRH_HIP(c, hipMemcpyAsync(c->pinned, c->ppr_state.p, sizeof(prh::PprState), hipMemcpyDeviceToHost, s));
  PRH_CHECK(copy_to_host(c, ranks, c->ppr_out.p, size_t(N) * k * 8));  // synchronises the stream
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
  if (g->updated)
    return fail(c, PRH_ERROR_UNSUPPORTED, "weighted solves need a graph as uploaded (this one was updated in place)");
  if (!g->out_off || !g->l_out_off) return fail(c, PRH_ERROR_UNSUPPORTED, "weighted solves need the graph's out-CSR");
  if (g->dead_ends)
    return fail(c, PRH_ERROR_INVALID_GRAPH, "weighted solves need every out-weight sum positive: %u dead ends",
                g->dead_ends);
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
    const float prep = gm->layout_prep.take() + gm->bin.prep.take();  // one-time preparation, as solve_impl reports it
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
    if (st) *st = prh_stats{n, t01 + t12 + prep, t01 + prep, 0.0f, t12, prep};
    return PRH_SUCCESS;
  }
  const float prep = const_cast<prh_graph *>(g)->layout_prep.take();
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
  if (st) *st = prh_stats{n, t01 + t12 + prep, t01 + prep, 0.0f, t12, prep};
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
  float prep = const_cast<prh_graph *>(g)->layout_prep.take();
  if (blocked) {
    prh_graph *gm = const_cast<prh_graph *>(g);
    PRH_CHECK(build_binned_plan_t(c, gm));
    prep += gm->bin_t.prep.take();
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
  if (st) *st = prh_stats{n, t01 + t12 + prep, t01 + prep, 0.0f, t12, prep};
  return PRH_SUCCESS;
}

}  // namespace
}  // namespace prhc

using namespace prhc;

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

}  // extern "C"
