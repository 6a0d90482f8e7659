// solve.hip -- the solver driver: restates pagerankInvokeCuda + pagerankLoopCuU (pagerankCuda.hxx:675-692,
// :713-849) for one HIP stream. Per repeat, partition (and initial ranks) = initialization time, affected marking =
// marking time, iterations = computation time, each measured with HIP events. Iterations run without a host
// round trip per iteration: convergence is detected on the device (LoopCtl), and the host polls the device flag
// only every check_interval iterations, so the iteration count and final ranks equal those of the per-iteration
// check.
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
static int load_initial_ranks(prh_context *c, const prh_graph *g, const double *q, bool host) {
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
               prh_stats *st, bool host, bool account_prep) {
  if (!c) return PRH_ERROR_INVALID_ARGUMENT;
  if (!g || !o || !valid_method(method)) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "null graph/options or bad method");
  if (g->ctx != c) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "graph belongs to another context");
  if (o->repeat < 1 || o->max_iterations < 0 || !(o->tolerance >= 0) || !(o->damping > 0 && o->damping < 1))
    return fail(c, PRH_ERROR_INVALID_ARGUMENT, "bad options (repeat >= 1, max_iterations >= 0, tolerance >= 0, 0 < damping < 1)");
  if (method != PRH_STATIC && !q) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "initial ranks required for this method");
  if (method == PRH_PRUNE_DYNAMIC_FRONTIER && g->dead_ends)
    return fail(c, PRH_ERROR_INVALID_GRAPH,
                "DF-P's closed-form update divides by the out-degree: %u vertices have none (dead ends)", g->dead_ends);
  const uint32_t N = g->N;
  prh_stats stats{};
  if (N == 0) {  // the reference returns an empty result for an empty graph (pagerankCuda.hxx:869)
    if (st) *st = stats;
    return PRH_SUCCESS;
  }
  if (!ranks) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "null ranks");
  const bool dynamic = method >= PRH_DYNAMIC_TRAVERSAL;
  const bool frontier = method >= PRH_DYNAMIC_FRONTIER;
  if (dynamic && !g->l_out_off) return fail(c, PRH_ERROR_UNSUPPORTED, "dynamic method needs the graph's out-CSR");
  const int L = o->max_iterations;
  const int variant = method == PRH_PRUNE_DYNAMIC_FRONTIER ? prh::kPrune
                      : method == PRH_DYNAMIC_FRONTIER    ? prh::kFrontier
                      : method == PRH_DYNAMIC_TRAVERSAL   ? prh::kTraversal
                                                          : prh::kPlain;
  hipStream_t s = c->stream;
  PRH_HIP(c, hipSetDevice(c->device));
  PRH_CHECK(ensure_workspace(c, N, dynamic, L));
  prh_graph *gm = const_cast<prh_graph *>(g);
  if (c->tuning.layout != g->layout && !g->updated) PRH_CHECK(relayout(c, gm, c->tuning.layout));
  const bool binned = c->tuning.blocking > 0;
  const bool window = !binned && c->tuning.window != 0 && g->layout == 1;
  const bool sliced = !binned && !window && variant == prh::kPlain && c->tuning.source_slices > 0 && g->layout == 0;
  if (!binned) PRH_CHECK(ensure_in_csr(c, gm));  // the other plans read the in-CSR in every iteration
  if (binned) PRH_CHECK(build_binned_plan(c, gm));
  else if (window) PRH_CHECK(build_window_plan(c, gm));
  else if (sliced) PRH_CHECK(build_sliced_plan(c, gm));
  else PRH_CHECK(build_rank_plan(c, gm));
  // One-time preparation of this graph not yet reported by a solve: its layout (built at upload) and the plan this
  // solve runs (built on first use), plus the extension schedule below.
  float prep = 0.0f;
  if (account_prep)
    prep = gm->layout_prep.take() + gm->update_prep.take() +
           (binned   ? gm->bin.prep.take()
            : window ? gm->win_prep.take()
            : sliced ? gm->sliced_prep.take()
                     : gm->plan_prep.take());
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
  const uint32_t stat_slots = binned   ? g->bin.plan.n_ac + g->bin.plan.n_multi
                              : window ? g->win.wgs
                                       : plan.blocks + (plan.n_chunked ? prh::finish_chunked_grid(plan.n_chunked) : 0);
  if (dir_auto) PRH_CHECK(ensure(c, c->ext_stats, 2 * size_t(std::max<uint32_t>(stat_slots, 1))));
  int *dir = frontier ? c->flags.p + 2 : nullptr;
  const prh::PullArgs pull{c->vaff.p, c->naff.p, g->l_in_off, g->l_in_src, layout_rows_above(g, 64),
                           layout_rows_above(g, 4), N, dir};
  for (int rep = 0; rep < o->repeat; ++rep) {
    if (q) PRH_CHECK(load_initial_ranks(c, g, q, host));  // untimed, as the reference's q copy (pagerankCuda.hxx:810)
    PRH_HIP(c, hipEventRecord(c->ev[0], s));
    // Initialisation. The degree classes the reference partitions here (pagerankCuda.hxx:811-814) are the row
    // ranges of the layout, fixed when the graph was prepared; the extension kernels' out-degree split is per solve.
    // The out-degree schedule serves the traversal levels and the row plan's frontier extension; the initial
    // extension runs from the batch list and the blocked plan extends inside its update.
    prh::ExtendArgs xo{};
    if (dynamic && (method == PRH_DYNAMIC_TRAVERSAL || (frontier && !binned))) {
      PRH_CHECK(out_schedule(c, gm, xo));
      if (account_prep) prep += gm->out_prep.take();
    }
    // The blocked and banded plans compute contributions from the ranks where they read them and keep none.
    const bool keeps = !binned && !window;
    PRH_HIP(c, prh::launch_init_ranks(c->rank[0].p, c->rank[1].p, keeps ? c->contrib[0].p : nullptr,
                                      keeps ? c->contrib[1].p : nullptr, g->l_outdeg, N,
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
      else if (window) {
        PRH_HIP(c, prh::launch_window_update(variant, A, g->win, err, done, l, o->tolerance, s));
        if (mark) PRH_HIP(c, hipEventRecord(mark, s));
      } else if (sliced) PRH_HIP(c, prh::launch_sliced_update(A, g->sliced, err, done, l, o->tolerance, s, mark));
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
  stats.initialization_time = float(ti / o->repeat) + prep;
  stats.marking_time = float(tm / o->repeat);
  stats.computation_time = float(tc / o->repeat);
  stats.time = float((ti + tm + tc) / o->repeat) + prep;
  stats.preparation_time = prep;
  if (st) *st = stats;
  return PRH_SUCCESS;
}

}  // namespace prhc
