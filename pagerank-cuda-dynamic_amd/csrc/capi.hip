// capi.hip -- the C ABI (include/pagerank_hip.h): options and tuning defaults, contexts, graph upload, the solve
// entry points, and the partition, marking and kernel-timing exports. The solver driver is in solve.hip, the layout
// and launch plans in plans.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "capi_internal.hpp"

namespace prhc {

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

Temps::Temps(prh_context *cc) : c(cc), mark(cc->arena_top) {}

Temps::~Temps() {
  for (auto &b : own) dev_free(c, b.first, b.second);
  c->arena_top = mark;
  if (mark == 0 && c->arena_want > c->arena.cap) {
    // the work that used the arena has to end before its memory is released
    if (c->arena.p && hipStreamSynchronize(c->stream) != hipSuccess) (void)hipGetLastError();
    release(c, c->arena);
    const std::string keep = c->error;  // the arena is an optimisation: a failed regrow reports nothing
    if (ensure(c, c->arena, std::min(c->arena_want, kArenaMax)) != PRH_SUCCESS) c->error = keep;
  }
}

int Temps::get_bytes(void **p, size_t bytes) {
  const size_t size = (bytes + kArenaAlign - 1) / kArenaAlign * kArenaAlign;
  if (bytes <= kArenaItem) {
    const size_t top = c->arena_top + size;
    c->arena_want = std::max(c->arena_want, top);
    if (c->arena.p && top <= c->arena.cap) {
      *p = c->arena.p + c->arena_top;
      c->arena_top = top;
      return PRH_SUCCESS;
    }
  }
  PRH_CHECK(dev_alloc(c, p, bytes));
  own.emplace_back(*p, bytes);
  return PRH_SUCCESS;
}

int copy_to_host(prh_context *c, void *dst, const void *src, size_t n) {
  hipStream_t s = c->stream;
  if (n <= kStagingBytes) {
    if (n) PRH_HIP(c, hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, s));
    PRH_HIP(c, hipStreamSynchronize(s));
    return PRH_SUCCESS;
  }
  if (!c->staging && hipHostMalloc(reinterpret_cast<void **>(&c->staging), kStagingBytes, hipHostMallocDefault) !=
                         hipSuccess) {
    (void)hipGetLastError();
    c->staging = nullptr;
    return fail(c, PRH_ERROR_OUT_OF_MEMORY, "pinned staging area (%zu bytes)", kStagingBytes);
  }
  for (size_t o = 0; o < n; o += kStagingBytes) {
    const size_t m = std::min(kStagingBytes, n - o);
    PRH_HIP(c, hipMemcpyAsync(c->staging, static_cast<const uint8_t *>(src) + o, m, hipMemcpyDeviceToHost, s));
    PRH_HIP(c, hipStreamSynchronize(s));
    std::memcpy(static_cast<uint8_t *>(dst) + o, c->staging, m);
  }
  return PRH_SUCCESS;
}

}  // namespace prhc

using namespace prhc;

extern "C" {

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
  t->layout = 0;
  t->window = 0;
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
  if (const char *e = std::getenv("PRH_GRAPH_CAPTURE")) c->capture = std::strcmp(e, "0") != 0;
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
  if (noalloc_setup(c) != PRH_SUCCESS) {
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
  release(c, c->upd_delta);
  release(c, c->upd_changed);
  release(c, c->upd_deg);
  release(c, c->upd_ranges);
  release(c, c->arena);
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
  if (c->staging) (void)hipHostFree(c->staging);
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
        t->blocking == 8192 || t->blocking == 16384) ||
      !(t->layout == 0 || t->layout == 1) || !(t->window == 0 || t->window == 1))
    return fail(c, PRH_ERROR_INVALID_ARGUMENT, "bad tuning");
  if (t->window && t->layout != 1)
    return fail(c, PRH_ERROR_INVALID_ARGUMENT, "bad tuning: the banded plan (window = 1) runs over rows in the "
                                                "caller's vertex order (layout = 1)");
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
      const int bits = *reinterpret_cast<int *>(c->pinned) & ~2;  // bit 2 (a dead end) is not an error
      if (bits)
        st = fail(c, PRH_ERROR_INVALID_GRAPH, "invalid graph (bits 0x%x:%s%s%s%s%s%s)", bits,
                  bits & 1 ? " in-offsets" : "", bits & 2 ? " dead-end" : "", bits & 4 ? " out-degree-mismatch" : "",
                  bits & 8 ? " in-id-range" : "", bits & 16 ? " out-id-range" : "", bits & 32 ? " out-offsets" : "");
    }
  }
  if (st == PRH_SUCCESS)
    for (uint32_t v = 0; v < N; ++v) g->dead_ends += outdeg[v] == 0;
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
  dev_free(c, g->l_in_src, std::max<size_t>(g->l_in_cap, 4) * 4);
  dev_free(c, g->l_outdeg, N * 4);
  dev_free(c, g->l_out_off, (N + 1) * 4);
  dev_free(c, g->l_out_dst, std::max<size_t>(g->l_cap, 4) * 4);
  dev_free(c, g->spare_in_off, (N + 1) * 4);
  dev_free(c, g->spare_in_src, std::max<size_t>(g->spare_in_cap, 4) * 4);
  dev_free(c, g->spare_out_off, (N + 1) * 4);
  dev_free(c, g->spare_out_dst, std::max<size_t>(g->spare_out_cap, 4) * 4);
  dev_free(c, g->net_del, g->net_del_cap * 8);
  dev_free(c, g->net_ins, g->net_ins_cap * 8);
  dev_free(c, g->net_counts, 8);
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
  return solve_impl(c, g, method, o, du, dv, nd, iu, ni, q, ranks, st, true, true);
}

int prh_solve_device(prh_context *c, const prh_graph *g, prh_method method, int partition, const prh_options *o,
                     const uint32_t *du, const uint32_t *dv, size_t nd, const uint32_t *iu, size_t ni,
                     const double *q, double *ranks, prh_stats *st) {
  (void)partition;
  return solve_impl(c, g, method, o, du, dv, nd, iu, ni, q, ranks, st, false, true);
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
  if (!g->l_out_off) return fail(c, PRH_ERROR_UNSUPPORTED, "marking needs the graph's out-CSR");
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
