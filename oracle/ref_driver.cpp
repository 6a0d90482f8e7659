// ref_driver.cpp -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
//
// Exposes the reference's own CPU PageRank path (puzzlef/pagerank-cuda-dynamic, inc/pagerank.hxx,
// inc/pagerankPrune.hxx, inc/bfs.hxx, inc/batch.hxx, inc/transpose.hxx) through a C ABI, so tests can pin the
// oracle against it and bench.py can time it as the "reference" CPU baseline. The reference headers are
// compiled where they lie (oracle/Makefile passes -I <reference>/inc); nothing of them is copied here, and the
// build output goes only to oracle/_ref/.
//
// Graphs come in as the out-CSR of x (dense ids 0..N-1, neighbours ascending) and are loaded into the
// reference's DiGraph<uint32_t>; y = x with the batch tidied and applied by the reference's own
// tidyBatchUpdateU / applyBatchUpdateOmpU (batch.hxx:201-247), and yt = transposeWithDegreeOmp(y)
// (transpose.hxx:128-134), exactly as main.cu:80-86 prepares them.
#define BUILD 0
#define OPENMP 1
#include <cstdint>
#include <cstring>
#include <omp.h>
#include "_main.hxx"
#include "Graph.hxx"
#include "update.hxx"
#include "duplicate.hxx"
#include "transpose.hxx"
#include "bfs.hxx"
#include "batch.hxx"
#include "pagerank.hxx"
#include "pagerankPrune.hxx"
#include "mtx.hxx"
#include "snap.hxx"

using K = uint32_t;
using V = double;
using G = DiGraph<K>;

namespace {

G loadGraph(uint32_t N, const uint32_t *off, const uint32_t *edg) {
  G x;
  x.respan(N);
  for (K u = 0; u < N; ++u) x.addVertex(u);
  for (K u = 0; u < N; ++u)
    for (uint32_t i = off[u]; i < off[u + 1]; ++i) x.addEdge(u, edg[i]);
  updateOmpU(x);
  return x;
}

vector<tuple<K, K>> loadEdges(const uint32_t *us, const uint32_t *vs, size_t n) {
  vector<tuple<K, K>> a;
  a.reserve(n);
  for (size_t i = 0; i < n; ++i) a.push_back({us[i], vs[i]});
  return a;
}

}  // namespace

extern "C" {

// mode: 0 static, 1 naive-dynamic, 2 dynamic traversal, 3 dynamic frontier, 4 dynamic frontier with pruning.
// async: the reference's ASYNC template switch (main.cu runs the dynamic CPU solvers with ASYNC=true).
// ranks: span(y) outputs. Returns 0 on success, -1 on bad arguments.
int ref_pagerank(int mode, int async, int threads, uint32_t N, const uint32_t *xoff, const uint32_t *xedg,
                 const uint32_t *delu, const uint32_t *delv, size_t ND, const uint32_t *insu,
                 const uint32_t *insv, size_t NI, const double *q, double tolerance, double frontierTolerance,
                 double pruneTolerance, double damping, int maxIterations, int repeat, double *ranks,
                 int *iterations, float *time_ms) {
  if (mode < 0 || mode > 4 || N == 0) return -1;
  if (threads > 0) omp_set_num_threads(threads);
  G x = loadGraph(N, xoff, xedg);
  auto xt = transposeWithDegreeOmp(x);
  auto deletions = loadEdges(delu, delv, ND);
  auto insertions = loadEdges(insu, insv, NI);
  auto y = duplicate(x);
  tidyBatchUpdateU(deletions, insertions, y);
  applyBatchUpdateOmpU(y, deletions, insertions);
  auto yt = transposeWithDegreeOmp(y);
  PagerankOptions<V> o(repeat, tolerance, frontierTolerance, pruneTolerance, damping, maxIterations);
  vector<V> qv;
  if (q) qv.assign(q, q + N);
  PagerankResult<V> res;
  if (mode != 0 && !q) return -1;
  switch (mode) {
    case 0: res = async ? pagerankStaticOmp<true>(yt, o) : pagerankStaticOmp<false>(yt, o); break;
    case 1: res = async ? pagerankNaiveDynamicOmp<true>(yt, &qv, o) : pagerankNaiveDynamicOmp<false>(yt, &qv, o); break;
    case 2:
      res = async ? pagerankDynamicTraversalOmp<true>(x, xt, y, yt, deletions, insertions, &qv, o)
                  : pagerankDynamicTraversalOmp<false>(x, xt, y, yt, deletions, insertions, &qv, o);
      break;
    case 3:
      res = async ? pagerankDynamicFrontierOmp<true>(x, xt, y, yt, deletions, insertions, &qv, o)
                  : pagerankDynamicFrontierOmp<false>(x, xt, y, yt, deletions, insertions, &qv, o);
      break;
    case 4:
      res = async ? pagerankPruneDynamicFrontierOmp<true>(x, xt, y, yt, deletions, insertions, &qv, o)
                  : pagerankPruneDynamicFrontierOmp<false>(x, xt, y, yt, deletions, insertions, &qv, o);
      break;
  }
  size_t S = res.ranks.size();
  for (size_t i = 0; i < N; ++i) ranks[i] = i < S ? res.ranks[i] : 0.0;
  *iterations = res.iterations;
  *time_ms = res.time;
  return 0;
}

// Affected flags of the reference's CPU marking on the tidied batch: kind 0 = dynamic frontier
// (pagerankAffectedFrontierOmpW, pagerank.hxx:662-676), kind 1 = dynamic traversal
// (pagerankAffectedTraversalOmpW, pagerank.hxx:562-578). vaff: N outputs (0/1).
int ref_affected(int kind, uint32_t N, const uint32_t *xoff, const uint32_t *xedg, const uint32_t *delu,
                 const uint32_t *delv, size_t ND, const uint32_t *insu, const uint32_t *insv, size_t NI,
                 uint8_t *vaff) {
  if (N == 0) return -1;
  G x = loadGraph(N, xoff, xedg);
  auto deletions = loadEdges(delu, delv, ND);
  auto insertions = loadEdges(insu, insv, NI);
  auto y = duplicate(x);
  tidyBatchUpdateU(deletions, insertions, y);
  applyBatchUpdateOmpU(y, deletions, insertions);
  vector<char> vis(max(x.span(), y.span()));
  if (kind == 0) pagerankAffectedFrontierOmpW(vis, x, y, deletions, insertions);
  else pagerankAffectedTraversalOmpW(vis, x, y, deletions, insertions);
  for (size_t i = 0; i < N; ++i) vaff[i] = vis[i] ? 1 : 0;
  return 0;
}

// Tidy and apply a batch with the reference's own routines (batch.hxx:201-247) and return y as an out-CSR.
// y_off: N+1 outputs; y_edg: capacity M_x + NI. Returns the edge count of y, or -1.
int64_t ref_apply_batch(uint32_t N, const uint32_t *xoff, const uint32_t *xedg, const uint32_t *delu,
                        const uint32_t *delv, size_t ND, const uint32_t *insu, const uint32_t *insv, size_t NI,
                        uint32_t *y_off, uint32_t *y_edg) {
  if (N == 0) return -1;
  G x = loadGraph(N, xoff, xedg);
  auto deletions = loadEdges(delu, delv, ND);
  auto insertions = loadEdges(insu, insv, NI);
  auto y = duplicate(x);
  tidyBatchUpdateU(deletions, insertions, y);
  applyBatchUpdateOmpU(y, deletions, insertions);
  uint32_t o = 0;
  for (K u = 0; u < N; ++u) {
    y_off[u] = o;
    y.forEachEdgeKey(u, [&](K v) { y_edg[o++] = v; });
  }
  y_off[N] = o;
  return int64_t(o);
}

// The reference's readers (mtx.hxx readMtxW, snap.hxx readTemporalW) into its DiGraph, exported as an out-CSR over
// the graph's existing vertices in ascending key order (the compaction pagerankInvokeCuda applies,
// pagerankCuda.hxx:756-763). Call with off = edg = NULL to get *n and *m, then again with arrays of n + 1 and m.
static int export_csr(const G &x, uint32_t *n, uint64_t *m, uint32_t *off, uint32_t *edg) {
  vector<K> keys;
  x.forEachVertexKey([&](auto u) { keys.push_back(u); });
  std::sort(keys.begin(), keys.end());
  vector<uint32_t> id(x.span(), 0);
  for (size_t i = 0; i < keys.size(); ++i) id[keys[i]] = uint32_t(i);
  *n = uint32_t(keys.size());
  *m = x.size();
  if (!off || !edg) return 0;
  uint64_t k = 0;
  for (size_t i = 0; i < keys.size(); ++i) {
    off[i] = uint32_t(k);
    vector<uint32_t> row;
    x.forEachEdgeKey(keys[i], [&](auto v) { row.push_back(id[v]); });
    std::sort(row.begin(), row.end());
    for (uint32_t v : row) edg[k++] = v;
  }
  off[keys.size()] = uint32_t(k);
  return 0;
}

int ref_read_mtx(const char *path, uint32_t *n, uint64_t *m, uint32_t *off, uint32_t *edg) {
  G x;
  readMtxW(x, path);
  return export_csr(x, n, m, off, edg);
}

int ref_read_temporal(const char *path, uint64_t rows, uint64_t lines, int symmetric, uint32_t *n, uint64_t *m,
                      uint32_t *off, uint32_t *edg) {
  G x;
  readTemporalW(x, path, false, symmetric != 0, rows, lines);
  return export_csr(x, n, m, off, edg);
}

}  // extern "C"
