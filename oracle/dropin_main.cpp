// dropin_main.cpp -- TEST INFRASTRUCTURE ONLY. The reference's own experiment calls (main.cu:88-143), compiled with
// the reference's DiGraph, batch and CPU headers (from <reference>/inc, where they lie) but with pagerankHip.hxx in
// place of pagerankCuda.hxx. Built by `make -C oracle ref` into oracle/_ref/dropin_main; it shows that callers
// written against the reference compile unchanged and links libpagerank_hip.so.
//
// usage: dropin_main <x_off.bin> <x_edg.bin> <batch.bin> <out_prefix>
//   The graph x (uint32 out-CSR) is loaded into DiGraph<uint32_t>; the batch [nd, ni, du..., dv..., iu..., iv...]
//   is tidied and applied with tidyBatchUpdateU / applyBatchUpdateOmpU; q = pagerankStaticOmp(xt) as in main.cu:65.
//   For every GPU solver it prints the L1 error against pagerankStaticOmp(yt, tolerance 1e-100), like main.cu:54-61,
//   and writes <out_prefix>.<name>.bin (float64 ranks over span).
#define BUILD 0
#define OPENMP 1
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <string>
#include "_main.hxx"
#include "Graph.hxx"
#include "update.hxx"
#include "duplicate.hxx"
#include "transpose.hxx"
#include "batch.hxx"
#include "pagerank.hxx"
#include "pagerankPrune.hxx"
#include "pagerankHip.hxx"

template <class T>
static vector<T> load(const char *path) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  size_t n = size_t(f.tellg()) / sizeof(T);
  vector<T> a(n);
  f.seekg(0);
  f.read(reinterpret_cast<char *>(a.data()), n * sizeof(T));
  return a;
}

int main(int argc, char **argv) {
  if (argc != 5) return 2;
  using K = uint32_t;
  using V = double;
  auto off = load<uint32_t>(argv[1]), edg = load<uint32_t>(argv[2]), b = load<uint32_t>(argv[3]);
  string out = argv[4];
  DiGraph<K> x;
  K N = K(off.size() - 1);
  x.respan(N);
  for (K u = 0; u < N; ++u) x.addVertex(u);
  for (K u = 0; u < N; ++u)
    for (uint32_t i = off[u]; i < off[u + 1]; ++i) x.addEdge(u, edg[i]);
  updateOmpU(x);
  auto xt = transposeWithDegreeOmp(x);
  size_t nd = b[0], ni = b[1];
  vector<tuple<K, K>> deletions, insertions;
  for (size_t i = 0; i < nd; ++i) deletions.push_back({b[2 + i], b[2 + nd + i]});
  for (size_t i = 0; i < ni; ++i) insertions.push_back({b[2 + 2 * nd + i], b[2 + 2 * nd + ni + i]});
  auto y = duplicate(x);
  tidyBatchUpdateU(deletions, insertions, y);
  applyBatchUpdateOmpU(y, deletions, insertions);
  auto yt = transposeWithDegreeOmp(y);
  auto r0 = pagerankStaticOmp(xt, PagerankOptions<V>(1, 1e-100));
  auto s0 = pagerankStaticOmp(yt, PagerankOptions<V>(1, 1e-100));
  auto q = r0.ranks;
  {
    std::ofstream f(out + ".q.bin", std::ios::binary);
    f.write(reinterpret_cast<const char *>(q.data()), q.size() * sizeof(V));
  }
  auto glog = [&](const char *name, const PagerankResult<V> &a) {
    std::printf("%s iterations=%d l1err=%.3e time=%.3f\n", name, a.iterations, l1NormDeltaOmp(a.ranks, s0.ranks),
                a.time);
    std::ofstream f(out + "." + name + ".bin", std::ios::binary);
    f.write(reinterpret_cast<const char *>(a.ranks.data()), a.ranks.size() * sizeof(V));
  };
  int repeat = 1;
  glog("static", pagerankStaticCuda<false>(y, yt, PagerankOptions<V>(repeat)));
  glog("static_partition", pagerankStaticCuda<true>(y, yt, PagerankOptions<V>(repeat)));
  glog("naive", pagerankNaiveDynamicCuda<true>(y, yt, &q, {repeat}));
  glog("frontier", pagerankDynamicFrontierCuda<true>(x, xt, y, yt, deletions, insertions, &q, {repeat}));
  glog("prune", pagerankPruneDynamicFrontierCuda<true>(x, xt, y, yt, deletions, insertions, &q, {repeat}));
  glog("traversal", pagerankDynamicTraversalCuda<true>(x, xt, y, yt, deletions, insertions, &q, {repeat}));
  return 0;
}
