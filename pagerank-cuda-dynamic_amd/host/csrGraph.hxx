// csrGraph.hxx -- a read-only CSR graph that satisfies the graph concept the reference's entry points take
// (the accessors of DiGraph, Graph.hxx:23-372, that pagerankCuda.hxx and csr.hxx use), without per-vertex
// containers, so graphs with billions of edges fit in host memory.
#pragma once
#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

// Vertex ids 0..span()-1, all present; neighbours ascending. V is the per-vertex value (e.g. out-degree for a
// transpose built by transposeWithDegree).
template <class K = uint32_t, class V = K>
class CsrDiGraph {
 public:
  using key_type = K;
  using vertex_value_type = V;

  CsrDiGraph() = default;
  CsrDiGraph(std::vector<uint32_t> offsets, std::vector<K> targets, std::vector<V> values = {})
      : off_(std::move(offsets)), edg_(std::move(targets)), val_(std::move(values)) {}

  size_t span() const noexcept { return off_.empty() ? 0 : off_.size() - 1; }
  size_t order() const noexcept { return span(); }
  size_t size() const noexcept { return edg_.size(); }
  bool empty() const noexcept { return span() == 0; }
  bool directed() const noexcept { return true; }
  bool hasVertex(K u) const noexcept { return size_t(u) < span(); }
  size_t degree(K u) const noexcept { return off_[u + 1] - off_[u]; }
  V vertexValue(K u) const noexcept { return val_.empty() ? V() : val_[u]; }

  template <class F>
  void forEachVertexKey(F f) const {
    for (K u = 0; size_t(u) < span(); ++u) f(u);
  }
  template <class F>
  void forEachVertex(F f) const {
    for (K u = 0; size_t(u) < span(); ++u) f(u, vertexValue(u));
  }
  template <class F>
  void forEachEdgeKey(K u, F f) const {
    for (uint32_t i = off_[u]; i < off_[u + 1]; ++i) f(edg_[i]);
  }

  const std::vector<uint32_t>& offsets() const noexcept { return off_; }
  const std::vector<K>& targets() const noexcept { return edg_; }

 private:
  std::vector<uint32_t> off_;
  std::vector<K> edg_;
  std::vector<V> val_;
};

// The transpose with out-degrees as vertex values (what transposeWithDegreeOmp gives, transpose.hxx:109-134).
template <class K, class V>
CsrDiGraph<K, K> transposeWithDegree(const CsrDiGraph<K, V>& x) {
  const size_t N = x.span();
  std::vector<uint32_t> off(N + 1, 0);
  for (K u = 0; size_t(u) < N; ++u) x.forEachEdgeKey(u, [&](K v) { ++off[size_t(v) + 1]; });
  for (size_t i = 0; i < N; ++i) off[i + 1] += off[i];
  std::vector<K> edg(x.size());
  std::vector<uint32_t> cur(off.begin(), off.end() - 1);
  for (K u = 0; size_t(u) < N; ++u) x.forEachEdgeKey(u, [&](K v) { edg[cur[v]++] = u; });  // ascending u per row
  std::vector<K> deg(N);
  for (K u = 0; size_t(u) < N; ++u) deg[u] = K(x.degree(u));
  return CsrDiGraph<K, K>(std::move(off), std::move(edg), std::move(deg));
}
