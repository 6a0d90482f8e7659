// pagerankHip.hxx -- drop-in replacement for the reference's inc/pagerankCuda.hxx on MI355X.
//
// Provides the reference's five GPU entry points with their names, template parameters and argument lists
// (pagerankCuda.hxx:864-988):
//   pagerankStaticCuda<PARTITION, FLAG>(x, xt, o)
//   pagerankNaiveDynamicCuda<PARTITION, FLAG>(x, xt, q, o)
//   pagerankDynamicTraversalCuda / pagerankDynamicFrontierCuda / pagerankPruneDynamicFrontierCuda
//       <PARTITION, FLAG>(x, xt, y, yt, deletions, insertions, q, o)
// plus the same templates under Hip names (pagerankStaticHip, ...). Callers written against the reference compile
// unchanged: replace #include "pagerankCuda.hxx" with #include "pagerankHip.hxx" and link libpagerank_hip.so.
//
// G and H follow the reference's graph concept (Graph.hxx): key_type, span(), order(), size(), empty(),
// forEachVertexKey(f), forEachEdgeKey(u, f) in ascending key order, degree(u), and for H vertexValue(u) = out-degree
// (transposeWithDegreeOmp, transpose.hxx:109-134). PagerankOptions<V> / PagerankResult<V> must be declared first:
// the reference's pagerank.hxx or this directory's pagerankTypes.hxx.
//
// Marshalling restates pagerankInvokeCuda (pagerankCuda.hxx:713-849): the existing vertices of xt (ascending) are
// compacted to 0..N-1 (vertexKeys, properties.hxx:26-33), the in-CSR of yt / out-CSR of y are built in that order
// (csr.hxx:127-247), q is gathered and the ranks scattered back over span() with 0 for absent ids. Differences:
// an unknown batch vertex is an error (the reference silently maps it to id 0, :768-774); device and runtime
// errors throw std::runtime_error (the reference exits the process, _cuda.hxx:136-150); V must be double.
#pragma once
#include <cstdint>
#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "pagerank_hip.h"

namespace prh_host {

// One engine context per host thread (device 0, or PRH_DEVICE); it keeps its device workspace between calls.
//
// Launch plan. Graphs of at least kAutotuneEdges edges run the plan the engine's autotuner picks for them
// (prh_autotune: on RMAT-20 and larger graphs, the propagation-blocked update), chosen once per graph size: a later
// graph with the same vertex count and an edge count within 1% of the tuned one (an updated graph of a dynamic
// series) keeps it. PRH_TUNING_CACHE names a JSON file that keeps the choices across processes. Smaller graphs, and
// every graph when PRH_DETERMINISTIC=1, run the row plan: its sums are in a fixed order, so repeated solves are bit
// for bit identical, where the blocked plan's LDS additions repeat to rounding only.
class ContextHandle {
 public:
  static constexpr uint64_t kAutotuneEdges = uint64_t(1) << 20;

  ContextHandle() {
    int dev = 0;
    if (const char* e = std::getenv("PRH_DEVICE")) dev = std::atoi(e);
    if (prh_context_create(dev, &ctx_) != PRH_SUCCESS) throw std::runtime_error("prh_context_create failed");
    const char* d = std::getenv("PRH_DETERMINISTIC");
    deterministic_ = d && std::atoi(d) != 0;
    if (const char* c = std::getenv("PRH_TUNING_CACHE")) cache_ = c;
    prh_default_tuning(&row_plan_);
  }
  ~ContextHandle() { prh_context_destroy(ctx_); }
  prh_context* get() const { return ctx_; }
  void check(int status) const {
    if (status != PRH_SUCCESS) throw std::runtime_error(std::string("pagerank_hip: ") + prh_context_last_error(ctx_));
  }
  // Sets the context's launch plan for graph g (n vertices, m edges) as described above.
  void choosePlan(prh_graph* g, uint32_t n, uint64_t m) {
    if (deterministic_ || m < kAutotuneEdges) {
      check(prh_context_set_tuning(ctx_, &row_plan_));
      return;
    }
    if (tuned_ && n == tuned_n_ && (m > tuned_m_ ? m - tuned_m_ : tuned_m_ - m) * 100 <= tuned_m_) {
      check(prh_context_set_tuning(ctx_, &tuned_plan_));
      return;
    }
    double ms = 0;
    int hit = 0;
    check(prh_context_set_tuning(ctx_, &row_plan_));
    check(prh_autotune(ctx_, g, cache_.empty() ? nullptr : cache_.c_str(), 10, &tuned_plan_, &ms, &hit));
    tuned_ = true;
    tuned_n_ = n;
    tuned_m_ = m;
  }

 private:
  prh_context* ctx_ = nullptr;
  bool deterministic_ = false;
  std::string cache_;
  prh_tuning row_plan_{}, tuned_plan_{};
  bool tuned_ = false;
  uint32_t tuned_n_ = 0;
  uint64_t tuned_m_ = 0;
};

inline ContextHandle& context() {
  thread_local ContextHandle handle;
  return handle;
}

// Dense index of every existing vertex key of xt (ascending), as vertexKeys + the key map of the reference.
template <class H>
struct KeyIndex {
  using K = typename H::key_type;
  std::vector<K> keys;         // dense id -> key
  std::vector<uint32_t> ids;   // key -> dense id, UINT32_MAX if absent
  bool identity = true;
  explicit KeyIndex(const H& xt) {
    keys.reserve(xt.order());
    ids.assign(xt.span(), UINT32_MAX);
    xt.forEachVertexKey([&](auto u) {
      ids[u] = uint32_t(keys.size());
      if (K(keys.size()) != K(u)) identity = false;
      keys.push_back(K(u));
    });
  }
  uint32_t id(K u) const {
    if (size_t(u) >= ids.size() || ids[u] == UINT32_MAX)
      throw std::invalid_argument("pagerank_hip: batch vertex is not a vertex of the graph");
    return ids[u];
  }
};

// CSR of a graph over the dense ids of `index` (offsets uint32, as the reference's O = uint32_t).
template <class G, class H>
void buildCsr(std::vector<uint32_t>& off, std::vector<uint32_t>& edg, const G& g, const KeyIndex<H>& index) {
  size_t N = index.keys.size(), M = 0;
  for (auto u : index.keys) M += g.degree(u);
  if (M >= (uint64_t(1) << 32)) throw std::length_error("pagerank_hip: more than 2^32-1 edges");
  off.resize(N + 1);
  edg.resize(M);
  size_t o = 0;
  for (size_t i = 0; i < N; ++i) {
    off[i] = uint32_t(o);
    g.forEachEdgeKey(index.keys[i], [&](auto v) { edg[o++] = index.identity ? uint32_t(v) : index.id(v); });
  }
  off[N] = uint32_t(o);
}

class GraphHandle {
 public:
  GraphHandle(prh_context* ctx, const std::vector<uint32_t>& in_off, const std::vector<uint32_t>& in_src,
              const std::vector<uint32_t>& outdeg, const std::vector<uint32_t>* out_off,
              const std::vector<uint32_t>* out_dst) {
    context().check(prh_graph_create(ctx, uint32_t(outdeg.size()), in_src.size(), in_off.data(), in_src.data(),
                                     outdeg.data(), out_off ? out_off->data() : nullptr,
                                     out_dst ? out_dst->data() : nullptr, &g_));
  }
  ~GraphHandle() { prh_graph_destroy(g_); }
  prh_graph* get() const { return g_; }

 private:
  prh_graph* g_ = nullptr;
};

// pagerankInvokeCuda (pagerankCuda.hxx:713-849) over the C ABI: y / yt are the updated graph and its transpose.
template <class V, class G, class H, class K>
PagerankResult<V> invoke(prh_method method, bool partition, const G& y, const H& yt,
                         const std::vector<std::tuple<K, K>>& deletions,
                         const std::vector<std::tuple<K, K>>& insertions, const std::vector<V>* q,
                         const PagerankOptions<V>& o) {
  static_assert(std::is_same<V, double>::value, "pagerank_hip computes FP64 ranks (V = double)");
  ContextHandle& c = context();
  const bool dynamic = method >= PRH_DYNAMIC_TRAVERSAL;
  KeyIndex<H> index(yt);
  const size_t N = index.keys.size(), S = yt.span();
  std::vector<uint32_t> in_off, in_src, outdeg(N), out_off, out_dst;
  buildCsr(in_off, in_src, yt, index);
  for (size_t i = 0; i < N; ++i) outdeg[i] = uint32_t(yt.vertexValue(index.keys[i]));
  if (dynamic) buildCsr(out_off, out_dst, y, index);
  std::vector<double> qc;
  if (q) {
    qc.resize(N);
    for (size_t i = 0; i < N; ++i) qc[i] = (*q)[index.keys[i]];
  }
  std::vector<uint32_t> delu, delv, insu;
  if (dynamic) {
    for (const auto& [u, v] : deletions) {
      delu.push_back(index.id(u));
      delv.push_back(index.id(v));
    }
    for (const auto& [u, v] : insertions) insu.push_back(index.id(u));
  }
  GraphHandle g(c.get(), in_off, in_src, outdeg, dynamic ? &out_off : nullptr, dynamic ? &out_dst : nullptr);
  c.choosePlan(g.get(), uint32_t(N), in_src.size());
  prh_options opt{o.repeat, o.tolerance, o.frontierTolerance, o.pruneTolerance, o.damping, o.maxIterations};
  prh_stats st{};
  std::vector<double> rc(N);
  c.check(prh_solve(c.get(), g.get(), method, partition ? 1 : 0, &opt, delu.data(), delv.data(), delu.size(),
                    insu.data(), insu.size(), q ? qc.data() : nullptr, rc.data(), &st));
  std::vector<V> r(S, V());
  for (size_t i = 0; i < N; ++i) r[index.keys[i]] = rc[i];
  return PagerankResult<V>(r, st.iterations, st.time, st.initialization_time, st.marking_time, st.computation_time);
}

}  // namespace prh_host

// ---- the reference's entry points (pagerankCuda.hxx:864-988) ----------------------------------------------------

template <bool PARTITION = false, class FLAG = char, class G, class H, class V>
inline PagerankResult<V> pagerankStaticCuda(const G& x, const H& xt, const PagerankOptions<V>& o) {
  using K = typename G::key_type;
  if (xt.empty()) return {};
  const std::vector<std::tuple<K, K>> none;
  return prh_host::invoke<V>(PRH_STATIC, PARTITION, x, xt, none, none, static_cast<const std::vector<V>*>(nullptr), o);
}

template <bool PARTITION = false, class FLAG = char, class G, class H, class V>
inline PagerankResult<V> pagerankNaiveDynamicCuda(const G& x, const H& xt, const std::vector<V>* q,
                                                  const PagerankOptions<V>& o) {
  using K = typename G::key_type;
  if (xt.empty()) return {};
  const std::vector<std::tuple<K, K>> none;
  return prh_host::invoke<V>(PRH_NAIVE_DYNAMIC, PARTITION, x, xt, none, none, q, o);
}

template <bool PARTITION = false, class FLAG = char, class G, class H, class K, class V>
inline PagerankResult<V> pagerankDynamicTraversalCuda(const G& x, const H& xt, const G& y, const H& yt,
                                                      const std::vector<std::tuple<K, K>>& deletions,
                                                      const std::vector<std::tuple<K, K>>& insertions,
                                                      const std::vector<V>* q, const PagerankOptions<V>& o) {
  if (xt.empty()) return {};
  return prh_host::invoke<V>(PRH_DYNAMIC_TRAVERSAL, PARTITION, y, yt, deletions, insertions, q, o);
}

template <bool PARTITION = false, class FLAG = char, class G, class H, class K, class V>
inline PagerankResult<V> pagerankDynamicFrontierCuda(const G& x, const H& xt, const G& y, const H& yt,
                                                     const std::vector<std::tuple<K, K>>& deletions,
                                                     const std::vector<std::tuple<K, K>>& insertions,
                                                     const std::vector<V>* q, const PagerankOptions<V>& o) {
  if (xt.empty()) return {};
  return prh_host::invoke<V>(PRH_DYNAMIC_FRONTIER, PARTITION, y, yt, deletions, insertions, q, o);
}

template <bool PARTITION = false, class FLAG = char, class G, class H, class K, class V>
inline PagerankResult<V> pagerankPruneDynamicFrontierCuda(const G& x, const H& xt, const G& y, const H& yt,
                                                          const std::vector<std::tuple<K, K>>& deletions,
                                                          const std::vector<std::tuple<K, K>>& insertions,
                                                          const std::vector<V>* q, const PagerankOptions<V>& o) {
  if (xt.empty()) return {};
  return prh_host::invoke<V>(PRH_PRUNE_DYNAMIC_FRONTIER, PARTITION, y, yt, deletions, insertions, q, o);
}

// ---- HIP-named aliases -----------------------------------------------------------------------------------------

template <bool PARTITION = false, class FLAG = char, class G, class H, class V>
inline PagerankResult<V> pagerankStaticHip(const G& x, const H& xt, const PagerankOptions<V>& o) {
  return pagerankStaticCuda<PARTITION, FLAG>(x, xt, o);
}

template <bool PARTITION = false, class FLAG = char, class G, class H, class V>
inline PagerankResult<V> pagerankNaiveDynamicHip(const G& x, const H& xt, const std::vector<V>* q,
                                                 const PagerankOptions<V>& o) {
  return pagerankNaiveDynamicCuda<PARTITION, FLAG>(x, xt, q, o);
}

template <bool PARTITION = false, class FLAG = char, class G, class H, class K, class V>
inline PagerankResult<V> pagerankDynamicTraversalHip(const G& x, const H& xt, const G& y, const H& yt,
                                                     const std::vector<std::tuple<K, K>>& deletions,
                                                     const std::vector<std::tuple<K, K>>& insertions,
                                                     const std::vector<V>* q, const PagerankOptions<V>& o) {
  return pagerankDynamicTraversalCuda<PARTITION, FLAG>(x, xt, y, yt, deletions, insertions, q, o);
}

template <bool PARTITION = false, class FLAG = char, class G, class H, class K, class V>
inline PagerankResult<V> pagerankDynamicFrontierHip(const G& x, const H& xt, const G& y, const H& yt,
                                                    const std::vector<std::tuple<K, K>>& deletions,
                                                    const std::vector<std::tuple<K, K>>& insertions,
                                                    const std::vector<V>* q, const PagerankOptions<V>& o) {
  return pagerankDynamicFrontierCuda<PARTITION, FLAG>(x, xt, y, yt, deletions, insertions, q, o);
}

template <bool PARTITION = false, class FLAG = char, class G, class H, class K, class V>
inline PagerankResult<V> pagerankPruneDynamicFrontierHip(const G& x, const H& xt, const G& y, const H& yt,
                                                         const std::vector<std::tuple<K, K>>& deletions,
                                                         const std::vector<std::tuple<K, K>>& insertions,
                                                         const std::vector<V>* q, const PagerankOptions<V>& o) {
  return pagerankPruneDynamicFrontierCuda<PARTITION, FLAG>(x, xt, y, yt, deletions, insertions, q, o);
}
