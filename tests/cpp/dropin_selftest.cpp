// dropin_selftest.cpp -- compiles a caller written against the reference's GPU entry points (pagerankCuda.hxx:864-988)
// against pagerankHip.hxx, with the CSR graph class of the package, and runs all five solvers on one graph pair.
//
// usage: dropin_selftest <x_off.bin> <x_edg.bin> <y_off.bin> <y_edg.bin> <batch.bin> <q.bin> <out_prefix>
//   *.bin are raw little-endian arrays (uint32 for graphs and the batch [nd, ni, du..., dv..., iu..., iv...],
//   float64 for q); writes <out_prefix>.<method>.bin (float64 ranks) and prints one line per method.
// usage: dropin_selftest --static <off.bin> <edg.bin> <repeats> <out_prefix>
//   static solves of one graph, `repeats` times: one line each with the times and the launch plan the drop-in's
//   context ran (its blocking width; 0 = the row plan), and the ranks of each solve in <out_prefix>.<i>.bin.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <fstream>
#include <stdexcept>
#include <tuple>
#include <vector>

#include "pagerankTypes.hxx"
#include "csrGraph.hxx"
#include "pagerankHip.hxx"

template <class T>
static std::vector<T> load(const char *path) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f) throw std::runtime_error(std::string("cannot open ") + path);
  size_t n = size_t(f.tellg()) / sizeof(T);
  std::vector<T> a(n);
  f.seekg(0);
  f.read(reinterpret_cast<char *>(a.data()), n * sizeof(T));
  return a;
}

static void save(const std::string &path, const std::vector<double> &a) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char *>(a.data()), a.size() * sizeof(double));
}

static int static_only(int argc, char **argv) {
  if (argc != 6) {
    std::fprintf(stderr, "usage: %s --static off edg repeats out_prefix\n", argv[0]);
    return 2;
  }
  using K = uint32_t;
  using V = double;
  CsrDiGraph<K> y(load<uint32_t>(argv[2]), load<K>(argv[3]));
  auto yt = transposeWithDegree(y);
  const int repeats = std::atoi(argv[4]);
  const std::string out = argv[5];
  PagerankOptions<V> o;
  for (int i = 0; i < repeats; ++i) {
    auto r = pagerankStaticCuda(y, yt, o);
    prh_tuning t{};
    prh_context_get_tuning(prh_host::context().get(), &t);
    std::printf("static[%d] iterations=%d time=%.3f init=%.3f comp=%.3f ms_per_iteration=%.4f blocking=%d\n", i,
                r.iterations, r.time, r.initializationTime, r.computationTime,
                r.computationTime / (r.iterations > 0 ? r.iterations : 1), t.blocking);
    save(out + "." + std::to_string(i) + ".bin", r.ranks);
  }
  return 0;
}

int main(int argc, char **argv) {
  if (argc > 1 && std::string(argv[1]) == "--static") return static_only(argc, argv);
  if (argc != 8) {
    std::fprintf(stderr, "usage: %s x_off x_edg y_off y_edg batch q out_prefix\n", argv[0]);
    return 2;
  }
  using K = uint32_t;
  using V = double;
  CsrDiGraph<K> x(load<uint32_t>(argv[1]), load<K>(argv[2]));
  CsrDiGraph<K> y(load<uint32_t>(argv[3]), load<K>(argv[4]));
  auto xt = transposeWithDegree(x);
  auto yt = transposeWithDegree(y);
  auto b = load<uint32_t>(argv[5]);
  size_t nd = b[0], ni = b[1];
  std::vector<std::tuple<K, K>> deletions, insertions;
  for (size_t i = 0; i < nd; ++i) deletions.push_back({b[2 + i], b[2 + nd + i]});
  for (size_t i = 0; i < ni; ++i) insertions.push_back({b[2 + 2 * nd + i], b[2 + 2 * nd + ni + i]});
  std::vector<V> q = load<V>(argv[6]);
  std::string out = argv[7];
  PagerankOptions<V> o;

  auto report = [&](const char *name, const PagerankResult<V> &r) {
    std::printf("%s iterations=%d time=%.3f init=%.3f mark=%.3f comp=%.3f\n", name, r.iterations, r.time,
                r.initializationTime, r.markingTime, r.computationTime);
    save(out + "." + name + ".bin", r.ranks);
  };
  report("static", pagerankStaticCuda<true>(y, yt, o));
  report("naive", pagerankNaiveDynamicCuda<true>(y, yt, &q, o));
  report("traversal", pagerankDynamicTraversalCuda<true>(x, xt, y, yt, deletions, insertions, &q, o));
  report("frontier", pagerankDynamicFrontierCuda<true>(x, xt, y, yt, deletions, insertions, &q, o));
  report("prune", pagerankPruneDynamicFrontierCuda<false>(x, xt, y, yt, deletions, insertions, &q, o));
  // The HIP-named aliases and an empty graph (the reference returns an empty result).
  CsrDiGraph<K> e;
  auto et = transposeWithDegree(e);
  auto r = pagerankStaticHip(e, et, o);
  std::printf("empty ranks=%zu iterations=%d\n", r.ranks.size(), r.iterations);
  // An unknown batch vertex is an error, not a silent remap to id 0.
  try {
    std::vector<std::tuple<K, K>> bad{{K(y.span() + 5), 0}};
    pagerankDynamicFrontierHip(x, xt, y, yt, bad, insertions, &q, o);
    std::printf("unknown-vertex: accepted\n");
  } catch (const std::invalid_argument &) {
    std::printf("unknown-vertex: rejected\n");
  }
  return 0;
}
