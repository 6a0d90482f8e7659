// autotune.hip -- the launch-plan autotuner (no counterpart in the reference, whose launch constants are fixed)
// and its JSON cache.
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

// Key of a (device, graph) pair: the device's architecture and compute-unit count, N, M and an FNV-1a hash of
// the graph's in-degree histogram (which, with the layout, fixes every launch plan).
std::string tuning_key(prh_context *c, const prh_graph *g) {
  hipDeviceProp_t prop{};
  if (hipGetDeviceProperties(&prop, c->device) != hipSuccess) (void)hipGetLastError();
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](uint64_t x) {
    for (int k = 0; k < 8; ++k) {
      h ^= (x >> (8 * k)) & 0xFF;
      h *= 1099511628211ull;
    }
  };
  mix(g->N);
  mix(g->M);
  // The in-degree histogram (the same in every vertex order, so the key survives a change of layout).
  std::vector<uint64_t> hist;
  for (uint32_t d : g->sorted_deg) {
    if (d >= hist.size()) hist.resize(size_t(d) + 1, 0);
    ++hist[d];
  }
  for (size_t d = 0; d < hist.size(); ++d)
    if (hist[d]) {
      mix(d);
      mix(hist[d]);
    }
  char buf[160];
  std::string arch(prop.gcnArchName);
  arch = arch.substr(0, arch.find(':'));
  std::snprintf(buf, sizeof(buf), "%s-cu%d-n%u-m%llu-%016llx", arch.c_str(), prop.multiProcessorCount, g->N,
                (unsigned long long)g->M, (unsigned long long)h);
  return buf;
}

constexpr int kTuningCount = 10;
const char *const kTuningFields[kTuningCount] = {"switch_degree", "huge_degree", "low_width", "check_interval",
                                                 "source_slices", "block_edges", "extension_direction", "blocking",
                                                 "layout", "window"};

void tuning_to_values(const prh_tuning &t, long long v[kTuningCount]) {
  v[0] = t.switch_degree;
  v[1] = t.huge_degree;
  v[2] = t.low_width;
  v[3] = t.check_interval;
  v[4] = t.source_slices;
  v[5] = t.block_edges;
  v[6] = t.extension_direction;
  v[7] = t.blocking;
  v[8] = t.layout;
  v[9] = t.window;
}

void values_to_tuning(const long long v[kTuningCount], prh_tuning &t) {
  t.switch_degree = uint32_t(v[0]);
  t.huge_degree = uint32_t(v[1]);
  t.low_width = int(v[2]);
  t.check_interval = int(v[3]);
  t.source_slices = int(v[4]);
  t.block_edges = uint32_t(v[5]);
  t.extension_direction = int(v[6]);
  t.blocking = int(v[7]);
  t.layout = int(v[8]);
  t.window = int(v[9]);
}

// The cache is a JSON object with one entry per line:
//   {
//   "<key>": {"switch_degree": 128, ..., "ms_per_iteration": 1.92},
//   ...
//   }
// Lines that do not parse as an entry are ignored (a hand-edited file degrades to a cache miss).
struct CacheEntry {
  std::string key;
  long long v[kTuningCount];
  double ms;
};

std::vector<CacheEntry> read_cache(const char *path) {
  std::vector<CacheEntry> out;
  FILE *f = std::fopen(path, "r");
  if (!f) return out;
  char line[1024];
  while (std::fgets(line, sizeof(line), f)) {
    const char *q0 = std::strchr(line, '"');
    const char *q1 = q0 ? std::strchr(q0 + 1, '"') : nullptr;
    if (!q1) continue;
    CacheEntry e;
    e.key.assign(q0 + 1, q1);
    bool ok = true;
    for (int k = 0; k < kTuningCount && ok; ++k) {
      const std::string pat = std::string("\"") + kTuningFields[k] + "\":";
      const char *p = std::strstr(q1, pat.c_str());
      ok = p && std::sscanf(p + pat.size(), "%lld", &e.v[k]) == 1;
    }
    const char *pm = std::strstr(q1, "\"ms_per_iteration\":");
    if (!ok || !pm || std::sscanf(pm + 19, "%lf", &e.ms) != 1) continue;
    out.push_back(e);
  }
  std::fclose(f);
  return out;
}

bool write_cache(const char *path, const std::vector<CacheEntry> &entries) {
  const std::string tmp = std::string(path) + ".tmp";
  FILE *f = std::fopen(tmp.c_str(), "w");
  if (!f) return false;
  std::fprintf(f, "{\n");
  for (size_t i = 0; i < entries.size(); ++i) {
    const CacheEntry &e = entries[i];
    std::fprintf(f, "\"%s\": {", e.key.c_str());
    for (int k = 0; k < kTuningCount; ++k) std::fprintf(f, "\"%s\": %lld, ", kTuningFields[k], e.v[k]);
    std::fprintf(f, "\"ms_per_iteration\": %.6f}%s\n", e.ms, i + 1 < entries.size() ? "," : "");
  }
  std::fprintf(f, "}\n");
  const bool ok = std::fclose(f) == 0;
  return ok && std::rename(tmp.c_str(), path) == 0;
}

}  // namespace
}  // namespace prhc

using namespace prhc;

extern "C" {

int prh_autotune(prh_context *c, const prh_graph *g, const char *cache_path, int iterations, prh_tuning *chosen,
                 double *ms_per_iteration, int *from_cache) {
  if (!c) return PRH_ERROR_INVALID_ARGUMENT;
  if (!g || g->ctx != c || iterations < 1) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "bad graph or iterations");
  if (from_cache) *from_cache = 0;
  PRH_HIP(c, hipSetDevice(c->device));
  const prh_tuning base = c->tuning;
  if (g->N == 0) {
    if (chosen) *chosen = base;
    if (ms_per_iteration) *ms_per_iteration = 0;
    return PRH_SUCCESS;
  }
  const std::string key = tuning_key(c, g);
  std::vector<CacheEntry> cache;
  if (cache_path) {
    cache = read_cache(cache_path);
    for (const CacheEntry &e : cache)
      if (e.key == key) {
        prh_tuning t = base;
        values_to_tuning(e.v, t);
        PRH_CHECK(prh_context_set_tuning(c, &t));
        if (chosen) *chosen = t;
        if (ms_per_iteration) *ms_per_iteration = e.ms;
        if (from_cache) *from_cache = 1;
        return PRH_SUCCESS;
      }
  }
  // Candidates on the degree-ordered layout: the row plan over switch degree x chunk threshold x block size, the
  // sliced plan over slice count x block size, and the propagation-blocked plan over its block width (graphs of 2^20
  // edges and more, where its streaming passes pay). On graphs whose in-degrees all stay at or below 64 (a road
  // network, a grid), also the row plan, the banded plan and the blocked plan on the caller's vertex order, which
  // keeps their neighbourhoods together (tried last: each layout change rebuilds the layout). A graph updated in
  // place keeps its layout, so its candidates all take that one. Each is timed on `iterations` static iterations
  // (tolerance 0) after one warm-up solve.
  const int fixed_layout = g->updated ? g->layout : -1;
  const uint32_t max_deg = *std::max_element(g->sorted_deg.begin(), g->sorted_deg.end());
  std::vector<prh_tuning> cand;
  for (uint32_t sw : {64u, 128u, 256u})
    for (uint32_t hu : {4096u, 16384u})
      for (uint32_t be : {2048u, 4096u}) {
        prh_tuning t = base;
        t.switch_degree = sw;
        t.huge_degree = hu;
        t.low_width = 0;
        t.source_slices = 0;
        t.block_edges = be;
        t.blocking = 0;
        cand.push_back(t);
      }
  for (int sl : {1, 8})
    for (uint32_t be : {2048u, 4096u}) {
      prh_tuning t = base;
      t.switch_degree = 128;
      t.huge_degree = 16384;
      t.low_width = 0;
      t.source_slices = sl;
      t.block_edges = be;
      t.blocking = 0;
      cand.push_back(t);
    }
  if (g->M >= (uint64_t(1) << 20))
    for (int bw : {2048, 4096, 8192, 16384}) {
      const uint64_t nb = (uint64_t(g->N) + bw - 1) / bw;
      if (nb * nb > (uint64_t(1) << 32)) continue;
      prh_tuning t = base;
      t.switch_degree = 64;
      t.huge_degree = 16384;
      t.low_width = 0;
      t.source_slices = 0;
      t.block_edges = 2048;
      t.blocking = bw;
      cand.push_back(t);
    }
  for (prh_tuning &t : cand) {
    t.layout = fixed_layout < 0 ? 0 : fixed_layout;
    t.window = 0;
  }
  if (max_deg <= 64 && fixed_layout == 1) {  // an updated graph in the caller's order: the banded plan too
    prh_tuning t = base;
    t.blocking = 0;
    t.source_slices = 0;
    t.layout = 1;
    t.window = 1;
    cand.push_back(t);
  }
  if (max_deg <= 64 && fixed_layout < 0) {
    {
      prh_tuning t = base;
      t.switch_degree = 64;
      t.huge_degree = 16384;
      t.low_width = 0;
      t.source_slices = 0;
      t.block_edges = 2048;
      t.blocking = 0;
      t.layout = 1;
      t.window = 1;
      cand.push_back(t);
    }
    // rows in the caller's order: lanes per row from the largest in-degree (0), or one or two lanes per row, which
    // give a wave 64 or 32 neighbouring rows whose gathers land close together on a mesh-like graph
    for (uint32_t lw : {0u, 1u, 2u})
      for (uint32_t be : {2048u, 4096u}) {
        prh_tuning t = base;
        t.switch_degree = 64;
        t.huge_degree = 16384;
        t.low_width = lw;
        t.source_slices = 0;
        t.block_edges = be;
        t.blocking = 0;
        t.layout = 1;
        t.window = 0;
        cand.push_back(t);
      }
    if (g->M >= (uint64_t(1) << 20))
      for (int bw : {2048, 4096, 8192, 16384}) {
        const uint64_t nb = (uint64_t(g->N) + bw - 1) / bw;
        if (nb * nb > (uint64_t(1) << 32)) continue;
        prh_tuning t = base;
        t.switch_degree = 64;
        t.huge_degree = 16384;
        t.low_width = 0;
        t.source_slices = 0;
        t.block_edges = 2048;
        t.blocking = bw;
        t.layout = 1;
        t.window = 0;
        cand.push_back(t);
      }
  }
  double *ranks = nullptr;
  PRH_CHECK(dev_alloc(c, reinterpret_cast<void **>(&ranks), size_t(g->N) * 8));
  prh_options o;
  prh_default_options(&o);
  o.tolerance = 0.0;
  int best = -1;
  double best_ms = 0.0;
  int status = PRH_SUCCESS;
  for (size_t i = 0; i < cand.size() && status == PRH_SUCCESS; ++i) {
    status = prh_context_set_tuning(c, &cand[i]);
    prh_stats st{};
    o.max_iterations = 2;
    if (status == PRH_SUCCESS) status = solve_impl(c, g, PRH_STATIC, &o, nullptr, nullptr, 0, nullptr, 0, nullptr,
                                                   ranks, &st, false, false);
    o.max_iterations = iterations;
    if (status == PRH_SUCCESS) status = solve_impl(c, g, PRH_STATIC, &o, nullptr, nullptr, 0, nullptr, 0, nullptr,
                                                   ranks, &st, false, false);
    if (status != PRH_SUCCESS) break;
    const double ms = double(st.computation_time) / std::max(st.iterations, 1);
    if (best < 0 || ms < best_ms) {
      best = int(i);
      best_ms = ms;
    }
  }
  dev_free(c, ranks, size_t(g->N) * 8);
  if (status != PRH_SUCCESS) {
    c->tuning = base;
    return status;
  }
  PRH_CHECK(prh_context_set_tuning(c, &cand[best]));
  if (chosen) *chosen = cand[best];
  if (ms_per_iteration) *ms_per_iteration = best_ms;
  if (cache_path) {
    CacheEntry e;
    e.key = key;
    tuning_to_values(cand[best], e.v);
    e.ms = best_ms;
    bool replaced = false;
    for (CacheEntry &old : cache)
      if (old.key == key) {
        old = e;
        replaced = true;
      }
    if (!replaced) cache.push_back(e);
    if (!write_cache(cache_path, cache)) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "cannot write %s", cache_path);
  }
  return PRH_SUCCESS;
}

}  // extern "C"
