// graphgen.cpp -- host-side graph preparation for the PageRank engine (C++17 + OpenMP).
//
// Builds the inputs the solvers take, CSR-natively (no per-vertex containers), so that the 1.9B-edge
// configuration fits in host memory:
//   * seeded RMAT and 2D-grid generators, plus a builder from an explicit edge list;
//   * the graph preparation of the reference's experiment (main.cu:168-170): duplicate edges removed, one
//     self-loop per vertex (selfLoop.hxx:59-79), out-neighbours and in-neighbours ascending (the key order of
//     the reference's DiGraph, _bitset.hxx:19-21), out-degree as the per-vertex value of the transpose
//     (transpose.hxx:109-134);
//   * random update batches (80% insertions of uniform pairs, 20% deletions of uniform existing non-self-loop
//     edges), tidied like tidyBatchUpdateU (batch.hxx:201-208) and applied like applyBatchUpdateOmpU
//     (batch.hxx:240-247).
// Random numbers: SplitMix64 (Steele, Lea and Flood, "Fast splittable pseudorandom number generators",
// OOPSLA 2014), one stream per sample index, so results do not depend on the thread count.
#include <omp.h>
#include <cstdio>
#include <cstdlib>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

namespace {

inline uint64_t splitmix64_next(uint64_t &x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Start state of the stream for sample i under seed s.
inline uint64_t stream_state(uint64_t seed, uint64_t i) {
  uint64_t x = seed ^ (i * 0xD1B54A32D192ED03ull);
  return splitmix64_next(x);
}

inline double uniform01(uint64_t &st) { return double(splitmix64_next(st) >> 11) * 0x1.0p-53; }

template <class T>
T *alloc_array(size_t n) {
  return static_cast<T *>(::operator new[](sizeof(T) * (n ? n : 1)));
}

}  // namespace

// A prepared graph: out-CSR and in-CSR with ascending neighbour lists, out-degrees, and the out-CSR edge id of
// every in-CSR slot (in_eid), used to align per-edge data (weights) between the two orders.
struct prhh_graph {
  uint32_t N = 0;
  uint64_t M = 0;
  uint32_t *out_off = nullptr, *out_edg = nullptr;
  uint32_t *in_off = nullptr, *in_edg = nullptr, *in_eid = nullptr;
  uint32_t *outdeg = nullptr;
  ~prhh_graph() {
    ::operator delete[](out_off);
    ::operator delete[](out_edg);
    ::operator delete[](in_off);
    ::operator delete[](in_edg);
    ::operator delete[](in_eid);
    ::operator delete[](outdeg);
  }
};

// A tidied batch update: deletions (del_u[i], del_v[i]) and insertions (ins_u[i], ins_v[i]), each sorted by
// (source, target) and unique.
struct prhh_batch {
  std::vector<uint32_t> del_u, del_v, ins_u, ins_v;
};

namespace {

// Exclusive prefix sum of counts[0..n) into off[0..n] (off[n] = total). Returns the total.
uint64_t exclusive_scan(const uint32_t *counts, uint64_t *off, uint32_t n) {
  int T = omp_get_max_threads();
  std::vector<uint64_t> part(size_t(T) + 1, 0);
  #pragma omp parallel num_threads(T)
  {
    int t = omp_get_thread_num(), nt = omp_get_num_threads();
    uint64_t b = uint64_t(n) * t / nt, e = uint64_t(n) * (t + 1) / nt, s = 0;
    for (uint64_t i = b; i < e; ++i) s += counts[i];
    part[t + 1] = s;
    #pragma omp barrier
    #pragma omp single
    for (int k = 0; k < nt; ++k) part[k + 1] += part[k];
    s = part[t];
    for (uint64_t i = b; i < e; ++i) { off[i] = s; s += counts[i]; }
  }
  off[n] = part[T];
  return part[T];
}

// Turn bucketed rows (row u holds raw[bstart[u] .. bstart[u]+bcount[u]) in any order, possibly with repeats)
// into a sorted, duplicate-free out-CSR, adding the self-loop (u,u) to every row when self_loops is set.
bool finalize_rows(prhh_graph *g, uint32_t N, uint32_t *raw, const uint64_t *bstart, uint32_t *bcount,
                   bool self_loops) {
  uint32_t *deg = alloc_array<uint32_t>(N);
  #pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t u = 0; u < int64_t(N); ++u) {
    uint32_t *b = raw + bstart[u], *e = b + bcount[u];
    std::sort(b, e);
    uint32_t *m = std::unique(b, e);
    uint32_t d = uint32_t(m - b);
    bcount[u] = d;  // now the unique count of the sorted prefix
    if (self_loops && !std::binary_search(b, m, uint32_t(u))) ++d;
    deg[u] = d;
  }
  std::unique_ptr<uint64_t[]> off(new uint64_t[size_t(N) + 1]);
  uint64_t M = exclusive_scan(deg, off.get(), N);
  if (M >= (uint64_t(1) << 32)) { ::operator delete[](deg); return false; }
  g->N = N;
  g->M = M;
  g->out_off = alloc_array<uint32_t>(size_t(N) + 1);
  g->out_edg = alloc_array<uint32_t>(M);
  g->outdeg = deg;
  #pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t u = 0; u < int64_t(N); ++u) {
    g->out_off[u] = uint32_t(off[u]);
    uint32_t *b = raw + bstart[u], *m = b + bcount[u];
    uint32_t *dst = g->out_edg + off[u];
    bool need = self_loops && !std::binary_search(b, m, uint32_t(u));
    for (uint32_t *p = b; p < m; ++p) {
      if (need && *p > uint32_t(u)) { *dst++ = uint32_t(u); need = false; }
      *dst++ = *p;
    }
    if (need) *dst++ = uint32_t(u);
  }
  g->out_off[N] = uint32_t(M);
  return true;
}

// Build the in-CSR (sources ascending) and in_eid from the out-CSR.
void build_transpose(prhh_graph *g) {
  uint32_t N = g->N;
  uint64_t M = g->M;
  std::unique_ptr<std::atomic<uint32_t>[]> cnt(new std::atomic<uint32_t>[N ? N : 1]);
  #pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < int64_t(N); ++v) cnt[v].store(0, std::memory_order_relaxed);
  #pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t u = 0; u < int64_t(N); ++u)
    for (uint32_t i = g->out_off[u]; i < g->out_off[u + 1]; ++i)
      cnt[g->out_edg[i]].fetch_add(1, std::memory_order_relaxed);
  uint32_t *indeg = alloc_array<uint32_t>(N);
  #pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < int64_t(N); ++v) indeg[v] = cnt[v].load(std::memory_order_relaxed);
  std::unique_ptr<uint64_t[]> off(new uint64_t[size_t(N) + 1]);
  exclusive_scan(indeg, off.get(), N);
  ::operator delete[](indeg);
  g->in_off = alloc_array<uint32_t>(size_t(N) + 1);
  g->in_edg = alloc_array<uint32_t>(M);
  g->in_eid = alloc_array<uint32_t>(M);
  #pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < int64_t(N); ++v) {
    g->in_off[v] = uint32_t(off[v]);
    cnt[v].store(0, std::memory_order_relaxed);
  }
  g->in_off[N] = uint32_t(M);
  #pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t u = 0; u < int64_t(N); ++u)
    for (uint32_t i = g->out_off[u]; i < g->out_off[u + 1]; ++i) {
      uint32_t v = g->out_edg[i];
      uint64_t p = off[v] + cnt[v].fetch_add(1, std::memory_order_relaxed);
      g->in_edg[p] = uint32_t(u);
      g->in_eid[p] = i;
    }
  // Sort each in-row by source (eids follow; sources are unique within a row).
  #pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t v = 0; v < int64_t(N); ++v) {
    uint32_t b = g->in_off[v], e = g->in_off[v + 1];
    uint32_t n = e - b;
    if (n < 2) continue;
    bool sorted = true;
    for (uint32_t k = b + 1; k < e && sorted; ++k) sorted = g->in_edg[k - 1] < g->in_edg[k];
    if (sorted) continue;
    std::vector<uint64_t> tmp(n);
    for (uint32_t k = 0; k < n; ++k) tmp[k] = (uint64_t(g->in_edg[b + k]) << 32) | g->in_eid[b + k];
    std::sort(tmp.begin(), tmp.end());
    for (uint32_t k = 0; k < n; ++k) {
      g->in_edg[b + k] = uint32_t(tmp[k] >> 32);
      g->in_eid[b + k] = uint32_t(tmp[k]);
    }
  }
}

// Phase lines on stderr when PRHH_PROGRESS is set (large builds take minutes).
void progress(const char *phase, uint64_t n) {
  static const bool on = std::getenv("PRHH_PROGRESS") != nullptr;
  if (!on) return;
  std::fprintf(stderr, "[graphgen] %s (%llu)\n", phase, (unsigned long long)n);
  std::fflush(stderr);
}

// Generic bucketed build from a sample function f(i) -> (u, v) for i in [0, S).
template <class F>
prhh_graph *build_from_samples(uint32_t N, uint64_t S, bool self_loops, F f) {
  std::unique_ptr<std::atomic<uint32_t>[]> cnt(new (std::nothrow) std::atomic<uint32_t>[N ? N : 1]);
  if (!cnt) return nullptr;
  #pragma omp parallel for schedule(static)
  for (int64_t u = 0; u < int64_t(N); ++u) cnt[u].store(0, std::memory_order_relaxed);
  #pragma omp parallel for schedule(static, 65536)
  for (int64_t i = 0; i < int64_t(S); ++i) {
    uint32_t u, v;
    f(uint64_t(i), u, v);
    cnt[u].fetch_add(1, std::memory_order_relaxed);
  }
  progress("sources counted", S);
  uint32_t *bcount = alloc_array<uint32_t>(N);
  #pragma omp parallel for schedule(static)
  for (int64_t u = 0; u < int64_t(N); ++u) {
    bcount[u] = cnt[u].load(std::memory_order_relaxed);
    cnt[u].store(0, std::memory_order_relaxed);
  }
  std::unique_ptr<uint64_t[]> bstart(new uint64_t[size_t(N) + 1]);
  exclusive_scan(bcount, bstart.get(), N);
  uint32_t *raw = alloc_array<uint32_t>(S);
  #pragma omp parallel for schedule(static, 65536)
  for (int64_t i = 0; i < int64_t(S); ++i) {
    uint32_t u, v;
    f(uint64_t(i), u, v);
    raw[bstart[u] + cnt[u].fetch_add(1, std::memory_order_relaxed)] = v;
  }
  cnt.reset();
  progress("samples bucketed by source", S);
  prhh_graph *g = new prhh_graph();
  bool ok = finalize_rows(g, N, raw, bstart.get(), bcount, self_loops);
  ::operator delete[](raw);
  ::operator delete[](bcount);
  if (!ok) { delete g; return nullptr; }
  progress("rows sorted and deduplicated", g->M);
  build_transpose(g);
  progress("transpose built", g->M);
  return g;
}

}  // namespace

extern "C" {

// RMAT graph: 2^scale vertices, round(edge_factor * 2^scale) samples; at each of the scale levels the
// quadrant (src bit, dst bit) is (0,0), (0,1), (1,0), (1,1) with probabilities a, b, c, 1-a-b-c.
// Duplicates are removed and, if self_loops, one self-loop is added per vertex.
prhh_graph *prhh_rmat(uint32_t scale, double edge_factor, double a, double b, double c, uint64_t seed,
                      int self_loops) {
  if (scale == 0 || scale > 31) return nullptr;
  uint32_t N = uint32_t(1) << scale;
  uint64_t S = uint64_t(edge_factor * double(N) + 0.5);
  const double ab = a + b, abc = a + b + c;
  auto f = [=](uint64_t i, uint32_t &u, uint32_t &v) {
    uint64_t st = stream_state(seed, i);
    uint32_t x = 0, y = 0;
    for (uint32_t l = 0; l < scale; ++l) {
      double p = uniform01(st);
      uint32_t bu = p >= ab, bv = (p >= a && p < ab) || p >= abc;
      x = (x << 1) | bu;
      y = (y << 1) | bv;
    }
    u = x;
    v = y;
  };
  return build_from_samples(N, S, self_loops != 0, f);
}

// Road-like grid: rows x cols vertices (id = i*cols + j); each right edge (i,j)-(i,j+1) and down edge
// (i,j)-(i+1,j) is kept with probability p, kept edges are symmetrized, and every vertex gets a self-loop.
prhh_graph *prhh_grid(uint32_t rows, uint32_t cols, double p, uint64_t seed) {
  uint64_t NN = uint64_t(rows) * cols;
  if (NN == 0 || NN >= (uint64_t(1) << 32)) return nullptr;
  uint32_t N = uint32_t(NN);
  auto keep = [=](uint64_t slot) {  // slot = 2*id + 0 (right) / 1 (down)
    uint64_t st = stream_state(seed, slot);
    return uniform01(st) < p;
  };
  prhh_graph *g = new prhh_graph();
  g->N = N;
  uint32_t *deg = alloc_array<uint32_t>(N);
  #pragma omp parallel for schedule(static)
  for (int64_t id = 0; id < int64_t(N); ++id) {
    uint32_t i = uint32_t(id / cols), j = uint32_t(id % cols), d = 1;
    if (i > 0 && keep(2 * uint64_t(id - cols) + 1)) ++d;
    if (j > 0 && keep(2 * uint64_t(id - 1))) ++d;
    if (j + 1 < cols && keep(2 * uint64_t(id))) ++d;
    if (i + 1 < rows && keep(2 * uint64_t(id) + 1)) ++d;
    deg[id] = d;
  }
  std::unique_ptr<uint64_t[]> off(new uint64_t[size_t(N) + 1]);
  g->M = exclusive_scan(deg, off.get(), N);
  g->outdeg = deg;
  g->out_off = alloc_array<uint32_t>(size_t(N) + 1);
  g->out_edg = alloc_array<uint32_t>(g->M);
  #pragma omp parallel for schedule(static)
  for (int64_t id = 0; id < int64_t(N); ++id) {
    uint32_t i = uint32_t(id / cols), j = uint32_t(id % cols);
    uint32_t *dst = g->out_edg + off[id];
    g->out_off[id] = uint32_t(off[id]);
    if (i > 0 && keep(2 * uint64_t(id - cols) + 1)) *dst++ = uint32_t(id - cols);
    if (j > 0 && keep(2 * uint64_t(id - 1))) *dst++ = uint32_t(id - 1);
    *dst++ = uint32_t(id);
    if (j + 1 < cols && keep(2 * uint64_t(id))) *dst++ = uint32_t(id + 1);
    if (i + 1 < rows && keep(2 * uint64_t(id) + 1)) *dst++ = uint32_t(id + cols);
  }
  g->out_off[N] = uint32_t(g->M);
  build_transpose(g);
  return g;
}

// Graph from an explicit edge list (src[i] -> dst[i], ids < N); duplicates removed; optional self-loops.
prhh_graph *prhh_from_edges(uint32_t N, uint64_t S, const uint32_t *src, const uint32_t *dst, int self_loops) {
  if (N == 0) return nullptr;
  for (uint64_t i = 0; i < S; ++i)
    if (src[i] >= N || dst[i] >= N) return nullptr;
  auto f = [=](uint64_t i, uint32_t &u, uint32_t &v) { u = src[i]; v = dst[i]; };
  return build_from_samples(N, S, self_loops != 0, f);
}

void prhh_free(prhh_graph *g) { delete g; }
uint32_t prhh_order(const prhh_graph *g) { return g->N; }
uint64_t prhh_size(const prhh_graph *g) { return g->M; }
const uint32_t *prhh_out_off(const prhh_graph *g) { return g->out_off; }
const uint32_t *prhh_out_edg(const prhh_graph *g) { return g->out_edg; }
const uint32_t *prhh_in_off(const prhh_graph *g) { return g->in_off; }
const uint32_t *prhh_in_edg(const prhh_graph *g) { return g->in_edg; }
const uint32_t *prhh_in_eid(const prhh_graph *g) { return g->in_eid; }
const uint32_t *prhh_outdeg(const prhh_graph *g) { return g->outdeg; }

// Random batch on x: round(size * ins_frac) insertions of uniform (u, v) pairs and the rest deletions of uniform
// existing non-self-loop edges, then tidied (deletions kept only if present, insertions only if absent, each
// sorted by (u, v) and unique), as tidyBatchUpdateU does (batch.hxx:152-208).
prhh_batch *prhh_random_batch(const prhh_graph *x, uint64_t size, double ins_frac, uint64_t seed) {
  if (!x || x->N == 0) return nullptr;
  uint64_t NI = uint64_t(double(size) * ins_frac + 0.5), ND = size - std::min(size, NI);
  uint32_t N = x->N;
  uint64_t M = x->M;
  std::vector<uint64_t> ins(NI), del(ND);
  #pragma omp parallel for schedule(static, 4096)
  for (int64_t i = 0; i < int64_t(NI); ++i) {
    uint64_t st = stream_state(seed, uint64_t(i));
    uint32_t u = uint32_t(uniform01(st) * N), v = uint32_t(uniform01(st) * N);
    ins[i] = (uint64_t(std::min(u, N - 1)) << 32) | std::min(v, N - 1);
  }
  #pragma omp parallel for schedule(static, 4096)
  for (int64_t i = 0; i < int64_t(ND); ++i) {
    uint64_t st = stream_state(seed ^ 0xA5A5A5A5DEADBEEFull, uint64_t(i));
    uint64_t key = ~uint64_t(0);
    for (int tries = 0; tries < 64; ++tries) {
      uint64_t e = std::min(uint64_t(uniform01(st) * double(M)), M - 1);
      const uint32_t *ub = std::upper_bound(x->out_off, x->out_off + N + 1, uint32_t(e));
      uint32_t u = uint32_t(ub - x->out_off - 1), v = x->out_edg[e];
      if (u != v) { key = (uint64_t(u) << 32) | v; break; }
    }
    del[i] = key;
  }
  del.erase(std::remove(del.begin(), del.end(), ~uint64_t(0)), del.end());
  auto exists = [&](uint64_t k) {
    uint32_t u = uint32_t(k >> 32), v = uint32_t(k);
    return std::binary_search(x->out_edg + x->out_off[u], x->out_edg + x->out_off[u + 1], v);
  };
  ins.erase(std::remove_if(ins.begin(), ins.end(), [&](uint64_t k) { return exists(k); }), ins.end());
  std::sort(ins.begin(), ins.end());
  ins.erase(std::unique(ins.begin(), ins.end()), ins.end());
  std::sort(del.begin(), del.end());
  del.erase(std::unique(del.begin(), del.end()), del.end());
  prhh_batch *b = new prhh_batch();
  for (uint64_t k : del) { b->del_u.push_back(uint32_t(k >> 32)); b->del_v.push_back(uint32_t(k)); }
  for (uint64_t k : ins) { b->ins_u.push_back(uint32_t(k >> 32)); b->ins_v.push_back(uint32_t(k)); }
  return b;
}

// Batch from explicit lists, tidied against x the same way.
prhh_batch *prhh_make_batch(const prhh_graph *x, const uint32_t *du, const uint32_t *dv, uint64_t nd,
                            const uint32_t *iu, const uint32_t *iv, uint64_t ni) {
  if (!x) return nullptr;
  auto exists = [&](uint32_t u, uint32_t v) {
    return u < x->N && std::binary_search(x->out_edg + x->out_off[u], x->out_edg + x->out_off[u + 1], v);
  };
  std::vector<uint64_t> del, ins;
  for (uint64_t i = 0; i < nd; ++i)
    if (exists(du[i], dv[i])) del.push_back((uint64_t(du[i]) << 32) | dv[i]);
  for (uint64_t i = 0; i < ni; ++i) {
    if (iu[i] >= x->N || iv[i] >= x->N) { return nullptr; }
    if (!exists(iu[i], iv[i])) ins.push_back((uint64_t(iu[i]) << 32) | iv[i]);
  }
  std::sort(del.begin(), del.end());
  del.erase(std::unique(del.begin(), del.end()), del.end());
  std::sort(ins.begin(), ins.end());
  ins.erase(std::unique(ins.begin(), ins.end()), ins.end());
  prhh_batch *b = new prhh_batch();
  for (uint64_t k : del) { b->del_u.push_back(uint32_t(k >> 32)); b->del_v.push_back(uint32_t(k)); }
  for (uint64_t k : ins) { b->ins_u.push_back(uint32_t(k >> 32)); b->ins_v.push_back(uint32_t(k)); }
  return b;
}

void prhh_batch_free(prhh_batch *b) { delete b; }
uint64_t prhh_batch_deletions(const prhh_batch *b) { return b->del_u.size(); }
uint64_t prhh_batch_insertions(const prhh_batch *b) { return b->ins_u.size(); }
const uint32_t *prhh_batch_del_u(const prhh_batch *b) { return b->del_u.data(); }
const uint32_t *prhh_batch_del_v(const prhh_batch *b) { return b->del_v.data(); }
const uint32_t *prhh_batch_ins_u(const prhh_batch *b) { return b->ins_u.data(); }
const uint32_t *prhh_batch_ins_v(const prhh_batch *b) { return b->ins_v.data(); }

// y = x with the (tidied) batch applied: deletions removed, then insertions added (applyBatchUpdateOmpU,
// batch.hxx:240-247); out-rows stay ascending, in-CSR rebuilt.
prhh_graph *prhh_apply_batch(const prhh_graph *x, const prhh_batch *b) {
  if (!x || !b) return nullptr;
  uint32_t N = x->N;
  size_t ND = b->del_u.size(), NI = b->ins_u.size();
  // Per-source ranges of the sorted batch lists.
  std::vector<uint64_t> dstart(size_t(N) + 1, 0), istart(size_t(N) + 1, 0);
  for (size_t i = 0; i < ND; ++i) ++dstart[b->del_u[i] + 1];
  for (size_t i = 0; i < NI; ++i) ++istart[b->ins_u[i] + 1];
  for (uint32_t u = 0; u < N; ++u) { dstart[u + 1] += dstart[u]; istart[u + 1] += istart[u]; }
  uint32_t *deg = alloc_array<uint32_t>(N);
  #pragma omp parallel for schedule(static)
  for (int64_t u = 0; u < int64_t(N); ++u)
    deg[u] = x->out_off[u + 1] - x->out_off[u] - uint32_t(dstart[u + 1] - dstart[u]) +
             uint32_t(istart[u + 1] - istart[u]);
  std::unique_ptr<uint64_t[]> off(new uint64_t[size_t(N) + 1]);
  uint64_t M = exclusive_scan(deg, off.get(), N);
  if (M >= (uint64_t(1) << 32)) { ::operator delete[](deg); return nullptr; }
  prhh_graph *y = new prhh_graph();
  y->N = N;
  y->M = M;
  y->outdeg = deg;
  y->out_off = alloc_array<uint32_t>(size_t(N) + 1);
  y->out_edg = alloc_array<uint32_t>(M);
  #pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t u = 0; u < int64_t(N); ++u) {
    y->out_off[u] = uint32_t(off[u]);
    const uint32_t *p = x->out_edg + x->out_off[u], *pe = x->out_edg + x->out_off[u + 1];
    const uint32_t *d = b->del_v.data() + dstart[u], *de = b->del_v.data() + dstart[u + 1];
    const uint32_t *q = b->ins_v.data() + istart[u], *qe = b->ins_v.data() + istart[u + 1];
    uint32_t *o = y->out_edg + off[u];
    while (p < pe || q < qe) {
      if (q == qe || (p < pe && *p < *q)) {
        while (d < de && *d < *p) ++d;
        if (!(d < de && *d == *p)) *o++ = *p;
        ++p;
      } else {
        *o++ = *q++;
      }
    }
  }
  y->out_off[N] = uint32_t(M);
  build_transpose(y);
  return y;
}


// ---- readers (mtx.hxx, snap.hxx) ---------------------------------------------------------------------------------
// Edge lists with 0-based ids: file vertex k (1-based, as the reference's DiGraph keeps it) becomes k - 1, which is
// where pagerankInvokeCuda's compaction of the existing vertices puts it (pagerankCuda.hxx:756-763, properties.hxx).

struct prhh_edges {
  std::vector<uint32_t> src, dst;
  uint64_t order = 0;  // vertices 1..order in the file
  int symmetric = 0;
};

struct prhh_temporal {
  FILE *f = nullptr;
  ~prhh_temporal() {
    if (f) std::fclose(f);
  }
};

namespace {

// Parses "u v ..." (1-based ids); false on a line that does not start with two integers.
bool parse_pair(const char *line, uint64_t &u, uint64_t &v) {
  char *end = nullptr;
  u = std::strtoull(line, &end, 10);
  if (end == line) return false;
  const char *p = end;
  v = std::strtoull(p, &end, 10);
  return end != p;
}

}  // namespace

// Matrix Market coordinate file as a directed edge list (readMtxDo, mtx.hxx:120-140): comment lines start with
// '%'; the banner's symmetry field "symmetric" or "skew-symmetric" adds each edge's mirror; the size line is
// "rows cols nnz"; body lines "u v [w]" stop at the first line without two integers. Returns NULL if the file
// cannot be read, is not "matrix coordinate", or names a vertex 0 or above max(rows, cols).
prhh_edges *prhh_read_mtx(const char *path) {
  FILE *f = std::fopen(path, "r");
  if (!f) return nullptr;
  std::unique_ptr<prhh_edges> e(new prhh_edges());
  char line[4096];
  char h[5][64] = {};
  bool banner = false, ok = true;
  uint64_t rows = 0, cols = 0, size = 0;
  while (std::fgets(line, sizeof(line), f)) {
    if (line[0] == '%') {
      if (line[1] == '%') {
        banner = std::sscanf(line, "%63s %63s %63s %63s %63s", h[0], h[1], h[2], h[3], h[4]) >= 4;
      }
      continue;
    }
    ok = banner && std::strcmp(h[1], "matrix") == 0 && std::strcmp(h[2], "coordinate") == 0 &&
         std::sscanf(line, "%llu %llu %llu", (unsigned long long *)&rows, (unsigned long long *)&cols,
                     (unsigned long long *)&size) == 3;
    break;
  }
  if (!ok) {
    std::fclose(f);
    return nullptr;
  }
  e->symmetric = std::strcmp(h[4], "symmetric") == 0 || std::strcmp(h[4], "skew-symmetric") == 0;
  e->order = std::max(rows, cols);
  e->src.reserve(size * (e->symmetric ? 2 : 1));
  e->dst.reserve(size * (e->symmetric ? 2 : 1));
  while (std::fgets(line, sizeof(line), f)) {
    uint64_t u, v;
    if (!parse_pair(line, u, v)) break;
    if (u == 0 || v == 0 || u > e->order || v > e->order) {
      std::fclose(f);
      return nullptr;
    }
    e->src.push_back(uint32_t(u - 1));
    e->dst.push_back(uint32_t(v - 1));
    if (e->symmetric) {
      e->src.push_back(uint32_t(v - 1));
      e->dst.push_back(uint32_t(u - 1));
    }
  }
  std::fclose(f);
  return e.release();
}

// SNAP temporal edge file read in consecutive pieces, as readTemporalDo does (snap.hxx:40-60): each call reads up to
// `lines` further lines "u v [t]" (1-based ids <= rows), adding mirrors when symmetric, and stops early at the end of
// the file or at a line without two integers. Returns NULL on an id 0 or above rows.
prhh_temporal *prhh_temporal_open(const char *path) {
  FILE *f = std::fopen(path, "r");
  if (!f) return nullptr;
  prhh_temporal *t = new prhh_temporal();
  t->f = f;
  return t;
}

prhh_edges *prhh_temporal_read(prhh_temporal *t, uint64_t rows, uint64_t lines, int symmetric) {
  if (!t || !t->f) return nullptr;
  std::unique_ptr<prhh_edges> e(new prhh_edges());
  e->order = rows;
  e->symmetric = symmetric != 0;
  char line[4096];
  for (; lines > 0 && std::fgets(line, sizeof(line), t->f); --lines) {
    uint64_t u, v;
    if (!parse_pair(line, u, v)) break;
    if (u == 0 || v == 0 || u > rows || v > rows) return nullptr;
    e->src.push_back(uint32_t(u - 1));
    e->dst.push_back(uint32_t(v - 1));
    if (symmetric) {
      e->src.push_back(uint32_t(v - 1));
      e->dst.push_back(uint32_t(u - 1));
    }
  }
  return e.release();
}

void prhh_temporal_close(prhh_temporal *t) { delete t; }

uint64_t prhh_edges_count(const prhh_edges *e) { return e ? e->src.size() : 0; }
uint64_t prhh_edges_order(const prhh_edges *e) { return e ? e->order : 0; }
int prhh_edges_symmetric(const prhh_edges *e) { return e ? e->symmetric : 0; }
const uint32_t *prhh_edges_src(const prhh_edges *e) { return e ? e->src.data() : nullptr; }
const uint32_t *prhh_edges_dst(const prhh_edges *e) { return e ? e->dst.data() : nullptr; }
void prhh_edges_free(prhh_edges *e) { delete e; }

}  // extern "C"
