// capi_internal.hpp -- host-side state of the C ABI (contexts, graphs) and the helpers its translation units share.
//
// The C ABI (include/pagerank_hip.h) is implemented by several units:
//   capi.hip      options, tuning, contexts, graph upload, the solve entry points, partition and marking exports;
//   plans.hip     the degree-ordered layout and the launch plans built from it (row, sliced, propagation-blocked),
//                 and the out-degree schedule of the extension kernels;
//   solve.hip     the solver driver (pagerankInvokeCuda + pagerankLoopCuU, pagerankCuda.hxx:675-692, :713-849);
//   forms.hip     batched personalized PageRank and the weighted forward / adjoint solves;
//   autotune.hip  the launch-plan autotuner and its cache;
//   session.hip   the streaming dynamic session.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../../include/pagerank_hip.h"
#include "prh_internal.hpp"

namespace prhc {

constexpr size_t kPinnedBytes = 4096;

constexpr size_t kStagingBytes = size_t(64) << 20;

// One-time preparation of a graph (its layout, a launch plan, the extension schedule): the wall time of its last
// build, and whether a solve has reported it yet. The first solve that uses it adds it to its initialization time,
// as the reference times its per-solve partition there (pagerankCuda.hxx:811-814).
struct Prep {
  float ms = 0.0f;
  bool reported = true;
  float take() {
    if (reported) return 0.0f;
    reported = true;
    return ms;
  }
};

// Sets p to the wall time from construction to destruction (a builder's scope), not yet reported.
struct PrepTimer {
  Prep &p;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  explicit PrepTimer(Prep &q) : p(q) {}
  ~PrepTimer() {
    p.ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    p.reported = false;
  }
};

template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t cap = 0;  // elements
};

}  // namespace prhc

struct prh_context {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string error;
  size_t limit = 0, used = 0;
  prh_tuning tuning{};
  prhc::DevBuf<double> rank[2], contrib[2];
  prhc::DevBuf<uint8_t> vaff, naff, buff;
  prhc::DevBuf<uint32_t> sched_in, sched_out, tiles, counts, batch, scratch_u32;
  prhc::DevBuf<unsigned long long> loop;  // [L] error bits, then one slot holding `done`
  prhc::DevBuf<int> flags;                // [0] validation bits, [1] traversal "changed", [2] extension direction
  prhc::DevBuf<double> ext_stats;         // per rank workgroup: pull and push costs of the next extension
  prhc::DevBuf<uint32_t> ext_chunks;      // initial extension: chunk counts and their scan per batch source
  prhc::DevBuf<uint8_t> ext_tmp;          // its scan's temporary storage
  // in-place graph updates: per-row degree change, changed-row flags, new degrees (kept between updates)
  prhc::DevBuf<int> upd_delta;
  prhc::DevBuf<uint8_t> upd_changed;
  prhc::DevBuf<uint32_t> upd_deg;
  prhc::DevBuf<uint4> upd_ranges;
  // Arena of the updates' small temporaries (kept between calls; see prhc::Temps): in use up to arena_top, grown to
  // arena_want when a call asked for more than it holds.
  prhc::DevBuf<uint8_t> arena;
  size_t arena_top = 0, arena_want = 0;
  // batched personalized PageRank
  prhc::DevBuf<double> ppr_rank[2], ppr_contrib[2], ppr_partial, ppr_mass, ppr_out;
  prhc::DevBuf<int32_t> ppr_row;
  prhc::DevBuf<uint32_t> ppr_tables;
  prhc::DevBuf<unsigned long long> ppr_err;
  prhc::DevBuf<prh::PprState> ppr_state;
  // weighted forward / adjoint
  // weighted forward / adjoint, on the degree-ordered layout ("l": layout order, else the caller's order)
  prhc::DevBuf<double> wt_x[2], wt_z[2], wt_W, wt_Wl, wt_term, wt_terml, wt_r, wt_rl, wt_gwl, wt_gsl, wt_outM, wt_outN;
  prhc::DevBuf<float> wt_wout, wt_wl;
  prhc::DevBuf<uint32_t> wt_chunks;       // heavy-row chunk tables of the weighted pull
  prhc::DevBuf<uint32_t> wt_sort, wt_sort_tmp;  // degree sort of its low rows
  prhc::DevBuf<double> wt_partial;        // their sums
  uint64_t *pinned = nullptr;       // small host-pinned readback area (kPinnedBytes)
  // Launch mode of the iteration loops that replay captured HIP graphs (batched PPR, the streaming session): true =
  // captured graphs; false (PRH_GRAPH_CAPTURE=0 in the environment at context creation) = the same launches issued
  // one by one, for tools that trace kernels.
  bool capture = true;
  uint8_t *staging = nullptr;  // host-pinned staging area of large device-to-host copies (kStagingBytes), on demand
  uint32_t *noalloc_scratch = nullptr;  // device address of noalloc.hip's fixed scratch on this device
  hipEvent_t ev[4] = {};
  bool ktiming = false;
  std::vector<hipEvent_t> kev;  // start/stop pairs per iteration
  uint64_t k_launches = 0;
  double k_ms = 0.0;
};

struct BinnedCache {
  int B = 0;
  bool valid = false;
  bool slot = false;  // the plan keeps the slot -> in-CSR-slot map (weighted solves, in-place updates)
  prhc::Prep prep;
  // In-place updates (prh_graph_update): the in-CSR the plan was built from (kept from the first update on), the
  // inverse of its slot map, and the overlay of inserted edges that are not slots of the plan (capacity ov_cap).
  uint32_t *base_off = nullptr, *base_src = nullptr, *slot_of = nullptr;
  uint64_t base_M = 0, base_cap = 0;  // base edges, and the capacity of base_src
  uint64_t slot_of_n = 0;             // entries of slot_of (the edges of the graph the plan was built from)
  unsigned long long *ov_key = nullptr;
  uint32_t *ov_off = nullptr;
  double *ovX = nullptr;
  size_t ov_cap = 0;
  prh::BinnedPlan plan{};
  std::vector<std::pair<void *, size_t>> bufs;  // its device allocations
};

struct prh_graph {
  prh_context *ctx = nullptr;
  uint32_t N = 0;
  uint64_t M = 0;
  // As uploaded (caller's ids): used by the batched and weighted solvers.
  uint32_t *in_off = nullptr, *in_src = nullptr, *outdeg = nullptr, *out_off = nullptr, *out_dst = nullptr;
  // Raw out-CSR slot of every slot of the layout's in-CSR and out-CSR, built on first weighted use.
  uint32_t *l_in_eid = nullptr, *l_out_eid = nullptr;
  // Degree-ordered layout used by the five solvers: layout row i is vertex perm[i] (in-degree descending, stable),
  // inv is the inverse; in-rows sorted by new source id.
  uint32_t *perm = nullptr, *inv = nullptr;
  uint32_t *l_in_off = nullptr, *l_in_src = nullptr, *l_outdeg = nullptr, *l_out_off = nullptr, *l_out_dst = nullptr;
  std::vector<uint32_t> sorted_deg;  // host: in-degree of layout row i (descending when layout == 0)
  int layout = 0;                    // 0: rows by in-degree, descending; 1: the caller's vertex order
  uint32_t dead_ends = 0;  // vertices with out-degree 0
  // In-place updates (prh_graph_update): once one has run, the caller-order arrays above are freed (the batched and
  // weighted solvers then refuse the graph); the layout's out-rows are kept ascending for the batch tidy.
  bool updated = false, out_sorted = false;
  // Arrays of an earlier CSR kept for the next update to write into (capacity in elements), so that a series of
  // updates does not allocate per batch.
  uint32_t *spare_in_off = nullptr, *spare_in_src = nullptr, *spare_out_off = nullptr, *spare_out_dst = nullptr;
  uint64_t spare_in_cap = 0, spare_out_cap = 0;
  uint64_t l_cap = 0;     // capacity (elements) of l_out_dst
  uint64_t l_in_cap = 0;  // capacity (elements) of l_in_src
  // An update that patches the blocked plan leaves the in-CSR unbuilt (the plan's iterations do not read it): while
  // in_stale, l_in_off and l_in_src are null and the in-CSR is the plan's base (bin.base_off, bin.base_src) less the
  // net_del_n edges of net_del plus the net_ins_n edges of net_ins (device arrays of keys row << 32 | column in
  // layout ids, ascending, relative to the base as long as it exists). ensure_in_csr builds it for whatever reads it.
  bool in_stale = false;
  unsigned long long *net_del = nullptr, *net_ins = nullptr;
  uint64_t net_del_n = 0, net_ins_n = 0, net_del_cap = 0, net_ins_cap = 0;
  uint32_t *net_counts = nullptr;  // device: the two counts written by the last composition
  prhc::Prep update_prep;  // update time not yet reported by a solve
  prhc::Prep layout_prep, plan_prep, sliced_prep, out_prep;
  // Rank-kernel plan for the current tuning: section table and the chunk tables of the heaviest rows.
  prh_tuning plan_tuning{};
  bool plan_valid = false;
  prh::RankPlan plan{};
  uint32_t *chunk_tables = nullptr;  // item_row, item_beg, item_end, row_first
  size_t chunk_tables_bytes = 0;
  double *chunk_partial = nullptr;
  size_t chunk_partial_bytes = 0;
  uint32_t n_chunked = 0;
  // Source-sliced plan of the plain update, built on first use for the current slices / chunk threshold.
  int sliced_slices = 0;
  uint32_t sliced_huge = 0;
  uint32_t sliced_block_edges = 0;
  bool sliced_valid = false;
  prh::SlicedPlan sliced{};
  std::vector<std::pair<void *, size_t>> sliced_bufs;  // its device allocations
  // Out-degree schedule of the layout's out-CSR for the extension kernels (built on first use): the stable
  // partition [<= 64 | <= kExtendChunk | heavier] and the heavier rows cut into chunks (beg, end, row).
  bool out_valid = false;
  uint32_t *out_order = nullptr, *out_chunks = nullptr;
  uint32_t out_n[3] = {0, 0, 0}, out_n_chunks = 0;
  // Propagation-blocked plans, built on first use for the current block width: of the layout in-CSR (the rank
  // update, the weighted forward) and of its out-CSR (the weighted adjoint).
  BinnedCache bin, bin_t;
  // Banded plan (rows in the caller's order), built on first use from the layout in-CSR's spread; an in-place update
  // keeps it (edges beyond its window are gathered directly).
  bool win_valid = false;
  prh::WindowPlan win{};
  prhc::Prep win_prep;
};

namespace prhc {

int fail(prh_context *c, int status, const char *fmt, ...) __attribute__((format(printf, 3, 4)));

#define PRH_HIP(ctx, expr)                                                                              \
  do {                                                                                                  \
    hipError_t e_ = (expr);                                                                             \
    if (e_ != hipSuccess) return prhc::fail(ctx, PRH_ERROR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

#define PRH_CHECK(expr)      \
  do {                       \
    int s_ = (expr);         \
    if (s_ != PRH_SUCCESS) return s_; \
  } while (0)

int dev_alloc(prh_context *c, void **p, size_t bytes);
void dev_free(prh_context *c, void *p, size_t bytes);

template <class T>
int ensure(prh_context *c, DevBuf<T> &b, size_t n) {
  if (b.cap >= n && b.p) return PRH_SUCCESS;
  if (b.p) {
    PRH_HIP(c, hipStreamSynchronize(c->stream));
    dev_free(c, b.p, b.cap * sizeof(T));
    b.p = nullptr;
    b.cap = 0;
  }
  void *p = nullptr;
  PRH_CHECK(dev_alloc(c, &p, n * sizeof(T)));
  b.p = static_cast<T *>(p);
  b.cap = n;
  return PRH_SUCCESS;
}

// Device temporaries of one call. A request of at most kArenaItem bytes is carved from the context's arena when it
// fits there; other requests are allocated here and freed when the Temps ends. Nested Temps end in reverse order of
// creation. When the outermost ends, an arena smaller than the largest total of small requests a call made is
// regrown (at most kArenaMax), so that repeated calls of one size allocate nothing.
constexpr size_t kArenaItem = size_t(64) << 20, kArenaMax = size_t(512) << 20, kArenaAlign = 256;

struct Temps {
  prh_context *c;
  size_t mark;
  std::vector<std::pair<void *, size_t>> own;
  explicit Temps(prh_context *cc);
  ~Temps();
  Temps(const Temps &) = delete;
  Temps &operator=(const Temps &) = delete;
  int get_bytes(void **p, size_t bytes);
  template <class T>
  int get(T **p, size_t n) {
    void *q = nullptr;
    PRH_CHECK(get_bytes(&q, std::max<size_t>(n * sizeof(T), 16)));
    *p = static_cast<T *>(q);
    return PRH_SUCCESS;
  }
};

template <class T>
void release(prh_context *c, DevBuf<T> &b) {
  dev_free(c, b.p, b.cap * sizeof(T));
  b.p = nullptr;
  b.cap = 0;
}

inline bool valid_method(int m) { return m >= PRH_STATIC && m <= PRH_PRUNE_DYNAMIC_FRONTIER; }

// Number of rows with in-degree > x in the descending degree list.
inline uint32_t rows_above(const std::vector<uint32_t> &deg, uint32_t x) {
  return uint32_t(std::lower_bound(deg.begin(), deg.end(), x, [](uint32_t a, uint32_t b) { return a > b; }) -
                  deg.begin());
}

// Rows [0, n) of a degree-ordered layout are those with in-degree > x; a layout in the caller's order has no such
// prefix (0).
inline uint32_t layout_rows_above(const prh_graph *g, uint32_t x) {
  return g->layout == 0 ? rows_above(g->sorted_deg, x) : 0u;
}

// Capture `body` (launches on the context stream only) into an executable graph.
template <class F>
int capture_graph(prh_context *c, F body, hipGraphExec_t *exec) {
  hipGraph_t graph = nullptr;
  PRH_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  const int st = body();
  const hipError_t e = hipStreamEndCapture(c->stream, &graph);
  if (st != PRH_SUCCESS || e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return st != PRH_SUCCESS ? st : fail(c, PRH_ERROR_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
  }
  const hipError_t ie = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ie != hipSuccess) return fail(c, PRH_ERROR_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(ie));
  return PRH_SUCCESS;
}

// Batch arrays staged in the context (host or device source), checked against N and mapped to layout ids.
struct DevBatch {
  const uint32_t *delu = nullptr, *delv = nullptr, *insu = nullptr;
  size_t ND = 0, NI = 0;
};

// Device-to-host copy of n bytes into caller (pageable) memory on the context's stream, synchronously: in chunks of
// kStagingBytes through the context's pinned staging area when n exceeds one chunk, else directly (capi.hip).
int copy_to_host(prh_context *c, void *dst, const void *src, size_t n);

// plans.hip
int build_layout(prh_context *c, prh_graph *g, const uint32_t *in_off, const uint32_t *outdeg, bool with_out);
// Rebuilds the graph's layout in `mode` (0 degree-ordered, 1 the caller's order) from its uploaded arrays and drops
// every plan built on the old one.
int relayout(prh_context *c, prh_graph *g, int mode);
// Builds the layout in-CSR of a graph whose last updates left it unbuilt (in_stale); no-op otherwise.
int ensure_in_csr(prh_context *c, prh_graph *g);
void free_sliced(prh_context *c, prh_graph *g);
int build_sliced_plan(prh_context *c, prh_graph *g);
// The propagation-blocked plan of a pull CSR (off, idx; N rows of degrees rowdeg, M slots) into bc, for the
// context's blocking width (kept: the slot -> CSR-slot map).
int build_binned(prh_context *c, uint32_t N, uint64_t M, const uint32_t *off, const uint32_t *idx,
                 const std::vector<uint32_t> &rowdeg, bool keep_slot, BinnedCache &bc);
void free_binned(prh_context *c, BinnedCache &bc);
void free_overlay(prh_context *c, BinnedCache &bc);
int build_binned_plan(prh_context *c, prh_graph *g, bool keep_slot = false);
int build_binned_plan_t(prh_context *c, prh_graph *g);
int build_rank_plan(prh_context *c, prh_graph *g);
int build_window_plan(prh_context *c, prh_graph *g);
int out_schedule(prh_context *c, prh_graph *g, prh::ExtendArgs &E);

// noalloc.hip
int noalloc_setup(prh_context *c);

// apply.hip
int csr_to_caller(prh_context *c, const uint32_t *perm_d, const uint32_t *inv_d, uint32_t N, const uint32_t *loff_d,
                  const uint32_t *lidx_d, uint64_t M, uint32_t *off, uint32_t *idx);

// solve.hip
int stage_batch(prh_context *c, const prh_graph *g, const uint32_t *du, const uint32_t *dv, size_t nd,
                const uint32_t *iu, size_t ni, bool host, DevBatch &b);
int mark_affected(prh_context *c, const prh_graph *g, int method, const DevBatch &b, const prh::ExtendArgs &xo);
int ensure_workspace(prh_context *c, uint32_t N, bool dynamic, int L);
int solve_impl(prh_context *c, const prh_graph *g, int method, const prh_options *o, const uint32_t *du,
               const uint32_t *dv, size_t nd, const uint32_t *iu, size_t ni, const double *q, double *ranks,
               prh_stats *st, bool host, bool account_prep);

}  // namespace prhc
