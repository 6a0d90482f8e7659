// prh_internal.hpp -- host-side declarations shared by the kernel translation units and the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace prh {

constexpr int kWave = 64;    // CDNA wavefront width
constexpr int kBlock = 256;  // every kernel here runs 4 waves per workgroup
constexpr int kWavesPerBlock = kBlock / kWave;

// Device view of a loop's convergence state (see LoopCtl in prh_device.hpp).
struct LoopState {
  unsigned long long *err;  // [capacity] per-iteration L-inf bits
  int *done;                // single int
};

// Solver variant bits, as the reference's template switches <DYNAMIC, FRONTIER, PRUNE>
// (pagerankCuda.hxx:87, :162, :675).
enum Variant : int {
  kPlain = 0,       // Static / Naive-dynamic: every vertex updated
  kTraversal = 1,   // DYNAMIC: unaffected vertices copied through (Dynamic Traversal)
  kFrontier = 2,    // DYNAMIC + FRONTIER (Dynamic Frontier)
  kPrune = 3,       // DYNAMIC + FRONTIER + PRUNE (Dynamic Frontier with Pruning)
};

// Schedule of one CSR: vertex ids stably partitioned by degree into [low | high | huge] (see partition kernels).
struct Schedule {
  const uint32_t *order;  // [N]
  uint32_t n_low, n_high, n_huge;
};

// The rank kernel works on the degree-ordered layout (layout.hip): rows sorted by in-degree, descending, so each
// degree class is a contiguous row range. A launch is a table of sections, each a block range doing one kind of
// work over a row (or item) range.
enum SectionKind : int {
  kGroupRows = 0,    // `width` lanes per row (1..32), rows split into contiguous per-block ranges
  kWaveRows = 1,     // one wave per row
  kChunkedRows = 2,  // one workgroup per fixed-size edge chunk of a row; a second launch finishes the rows
};

struct RankSection {
  int kind;
  int width;                 // lanes per row for kGroupRows
  uint32_t begin, end;       // rows (kGroupRows, kWaveRows) or items (kChunkedRows)
  uint32_t block_begin;      // first block of the section
  uint32_t per_block;        // rows or items per block
};

constexpr int kMaxSections = 8;

struct RankPlan {
  RankSection sec[kMaxSections];
  int nsec;
  uint32_t blocks;
  // Chunked rows (the heaviest in-degrees, rows 0..n_chunked-1): item i covers edges [item_beg[i], item_end[i]) of
  // row item_row[i] and leaves its sum in partial[i]; row r owns items [row_first[r], row_first[r+1]).
  const uint32_t *item_row, *item_beg, *item_end, *row_first;
  double *partial;
  uint32_t n_chunked;
};

struct RankArgs {
  double *a;          // new ranks
  const double *r;    // old ranks
  double *ca;         // new contributions a[v]/deg[v]
  const double *cr;   // old contributions r[v]/deg[v]
  uint8_t *naff;      // neighbour-affected flags (FRONTIER)
  uint8_t *vaff;      // vertex-affected flags (DYNAMIC)
  const uint32_t *off;  // in-CSR offsets [N+1] (degree-ordered layout)
  const uint32_t *src;  // in-CSR sources [M]
  const uint32_t *deg;  // out-degree [N]
  double C0, P, D, C;  // teleport share (1-P)/N, damping, frontier and prune tolerances
};

struct ExtendArgs {
  uint8_t *vaff;         // flags to set
  const uint8_t *naff;   // vertices whose out-neighbours get marked
  const uint32_t *off;   // out-CSR offsets [N+1]
  const uint32_t *dst;   // out-CSR targets [M]
  Schedule sched;        // out-degree partition
  uint8_t *fresh;        // optional: also set for newly marked vertices (traversal levels)
  int *changed;          // optional: set to 1 when a vertex is newly marked
};

// ---- partition.hip ----
// Stable 3-way partition of [0,N) by degree d = off[v+1]-off[v]: d <= t1, t1 < d <= t2, d > t2.
// tiles: scratch of 3 * partition_tiles(N) uint32; counts_dev: 3 uint32 outputs (device).
uint32_t partition_tiles(uint32_t N);
hipError_t launch_partition(const uint32_t *off, uint32_t N, uint32_t t1, uint32_t t2, uint32_t *order,
                            uint32_t *tiles, uint32_t *counts_dev, hipStream_t s);

// ---- rank.hip ----
hipError_t launch_init_ranks(double *r, double *a, double *cr, double *ca, const uint32_t *deg, uint32_t N,
                             bool from_r, hipStream_t s);
// Rank update of iteration l: the rank kernel, then (if the plan has chunked rows) their combining kernel. `mark`,
// when non-null, is recorded between the two, so that it times the rank kernel alone.
hipError_t launch_rank_update(int variant, const RankArgs &A, const RankPlan &plan, unsigned long long *err,
                              int *done, int l, double tol, hipStream_t s, hipEvent_t mark = nullptr);
hipError_t launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s);
hipError_t launch_fill_u8(uint8_t *p, uint8_t v, size_t n, hipStream_t s);
hipError_t launch_fill_u64(unsigned long long *p, unsigned long long v, size_t n, hipStream_t s);
hipError_t launch_fill_i32(int *p, int v, size_t n, hipStream_t s);
hipError_t launch_copy_f64(double *dst, const double *src, size_t n, hipStream_t s);
hipError_t launch_validate_graph(const uint32_t *in_off, const uint32_t *in_src, const uint32_t *outdeg,
                                 const uint32_t *out_off, const uint32_t *out_dst, uint32_t N, uint64_t M,
                                 int *flags, hipStream_t s);
hipError_t launch_validate_ids(const uint32_t *ids, size_t n, uint32_t N, int *flags, int bit, hipStream_t s);
hipError_t launch_check_finite(const double *x, size_t n, int *flags, int bit, hipStream_t s);
// beg[i] = off[ids[i]], end[i] = off[ids[i] + 1]
hipError_t launch_gather_ranges(const uint32_t *off, const uint32_t *ids, size_t n, uint32_t *beg, uint32_t *end,
                                hipStream_t s);

// ---- ppr.hip ----
constexpr int kPprCols = 64;  // teleport columns per batch (one per lane)

struct PprState {
  int iter;              // index of the iteration the next launch performs
  int done[kPprCols];    // 0 while column j runs, else its number of iterations
};

struct PprArgs {
  double *rank[2];                 // [N][64] ping-pong ranks
  double *contrib[2];              // [N][64] ping-pong rank / out-degree
  const uint32_t *off, *src, *deg;  // in-CSR and out-degrees
  const int32_t *seed_row;          // [N]: row of v in seed_mass, or -1
  const double *seed_mass;          // [rows][64]: s_j[v]
  const uint32_t *light;            // vertices with in-degree <= chunk
  uint32_t n_light;
  const uint32_t *chunk_beg, *chunk_end;  // edge ranges of the heavy vertices' chunks, heavy vertex by vertex
  uint32_t n_chunks;
  const uint32_t *heavy_v, *heavy_first;  // heavy vertices; chunks of heavy_v[h] are [heavy_first[h], heavy_first[h+1])
  uint32_t n_heavy;
  double *partial;                  // [n_chunks][64]
  PprState *state;
  unsigned long long *err;          // [max_iter][64] per-column L-inf change bits
  double P, tol;
  int k;                            // columns in use (1..64)
  int max_iter;
  uint32_t N;
};

hipError_t launch_ppr_reset(const PprArgs &a, hipStream_t s);
hipError_t launch_ppr_init(const PprArgs &a, hipStream_t s);
hipError_t launch_ppr_iteration(const PprArgs &a, hipStream_t s);  // identical launches every iteration
hipError_t launch_ppr_collect(const PprArgs &a, double *out, hipStream_t s);

// ---- weighted.hip ----
struct WeightedArgs {
  const uint32_t *off, *idx;  // pull CSR: in-CSR (forward) or out-CSR (adjoint)
  const float *w;             // FP32 weights aligned with idx
  const double *z;            // gathered values: r/W (forward) or lambda (adjoint)
  double *x_new, *z_new;      // new values and new gathered values
  const double *x_old;        // previous values (for the L-inf change)
  const double *term;         // teleport s (forward, scaled by 1-P in the kernel) or g (adjoint)
  const double *W;            // out-weight sums
  Schedule sched;             // degree schedule of the pull CSR (two-way)
  double P;
};

hipError_t launch_weighted_step(bool adjoint, const WeightedArgs &A, int width, unsigned long long *err, int *done,
                                int l, double tol, hipStream_t s);
hipError_t launch_out_weight_sum(const uint32_t *out_off, const float *w, uint32_t N, double *W, hipStream_t s);
hipError_t launch_edge_map(const uint32_t *in_off, const uint32_t *in_src, const uint32_t *out_off,
                           const uint32_t *out_dst, uint32_t N, uint32_t *in_eid, int *missing, hipStream_t s);
hipError_t launch_permute_weights(const float *w_out, const uint32_t *in_eid, size_t M, float *w_in, hipStream_t s);
hipError_t launch_weighted_init(const double *start, const double *W, uint32_t N, bool adjoint, double *x, double *z,
                                hipStream_t s);
hipError_t launch_weighted_grad(const uint32_t *out_off, const uint32_t *out_dst, const double *W, const double *r,
                                const double *lam, const double *g, double P, uint32_t N, double *grad_w,
                                double *grad_s, hipStream_t s);

// ---- layout.hip ----
hipError_t launch_invert(const uint32_t *perm, uint32_t N, uint32_t *inv, hipStream_t s);
hipError_t launch_relabel_rows(const uint32_t *off, const uint32_t *idx, const uint32_t *perm, const uint32_t *inv,
                               const uint32_t *new_off, uint32_t N, uint32_t *out, hipStream_t s);
// Segmented ascending sort of each row (rocPRIM); call with tmp = nullptr to size tmp_bytes.
hipError_t sort_rows(void *tmp, size_t &tmp_bytes, const uint32_t *in, uint32_t *out, size_t M, uint32_t N,
                     const uint32_t *off, hipStream_t s);
hipError_t launch_gather_u32(const uint32_t *src, const uint32_t *idx, size_t n, uint32_t *dst, hipStream_t s);
hipError_t launch_gather_f64(const double *src, const uint32_t *idx, size_t n, double *dst, hipStream_t s);
hipError_t launch_scatter_f64(const double *src, const uint32_t *idx, size_t n, double *dst, hipStream_t s);
hipError_t launch_scatter_u8(const uint8_t *src, const uint32_t *idx, size_t n, uint8_t *dst, hipStream_t s);
hipError_t launch_map_ids(uint32_t *ids, size_t n, const uint32_t *inv, hipStream_t s);

// ---- frontier.hip ----
hipError_t launch_mark_partial(uint8_t *naff, uint8_t *vaff, const uint32_t *delu, const uint32_t *delv,
                               size_t ND, const uint32_t *insu, size_t NI, hipStream_t s);
// Extension; when err/done are non-null the kernel honours the iteration control of iteration l.
hipError_t launch_extend(const ExtendArgs &A, unsigned long long *err, int *done, int l, double tol,
                         hipStream_t s);

}  // namespace prh
