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
  // Weighted solves through the propagation-blocked plan (no reference counterpart): the forward
  // r_v = (1-P) s_v + P sum_u w_uv r_u / W_u and the adjoint lam_u = g_u + (P / W_u) sum_v w_uv lam_v.
  kWeightedForward = 4,
  kWeightedAdjoint = 5,
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
  // Frontier variants, optional: per workgroup, [2 b] = in-edges of the rows left unaffected (the cost of pulling
  // the next extension) and [2 b + 1] = out-edges of the rows put in the frontier (the cost of pushing it).
  double *stats;
  // Weighted variants: the teleport share s (forward) or the seed g (adjoint), and the out-weight sums W; there a
  // holds the new values, ca the new gathered values (a / W forward, a adjoint), cr the old ones.
  const double *term = nullptr, *W = nullptr;
};

struct ExtendArgs {
  uint8_t *vaff;         // flags to set
  const uint8_t *naff;   // vertices whose out-neighbours get marked
  const uint32_t *off;   // out-CSR offsets [N+1]
  const uint32_t *dst;   // out-CSR targets [M]
  Schedule sched;        // out-degree partition
  uint8_t *fresh;        // optional: also set for newly marked vertices (traversal levels)
  int *changed;          // optional: set to 1 when a vertex is newly marked
  const int *direction;  // optional: the push runs only while *direction == 0 (see launch_direction)
  // optional: the huge class (sched.n_huge rows after the high ones) as chunks of at most kExtendChunk edges,
  // chunk k covering edges [chunk_beg[k], chunk_end[k]) of row chunk_row[k]; without it huge rows take a wave each
  const uint32_t *chunk_beg = nullptr, *chunk_end = nullptr, *chunk_row = nullptr;
  uint32_t n_chunks = 0;
};
constexpr uint32_t kExtendChunk = 4096;  // out-degree above which a row's extension is cut into chunks

// The pull form of the extension: every unaffected row w gets vaff[w] = 1 if one of its in-neighbours is in the
// frontier. Same flags as the push. Rows [0, n_wave) take a wave each, [n_wave, n_group) 8 lanes, the rest 1 lane.
struct PullArgs {
  uint8_t *vaff;
  const uint8_t *naff;
  const uint32_t *off, *src;  // in-CSR (degree-ordered layout)
  uint32_t n_wave, n_group, N;
  const int *direction;       // runs only while *direction == 1
};

// Source-sliced plan of the plain update (Static / Naive-dynamic). Source u belongs to slice
// (u >> 4) & (slices - 1): 16 FP64 contributions = one 128-byte line, so every line of the contribution array
// belongs to exactly one slice. Workgroup b of the gather launch works on slice b % slices; with 8 slices that is
// the XCD the workgroup is dispatched to (round-robin), so each XCD's L2 caches only its own eighth of the lines.
//   gather launch: per slice, one partial sum per (row, slice) item over the row's sources in that slice
//                  (chunk items for rows above the chunk threshold);
//   finish launch: per row, the sum of its (at most 8) partials in slice order, then the update and the L-inf.
constexpr int kMaxSlices = 8;
constexpr uint32_t kSliceLineShift = 4;

struct SlicedPlan {
  RankSection sec[kMaxSlices][kMaxSections];  // per slice: chunk section, then item bands (kWaveRows/kGroupRows)
  int nsec[kMaxSlices];
  int slices;                       // 1, 2, 4 or 8
  uint32_t blocks_per_slice;        // gather grid = slices * blocks_per_slice
  const uint32_t *src;              // [M] sources, slice-major (chunked rows first, then slice 0 items, slice 1 ...)
  const uint32_t *item_off;         // [n_items + 1]: item k covers src[item_off[k] .. item_off[k+1])
  double *partial;                  // [n_items]
  const uint32_t *chunk_list;       // chunk ids grouped by slice (indexed by kChunkedRows sections)
  const uint32_t *chunk_beg, *chunk_end;  // [n_chunks] by chunk id: row-major (row, slice, part)
  double *chunk_partial;            // [n_chunks]
  const uint32_t *row_chunk_first;  // [n_huge + 1]
  uint32_t n_huge;                  // rows 0..n_huge-1 are summed from chunks
  const uint8_t *mask;              // [N]: bit x = row has an item in slice x (0 for chunked rows)
  const uint32_t *wbase;            // [ceil((N - n_huge) / 64) * slices]: item id of each 64-row group's first item
  uint32_t N;
};

// Edge and item bases of the slices in the slice-major arrays.
struct SliceBases {
  uint32_t edge[kMaxSlices];  // first edge of slice x's items in SlicedPlan::src
  uint32_t item[kMaxSlices];  // first item id of slice x
};

// Propagation-blocked plan of the rank update (every variant; binned.hip). Layout rows are cut into runs
// of B vertices: source blocks and destination bins, both numbered s, b = vertex / B. Slots (one per edge) are
// ordered bin-major (bin, source block, destination, source); X holds the slots' contributions block-major (block,
// bin, same order within a tile), so each nonempty tile is one run, consecutive in both orders.
//   expand item     (s, k0, k1, 0): X entries k0..k1-1, all leaving block s;
//   accumulate item (b, j0, j1, row): slots j0..j1-1, all entering bin b; row = kWholeBin when the item is the
//                   bin's only one (it finishes the bin), else its row of B partial sums;
//   finish unit     (b, first row, rows, first vertex | mode << 31): part of a bin with several items, finished
//                   by finish_bins_kernel (mode 0: 1024 vertices, a thread each; 1: 64 vertices, 16 waves each).
constexpr int kBinBlock = 1024;
constexpr uint32_t kWholeBin = 0xFFFFFFFFu;

struct BinnedPlan {
  int B;                   // vertices per block and bin: 1024, 2048, 4096, 8192 or 16384
  uint32_t N;
  const uint16_t *src16;   // [Q] X order: source - s * B
  const uint16_t *dst16;   // [M padded to 64] slot order: destination - b * B, | 0x8000 pair start, | 0x4000 run start
  double *X;               // [Q] one entry per (source, destination bin) pair, block-major
  const uint32_t *run_off; // [n_runs] per run (a tile's pairs, runs in slot order): X entry - pair number, mod 2^32
  const uint4 *grp;        // [ceil(M / 64)] of slot 64 i: run r, run_off[r], run_off[r + 1], pair number
  uint32_t n_runs;
  const uint4 *ex_item;    // [n_ex]
  const uint4 *ac_item;    // [n_ac]
  const uint4 *multi;      // [n_multi] finish units
  double *partial;         // [partial rows][B]
  uint32_t n_ex, n_ac, n_multi;
  uint64_t Q;              // pairs
  const uint32_t *slot;    // [M] CSR slot of slot j (the in-CSR slot at build time), else null
  const float *ws;         // [M] weighted solves: the edge weight of slot j, applied in accumulate; else null
  // Overlay of an in-place updated graph (prh_graph_update, the streaming session): edges inserted since the build
  // that are not slots of the plan, keyed (destination << 32 | source) in layout ids, sorted; the entries of bin b
  // are [ov_off[b], ov_off[b + 1]), ov_off[bins] of them in all; ovX[i] holds entry i's value for the current
  // iteration. ov_n: 0 = no overlay, else its capacity (the kernels' grid). Deleted edges that are slots of the plan
  // carry the dead flag in dst16 instead (blocks of at most 8192 vertices).
  uint32_t ov_n;
  const unsigned long long *ov_key;
  const uint32_t *ov_off;  // [bins + 1]
  double *ovX;
};

// Banded plan (window.hip): rows in the caller's vertex order, workgroup b sweeping rows [b N / wgs, (b + 1) N / wgs)
// in chunks of kBinBlock rows through a ring of S contributions in LDS that holds the window [c0 - H, c0 + R + H) of
// chunk c0 (S >= kBinBlock + 2H); sources outside it are gathered from the ranks.
constexpr uint32_t kWindowRingSmall = 10112;  // 79 KiB of LDS: two workgroups per CU, H <= 4544
constexpr uint32_t kWindowRingLarge = 20224;  // 158 KiB: one workgroup per CU, H <= 9600
struct WindowPlan {
  uint32_t N;
  uint32_t H;    // rows the window reaches on each side of a chunk
  uint32_t S;    // ring entries: kWindowRingSmall or kWindowRingLarge
  uint32_t wgs;  // workgroups
};

// ---- partition.hip ----
// Stable 3-way partition of [0,N) by degree d = off[v+1]-off[v]: d <= t1, t1 < d <= t2, d > t2.
// tiles: scratch of 3 * partition_tiles(N) uint32; counts_dev: 3 uint32 outputs (device).
constexpr uint32_t kPartitionTile = 2048;  // vertices per tile: partition_tiles(N) = ceil(N / kPartitionTile)
uint32_t partition_tiles(uint32_t N);
hipError_t launch_partition(const uint32_t *off, uint32_t N, uint32_t t1, uint32_t t2, uint32_t *order,
                            uint32_t *tiles, uint32_t *counts_dev, hipStream_t s);

// ---- rank.hip ----
hipError_t launch_init_ranks(double *r, double *a, double *cr, double *ca, const uint32_t *deg, uint32_t N,
                             bool from_r, hipStream_t s);
// Workgroups of the chunked rows' combining launch.
inline uint32_t finish_chunked_grid(uint32_t n_chunked) {
  const uint32_t g = (n_chunked + kWavesPerBlock - 1) / kWavesPerBlock;
  return g > 4096 ? 4096 : g;
}
// Rank update of iteration l: the rank kernel, then (if the plan has chunked rows) their combining kernel. `mark`,
// when non-null, is recorded between the two, so that it times the rank kernel alone.
hipError_t launch_rank_update(int variant, const RankArgs &A, const RankPlan &plan, unsigned long long *err,
                              int *done, int l, double tol, hipStream_t s, hipEvent_t mark = nullptr);
// Plain update through the source-sliced plan: the gather launch, then the finish launch; `mark` (optional) is
// recorded between them.
hipError_t launch_sliced_update(const RankArgs &A, const SlicedPlan &plan, unsigned long long *err, int *done, int l,
                                double tol, hipStream_t s, hipEvent_t mark = nullptr);
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

// ---- binned.hip ----
// One iteration through the propagation-blocked plan: expand, accumulate, and (with split bins) finish. `extend`
// (frontier variants, every iteration after the first): the extension of the previous iteration's frontier
// (A.naff) happens inside this update, so no separate extension launch follows it. `mark`, when non-null, is
// recorded after the last kernel.
hipError_t launch_binned_update(int variant, const RankArgs &A, const BinnedPlan &P, unsigned long long *err,
                                int *done, int l, double tol, bool extend, hipStream_t s, hipEvent_t mark = nullptr);
// Rank update of iteration l through the banded plan (window.hip); unweighted variants only.
hipError_t launch_window_update(int variant, const RankArgs &A, const WindowPlan &W, unsigned long long *err, int *done,
                                int l, double tol, hipStream_t s);
// *out = the largest |source - destination| over the in-CSR's edges.
hipError_t launch_max_spread(const uint32_t *off, const uint32_t *src, uint32_t N, uint32_t *out, hipStream_t s);
// In-place update of a plan (see binned.hip): dead flags and fresh overlay entries from a batch's tidied in-CSR keys
// (the current overlay: *ov_count sorted keys), the next overlay's candidates (ov_cap + fresh_cap of them, the live
// count added to *kept), and the overlay's per-bin offsets over the first min(*count, cap) sorted keys (a count above
// cap sets bit in *flags).
hipError_t launch_patch_slots(const unsigned long long *keys, const uint32_t *count, uint32_t cap, bool dead,
                              const uint32_t *base_off, const uint32_t *base_src, const uint32_t *slot_of,
                              uint16_t *dst16, const unsigned long long *ov_key, const uint32_t *ov_count,
                              uint8_t *ov_gone, unsigned long long *fresh, hipStream_t s);
hipError_t launch_overlay_candidates(const unsigned long long *ov_key, const uint32_t *ov_count, uint32_t ov_cap,
                                     const uint8_t *ov_gone, const unsigned long long *fresh, uint32_t fresh_cap,
                                     unsigned long long *cand, uint32_t *kept, hipStream_t s);
hipError_t launch_overlay_bounds(const unsigned long long *keys, const uint32_t *count, uint32_t cap, uint32_t B,
                                 uint32_t nb, uint32_t *off, int *flags, int bit, hipStream_t s);
hipError_t launch_copy_u64(unsigned long long *dst, const unsigned long long *src, size_t n, hipStream_t s);
// ws[j] = w[eid[slot[j]]]: edge weights (raw out-CSR order) carried to slot order through the layout edge map.
hipError_t launch_binned_weights(const float *w, const uint32_t *eid, const uint32_t *slot, size_t M, float *ws,
                                 hipStream_t s);
// Build steps. key[e] = bin(row) * nblocks + block(src[e]), row[e] = row of in-CSR slot e.
hipError_t launch_bin_keys(const uint32_t *off, const uint32_t *src, uint32_t N, uint32_t B, uint32_t nblocks,
                           uint32_t *key, uint32_t *row, hipStream_t s);
// For slot j (in-CSR slot perm2[j], slots sorted by tile key key2s, then source): dst16[j] with its flags, pf[j] =
// it starts a pair.
hipError_t launch_pair_slots(const uint32_t *perm2, const uint32_t *src, const uint32_t *row, const uint32_t *key2s,
                             size_t M, uint32_t B, uint16_t *dst16, uint32_t *pf, hipStream_t s);
// For the pair q = pid[j] of each pair-starting slot j: psrc[q], pkey2[q] (tile key), key1[q] (source block).
hipError_t launch_pair_fill(const uint32_t *pf, const uint32_t *pid, const uint32_t *perm2, const uint32_t *src,
                            const uint32_t *key2s, size_t M, uint32_t B, uint32_t *psrc, uint32_t *pkey2,
                            uint32_t *key1, hipStream_t s);
// flag[j] = entry j starts a run (its sorted tile key differs from entry j-1's).
hipError_t launch_run_flags(const uint32_t *key2s, size_t M, uint32_t *flag, hipStream_t s);
// For X entry k (pair perm1[k]): src16[k], inv[perm1[k]] = k.
hipError_t launch_pair_expand(const uint32_t *perm1, const uint32_t *psrc, size_t Q, uint32_t B, uint16_t *src16,
                              uint32_t *inv, hipStream_t s);
// run_off from the pairs' run flags rf, their exclusive scan rid and the pairs' X entries inv.
hipError_t launch_pair_runs(const uint32_t *rf, const uint32_t *rid, const uint32_t *inv, size_t Q,
                            uint32_t *run_off, hipStream_t s);
// grp[i] = (run r, run_off[r], run_off[r + 1], pair) of slot 64 i.
hipError_t launch_group_table(const uint32_t *pf, const uint32_t *pid, const uint32_t *rf, const uint32_t *rid,
                              const uint32_t *run_off, uint32_t n_runs, size_t M, uint4 *grp, hipStream_t s);
// start[s] = lower bound of s in the sorted key[0..M), s = 0..n.
hipError_t launch_lower_bounds(const uint32_t *key, size_t M, uint32_t n, uint32_t *start, hipStream_t s);
// Stable radix sort of (key, value) on key bits [0, end_bit); val_in null means values 0, 1, 2, ...
hipError_t sort_pairs_u32(void *tmp, size_t &tmp_bytes, const uint32_t *key_in, uint32_t *key_out,
                          const uint32_t *val_in, uint32_t *val_out, size_t n, int end_bit, hipStream_t s);

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
constexpr int kWeightedBands = 6;  // lanes per low row: 32, 16, 8, 4, 2, 1

struct WeightedArgs {
  const uint32_t *off, *idx;  // pull CSR: in-CSR (forward) or out-CSR (adjoint)
  const float *w;             // FP32 weights aligned with idx
  const double *z;            // gathered values: r/W (forward) or lambda (adjoint)
  double *x_new, *z_new;      // new values and new gathered values
  const double *x_old;        // previous values (for the L-inf change)
  const double *term;         // teleport s (forward, scaled by 1-P in the kernel) or g (adjoint)
  const double *W;            // out-weight sums
  Schedule sched;             // degree schedule of the pull CSR: [<= 64 | <= kWeightedChunk | heavier]
  double P;
  // Rows of the third class are split into chunks of at most kWeightedChunk edges: chunk i covers edges
  // [chunk_beg[i], chunk_end[i]) and leaves its sum in partial[i]; heavy row h (order[n_low + n_high + h]) owns
  // chunks [heavy_first[h], heavy_first[h + 1]).
  const uint32_t *chunk_beg, *chunk_end, *heavy_first;
  double *partial;
  uint32_t n_chunks;
  // The low rows (order[0, n_low)) sorted by degree, descending, in bands of equal lanes per row: band b covers
  // order[band[b].begin, band[b].end) with band[b].width lanes per row, on blocks from band[b].block_begin,
  // band[b].per_block rows each (kind unused).
  RankSection band[kWeightedBands];
  int n_bands;
};

constexpr uint32_t kWeightedChunk = 4096;

// One Jacobi step: the chunk sums, the wave rows, the low-row bands, then (with chunks) the heavy rows' finish.
hipError_t launch_weighted_step(bool adjoint, const WeightedArgs &A, unsigned long long *err, int *done, int l,
                                double tol, hipStream_t s);
// key[i] = 64 - degree of row order[i], for the first n rows (degrees <= 64).
hipError_t launch_degree_keys(const uint32_t *off, const uint32_t *order, uint32_t n, uint32_t *key, hipStream_t s);
hipError_t launch_out_weight_sum(const uint32_t *out_off, const float *w, uint32_t N, double *W, hipStream_t s);
// eid[k] = raw out-CSR slot of slot k of a degree-ordered layout CSR (transpose: its in-CSR, else its out-CSR).
hipError_t launch_layout_edge_map(const uint32_t *l_off, const uint32_t *l_idx, const uint32_t *perm,
                                  const uint32_t *out_off, const uint32_t *out_dst, uint32_t N, bool transpose,
                                  uint32_t *eid, int *missing, hipStream_t s);
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
hipError_t launch_iota_u32(uint32_t *p, uint32_t n, hipStream_t s);  // p[i] = i
hipError_t launch_gather_f64(const double *src, const uint32_t *idx, size_t n, double *dst, hipStream_t s);
hipError_t launch_scatter_f64(const double *src, const uint32_t *idx, size_t n, double *dst, hipStream_t s);
hipError_t launch_scatter_u8(const uint8_t *src, const uint32_t *idx, size_t n, uint8_t *dst, hipStream_t s);
hipError_t launch_map_ids(uint32_t *ids, size_t n, const uint32_t *inv, hipStream_t s);
// Sliced layout. cnt: [slices][N] edges of row v in slice x.
hipError_t launch_slice_count(const uint32_t *off, const uint32_t *src, uint32_t N, int slices, uint32_t *cnt,
                              hipStream_t s);
// For slice x: has[v] = (v >= n_huge && cnt[x][v] > 0), ecnt[v] = (v >= n_huge ? cnt[x][v] : 0), for v < N, and
// has[N] = ecnt[N] = 0 (so that exclusive scans of N + 1 values end with the totals).
hipError_t launch_slice_flags(const uint32_t *cnt_x, uint32_t N, uint32_t n_huge, uint32_t *has, uint32_t *ecnt,
                              hipStream_t s);
// Exclusive prefix sum of n uint32 values (rocPRIM); call with tmp = nullptr to size tmp_bytes.
hipError_t scan_u32(void *tmp, size_t &tmp_bytes, const uint32_t *in, uint32_t *out, size_t n, hipStream_t s);
// Writes the slice-major sources, the item offsets of every (row, slice) item and the row masks.
// item / eoff: [slices][N + 1] exclusive scans of has / ecnt.
hipError_t launch_slice_scatter(const uint32_t *off, const uint32_t *src, uint32_t N, uint32_t n_huge, int slices,
                                const uint32_t *cnt, const uint32_t *item, const uint32_t *eoff, const SliceBases &b,
                                uint32_t *s_src, uint32_t *item_off, uint8_t *mask, hipStream_t s);
hipError_t launch_slice_wbase(const uint32_t *item, uint32_t N, uint32_t n_huge, int slices, const SliceBases &b,
                              uint32_t *wbase, hipStream_t s);

// ---- stream.hip (device-side batch application; see the file's header) ----
hipError_t launch_batch_keys(const uint32_t *bu, const uint32_t *bv, const uint32_t *count, uint32_t cap,
                             const uint32_t *inv, uint32_t N, unsigned long long *in_keys,
                             unsigned long long *out_keys, int *flags, int bad_bit, hipStream_t s);
hipError_t sort_keys_u64(void *tmp, size_t &tmp_bytes, const unsigned long long *in, unsigned long long *out,
                         uint32_t n, hipStream_t s);
// out = the merge of ascending a[0, na) and b[0, nb) (rocPRIM)
hipError_t merge_keys_u64(void *tmp, size_t &tmp_bytes, const unsigned long long *a, uint32_t na,
                          const unsigned long long *b, uint32_t nb, unsigned long long *out, hipStream_t s);
// valid[j] = 1 when a[j] is not in b (both ascending)
hipError_t launch_absent_from(const unsigned long long *a, uint32_t na, const unsigned long long *b, uint32_t nb,
                              uint32_t *valid, hipStream_t s);
// in-CSR keys (target << 32 | source) of the first *count out-CSR keys, kNoKey for the rest of the cap entries
hipError_t launch_transpose_keys(const unsigned long long *in, const uint32_t *count, uint32_t cap,
                                unsigned long long *out, hipStream_t s);
hipError_t launch_batch_valid(const unsigned long long *keys, uint32_t cap, const uint32_t *off, const uint32_t *idx,
                              bool want_exists, uint32_t *valid, hipStream_t s);
hipError_t launch_batch_compact(const unsigned long long *keys, const uint32_t *valid, const uint32_t *pos,
                                uint32_t cap, unsigned long long *out, uint32_t *count, hipStream_t s);
hipError_t launch_row_ranges(const unsigned long long *dkeys, const uint32_t *nd, const unsigned long long *ikeys,
                             const uint32_t *ni, uint32_t cap_total, int *delta, uint8_t *changed, uint4 *ranges,
                             hipStream_t s);
hipError_t launch_new_degrees(const uint32_t *off, const int *delta, uint32_t N, uint32_t *deg, int *flags,
                              int empty_bit, hipStream_t s);
// units: 4 * max_units + 1 words (max_units >= changed rows + 1); pos: N words, the exclusive scan of changed
// (scan_u8); max_edges: the old edge count or a bound on it (sizes the grid); new_cap: entries of new_idx (no store
// goes past it).
hipError_t scan_u8(void *tmp, size_t &tmp_bytes, const uint8_t *in, uint32_t *out, size_t n, hipStream_t s);
hipError_t launch_rewrite(const uint32_t *old_off, const uint32_t *old_idx, const uint32_t *new_off,
                          const uint8_t *changed, const uint32_t *pos, uint32_t *units, const uint4 *ranges,
                          const unsigned long long *dkeys, const unsigned long long *ikeys, uint32_t N,
                          uint32_t max_units, uint64_t max_edges, uint64_t new_cap, uint32_t *new_idx,
                          hipStream_t s);
hipError_t launch_insert(const uint32_t *old_off, const uint32_t *old_idx, const uint32_t *new_off,
                         const uint4 *ranges, const unsigned long long *dkeys, const unsigned long long *ikeys,
                         const uint32_t *ni, uint32_t cap, uint32_t *new_idx, hipStream_t s);
hipError_t launch_degrees(const uint32_t *off, uint32_t N, uint32_t *deg, hipStream_t s);
hipError_t launch_refresh_chunks(const uint32_t *item_row, const uint32_t *row_first, uint32_t n_items,
                                 const uint32_t *off, uint32_t chunk, uint32_t *item_beg, uint32_t *item_end,
                                 hipStream_t s);
hipError_t launch_mark_batch(uint8_t *naff, uint8_t *vaff, const unsigned long long *dkeys, const uint32_t *nd,
                             const unsigned long long *ikeys, const uint32_t *ni, uint32_t cap_total, hipStream_t s);
hipError_t launch_advance(int *l, int *done, const unsigned long long *err, double tol, int max_iter, hipStream_t s);
// Device-driven traversal closure (lvl: [0] level, [1] stop, [2], [3] changed flags of even and odd levels): the
// step before a level stops the closure when the previous level marked nothing, else clears this level's flag; the
// advance after it moves to the next level unless stopped.
hipError_t launch_closure_step(int *lvl, int parity, hipStream_t s);
hipError_t launch_closure_advance(int *lvl, hipStream_t s);
hipError_t launch_copy_u8(uint8_t *dst, const uint8_t *src, size_t n, hipStream_t s);

// ---- frontier.hip ----
hipError_t launch_mark_partial(uint8_t *naff, uint8_t *vaff, const uint32_t *delu, const uint32_t *delv,
                               size_t ND, const uint32_t *insu, size_t NI, hipStream_t s);
// Extension; when err/done are non-null the kernel honours the iteration control of iteration l.
// Initial extension from the batch list: vaff[w] = 1 for every out-neighbour w of every batch source (deletion
// sources delu, insertion sources insu), each source once: naff (1 for the sources, from mark_partial; its
// allocation a multiple of 4 bytes) is set to 2 for a source when an entry claims it. cnt and start hold ND + NI + 1
// entries. Call first with tmp = null for the scan's temporary size.
hipError_t launch_extend_batch(uint8_t *vaff, uint8_t *naff, const uint32_t *delu, size_t ND, const uint32_t *insu,
                               size_t NI, const uint32_t *off, const uint32_t *dst, uint32_t *cnt, uint32_t *start,
                               void *tmp, size_t &tmp_bytes, hipStream_t s);
// The same from a tidied batch of out-CSR keys (source << 32 | target) with device counts: deletions dk[0, *nd),
// insertions ik[0, *ni), at most cap each; cnt and start hold 2 cap + 1 entries (the streaming session's marking).
hipError_t launch_extend_batch_keys(uint8_t *vaff, uint8_t *naff, const unsigned long long *dk, const uint32_t *nd,
                                    const unsigned long long *ik, const uint32_t *ni, uint32_t cap,
                                    const uint32_t *off, const uint32_t *dst, uint32_t *cnt, uint32_t *start,
                                    void *tmp, size_t &tmp_bytes, hipStream_t s);
hipError_t launch_extend(const ExtendArgs &A, unsigned long long *err, int *done, int l, double tol,
                         hipStream_t s);
// Sums the per-workgroup costs the rank launches left in stats[0 .. 2 n) and sets *direction to 1 (pull) when
// pulling is cheaper, else 0 (push). Honours iteration l's control like the extension.
hipError_t launch_direction(const double *stats, uint32_t n, int *direction, unsigned long long *err, int *done, int l,
                            double tol, hipStream_t s);
hipError_t launch_extend_pull(const PullArgs &A, unsigned long long *err, int *done, int l, double tol,
                              hipStream_t s);

}  // namespace prh
