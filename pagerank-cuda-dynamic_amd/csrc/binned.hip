// binned.hip -- the propagation-blocked rank update (every solver variant) and the build of its layout.
//
// The pull kernel gathers cr[u] = r[u]/deg[u] at random; every gather from a source outside the L2-resident hot
// set costs a 128-byte line fill for 8 useful bytes. Propagation blocking (Beamer, Asanovic and Patterson,
// "Reducing PageRank Communication via Propagation Blocking", IPDPS 2017) moves that randomness into LDS. Vertices
// are cut into runs of B consecutive layout rows: source blocks on one side, destination bins on the other. Edges
// from one source into one bin share a single entry of X, the pair's FP64 contribution (on a power-law graph a
// source's edges crowd into the bins of the high in-degree vertices: on RMAT-24 with B = 8192 there are 0.38
// pairs per edge). Each iteration makes two streaming passes through X:
//   expand      one workgroup per item of a source block: the block's B contributions r/deg are computed into LDS
//               from the ranks and out-degrees, and every pair leaving the block copies its source's contribution
//               into its entry of X (the plan keeps no contribution array: 8 bytes per vertex and iteration less
//               than writing one in finish and reading it back);
//   accumulate  one workgroup per item of a destination bin: the bin's slots (one per edge) are read in order, each
//               takes its pair's entry of X and adds it into one of B LDS sums (ds_add_f64); a bin covered by one
//               item is finished right there, a split bin leaves one row of B partial sums per item;
//   finish      split bins only: a vertex's partials are added in item order, then the update.
// Slot order is bin-major: bin b, then source block s (the (s, b) tile), then source, then destination. A tile's
// pairs are one run of consecutive X entries; X is stored block-major (block s, then bin b, then source), so that
// expand writes it sequentially (pairs of entries, 16 bytes a lane) with a 16-bit local source id per entry.
// Accumulate finds a slot's entry without a per-slot index: each slot's 16-bit local destination id carries two
// flags, "starts a pair" and "starts a run", and a table gives, per 64 slots, the pair and run of the first slot
// (with the X offsets of that run and the next); a wave counts the flags with ballots and mbcnt, and the run
// offset turns a pair number into its X entry.
// Per edge and iteration: accumulate reads the 2-byte slot; per pair, expand reads 2 bytes and writes 8 and
// accumulate reads the 8 back (repeated loads of one entry by neighbouring lanes share a cache line).
// LDS additions are made in arrival order, so the sums repeat only to rounding, not bit for bit. Every vertex gets
// its full in-sum; the update from it is the other plans' (update.hpp), so the dynamic variants carry unaffected
// vertices over and set their prune and frontier flags as the row kernel does.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "prh_device.hpp"
#include "prh_internal.hpp"
#include "update.hpp"

namespace prh {
namespace {

constexpr int kBinWaves = kBinBlock / kWave;
typedef double d2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t ld_nt(const uint32_t *p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ uint32_t ld_nt(const uint16_t *p) { return __builtin_nontemporal_load(p); }

// Vertex v of a bin from its in-sum: updated (update.hpp), or carried over when a dynamic variant left it
// unaffected; `extend` (frontier variants: an in-neighbour was in the last frontier) marks it affected first. Its
// extension costs are added to `cost` (frontier variants with A.stats). Returns the L-inf term.
template <int VARIANT>
__device__ __forceinline__ double bin_finish(const RankArgs &A, uint32_t v, double cv, CostAcc &cost, bool extend) {
  if constexpr (VARIANT == kWeightedForward) {  // r = (1-P) s + P sum, gathered value r / W
    const double av = fma(A.P, cv, (1 - A.P) * A.term[v]);
    A.a[v] = av;
    A.ca[v] = av / A.W[v];
    return fabs(av - A.r[v]);
  } else if constexpr (VARIANT == kWeightedAdjoint) {  // lam = g + (P / W) sum, gathered value lam
    const double av = fma(A.P / A.W[v], cv, A.term[v]);
    A.a[v] = av;
    A.ca[v] = av;
    return fabs(av - A.r[v]);
  }
  if constexpr (Traits<VARIANT>::frontier)
    if (extend) A.vaff[v] = 1;
  // The unweighted variants keep no contributions here (expand computes r/deg): only DF-P's closed form needs the
  // out-degree, and its own old contribution r/d.
  const double rv = A.r[v];
  RowFlags f;
  double e = 0.0;
  if (affected<VARIANT>(A, v)) {
    const uint32_t d = Traits<VARIANT>::prune ? A.deg[v] : 1u;
    e = finish_vertex<VARIANT, false>(A, v, cv, d, rv, Traits<VARIANT>::prune ? rv / d : 0.0, f);
  } else {
    copy_vertex<VARIANT, false>(A, v, rv, 0.0);
    f.affected = false;
  }
  if constexpr (Traits<VARIANT>::frontier)
    if (A.stats) cost.add(f, A.off[v + 1] - A.off[v], A.deg[v]);
  return e;
}

// naff (frontier variants after their first iteration): the previous iteration's frontier flags, carried as the
// sign of the contribution (contributions are positive), so that accumulate can extend the frontier.
// Contribution of vertex v: r[v]/deg[v] (FROM_RANKS: the unweighted variants, whose finish keeps no
// contributions) or the kept value cr[v] (the weighted variants: a/W forward, a adjoint).
template <bool FROM_RANKS>
__device__ __forceinline__ d2 contrib2(const double *cr, const double *r, const uint32_t *deg, uint32_t v) {
  if constexpr (FROM_RANKS) {
    const d2 x = *reinterpret_cast<const d2 *>(r + v);
    const uint2 d = *reinterpret_cast<const uint2 *>(deg + v);
    return d2{x.x / d.x, x.y / d.y};
  } else {
    return *reinterpret_cast<const d2 *>(cr + v);
  }
}

template <bool FROM_RANKS>
__device__ __forceinline__ double contrib1(const double *cr, const double *r, const uint32_t *deg, uint32_t v) {
  if constexpr (FROM_RANKS) return r[v] / deg[v];
  else return cr[v];
}

template <int B, bool FROM_RANKS>
__global__ __launch_bounds__(kBinBlock) void expand_kernel(BinnedPlan P, const double *cr, const double *r,
                                                           const uint32_t *deg, const uint8_t *naff, LoopCtl ctl) {
  if (!iteration_begin(ctl)) return;
  if (iter_index(ctl) == 0) naff = nullptr;  // the first iteration extends nothing (the marking did)
  __shared__ double c[B];
  const uint4 it = P.ex_item[blockIdx.x];  // block, first entry, end entry
  const uint32_t base = it.x * uint32_t(B);
  const uint32_t n = min(uint32_t(B), P.N - base);
  // Pairs of contributions a lane (base is a multiple of B, so the pairs are 16-byte aligned); an odd n leaves one.
  if (naff) {
    const uint16_t *f2 = reinterpret_cast<const uint16_t *>(naff + base);
    for (uint32_t i = threadIdx.x; 2 * i + 1 < n; i += kBinBlock) {
      const d2 v = contrib2<FROM_RANKS>(cr, r, deg, base + 2 * i);
      const uint32_t f = f2[i];
      c[2 * i] = (f & 0xFFu) ? -v.x : v.x;
      c[2 * i + 1] = (f >> 8) ? -v.y : v.y;
    }
  } else {
    for (uint32_t i = threadIdx.x; 2 * i + 1 < n; i += kBinBlock) {
      const d2 v = contrib2<FROM_RANKS>(cr, r, deg, base + 2 * i);
      c[2 * i] = v.x;
      c[2 * i + 1] = v.y;
    }
  }
  if ((n & 1u) && threadIdx.x == 0) {
    const double x = contrib1<FROM_RANKS>(cr, r, deg, base + n - 1);
    c[n - 1] = naff && naff[base + n - 1] ? -x : x;
  }
  __syncthreads();
  // Pairs of entries: thread t takes pairs p = t + m * kBinBlock of the even-aligned range (16-byte stores, with the
  // non-temporal hint: X is read back only by the next kernel, and keeping it out of L2 helps both); a lone first
  // or last entry is written on its own.
  uint32_t k0 = it.y, k1 = it.z;
  if (k0 & 1u) {
    if (threadIdx.x == 0) P.X[k0] = c[ld_nt(P.src16 + k0)];
    ++k0;
  }
  if (k1 > k0 && (k1 & 1u)) {
    if (threadIdx.x == 0) P.X[k1 - 1] = c[ld_nt(P.src16 + k1 - 1)];
    --k1;
  }
  const uint32_t *s32 = reinterpret_cast<const uint32_t *>(P.src16);
  d2 *x2 = reinterpret_cast<d2 *>(P.X);
  const uint32_t p0 = k0 / 2, p1 = k1 / 2;
  uint32_t p = p0 + threadIdx.x;
  // Eight source-id loads in flight per lane before their stores (four: 4351 against 4229 us per update at 2.04B
  // edges).
  for (; p + 7 * kBinBlock < p1; p += 8 * kBinBlock) {
    uint32_t a[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) a[u] = __builtin_nontemporal_load(s32 + p + u * kBinBlock);
#pragma unroll
    for (int u = 0; u < 8; ++u)
      __builtin_nontemporal_store(d2{c[a[u] & 0xFFFFu], c[a[u] >> 16]}, x2 + p + u * kBinBlock);
  }
  for (; p < p1; p += kBinBlock) {
    const uint32_t a = __builtin_nontemporal_load(s32 + p);
    __builtin_nontemporal_store(d2{c[a & 0xFFFFu], c[a >> 16]}, x2 + p);
  }
}

constexpr uint32_t kPairStart = 0x8000u, kRunStart = 0x4000u;
// A slot whose edge an in-place update deleted (bins of at most 8192 vertices, whose local ids take 13 bits): it
// keeps its place, and its pair and run flags, but adds nothing.
constexpr uint32_t kDeadSlot = 0x2000u;
// Groups of 64 slots a wave of accumulate has in flight: 3 for Static / Naive-dynamic (the update at 2.04B edges:
// 4189 us; with 2 groups 4231 us, and 4 were slower than 2), 2 for the others (DF-P's accumulate took 4400 us per
// iteration at 2.04B edges with 3 groups, 3187 us with 2).
template <int VARIANT>
constexpr int acc_groups() {
  return VARIANT == kPlain ? 3 : 2;
}

// X entry of the pair of slot g + lane, for a group of 64 slots starting at a multiple of 64, from t (the group
// table: run r0 of slot g, run_off of runs r0 and r0 + 1, pair q0 of slot g) and d (the slot's 16-bit local
// destination id with its flags). Pairs are numbered in slot order; a run's pairs are consecutive in both orders,
// so pair q of run r sits at X entry q + run_off[r]. Lane L's pair is q0 plus the pair starts among lanes 1..L, its
// run r0 plus the run starts among them (mbcnt of the ballots shifted down by one lane); only lanes past a second
// run start in the group load run_off. All lanes must be active (slots past M read as 0: no flags).
__device__ __forceinline__ uint32_t pair_entry(const BinnedPlan &P, uint4 t, uint32_t d) {
  const unsigned long long pairs = __ballot((d & kPairStart) != 0) >> 1;
  const unsigned long long runs = __ballot((d & kRunStart) != 0) >> 1;
  const uint32_t r0 = __builtin_amdgcn_readfirstlane(t.x), off0 = __builtin_amdgcn_readfirstlane(t.y);
  const uint32_t off1 = __builtin_amdgcn_readfirstlane(t.z), q0 = __builtin_amdgcn_readfirstlane(t.w);
  const uint32_t q = __builtin_amdgcn_mbcnt_hi(uint32_t(pairs >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(pairs), q0));
  const uint32_t c = __builtin_amdgcn_mbcnt_hi(uint32_t(runs >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(runs), 0u));
  uint32_t off = c == 0 ? off0 : off1;
  if (c > 1) off = P.run_off[r0 + c];
  return q + off;
}

// Adds one wave's 64 slots into the LDS sums. All lanes must be active; lanes outside the item carry ok = false.
// A wave whose slots all go to one destination (a hub bin: many sources, one destination) is summed first and
// added once. EXT (frontier variants): a negative value comes from a frontier source; its destination's flag in ext
// is set and its magnitude added.
template <bool EXT>
__device__ __forceinline__ void add_slots(double *sum, uint8_t *ext, uint32_t d, double x, bool ok) {
  bool neg = false, any_neg = false;
  if constexpr (EXT) {
    neg = ok && x < 0.0;
    any_neg = __ballot(neg) != 0;
    if (any_neg) x = fabs(x);
  }
  const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
  if (__ballot(d == d0 || !ok) == ~0ull) {
    const double t = group_sum<kWave>(ok ? x : 0.0);
    if ((threadIdx.x & (kWave - 1)) == 0) {
      atomicAdd(&sum[d0], t);
      if (EXT && any_neg) ext[d0] = 1;  // one flag store for the wave's one destination
    }
    return;
  }
  if (ok) {
    atomicAdd(&sum[d], x);
    // Only the first flagged slot of a destination stores the flag: an LDS byte load costs far less than a byte
    // store (RMAT-26, DF-P: accumulate 4636 -> 3187 us per iteration; starting each bin's flags from the affected
    // flags as well, to spare those stores too, took 4986 us).
    if (EXT && neg && !ext[d]) ext[d] = 1;
  }
}

// Frontier variants: a vertex with a frontier in-neighbour (flag set in ext) is marked affected before its update,
// which is the reference's extension step (pagerankAffectedExtend*CukU, pagerankCuda.hxx:313-385) done by pull over
// the slots already being read; a split bin's partial rows carry the flag as their sign (a flagged sum is positive).
// Weighted variants multiply each slot's value by its edge's weight (P.ws, slot order).
template <int B, int VARIANT>
__global__ __launch_bounds__(kBinBlock) void accumulate_kernel(RankArgs A, BinnedPlan P, LoopCtl ctl) {
  if (!iteration_begin(ctl)) return;
  constexpr int kAccGroups = acc_groups<VARIANT>();
  constexpr bool EXT = Traits<VARIANT>::frontier;
  constexpr bool WEIGHTED = Traits<VARIANT>::weighted;
  constexpr uint32_t kMask = uint32_t(B) - 1;  // local destination id
  constexpr bool kDeadBit = B <= 8192;
  __shared__ double sum[B];
  __shared__ uint8_t ext[EXT ? B : 1];
  __shared__ double lds_max[kBinWaves];
  __shared__ double lds_cost[2 * kBinWaves];
  const uint4 it = P.ac_item[blockIdx.x];  // bin, first slot, end slot, partial row (kWholeBin: none)
  for (uint32_t i = threadIdx.x; i < uint32_t(B); i += kBinBlock) {
    sum[i] = 0.0;
    if constexpr (EXT) ext[i] = 0;
  }
  __syncthreads();
  const int w = threadIdx.x / kWave;
  const uint32_t lane = threadIdx.x & (kWave - 1);
  // Groups of 64 slots aligned to multiples of 64 (the slot array is padded to a multiple of 64), dealt to the
  // waves in turn, kAccGroups groups a batch; slots outside the item are masked. Software-pipelined: a batch's
  // X loads are issued, then the next batch's table and slot loads, so that the wait for X overlaps them. Loads of
  // groups past the item's end read the batch's first group again (always in bounds) and are not added.
  auto weight = [&](uint32_t j, bool ok) { return WEIGHTED && ok ? double(P.ws[j]) : 1.0; };
  uint32_t g = __builtin_amdgcn_readfirstlane((it.y & ~63u) + uint32_t(w) * kWave);  // wave-uniform
  if (g < it.z) {
    uint4 t[kAccGroups];
    uint32_t d[kAccGroups];
    auto fetch = [&](uint32_t gb, uint32_t gv) {  // batch at gb; gv: a group start below it.z
#pragma unroll
      for (int u = 0; u < kAccGroups; ++u) {
        const uint32_t gu = gb + u * kBinBlock < it.z ? gb + u * kBinBlock : gv;
        t[u] = P.grp[gu >> 6];
        d[u] = ld_nt(P.dst16 + gu + lane);
      }
    };
    fetch(g, g);
    for (; g < it.z; g += kAccGroups * kBinBlock) {
      uint32_t dc[kAccGroups], k[kAccGroups];
      double x[kAccGroups];
#pragma unroll
      for (int u = 0; u < kAccGroups; ++u) {
        dc[u] = d[u];
        k[u] = pair_entry(P, t[u], d[u]);
      }
#pragma unroll
      for (int u = 0; u < kAccGroups; ++u) x[u] = P.X[k[u]];
      fetch(g + kAccGroups * kBinBlock, g);
#pragma unroll
      for (int u = 0; u < kAccGroups; ++u) {
        const uint32_t gu = g + u * kBinBlock;
        if (gu >= it.z) break;
        const bool ok = gu + lane >= it.y && gu + lane < it.z && !(kDeadBit && (dc[u] & kDeadSlot));
        add_slots<EXT>(sum, ext, dc[u] & kMask, x[u] * weight(gu + lane, ok), ok);
      }
    }
  }
  // Edges an in-place update inserted outside the plan (the overlay, sorted by destination): the bin's first item
  // adds them, each with its source's value that overlay_expand_kernel wrote this iteration (negative: a frontier
  // source, as in X).
  if (P.ov_n && (blockIdx.x == 0 || P.ac_item[blockIdx.x - 1].x != it.x)) {
    const uint32_t o1 = P.ov_off[it.x + 1];
    for (uint32_t i = P.ov_off[it.x] + threadIdx.x; i < o1; i += kBinBlock) {
      double x = P.ovX[i];
      const uint32_t d = uint32_t(P.ov_key[i] >> 32) & kMask;
      if constexpr (EXT) {
        if (x < 0.0) {
          ext[d] = 1;
          x = -x;
        }
      }
      atomicAdd(&sum[d], x);
    }
  }
  __syncthreads();
  const uint32_t b0 = it.x * uint32_t(B);
  const uint32_t nb = min(uint32_t(B), P.N - b0);
  CostAcc cost;
  double emax = 0.0;
  if (it.w == kWholeBin) {
    for (uint32_t i = threadIdx.x; i < nb; i += kBinBlock)
      emax = fmax(emax, bin_finish<VARIANT>(A, b0 + i, sum[i], cost, EXT && ext[EXT ? i : 0]));
  } else {
    double *p = P.partial + size_t(it.w) * B;
    for (uint32_t i = threadIdx.x; i < nb; i += kBinBlock) p[i] = EXT && ext[EXT ? i : 0] ? -sum[i] : sum[i];
  }
  wg_store_costs<VARIANT, kBinWaves>(A, cost, blockIdx.x, lds_cost);
  emax = wg_max<kBinWaves>(emax, lds_max);
  if (threadIdx.x == 0) atomic_max_nonneg(&ctl.err[iter_index(ctl)], emax);
}

// Split bins, one workgroup per unit (bin, first partial row, rows, first vertex | mode << 31):
//   mode 0: vertices first .. first + 1023 of the bin, a thread per vertex adding all its partial rows;
//   mode 1: vertices first .. first + 63, wave w adding rows w, w + 16, ... of its lane's vertex, then the 16 wave
//           sums in wave order (bins with many rows).
template <int VARIANT>
__global__ __launch_bounds__(kBinBlock) void finish_bins_kernel(RankArgs A, BinnedPlan P, LoopCtl ctl) {
  if (!iteration_begin(ctl)) return;
  __shared__ double part[kBinWaves][kWave];
  __shared__ double lds_max[kBinWaves];
  __shared__ double lds_cost[2 * kBinWaves];
  CostAcc cost;
  const uint4 m = P.multi[blockIdx.x];
  const bool split = (m.w >> 31) != 0;
  const uint32_t first = m.w & 0x7FFFFFFFu;
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  const uint32_t i = first + (split ? uint32_t(lane) : threadIdx.x);  // vertex offset in the bin
  const uint32_t v = m.x * uint32_t(P.B) + i;
  const bool ok = i < uint32_t(P.B) && v < P.N;
  const double *p = P.partial + size_t(m.y) * P.B + i;
  double emax = 0.0;
  // Partial rows carry the frontier flag as their sign (frontier variants): add magnitudes, OR the signs. Other
  // variants add them as they are (the weighted adjoint's sums take either sign).
  bool flag = false;
  auto take = [&](double x) {
    if constexpr (Traits<VARIANT>::frontier) {
      flag |= x < 0.0;
      return fabs(x);
    } else {
      return x;
    }
  };
  if (!split) {
    if (ok) {
      double s0 = 0.0, s1 = 0.0;
      uint32_t k = 0;
      for (; k + 1 < m.z; k += 2) {
        s0 += take(p[size_t(k) * P.B]);
        s1 += take(p[size_t(k + 1) * P.B]);
      }
      if (k < m.z) s0 += take(p[size_t(k) * P.B]);
      emax = bin_finish<VARIANT>(A, v, s0 + s1, cost, flag);
    }
  } else {
    double s0 = 0.0, s1 = 0.0;
    if (ok) {
      uint32_t k = w;
      for (; k + kBinWaves < m.z; k += 2 * kBinWaves) {
        s0 += take(p[size_t(k) * P.B]);
        s1 += take(p[size_t(k + kBinWaves) * P.B]);
      }
      if (k < m.z) s0 += take(p[size_t(k) * P.B]);
    }
    part[w][lane] = flag ? -(s0 + s1) : s0 + s1;  // a flagged sum is positive
    __syncthreads();
    if (w == 0 && ok) {
      double t = 0.0;
      flag = false;
      for (int q = 0; q < kBinWaves; ++q) t += take(part[q][lane]);
      emax = bin_finish<VARIANT>(A, v, t, cost, flag);
    }
  }
  wg_store_costs<VARIANT, kBinWaves>(A, cost, P.n_ac + blockIdx.x, lds_cost);
  emax = wg_max<kBinWaves>(emax, lds_max);
  if (threadIdx.x == 0) atomic_max_nonneg(&ctl.err[iter_index(ctl)], emax);
}

// The overlay's values for this iteration: entry i takes its source's contribution, negated when the source was in
// the last frontier (naff, frontier variants after their first iteration), as expand does for X. It runs between
// expand and accumulate: accumulate overwrites naff with the new frontier. The entry count is ov_off[bins].
template <bool FROM_RANKS>
__global__ __launch_bounds__(kBlock) void overlay_expand_kernel(BinnedPlan P, const double *cr, const double *r,
                                                                const uint32_t *deg, const uint8_t *naff,
                                                                LoopCtl ctl) {
  if (!iteration_begin(ctl)) return;
  if (iter_index(ctl) == 0) naff = nullptr;
  const uint32_t n = P.ov_off[(P.N + P.B - 1) / P.B];
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const uint32_t u = uint32_t(P.ov_key[i]);
    const double c = contrib1<FROM_RANKS>(cr, r, deg, u);
    P.ovX[i] = naff && naff[u] ? -c : c;
  }
}

template <int B, int VARIANT>
hipError_t binned_update(const RankArgs &A, const BinnedPlan &P, const LoopCtl &ctl, bool extend, hipStream_t s,
                         hipEvent_t mark) {
  const uint8_t *naff = Traits<VARIANT>::frontier && extend ? A.naff : nullptr;
  constexpr bool kFromRanks = !Traits<VARIANT>::weighted;
  if (P.n_ex) expand_kernel<B, kFromRanks><<<P.n_ex, kBinBlock, 0, s>>>(P, A.cr, A.r, A.deg, naff, ctl);
  if (P.ov_n) {
    const uint32_t g = (P.ov_n + kBlock - 1) / kBlock;
    overlay_expand_kernel<kFromRanks><<<g < 4096 ? g : 4096, kBlock, 0, s>>>(P, A.cr, A.r, A.deg, naff, ctl);
  }
  accumulate_kernel<B, VARIANT><<<P.n_ac, kBinBlock, 0, s>>>(A, P, ctl);
  if (P.n_multi) finish_bins_kernel<VARIANT><<<P.n_multi, kBinBlock, 0, s>>>(A, P, ctl);
  if (mark) {
    const hipError_t e = hipEventRecord(mark, s);
    if (e != hipSuccess) return e;
  }
  return hipGetLastError();
}

// ---- build ----

inline uint32_t grid(size_t n, size_t per, uint32_t cap) {
  const size_t g = (n + per - 1) / per;
  return uint32_t(g < 1 ? 1 : (g > cap ? cap : g));
}

// key[e] = bin(v) * nblocks + block(src[e]) and row[e] = v for every in-CSR slot e of row v; one wave per row.
__global__ __launch_bounds__(kBlock) void bin_keys_kernel(const uint32_t *off, const uint32_t *src, uint32_t N,
                                                          uint32_t B, uint32_t nblocks, uint32_t *key, uint32_t *row) {
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  for (uint32_t v = blockIdx.x * kWavesPerBlock + w; v < N; v += gridDim.x * kWavesPerBlock) {
    const uint32_t kb = (v / B) * nblocks;
    for (uint32_t e = off[v] + lane; e < off[v + 1]; e += kWave) {
      key[e] = kb + src[e] / B;
      row[e] = v;
    }
  }
}

// Slot j holds in-CSR slot e = perm2[j] (slots sorted by tile, then source, then destination): its local
// destination id with the pair and run start flags, and pf[j] = 1 where it starts a pair.
__global__ __launch_bounds__(kBlock) void pair_slots_kernel(const uint32_t *perm2, const uint32_t *src,
                                                            const uint32_t *row, const uint32_t *key2s, size_t M,
                                                            uint32_t B, uint16_t *dst16, uint32_t *pf) {
  for (size_t j = size_t(blockIdx.x) * kBlock + threadIdx.x; j < M; j += size_t(gridDim.x) * kBlock) {
    const uint32_t e = perm2[j], u = src[e];
    const bool run = j == 0 || key2s[j] != key2s[j - 1];
    const bool pair = run || src[perm2[j - 1]] != u;
    dst16[j] = uint16_t((row[e] % B) | (pair ? kPairStart : 0u) | (run ? kRunStart : 0u));
    pf[j] = pair ? 1u : 0u;
  }
}

// Pair q = pid[j] of every pair-starting slot j (pid: exclusive scan of pf): its source psrc[q], tile key pkey2[q]
// and source block key1[q].
__global__ __launch_bounds__(kBlock) void pair_fill_kernel(const uint32_t *pf, const uint32_t *pid,
                                                           const uint32_t *perm2, const uint32_t *src,
                                                           const uint32_t *key2s, size_t M, uint32_t B,
                                                           uint32_t *psrc, uint32_t *pkey2, uint32_t *key1) {
  for (size_t j = size_t(blockIdx.x) * kBlock + threadIdx.x; j < M; j += size_t(gridDim.x) * kBlock) {
    if (!pf[j]) continue;
    const uint32_t q = pid[j], u = src[perm2[j]];
    psrc[q] = u;
    pkey2[q] = key2s[j];
    key1[q] = u / B;
  }
}

// flag[j] = 1 where entry j starts a run (a new tile in the sorted tile keys).
__global__ __launch_bounds__(kBlock) void run_flags_kernel(const uint32_t *key2s, size_t M, uint32_t *flag) {
  for (size_t j = size_t(blockIdx.x) * kBlock + threadIdx.x; j < M; j += size_t(gridDim.x) * kBlock)
    flag[j] = (j == 0 || key2s[j] != key2s[j - 1]) ? 1u : 0u;
}

// X entry k holds pair perm1[k]: src16[k] = its local source id, inv[perm1[k]] = k.
__global__ __launch_bounds__(kBlock) void pair_expand_kernel(const uint32_t *perm1, const uint32_t *psrc, size_t Q,
                                                             uint32_t B, uint16_t *src16, uint32_t *inv) {
  for (size_t k = size_t(blockIdx.x) * kBlock + threadIdx.x; k < Q; k += size_t(gridDim.x) * kBlock) {
    const uint32_t q = perm1[k];
    src16[k] = uint16_t(psrc[q] % B);
    inv[q] = uint32_t(k);
  }
}

// run_off[rid[q]] = inv[q] - q for every run-starting pair q (rid: exclusive scan of the run flags rf): pair p of
// that run is X entry p + run_off (modulo 2^32).
__global__ __launch_bounds__(kBlock) void pair_runs_kernel(const uint32_t *rf, const uint32_t *rid,
                                                           const uint32_t *inv, size_t Q, uint32_t *run_off) {
  for (size_t q = size_t(blockIdx.x) * kBlock + threadIdx.x; q < Q; q += size_t(gridDim.x) * kBlock)
    if (rf[q]) run_off[rid[q]] = inv[q] - uint32_t(q);
}

// grp[i] = (run r, run_off[r], run_off[r + 1] (0 past the last run), pair) of slot 64 i.
__global__ __launch_bounds__(kBlock) void group_table_kernel(const uint32_t *pf, const uint32_t *pid,
                                                             const uint32_t *rf, const uint32_t *rid,
                                                             const uint32_t *run_off, uint32_t n_runs, size_t M,
                                                             uint4 *grp) {
  const size_t G = (M + 63) / 64;
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < G; i += size_t(gridDim.x) * kBlock) {
    const size_t j = i * 64;
    const uint32_t q = pid[j] + pf[j] - 1;  // slot 0 starts pair 0, so q >= 0
    const uint32_t r = rid[q] + rf[q] - 1;
    grp[i] = make_uint4(r, run_off[r], r + 1 < n_runs ? run_off[r + 1] : 0u, q);
  }
}

// ws[j] = w[eid[slot[j]]] (weighted solves: the weight of each slot's edge, in slot order).
__global__ __launch_bounds__(kBlock) void binned_weights_kernel(const float *w, const uint32_t *eid,
                                                                const uint32_t *slot, size_t M, float *ws) {
  for (size_t k = size_t(blockIdx.x) * kBlock + threadIdx.x; k < M; k += size_t(gridDim.x) * kBlock)
    ws[k] = w[eid[slot[k]]];
}

// start[s] = first k with key[k] >= s, for s = 0 .. n (key sorted ascending).
__global__ __launch_bounds__(kBlock) void lower_bounds_kernel(const uint32_t *key, size_t M, uint32_t n,
                                                              uint32_t *start) {
  const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
  if (s > n) return;
  size_t lo = 0, hi = M;
  while (lo < hi) {
    const size_t mid = lo + (hi - lo) / 2;
    if (key[mid] < s) lo = mid + 1;
    else hi = mid;
  }
  start[s] = uint32_t(lo);
}

// ---- in-place update of a plan (prh_graph_update) ----

constexpr unsigned long long kNoKey = ~0ull;

// First index in a[lo, hi) with a[i] >= x.
template <class T>
__device__ __forceinline__ uint64_t lower_index(const T *a, uint64_t lo, uint64_t hi, T x) {
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (a[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// One tidied batch edge per thread (keys: destination << 32 | source, layout ids; count: device word). The edge is
// looked up in the base in-CSR (the CSR the plan was built from, rows ascending). A deletion (dead = 1) found there
// sets its slot's dead flag, one not found there is an overlay entry, marked gone; an insertion (dead = 0) found
// there (a base edge deleted earlier) clears the flag, one not found there becomes a new overlay entry (fresh[i]).
__global__ __launch_bounds__(kBlock) void patch_slots_kernel(const unsigned long long *keys, const uint32_t *count,
                                                             uint32_t cap, int dead, const uint32_t *base_off,
                                                             const uint32_t *base_src, const uint32_t *slot_of,
                                                             uint16_t *dst16, const unsigned long long *ov_key,
                                                             const uint32_t *ov_count, uint8_t *ov_gone,
                                                             unsigned long long *fresh) {
  const uint32_t n = *count, ov_n = *ov_count;
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < cap; i += gridDim.x * kBlock) {
    unsigned long long f = kNoKey;
    if (i < n) {
      const unsigned long long k = keys[i];
      const uint32_t v = uint32_t(k >> 32), u = uint32_t(k);
      const uint32_t b = base_off[v], e = base_off[v + 1];
      const uint32_t at = uint32_t(lower_index(base_src, b, e, u));
      if (at < e && base_src[at] == u) {
        const uint32_t j = slot_of[at];
        dst16[j] = uint16_t(dead ? (dst16[j] | kDeadSlot) : (dst16[j] & ~kDeadSlot));
      } else if (dead) {
        const uint64_t o = lower_index(ov_key, 0, ov_n, k);
        if (o < ov_n && ov_key[o] == k) ov_gone[o] = 1;
      } else {
        f = k;
      }
    }
    if (!dead) fresh[i] = f;
  }
}

// The next overlay's candidates, cand[0 .. ov_cap + fresh_cap): the current entries (ov_count of the ov_cap slots)
// not gone, then the fresh insertions (kNoKey for none); *kept counts the live ones.
__global__ __launch_bounds__(kBlock) void overlay_candidates_kernel(const unsigned long long *ov_key,
                                                                    const uint32_t *ov_count, uint32_t ov_cap,
                                                                    const uint8_t *ov_gone,
                                                                    const unsigned long long *fresh, uint32_t fresh_cap,
                                                                    unsigned long long *cand, uint32_t *kept) {
  const uint32_t ov_n = *ov_count;
  uint32_t live = 0;
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < ov_cap + fresh_cap; i += gridDim.x * kBlock) {
    const unsigned long long k = i < ov_cap ? (i < ov_n && !ov_gone[i] ? ov_key[i] : kNoKey) : fresh[i - ov_cap];
    cand[i] = k;
    live += k != kNoKey;
  }
  if (live) atomicAdd(kept, live);
}

// off[b] = first overlay entry with destination >= b B, for b = 0 .. nb, over the first min(*count, cap) sorted keys
// (off[nb] is the overlay's size); a count above cap sets bit in *flags.
__global__ __launch_bounds__(kBlock) void overlay_bounds_kernel(const unsigned long long *keys, const uint32_t *count,
                                                                uint32_t cap, uint32_t B, uint32_t nb, uint32_t *off,
                                                                int *flags, int bit) {
  const uint32_t b = blockIdx.x * kBlock + threadIdx.x;
  if (b > nb) return;
  const uint32_t c = *count, n = c < cap ? c : cap;
  if (b == 0 && c > cap && flags) atomicOr(flags, bit);
  off[b] = uint32_t(lower_index(keys, 0, n, (unsigned long long)b * B << 32));
}

__global__ __launch_bounds__(kBlock) void copy_u64_kernel(unsigned long long *dst, const unsigned long long *src,
                                                          size_t n) {
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += size_t(gridDim.x) * kBlock) dst[i] = src[i];
}

template <int B>
hipError_t binned_update_variant(int variant, const RankArgs &A, const BinnedPlan &P, const LoopCtl &ctl,
                                 bool extend, hipStream_t s, hipEvent_t mark) {
  switch (variant) {
    case kPlain: return binned_update<B, kPlain>(A, P, ctl, extend, s, mark);
    case kTraversal: return binned_update<B, kTraversal>(A, P, ctl, extend, s, mark);
    case kFrontier: return binned_update<B, kFrontier>(A, P, ctl, extend, s, mark);
    case kPrune: return binned_update<B, kPrune>(A, P, ctl, extend, s, mark);
    case kWeightedForward: return binned_update<B, kWeightedForward>(A, P, ctl, false, s, mark);
    case kWeightedAdjoint: return binned_update<B, kWeightedAdjoint>(A, P, ctl, false, s, mark);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

hipError_t launch_binned_update(int variant, const RankArgs &A, const BinnedPlan &P, unsigned long long *err,
                                int *done, int l, double tol, bool extend, hipStream_t s, hipEvent_t mark) {
  const LoopCtl ctl{err, done, l, tol};
  switch (P.B) {
    case 1024: return binned_update_variant<1024>(variant, A, P, ctl, extend, s, mark);
    case 2048: return binned_update_variant<2048>(variant, A, P, ctl, extend, s, mark);
    case 4096: return binned_update_variant<4096>(variant, A, P, ctl, extend, s, mark);
    case 8192: return binned_update_variant<8192>(variant, A, P, ctl, extend, s, mark);
    case 16384: return binned_update_variant<16384>(variant, A, P, ctl, extend, s, mark);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_patch_slots(const unsigned long long *keys, const uint32_t *count, uint32_t cap, bool dead,
                              const uint32_t *base_off, const uint32_t *base_src, const uint32_t *slot_of,
                              uint16_t *dst16, const unsigned long long *ov_key, const uint32_t *ov_count,
                              uint8_t *ov_gone, unsigned long long *fresh, hipStream_t s) {
  if (!cap) return hipSuccess;
  patch_slots_kernel<<<grid(cap, kBlock, 4096), kBlock, 0, s>>>(keys, count, cap, dead ? 1 : 0, base_off, base_src,
                                                                 slot_of, dst16, ov_key, ov_count, ov_gone, fresh);
  return hipGetLastError();
}

hipError_t launch_overlay_candidates(const unsigned long long *ov_key, const uint32_t *ov_count, uint32_t ov_cap,
                                     const uint8_t *ov_gone, const unsigned long long *fresh, uint32_t fresh_cap,
                                     unsigned long long *cand, uint32_t *kept, hipStream_t s) {
  if (!(ov_cap + fresh_cap)) return hipSuccess;
  overlay_candidates_kernel<<<grid(size_t(ov_cap) + fresh_cap, kBlock, 4096), kBlock, 0, s>>>(
      ov_key, ov_count, ov_cap, ov_gone, fresh, fresh_cap, cand, kept);
  return hipGetLastError();
}

hipError_t launch_overlay_bounds(const unsigned long long *keys, const uint32_t *count, uint32_t cap, uint32_t B,
                                 uint32_t nb, uint32_t *off, int *flags, int bit, hipStream_t s) {
  overlay_bounds_kernel<<<(nb + kBlock) / kBlock, kBlock, 0, s>>>(keys, count, cap, B, nb, off, flags, bit);
  return hipGetLastError();
}

hipError_t launch_copy_u64(unsigned long long *dst, const unsigned long long *src, size_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  copy_u64_kernel<<<grid(n, kBlock, 4096), kBlock, 0, s>>>(dst, src, n);
  return hipGetLastError();
}

hipError_t launch_bin_keys(const uint32_t *off, const uint32_t *src, uint32_t N, uint32_t B, uint32_t nblocks,
                           uint32_t *key, uint32_t *row, hipStream_t s) {
  if (!N) return hipSuccess;
  bin_keys_kernel<<<grid(N, kWavesPerBlock, 16384), kBlock, 0, s>>>(off, src, N, B, nblocks, key, row);
  return hipGetLastError();
}

hipError_t launch_pair_slots(const uint32_t *perm2, const uint32_t *src, const uint32_t *row, const uint32_t *key2s,
                             size_t M, uint32_t B, uint16_t *dst16, uint32_t *pf, hipStream_t s) {
  if (!M) return hipSuccess;
  pair_slots_kernel<<<grid(M, kBlock, 16384), kBlock, 0, s>>>(perm2, src, row, key2s, M, B, dst16, pf);
  return hipGetLastError();
}

hipError_t launch_pair_fill(const uint32_t *pf, const uint32_t *pid, const uint32_t *perm2, const uint32_t *src,
                            const uint32_t *key2s, size_t M, uint32_t B, uint32_t *psrc, uint32_t *pkey2,
                            uint32_t *key1, hipStream_t s) {
  if (!M) return hipSuccess;
  pair_fill_kernel<<<grid(M, kBlock, 16384), kBlock, 0, s>>>(pf, pid, perm2, src, key2s, M, B, psrc, pkey2, key1);
  return hipGetLastError();
}

hipError_t launch_run_flags(const uint32_t *key2s, size_t M, uint32_t *flag, hipStream_t s) {
  if (!M) return hipSuccess;
  run_flags_kernel<<<grid(M, kBlock, 16384), kBlock, 0, s>>>(key2s, M, flag);
  return hipGetLastError();
}

hipError_t launch_pair_expand(const uint32_t *perm1, const uint32_t *psrc, size_t Q, uint32_t B, uint16_t *src16,
                              uint32_t *inv, hipStream_t s) {
  if (!Q) return hipSuccess;
  pair_expand_kernel<<<grid(Q, kBlock, 16384), kBlock, 0, s>>>(perm1, psrc, Q, B, src16, inv);
  return hipGetLastError();
}

hipError_t launch_pair_runs(const uint32_t *rf, const uint32_t *rid, const uint32_t *inv, size_t Q,
                            uint32_t *run_off, hipStream_t s) {
  if (!Q) return hipSuccess;
  pair_runs_kernel<<<grid(Q, kBlock, 16384), kBlock, 0, s>>>(rf, rid, inv, Q, run_off);
  return hipGetLastError();
}

hipError_t launch_group_table(const uint32_t *pf, const uint32_t *pid, const uint32_t *rf, const uint32_t *rid,
                              const uint32_t *run_off, uint32_t n_runs, size_t M, uint4 *grp, hipStream_t s) {
  if (!M) return hipSuccess;
  group_table_kernel<<<grid((M + 63) / 64, kBlock, 16384), kBlock, 0, s>>>(pf, pid, rf, rid, run_off, n_runs, M,
                                                                            grp);
  return hipGetLastError();
}

hipError_t launch_binned_weights(const float *w, const uint32_t *eid, const uint32_t *slot, size_t M, float *ws,
                                 hipStream_t s) {
  if (!M) return hipSuccess;
  binned_weights_kernel<<<grid(M, kBlock, 16384), kBlock, 0, s>>>(w, eid, slot, M, ws);
  return hipGetLastError();
}

hipError_t launch_lower_bounds(const uint32_t *key, size_t M, uint32_t n, uint32_t *start, hipStream_t s) {
  lower_bounds_kernel<<<(n + kBlock) / kBlock, kBlock, 0, s>>>(key, M, n, start);
  return hipGetLastError();
}

hipError_t sort_pairs_u32(void *tmp, size_t &tmp_bytes, const uint32_t *key_in, uint32_t *key_out,
                          const uint32_t *val_in, uint32_t *val_out, size_t n, int end_bit, hipStream_t s) {
  if (val_in)
    return rocprim::radix_sort_pairs(tmp, tmp_bytes, key_in, key_out, val_in, val_out, n, 0, unsigned(end_bit), s);
  return rocprim::radix_sort_pairs(tmp, tmp_bytes, key_in, key_out, rocprim::counting_iterator<uint32_t>(0u), val_out,
                                   n, 0, unsigned(end_bit), s);
}

}  // namespace prh
