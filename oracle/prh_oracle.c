/*
 * prh_oracle.c -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * A CPU restatement of the PageRank path of puzzlef/pagerank-cuda-dynamic:
 *   - orc_partition / orc_mark_* / orc_pagerank_cuda restate the GPU path of inc/pagerankCuda.hxx
 *     sequentially, with the same iteration structure, flag rules and summation order
 *     (thread kernel: sequential ascending sum; block kernel: 256 strided partial sums + tree);
 *   - orc_static_omp restates the OpenMP static path of inc/pagerank.hxx (the CPU baseline "port");
 *   - orc_pagerank_async restates the asynchronous OpenMP solvers that the reference's experiment runs on dynamic
 *     graphs (ASYNC=true: ranks updated in place), the CPU baseline of the dynamic methods;
 *   - orc_power_iteration is a compensated-sum FP64 power iteration run to stagnation (the "true" fixed point);
 *   - orc_ppr / orc_weighted_* are the FP64 oracles for the personalized (batched) and adjoint forms.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library. The product library
 * (libpagerank_hip.so) neither links nor calls it.
 *
 * Layout is the reference's device layout (pagerankCuda.hxx:756-775): dense vertex ids 0..N-1,
 * uint32 CSR of the transposed graph (xtoff[N+1], xtedg[M] = in-neighbours ascending, xtdat[N] = out-degree)
 * and, for dynamic solves, the uint32 CSR of the graph itself (xoff[N+1], xedg[M] = out-neighbours ascending).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

enum {
  ORC_STATIC = 0,
  ORC_NAIVE_DYNAMIC = 1,
  ORC_DYNAMIC_TRAVERSAL = 2,
  ORC_DYNAMIC_FRONTIER = 3,
  ORC_PRUNE_DYNAMIC_FRONTIER = 4
};

/* Switch degree of the reference (pagerankCuda.hxx:93, :169, :319, :362, :627). */
#define ORC_SWITCH_DEGREE 64u
/* Block size of the reference's block-per-vertex kernels (BLOCK_LIMIT_REDUCE_CUDA, _cuda.hxx:68). */
#define ORC_BLOCK 256

/* Stable partition into degree <= d (ascending) ++ degree > d (ascending); returns the partition point.
 * Follows pagerankPartitionVerticesCuW, pagerankCuda.hxx:625-638 (test / exclusive scan / fill, twice). */
uint32_t orc_partition(uint32_t N, const uint32_t *off, uint32_t d, uint32_t *xpar) {
  uint32_t n = 0;
  for (uint32_t v = 0; v < N; ++v)
    if (off[v + 1] - off[v] <= d) xpar[n++] = v;
  uint32_t NP = n;
  for (uint32_t v = 0; v < N; ++v)
    if (off[v + 1] - off[v] > d) xpar[n++] = v;
  return NP;
}

/* Sum of r[u]/deg[u] over in(v)[i::DI], ascending; pagerankCalculateContribCud, pagerankCuda.hxx:32-43. */
static double orc_contrib(const uint32_t *xtoff, const uint32_t *xtdat, const uint32_t *xtedg, const double *r,
                          uint32_t v, size_t i, size_t DI) {
  size_t EO = xtoff[v], EN = (size_t)xtoff[v + 1] - xtoff[v];
  double a = 0.0;
  for (; i < EN; i += DI) {
    uint32_t u = xtedg[EO + i];
    uint32_t d = xtdat[u];
    a += r[u] / d;
  }
  return a;
}

/* Block-of-256 contribution: strided partial sums then the reverse binary tree of
 * sumValuesBlockReduceCudU (_cuda.hxx:542-551), as used at pagerankCuda.hxx:177-179. */
static double orc_contrib_block(const uint32_t *xtoff, const uint32_t *xtdat, const uint32_t *xtedg,
                                const double *r, uint32_t v) {
  double cache[ORC_BLOCK];
  for (int t = 0; t < ORC_BLOCK; ++t) cache[t] = orc_contrib(xtoff, xtdat, xtedg, r, v, (size_t)t, ORC_BLOCK);
  size_t n = ORC_BLOCK;
  while (n > 1) {
    size_t dn = (n + 1) / 2;
    for (size_t i = 0; i < n / 2; ++i) cache[i] += cache[dn + i];
    n = dn;
  }
  return cache[0];
}

/* One vertex of pagerankUpdateRanks{Thread,Block}CukU (pagerankCuda.hxx:94-107, :170-186).
 * The device path computes C0 + P*x as a fused multiply-add; the oracle states it with fma(). */
static void orc_update_vertex(int dynamic, int frontier, int prune, uint32_t v, double *a, uint8_t *naff,
                              uint8_t *vaff, const double *r, const uint32_t *xtoff, const uint32_t *xtdat,
                              const uint32_t *xtedg, double C0, double P, double D, double C) {
  size_t EN = (size_t)xtoff[v + 1] - xtoff[v];
  if (dynamic && !vaff[v]) {
    a[v] = r[v];
    return;
  }
  uint32_t d = xtdat[v];
  double rv = r[v];
  double cv = EN <= ORC_SWITCH_DEGREE ? orc_contrib(xtoff, xtdat, xtedg, r, v, 0, 1)
                                      : orc_contrib_block(xtoff, xtdat, xtedg, r, v);
  double av = prune ? fma(P, cv - rv / d, C0) * (1.0 / (1.0 - (P / d))) : fma(P, cv, C0);
  a[v] = av;
  double ev = fabs(av - rv);
  double rel = ev / fmax(rv, av);
  if (frontier && prune && rel <= C) vaff[v] = 0;
  if (frontier && rel > D) naff[v] = 1;
}

/* vaff[w] = 1 for w in out(u), for every u with naff[u]; pagerankAffectedExtend{Thread,Block}CukU with
 * pagerankMarkNeighborsCudU (pagerankCuda.hxx:55-63, :313-366). */
static void orc_extend(uint8_t *vaff, const uint8_t *naff, uint32_t N, const uint32_t *xoff, const uint32_t *xedg) {
  for (uint32_t u = 0; u < N; ++u) {
    if (!naff[u]) continue;
    for (uint32_t i = xoff[u]; i < xoff[u + 1]; ++i) vaff[xedg[i]] = 1;
  }
}

/* DF initial marking: vaff = out_y(U_del u U_ins) u V_del, naff = U_del u U_ins.
 * pagerankAffectedFrontierCuW + pagerankAffectedFrontierPartialCukU, pagerankCuda.hxx:272-281, :405-412. */
void orc_mark_frontier(uint32_t N, const uint32_t *xoff, const uint32_t *xedg, const uint32_t *delu,
                       const uint32_t *delv, size_t ND, const uint32_t *insu, size_t NI, uint8_t *vaff,
                       uint8_t *naff) {
  memset(vaff, 0, N);
  memset(naff, 0, N);
  for (size_t i = 0; i < ND + NI; ++i) {
    uint32_t u = i < ND ? delu[i] : insu[i - ND];
    naff[u] = 1;
    if (i < ND) vaff[delv[i]] = 1;
  }
  orc_extend(vaff, naff, N, xoff, xedg);
}

/* DT marking: DF initial marking, then BFS levels over out_y until the affected count stops changing.
 * pagerankAffectedTraversalCuW, pagerankCuda.hxx:438-456. */
void orc_mark_traversal(uint32_t N, const uint32_t *xoff, const uint32_t *xedg, const uint32_t *delu,
                        const uint32_t *delv, size_t ND, const uint32_t *insu, size_t NI, uint8_t *vaff) {
  uint8_t *naff = (uint8_t *)malloc(N ? N : 1), *buff = (uint8_t *)malloc(N ? N : 1);
  orc_mark_frontier(N, xoff, xedg, delu, delv, ND, insu, NI, vaff, naff);
  memcpy(naff, vaff, N);
  uint64_t count = 0, countNew = 0;
  for (uint32_t v = 0; v < N; ++v) count += vaff[v];
  for (;;) {
    memset(buff, 0, N);
    orc_extend(buff, naff, N, xoff, xedg);
    for (uint32_t v = 0; v < N; ++v) vaff[v] |= buff[v];
    uint8_t *t = naff; naff = buff; buff = t;
    countNew = 0;
    for (uint32_t v = 0; v < N; ++v) countNew += vaff[v];
    if (countNew == count) break;
    count = countNew;
  }
  free(naff);
  free(buff);
}

/* Sequential restatement of pagerankInvokeCuda + pagerankLoopCuU (pagerankCuda.hxx:675-692, :713-849) for the
 * five solvers (entry points :864-988). Returns 0 on success, -1 on bad arguments.
 * ranks: N outputs (dense ids). vaff_marked (nullable): the affected flags right after marking. */
int orc_pagerank_cuda(int mode, uint32_t N, const uint32_t *xtoff, const uint32_t *xtedg, const uint32_t *xtdat,
                      const uint32_t *xoff, const uint32_t *xedg, const uint32_t *delu, const uint32_t *delv,
                      size_t ND, const uint32_t *insu, size_t NI, const double *q, double P, double E, int L,
                      double D, double C, double *ranks, int *iterations, uint8_t *vaff_marked) {
  if (mode < ORC_STATIC || mode > ORC_PRUNE_DYNAMIC_FRONTIER) return -1;
  if (N == 0) { *iterations = 0; return 0; }
  int dynamic = mode >= ORC_DYNAMIC_TRAVERSAL;
  int frontier = mode >= ORC_DYNAMIC_FRONTIER;
  int prune = mode == ORC_PRUNE_DYNAMIC_FRONTIER;
  if (mode != ORC_STATIC && !q) return -1;
  if (dynamic && (!xoff || !xedg)) return -1;
  double *a = (double *)malloc(sizeof(double) * N), *r = (double *)malloc(sizeof(double) * N);
  uint8_t *vaff = (uint8_t *)calloc(N, 1), *naff = (uint8_t *)calloc(N, 1);
  /* Initial ranks (pagerankCuda.hxx:816-819): a = r = q, or 1/N; the static solver has no q (:872). */
  if (mode == ORC_STATIC) q = NULL;
  for (uint32_t v = 0; v < N; ++v) r[v] = a[v] = q ? q[v] : 1.0 / N;
  /* Marking (pagerankCuda.hxx:821, the fm lambdas :926-928, :955-957, :984-986). With an empty batch the
   * reference leaves vaff uninitialised; here it is all-zero (nothing affected). */
  if (dynamic && (ND > 0 || NI > 0)) {
    if (mode == ORC_DYNAMIC_TRAVERSAL) orc_mark_traversal(N, xoff, xedg, delu, delv, ND, insu, NI, vaff);
    else orc_mark_frontier(N, xoff, xedg, delu, delv, ND, insu, NI, vaff, naff);
  }
  if (vaff_marked) memcpy(vaff_marked, vaff, N);
  /* Iterations (pagerankLoopCuU, pagerankCuda.hxx:676-692). */
  const double C0 = (1 - P) / N;
  int l = 0;
  while (l < L) {
    if (frontier) memset(naff, 0, N);
    for (uint32_t v = 0; v < N; ++v)
      orc_update_vertex(dynamic, frontier, prune, v, a, naff, vaff, r, xtoff, xtdat, xtedg, C0, P, D, C);
    ++l;
    double el = 0.0;
    for (uint32_t v = 0; v < N; ++v) el = fmax(el, fabs(a[v] - r[v]));
    if (frontier) orc_extend(vaff, naff, N, xoff, xedg);
    double *t = a; a = r; r = t;
    if (el < E) break;
  }
  memcpy(ranks, r, sizeof(double) * N);
  *iterations = l;
  free(a); free(r); free(vaff); free(naff);
  return 0;
}

/* OpenMP static PageRank on the CPU: pagerankStaticOmp (pagerank.hxx:473-482) through pagerankInvokeOmp
 * (:412-440), pagerankInitializeRanksOmp (:303-313), pagerankUpdateRanksSyncOmp (:216-227, schedule(dynamic,2048)),
 * pagerankUpdateRank (:131-141) and liNormDeltaOmp (_vector.hxx:1266). The arithmetic is the host's:
 * plain a = C0 + P*c. time_ms covers initialisation and iterations, like PagerankResult.time. */
int orc_static_omp(uint32_t N, const uint32_t *xtoff, const uint32_t *xtedg, const uint32_t *xtdat, double P,
                   double E, int L, int threads, double *ranks, int *iterations, double *time_ms) {
  if (N == 0) { *iterations = 0; *time_ms = 0; return 0; }
  if (threads > 0) omp_set_num_threads(threads);
  double *a = (double *)malloc(sizeof(double) * N), *r = (double *)malloc(sizeof(double) * N);
  double t0 = omp_get_wtime();
  #pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < (int64_t)N; ++v) r[v] = a[v] = 1.0 / N;
  const double C0 = (1 - P) / N;
  int l = 0;
  while (l < L) {
    #pragma omp parallel for schedule(dynamic, 2048)
    for (int64_t v = 0; v < (int64_t)N; ++v) {
      double cv = 0.0;
      for (uint32_t i = xtoff[v]; i < xtoff[v + 1]; ++i) {
        uint32_t u = xtedg[i];
        cv += r[u] / xtdat[u];
      }
      volatile double pc = P * cv; /* keep the host's separate multiply and add */
      a[v] = C0 + pc;
    }
    ++l;
    double el = 0.0;
    #pragma omp parallel for schedule(static) reduction(max : el)
    for (int64_t v = 0; v < (int64_t)N; ++v) {
      double e = fabs(a[v] - r[v]);
      if (e > el) el = e;
    }
    double *t = a; a = r; r = t;
    if (el < E) break;
  }
  double t1 = omp_get_wtime();
  memcpy(ranks, r, sizeof(double) * N);
  *iterations = l;
  *time_ms = (t1 - t0) * 1e3;
  free(a); free(r);
  return 0;
}

/* The reference's asynchronous OpenMP solvers (template switch ASYNC=true), as main.cu:90-145 runs them:
 *   mode 0 pagerankStaticOmp<true> (pagerank.hxx:473-482), 1 pagerankNaiveDynamicOmp<true> (:517-525),
 *   2 pagerankDynamicTraversalOmp<true> (:619-627), 3 pagerankDynamicFrontierOmp<true> (:718-727),
 *   4 pagerankPruneDynamicFrontierOmp<true> (pagerankPrune.hxx:300-307 via pagerankPruneInvokeOmp, :211-246).
 * Ranks live in one array updated in place, vertex by vertex (pagerankUpdateRanksAsyncOmp, pagerank.hxx:240-253;
 * pagerankPruneUpdateRanksAsyncOmp, pagerankPrune.hxx:140-158) under schedule(dynamic, 2048), and a sweep's L-inf
 * is the largest |old - new| it saw; the loop stops after the first sweep below E (pagerankInvokeOmp, :412-440).
 * With several threads the in-place reads race, as in the reference; with one thread the order is 0..N-1 and the
 * result is deterministic. Marking: DT the BFS closure (the same set as pagerankAffectedTraversalOmpW,
 * :562-578), DF and DF-P the out-neighbours of the batch sources and the deleted edges' targets (the set of
 * pagerankAffectedFrontierOmpW, :662-676). During the sweeps DF marks every out-neighbour of a vertex whose
 * relative change exceeds D (:725); DF-P uses the closed-form self-loop update, prunes at <= C and marks
 * out-neighbours other than itself that are not yet affected (pagerankPrune.hxx:29-39, :148-155).
 * C0 = (1 - P) / N: the reference's N is the vertex count, and its DF-P span equals it for dense ids.
 * xt*: in-CSR of the updated graph with out-degrees; yoff/yedg: its out-CSR. time_ms covers initialisation,
 * marking and the sweeps, like PagerankResult.time. Returns 0, or -1 on bad arguments. */
int orc_pagerank_async(int mode, uint32_t N, const uint32_t *xtoff, const uint32_t *xtedg, const uint32_t *xtdat,
                       const uint32_t *yoff, const uint32_t *yedg, const uint32_t *delu, const uint32_t *delv,
                       size_t ND, const uint32_t *insu, size_t NI, const double *q, double P, double E, int L,
                       double D, double C, int threads, double *ranks, int *iterations, double *time_ms) {
  if (mode < 0 || mode > 4 || (mode != ORC_STATIC && !q)) return -1;
  if (N == 0) { *iterations = 0; *time_ms = 0; return 0; }
  if (threads > 0) omp_set_num_threads(threads);
  double *r = (double *)malloc(sizeof(double) * N);
  uint8_t *vaff = (uint8_t *)malloc(N), *naff = (uint8_t *)malloc(N);
  double t0 = omp_get_wtime();
  #pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < (int64_t)N; ++v) r[v] = mode == ORC_STATIC ? 1.0 / N : q[v];
  if (mode == ORC_DYNAMIC_TRAVERSAL) orc_mark_traversal(N, yoff, yedg, delu, delv, ND, insu, NI, vaff);
  else if (mode >= ORC_DYNAMIC_FRONTIER) orc_mark_frontier(N, yoff, yedg, delu, delv, ND, insu, NI, vaff, naff);
  else memset(vaff, 1, N);
  const double C0 = (1 - P) / N;
  int l = 0;
  while (l < L) {
    double el = 0.0;
    #pragma omp parallel for schedule(dynamic, 2048) reduction(max : el)
    for (int64_t v = 0; v < (int64_t)N; ++v) {
      if (!vaff[v]) continue;
      const double rv = r[v];
      double cv = 0.0;
      for (uint32_t i = xtoff[v]; i < xtoff[v + 1]; ++i) {
        uint32_t u = xtedg[i];
        cv += r[u] / xtdat[u];
      }
      double av;
      if (mode == ORC_PRUNE_DYNAMIC_FRONTIER) {
        const uint32_t d = xtdat[v];
        volatile double pc = P * (cv - rv / d); /* the host's separate multiply and add */
        av = (C0 + pc) * (1 / (1 - (P / d)));
      } else {
        volatile double pc = P * cv;
        av = C0 + pc;
      }
      r[v] = av;
      const double ev = fabs(rv - av);
      if (ev > el) el = ev;
      const double rel = ev / (rv > av ? rv : av);
      if (mode == ORC_DYNAMIC_FRONTIER) {
        if (rel > D)
          for (uint32_t i = yoff[v]; i < yoff[v + 1]; ++i) vaff[yedg[i]] = 1;
      } else if (mode == ORC_PRUNE_DYNAMIC_FRONTIER) {
        if (rel <= C) vaff[v] = 0;
        if (rel > D)
          for (uint32_t i = yoff[v]; i < yoff[v + 1]; ++i) {
            uint32_t w = yedg[i];
            if (w != (uint32_t)v && !vaff[w]) vaff[w] = 1;
          }
      }
    }
    ++l;
    if (el < E) break;
  }
  double t1 = omp_get_wtime();
  memcpy(ranks, r, sizeof(double) * N);
  *iterations = l;
  *time_ms = (t1 - t0) * 1e3;
  free(r); free(vaff); free(naff);
  return 0;
}

/* Neumaier-compensated sum of x over [0,n). */
typedef struct { double s, c; } orc_ksum;
static inline void orc_kadd(orc_ksum *k, double x) {
  double t = k->s + x;
  if (fabs(k->s) >= fabs(x)) k->c += (k->s - t) + x;
  else k->c += (x - t) + k->s;
  k->s = t;
}

/* FP64 power iteration with compensated sums, run until the L-inf change stops decreasing (stagnation) or
 * max_iter. Teleport s (nullable: uniform 1/N). Used as the "true" fixed point for L-inf parity checks. */
int orc_power_iteration(uint32_t N, const uint32_t *xtoff, const uint32_t *xtedg, const uint32_t *xtdat,
                        const double *s, double P, int max_iter, double *ranks, int *iterations) {
  if (N == 0) { *iterations = 0; return 0; }
  double *a = (double *)malloc(sizeof(double) * N), *r = (double *)malloc(sizeof(double) * N);
  for (uint32_t v = 0; v < N; ++v) r[v] = s ? s[v] : 1.0 / N;
  double best = INFINITY;
  int stale = 0, l = 0;
  for (l = 0; l < max_iter;) {
    #pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t v = 0; v < (int64_t)N; ++v) {
      orc_ksum k = {0.0, 0.0};
      for (uint32_t i = xtoff[v]; i < xtoff[v + 1]; ++i) orc_kadd(&k, r[xtedg[i]] / xtdat[xtedg[i]]);
      double tv = (1 - P) * (s ? s[v] : 1.0 / N);
      a[v] = tv + P * (k.s + k.c);
    }
    ++l;
    double el = 0.0;
    for (uint32_t v = 0; v < N; ++v) el = fmax(el, fabs(a[v] - r[v]));
    double *t = a; a = r; r = t;
    if (el == 0.0) break;
    if (el < best) { best = el; stale = 0; }
    else if (++stale >= 8) break;
  }
  memcpy(ranks, r, sizeof(double) * N);
  *iterations = l;
  free(a); free(r);
  return 0;
}

/* Personalized PageRank for one teleport column s: r0 = s, r <- (1-P) s + P * sum_{u in in(v)} r[u]/deg[u],
 * stop at the first iteration whose L-inf change is < E (same rule as pagerankLoopCuU). */
int orc_ppr(uint32_t N, const uint32_t *xtoff, const uint32_t *xtedg, const uint32_t *xtdat, const double *s,
            double P, double E, int L, double *ranks, int *iterations) {
  if (N == 0) { *iterations = 0; return 0; }
  double *a = (double *)malloc(sizeof(double) * N), *r = (double *)malloc(sizeof(double) * N);
  for (uint32_t v = 0; v < N; ++v) r[v] = s[v];
  int l = 0;
  while (l < L) {
    /* Vertices are independent (each sums its in-edges in ascending order), so the OpenMP loops give the same bits
     * for any thread count; the max reduction is order-independent. */
    #pragma omp parallel for schedule(dynamic, 2048)
    for (int64_t v = 0; v < (int64_t)N; ++v) {
      double cv = orc_contrib(xtoff, xtdat, xtedg, r, (uint32_t)v, 0, 1);
      a[v] = fma(P, cv, (1 - P) * s[v]);
    }
    ++l;
    double el = 0.0;
    #pragma omp parallel for schedule(static) reduction(max : el)
    for (int64_t v = 0; v < (int64_t)N; ++v) el = fmax(el, fabs(a[v] - r[v]));
    double *t = a; a = r; r = t;
    if (el < E) break;
  }
  memcpy(ranks, r, sizeof(double) * N);
  *iterations = l;
  free(a); free(r);
  return 0;
}

/* Weighted PageRank forward: P[u,v] = w_uv / W_u with W_u = sum_v w_uv (FP64 sum of FP32 weights).
 * r0 = s; r_v <- (1-P) s_v + P * sum_{u in in(v)} w_uv * (r_u / W_u). w_in is aligned with xtedg. */
int orc_weighted_forward(uint32_t N, const uint32_t *xtoff, const uint32_t *xtedg, const float *w_in,
                         const double *W, const double *s, double P, double E, int L, double *ranks,
                         int *iterations) {
  if (N == 0) { *iterations = 0; return 0; }
  double *a = (double *)malloc(sizeof(double) * N), *r = (double *)malloc(sizeof(double) * N);
  double *y = (double *)malloc(sizeof(double) * N);
  for (uint32_t v = 0; v < N; ++v) r[v] = s[v];
  int l = 0;
  while (l < L) {
    #pragma omp parallel for schedule(static)
    for (int64_t u = 0; u < (int64_t)N; ++u) y[u] = r[u] / W[u];
    #pragma omp parallel for schedule(dynamic, 2048)
    for (int64_t v = 0; v < (int64_t)N; ++v) {
      double cv = 0.0;
      for (uint32_t i = xtoff[v]; i < xtoff[v + 1]; ++i) cv += (double)w_in[i] * y[xtedg[i]];
      a[v] = fma(P, cv, (1 - P) * s[v]);
    }
    ++l;
    double el = 0.0;
    #pragma omp parallel for schedule(static) reduction(max : el)
    for (int64_t v = 0; v < (int64_t)N; ++v) el = fmax(el, fabs(a[v] - r[v]));
    double *t = a; a = r; r = t;
    if (el < E) break;
  }
  memcpy(ranks, r, sizeof(double) * N);
  *iterations = l;
  free(a); free(r); free(y);
  return 0;
}

/* Adjoint of the weighted forward: solves (I - P * Pm) lam = g, i.e. lam_u <- g_u + (P / W_u) * sum_{v in out(u)}
 * w_uv * lam_v, pulled over the out-CSR (xoff, xedg) with w_out aligned with xedg. lam0 = g. */
int orc_weighted_adjoint(uint32_t N, const uint32_t *xoff, const uint32_t *xedg, const float *w_out,
                         const double *W, const double *g, double P, double E, int L, double *lam,
                         int *iterations) {
  if (N == 0) { *iterations = 0; return 0; }
  double *a = (double *)malloc(sizeof(double) * N), *r = (double *)malloc(sizeof(double) * N);
  for (uint32_t v = 0; v < N; ++v) r[v] = g[v];
  int l = 0;
  while (l < L) {
    #pragma omp parallel for schedule(dynamic, 2048)
    for (int64_t u = 0; u < (int64_t)N; ++u) {
      double cu = 0.0;
      for (uint32_t i = xoff[u]; i < xoff[u + 1]; ++i) cu += (double)w_out[i] * r[xedg[i]];
      a[u] = fma(P / W[u], cu, g[u]);
    }
    ++l;
    double el = 0.0;
    #pragma omp parallel for schedule(static) reduction(max : el)
    for (int64_t v = 0; v < (int64_t)N; ++v) el = fmax(el, fabs(a[v] - r[v]));
    double *t = a; a = r; r = t;
    if (el < E) break;
  }
  memcpy(lam, r, sizeof(double) * N);
  *iterations = l;
  free(a); free(r);
  return 0;
}

/* Gradient assembly: dL/ds = (1-P) lam; dL/dw_uv = (P r_u / W_u) * (lam_v - (lam_u - g_u)/P) for each edge
 * (u,v) of the out-CSR (grad_w aligned with xedg). The inner sum_v' P[u,v'] lam_v' is recovered from the adjoint
 * fixed point as (lam_u - g_u)/P rather than stored. */
void orc_weighted_grad(uint32_t N, const uint32_t *xoff, const uint32_t *xedg, const double *W, const double *r,
                       const double *lam, const double *g, double P, double *grad_w, double *grad_s) {
  #pragma omp parallel for schedule(dynamic, 2048)
  for (int64_t u = 0; u < (int64_t)N; ++u) {
    grad_s[u] = (1 - P) * lam[u];
    double su = (lam[u] - g[u]) / P;
    double fu = P * r[u] / W[u];
    for (uint32_t i = xoff[u]; i < xoff[u + 1]; ++i) grad_w[i] = fu * (lam[xedg[i]] - su);
  }
}
