"""Experiment runner on synthetic graphs and SNAP temporal files, following the reference's protocol and log format
(main.cu:47-148).

    python -m pagerank_amd.experiment dynamic --scale 22 --edge-factor 16 --fractions 1e-5,1e-4,1e-3 --batches 3
    python -m pagerank_amd.experiment temporal --file sx-stackoverflow.txt --rows 26019770 --size 63497050 \\
        --fractions 1e-5 --batches 5
    python -m pagerank_amd.experiment stream --scale 22 --fractions 1e-4 --batches 16
    python -m pagerank_amd.experiment ppr --scale 22 --columns 64
    python -m pagerank_amd.experiment adjoint --scale 20

dynamic: as runExperiment (main.cu:47-148). A stream of update batches is applied to the graph one after another
(80% insertions of uniform random vertex pairs, 20% deletions of uniform random existing non-self-loop edges,
tidied as tidyBatchUpdateU does). Each batch is applied to the device graph in place (prh_graph_update: both CSRs
rewritten, the propagation-blocked plan patched), or with --upload the host-built updated graph is uploaded (a new
layout and plan). After each batch every GPU method runs on the updated graph: Static, and ND / DT / DF / DF-P, each
warm-started from its own result of the previous batch (main.cu:102-143). Every method's time includes the batch's
preparation (the update, or the upload's layout and plan build), which each of them needs. Each result is compared
with a static solve of the same graph at tolerance 1e-100, which runs the full 500 iterations (main.cu:88). The
reference computes that solve on the CPU; here it runs on the GPU (same Jacobi iteration, FP64). One line per method
is printed in glog's format (main.cu:56-60), with the batch's deletion and insertion fractions of |E|. A summary of
the speedups over Static follows.

temporal: the same loop on the reference's input (main.cu:157-174): the first 90% of a SNAP temporal file's lines
make the base graph (self-loop on every vertex), and each batch is the next fraction*size lines as insertions,
tidied against the current graph. --rows and --size are the reference's arguments 2 and 3 (main.sh:46-50).

stream: 16 (--batches) consecutive DF-P batches of one size (--fractions, first value) through the streaming session
(prh_stream: batch applied on the device, apply and iterations replayed from captured HIP graphs), against the same
batches as separate calls (host-side apply, upload of the new graph, DF-P solve), against the DF-P solve times
alone, and against in-place updates of one device graph with a DF-P solve each (prh_graph_update, not captured).
The final ranks of every path are compared with the separate calls'.

ppr: one batched personalized solve of k teleport columns, each with a seed set of uniform random size 1..1000,
against k single-column solves and k plain static solves on the same graph.

adjoint: weighted forward solve, adjoint solve with gradient assembly, and a plain static solve on the same graph.
FP32 weights are uniform in [0.5, 1.5); the loss is L = sum_v c_v r_v^2 with c_v ~ N(0, 1).

--csv FILE appends every measurement as a CSV row.
"""
from __future__ import annotations

import argparse
import csv
import sys
import time

import numpy as np

from . import engine
from .engine import Context, PagerankOptions, solve
from .graphs import Batch, HostGraph, TemporalReader, apply_batch

METHODS = [
    (engine.NAIVE_DYNAMIC, "pagerankNaiveDynamicHip"),
    (engine.DYNAMIC_TRAVERSAL, "pagerankDynamicTraversalHip"),
    (engine.DYNAMIC_FRONTIER, "pagerankDynamicFrontierHip"),
    (engine.PRUNE_DYNAMIC_FRONTIER, "pagerankPruneDynamicFrontierHip"),
]


def log(msg: str) -> None:
    print(f"# {msg}", flush=True)


def l1_error(a: np.ndarray, b: np.ndarray) -> float:
    return float(np.abs(a - b).sum())


def glog(ans, ref: np.ndarray, technique: str, threads: int, deletionsf: float, insertionsf: float,
         batch_index: int) -> str:
    """One result line in the reference's format (main.cu:56-60)."""
    return (f"{{-{deletionsf:.3e}/+{insertionsf:.3e} batchf, {batch_index:03d} batchi, {threads:03d} threads}} -> "
            f"{{{ans.time:09.1f}ms, {ans.initializationTime:09.1f}ms init, {ans.markingTime:09.1f}ms mark, "
            f"{ans.computationTime:09.1f}ms comp, {ans.iterations:03d} iter, {l1_error(ans.ranks, ref):.2e} err}} "
            f"{technique}")


class CsvLog:
    FIELDS = ["mode", "graph", "vertices", "edges", "batch_fraction", "batch_index", "technique", "time_ms",
              "init_ms", "mark_ms", "comp_ms", "iterations", "l1_error", "extra"]

    def __init__(self, path: str | None):
        self.f = open(path, "a", newline="") if path else None
        self.w = csv.DictWriter(self.f, fieldnames=self.FIELDS) if self.f else None
        if self.f and self.f.tell() == 0:
            self.w.writeheader()

    def row(self, **kw):
        if self.w:
            self.w.writerow({k: kw.get(k, "") for k in self.FIELDS})
            self.f.flush()


def make_rmat(scale: int, edge_factor: float, seed: int) -> tuple[HostGraph, str]:
    t0 = time.time()
    g = HostGraph.rmat(scale, edge_factor, 0.57, 0.19, 0.19, seed=seed)
    name = f"rmat{scale}x{edge_factor:g}"
    log(f"graph {name}: {g.N} vertices, {g.M} edges (dedup, self-loops), built in {time.time() - t0:.1f} s")
    return g, name


def autotune(ctx: Context, g: HostGraph, args) -> None:
    if args.no_autotune:
        return
    yg = ctx.upload(g, with_out=False)
    t, ms, hit = ctx.autotune(yg, args.tuning_cache, iterations=10)
    yg.close()
    log(f"autotune: switch {t.switch_degree}, chunk {t.huge_degree}, slices {t.source_slices}, "
        f"block edges {t.block_edges}, blocking {t.blocking}: {ms:.3f} ms per static iteration"
        f"{' (cache)' if hit else ''}")


def _with_prep(a, prep: float):
    """A copy of result a that also carries the batch's preparation (the update of the graph, or the layout and plan
    of a fresh upload) in its initialization time and time."""
    return engine.PagerankResult(a.ranks, a.iterations, a.time + prep, a.initializationTime + prep, a.markingTime,
                                 a.computationTime, prep)


def run_batches(ctx: Context, x: HostGraph, name: str, frac: float, next_batch, args, out: CsvLog,
                summary: list) -> None:
    """runExperiment's loop (main.cu:79-146) on one batch source: next_batch(x, bi) gives the tidied batch bi
    against the current graph x (None ends the stream early). Every method runs on each updated graph and is
    compared with a static solve of it at tolerance 1e-100 (on the GPU; the reference runs that solve on the CPU).

    The updated graph is made on the device by prh_graph_update (the blocked plan patched in place), or, with
    --upload, built on the host and uploaded (a new layout and plan). Either way that preparation is paid once per
    batch, by whichever solve runs first on the new graph; every method's reported time includes it (each needs the
    updated graph), so the speedups over Static count it on both sides."""
    o = PagerankOptions(repeat=args.repeat)
    exact = PagerankOptions(repeat=1, tolerance=1e-100)
    xg = ctx.upload(x, with_out=not args.upload)
    r0 = solve(xg, engine.STATIC, exact).ranks
    if args.upload:
        xg.close()
    prev = {m: r0 for m, _ in METHODS}
    for bi in range(args.batches):
        t0 = time.time()
        b = next_batch(x, bi)
        if b is None:
            log(f"input ended before batch {bi}")
            break
        y = apply_batch(x, b)
        if args.upload:
            yg = ctx.upload(y)
        else:
            yg = xg
            yg.update(b)
        s0 = solve(yg, engine.STATIC, exact)
        prep = s0.preparationTime
        log(f"batch {bi} of {frac:g}|E|: -{b.deletions}/+{b.insertions} edges, y has {y.M} edges "
            f"({time.time() - t0:.1f} s on the host to make the batch and y); "
            f"{'upload + layout + plan' if args.upload else 'device update'}: {prep:.3f} ms, added to every method")
        res = {engine.STATIC: _with_prep(solve(yg, engine.STATIC, o), prep)}
        for m, _ in METHODS:
            res[m] = _with_prep(solve(yg, m, o, batch=b, q=prev[m]), prep)
        for m, tech in [(engine.STATIC, "pagerankStaticHip")] + METHODS:
            a = res[m]
            print(glog(a, s0.ranks, tech, 1, b.deletions / x.M, b.insertions / x.M, bi), flush=True)
            out.row(mode=args.mode, graph=name, vertices=y.N, edges=y.M, batch_fraction=frac, batch_index=bi,
                    technique=tech, time_ms=a.time, init_ms=a.initializationTime, mark_ms=a.markingTime,
                    comp_ms=a.computationTime, iterations=a.iterations, l1_error=l1_error(a.ranks, s0.ranks),
                    extra=f"deletions={b.deletions};insertions={b.insertions};preparation_ms={prep:.3f};"
                          f"{'upload' if args.upload else 'update'}")
        for m, _ in METHODS:
            prev[m] = res[m].ranks
        summary.append((frac, bi, {m: res[m].time for m in res}, prep))
        if args.upload:
            yg.close()
        x = y
    if not args.upload:
        xg.close()


def print_summary(summary: list) -> None:
    log("speedup over GPU Static (time of Static / time of the method, each with the batch's preparation), mean over "
        "batches:")
    for frac in sorted({s[0] for s in summary}):
        rows = [s[2] for s in summary if s[0] == frac]
        prep = np.mean([s[3] for s in summary if s[0] == frac])
        st = np.mean([r[engine.STATIC] for r in rows])
        parts = [f"preparation {prep:.3f} ms", f"Static {st:.3f} ms"]
        for m, tech in METHODS:
            t = np.mean([r[m] for r in rows])
            parts.append(f"{tech.replace('pagerank', '').replace('Hip', '')} {t:.3f} ms ({st / t:.2f}x)")
        log(f"batch {frac:g}|E|: " + ", ".join(parts))


def run_dynamic(args, out: CsvLog) -> int:
    x0, name = make_rmat(args.scale, args.edge_factor, args.seed)
    ctx = Context(0)
    autotune(ctx, x0, args)
    summary = []
    for frac in [float(f) for f in args.fractions.split(",")]:
        def random_batch(x, bi, frac=frac):
            return Batch.random(x, max(1, int(frac * x.M)), 0.8, seed=args.seed * 1000 + bi + int(frac * 1e9))
        run_batches(ctx, x0, name, frac, random_batch, args, out, summary)
    print_summary(summary)
    return 0


def run_temporal(args, out: CsvLog) -> int:
    """main.cu:157-174 on a SNAP temporal file: the first 90% of --size lines (vertices 1..--rows) form the base
    graph, a self-loop is added to every vertex, and each batch is the next fraction*size lines, as insertions
    tidied against the current graph (main.cu:84-87)."""
    if not args.file or not args.rows or not args.size:
        raise SystemExit("temporal mode needs --file, --rows and --size")
    frac = float(args.fractions.split(",")[0])
    t0 = time.time()
    rd = TemporalReader(args.file, args.rows)
    su, sv = rd.read(int(0.90 * args.size))
    x0 = HostGraph.from_edges(args.rows, su, sv, self_loops=True)
    name = args.file.rsplit("/", 1)[-1]
    log(f"graph {name}: {x0.N} vertices, {x0.M} edges (90% of {args.size} lines, self-loops), "
        f"read in {time.time() - t0:.1f} s")
    ctx = Context(0)
    autotune(ctx, x0, args)
    lines = max(1, int(frac * args.size))

    def next_batch(x, bi):
        iu, iv = rd.read(lines)
        if len(iu) == 0:
            return None
        return Batch.tidy(x, np.zeros((0, 2), np.uint32), np.stack([iu, iv], axis=1))

    summary = []
    run_batches(ctx, x0, name, frac, next_batch, args, out, summary)
    rd.close()
    print_summary(summary)
    return 0


def run_stream(args, out: CsvLog) -> int:
    x, name = make_rmat(args.scale, args.edge_factor, args.seed)
    frac = float(args.fractions.split(",")[0])
    size = max(1, int(frac * x.M))
    ctx = Context(0)
    autotune(ctx, x, args)  # the session, the separate calls and the in-place path all run the tuned plan
    xg = ctx.upload(x)
    q0 = solve(xg, engine.STATIC, PagerankOptions()).ranks
    seq = []
    cur = x
    for i in range(args.batches):
        b = Batch.random(cur, size, 0.8, seed=args.seed * 7 + i)
        y = apply_batch(cur, b)
        seq.append((b, y))
        cur = y
    log(f"{args.batches} batches of {frac:g}|E| ({size} edges each) prepared")
    # streaming session: one warm-up pass (graph capture) on a throwaway session, then the timed pass
    st = engine.Stream(xg, engine.PRUNE_DYNAMIC_FRONTIER, max_batch=size + 16, initial_ranks=q0)
    st.update(seq[0][0], want_ranks=False)
    st.close()
    st = engine.Stream(xg, engine.PRUNE_DYNAMIC_FRONTIER, max_batch=size + 16, initial_ranks=q0)
    t0 = time.perf_counter()
    stream_res = [st.update(b, want_ranks=False) for b, _ in seq]
    stream_wall = (time.perf_counter() - t0) * 1e3
    final_stream = st.ranks()
    st.close()
    if args.stream_only:  # the session alone (for a kernel trace of it)
        s_iters = sum(r.iterations for r in stream_res)
        log(f"stream: {stream_wall:.2f} ms wall for {args.batches} batches ({stream_wall / args.batches:.3f} ms/batch), "
            f"device apply+mark {sum(r.initializationTime for r in stream_res):.2f} ms, iterations "
            f"{sum(r.computationTime for r in stream_res):.2f} ms, {s_iters} iterations")
        out.row(mode="stream", graph=name, vertices=x.N, edges=x.M, batch_fraction=frac, technique="dfp_stream",
                time_ms=stream_wall, iterations=s_iters, extra=f"batches={args.batches}")
        return 0
    # separate calls: host apply (prepared above, not timed), upload, DF-P solve
    q = q0
    t0 = time.perf_counter()
    sep_solve_ms, sep_iters, worst = 0.0, 0, 0.0
    results = []
    for b, y in seq:
        yg = ctx.upload(y)
        r = solve(yg, engine.PRUNE_DYNAMIC_FRONTIER, PagerankOptions(), batch=b, q=q)
        yg.close()
        sep_solve_ms += r.time - r.preparationTime  # the new graph's layout and plan build left out
        sep_iters += r.iterations
        q = r.ranks
        results.append(r)
    sep_wall = (time.perf_counter() - t0) * 1e3
    # in-place updates of one device graph (prh_graph_update) and a DF-P solve each, not captured
    ug = ctx.upload(x)
    solve(ug, engine.STATIC, PagerankOptions())  # builds the plan the updates patch (its time is not counted)
    qu = q0
    t0 = time.perf_counter()
    upd_ms, upd_iters = 0.0, 0
    for b, _ in seq:
        upd_ms += ug.update(b)
        r = solve(ug, engine.PRUNE_DYNAMIC_FRONTIER, PagerankOptions(), batch=b, q=qu)
        upd_iters += r.iterations
        qu = r.ranks
    upd_wall = (time.perf_counter() - t0) * 1e3
    ug.close()
    worst_upd = float(np.abs(qu - q).max())
    worst = float(np.abs(final_stream - q).max())
    s_iters = sum(r.iterations for r in stream_res)
    s_apply = sum(r.initializationTime for r in stream_res)
    s_comp = sum(r.computationTime for r in stream_res)
    log(f"stream: {stream_wall:.2f} ms wall for {args.batches} batches ({stream_wall / args.batches:.3f} ms/batch); "
        f"device apply+mark {s_apply:.2f} ms, iterations {s_comp:.2f} ms, {s_iters} iterations")
    log(f"separate calls: {sep_wall:.2f} ms wall (upload + layout + solve; host-side apply not counted), "
        f"DF-P solves alone (without layout and plan builds) {sep_solve_ms:.2f} ms, {sep_iters} iterations; stream speedup {sep_wall / stream_wall:.2f}x "
        f"(vs solves alone {sep_solve_ms / stream_wall:.2f}x); L-inf(final stream - final separate) {worst:.2e}")
    log(f"in-place updates: {upd_wall:.2f} ms wall (update {upd_ms:.2f} ms + DF-P solve per batch), {upd_iters} "
        f"iterations; stream speedup {upd_wall / stream_wall:.2f}x; L-inf(final update path - final separate) "
        f"{worst_upd:.2e}")
    for tech, t, its in (("dfp_stream", stream_wall, s_iters), ("dfp_separate_calls", sep_wall, sep_iters),
                         ("dfp_solves_only", sep_solve_ms, sep_iters), ("dfp_inplace_updates", upd_wall, upd_iters)):
        out.row(mode="stream", graph=name, vertices=x.N, edges=x.M, batch_fraction=frac, technique=tech, time_ms=t,
                iterations=its, extra=f"batches={args.batches};linf={worst:.3e}")
    return 0


def seed_sets(n: int, k: int, rng: np.random.Generator, max_size: int = 1000) -> list[np.ndarray]:
    return [rng.choice(n, size=int(rng.integers(1, max_size + 1)), replace=False).astype(np.uint32)
            for _ in range(k)]


def run_ppr(args, out: CsvLog) -> int:
    g, name = make_rmat(args.scale, args.edge_factor, args.seed)
    ctx = Context(0)
    autotune(ctx, g, args)  # the plain static solves compared against use the best launch plan
    yg = ctx.upload(g, with_out=False)
    rng = np.random.default_rng(args.seed)
    sets = seed_sets(g.N, args.columns, rng)
    o = PagerankOptions(repeat=args.repeat)
    engine.ppr_batched(yg, sets[:2], PagerankOptions())  # warm-up
    batched = engine.ppr_batched(yg, sets, o)
    single_ms, single_iters, worst = 0.0, 0, 0.0
    for j, s in enumerate(sets):
        r = engine.ppr_batched(yg, [s], o)
        single_ms += r.time
        single_iters += int(r.iterations[0])
        worst = max(worst, float(np.abs(r.ranks[:, 0] - batched.ranks[:, j]).max()))
    engine.pagerankStaticHip(yg, PagerankOptions())  # warm-up: its time carries the one-time layout and plan build
    static = engine.pagerankStaticHip(yg, o)
    k = len(sets)
    cols_iters = int(batched.iterations.sum())
    log(f"batched {k} columns: {batched.time:.3f} ms, {cols_iters} column-iterations "
        f"({g.M * cols_iters / (batched.time / 1e3) / 1e9:.1f} GEdges/s/iter per column)")
    log(f"{k} single-column solves: {single_ms:.3f} ms ({g.M * single_iters / (single_ms / 1e3) / 1e9:.1f} "
        f"GEdges/s/iter); batched speedup {single_ms / batched.time:.2f}x; max |batched - single| {worst:.2e}")
    log(f"{k} plain static solves (estimated from one): {k * static.time:.3f} ms; batched speedup "
        f"{k * static.time / batched.time:.2f}x")
    for tech, t, its, extra in (("ppr_batched", batched.time, cols_iters, f"columns={k}"),
                                ("ppr_single_x{}".format(k), single_ms, single_iters, f"linf_vs_batched={worst:.3e}"),
                                ("static_x{}".format(k), k * static.time, k * static.iterations, "")):
        out.row(mode="ppr", graph=name, vertices=g.N, edges=g.M, technique=tech, time_ms=t, iterations=its,
                extra=extra)
    return 0


def run_adjoint(args, out: CsvLog) -> int:
    import torch  # before the engine library loads: torch's HIP runtime must serve the process

    from .autograd import pagerank as pagerank_autograd
    g, name = make_rmat(args.scale, args.edge_factor, args.seed)
    ctx = Context(0)
    autotune(ctx, g, args)  # the plain static solve compared against uses the best launch plan
    yg = ctx.upload(g)
    rng = np.random.default_rng(args.seed)
    w = rng.uniform(0.5, 1.5, g.M).astype(np.float32)
    s = np.full(g.N, 1.0 / g.N)
    c = rng.standard_normal(g.N)
    o = PagerankOptions(repeat=args.repeat)
    engine.weighted_pagerank(yg, w, s, PagerankOptions())  # warm-up
    fw = engine.weighted_pagerank(yg, w, s, o)
    engine.weighted_adjoint(yg, w, fw.ranks, 2 * c * fw.ranks, PagerankOptions())  # warm-up (builds its plan)
    t0 = time.perf_counter()
    reps = args.repeat
    for _ in range(reps):
        lam, gw, gs, it = engine.weighted_adjoint(yg, w, fw.ranks, 2 * c * fw.ranks, o)
    adj_ms = (time.perf_counter() - t0) / reps * 1e3  # host wall time including the host<->device copies
    # The same pair through torch.autograd on device tensors (no host copies): forward, then backward timed alone.
    dev = torch.device("cuda", 0)
    wt = torch.tensor(w, device=dev, requires_grad=True)
    st_ = torch.tensor(s, device=dev, requires_grad=True)
    ct = torch.tensor(c, device=dev)
    fwd_ms = bwd_ms = 0.0
    for rep in range(reps + 1):  # the first pass warms up
        wt.grad = st_.grad = None
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        rt = pagerank_autograd(wt, st_, yg, o)
        loss = (ct * rt * rt).sum()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        if rep:
            fwd_ms += (t2 - t1) * 1e3 / reps
            bwd_ms += (t3 - t2) * 1e3 / reps
    engine.pagerankStaticHip(yg, PagerankOptions())  # warm-up (one-time layout and plan build)
    static = engine.pagerankStaticHip(yg, o)
    log(f"forward weighted solve: {fw.time:.3f} ms, {fw.iterations} iterations "
        f"({g.M * fw.iterations / (fw.computationTime / 1e3) / 1e9:.1f} GEdges/s/iter)")
    log(f"adjoint solve + gradient assembly: {adj_ms:.3f} ms wall incl. copies, {it} iterations")
    log(f"torch.autograd on device tensors: forward {fwd_ms:.3f} ms, backward (adjoint + gradients) {bwd_ms:.3f} ms "
        f"wall")
    log(f"plain static solve: {static.time:.3f} ms, {static.iterations} iterations; forward+adjoint / static "
        f"{(fw.time + adj_ms) / static.time:.2f}x")
    for tech, t, its in (("weighted_forward", fw.time, fw.iterations), ("weighted_adjoint_wall", adj_ms, it),
                         ("autograd_forward_wall", fwd_ms, fw.iterations), ("autograd_backward_wall", bwd_ms, it),
                         ("static", static.time, static.iterations)):
        out.row(mode="adjoint", graph=name, vertices=g.N, edges=g.M, technique=tech, time_ms=t, iterations=its)
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("mode", choices=["dynamic", "temporal", "stream", "ppr", "adjoint"])
    ap.add_argument("--file", default=None, help="temporal mode: SNAP temporal edge file (1-based ids)")
    ap.add_argument("--rows", type=int, default=0, help="temporal mode: number of vertices")
    ap.add_argument("--size", type=int, default=0, help="temporal mode: number of lines (temporal edges)")
    ap.add_argument("--scale", type=int, default=20)
    ap.add_argument("--edge-factor", type=float, default=16.0)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--fractions", default="1e-5,1e-4,1e-3")
    ap.add_argument("--batches", type=int, default=3)
    ap.add_argument("--repeat", type=int, default=5)
    ap.add_argument("--upload", action="store_true",
                    help="dynamic, temporal: build each updated graph on the host and upload it (a new layout and plan "
                         "per batch) instead of updating the device graph in place")
    ap.add_argument("--columns", type=int, default=64)
    ap.add_argument("--stream-only", action="store_true",
                    help="stream mode: run the streaming session alone (no separate-call or in-place comparison)")
    ap.add_argument("--csv", default=None)
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch plan (dynamic, temporal, ppr, adjoint)")
    ap.add_argument("--tuning-cache", default=None, help="the autotuner's JSON cache")
    args = ap.parse_args(argv)
    out = CsvLog(args.csv)
    return {"dynamic": run_dynamic, "temporal": run_temporal, "stream": run_stream, "ppr": run_ppr, "adjoint": run_adjoint}[args.mode](args, out)


if __name__ == "__main__":
    sys.exit(main())
