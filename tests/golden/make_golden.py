#!/usr/bin/env python3
"""Writes tests/golden/*.npz from the reference's own CPU path (oracle/_ref/libprh_ref.so, compiled from the
reference headers by oracle/Makefile). Run in the container that has the reference sources:

    make -C oracle ref && python tests/golden/make_golden.py

Each fixture holds a small graph x (out-CSR), a tidied batch, and what the reference computes on them:
  static_*  pagerankStaticOmp (synchronous)             on y = x + batch
  ND/DT/DF/DFP_*  the synchronous CPU dynamic solvers   on (x, y, batch, q = static ranks of x)
  async_{static,ND,DT,DF,DFP}_*  the asynchronous OpenMP solvers (ASYNC=true, as main.cu runs the dynamic ones)
                  on one thread, where their in-place sweep order is deterministic
  affected_df / affected_dt  pagerankAffected{Frontier,Traversal}OmpW flags
  y_off / y_edg  tidyBatchUpdateU + applyBatchUpdateOmpU result
The fixtures are data only (inputs and expected outputs)."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
sys.path.insert(0, str(REPO / "pagerank-cuda-dynamic_amd"))
sys.path.insert(0, str(REPO / "oracle"))

import oracle  # noqa: E402
from pagerank_amd import Batch, HostGraph  # noqa: E402

CASES = {
    "rmat8": dict(kind="rmat", scale=8, ef=8, seed=1, batch=40, bseed=3),
    "rmat10": dict(kind="rmat", scale=10, ef=16, seed=42, batch=130, bseed=5),
    "grid24": dict(kind="grid", rows=24, cols=24, seed=7, batch=30, bseed=9),
}


def build(case):
    if case["kind"] == "rmat":
        return HostGraph.rmat(case["scale"], case["ef"], seed=case["seed"])
    return HostGraph.grid(case["rows"], case["cols"], 0.62, seed=case["seed"])


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/libprh_ref.so is missing: run `make -C oracle ref` first")
    for name, case in CASES.items():
        x = build(case)
        b = Batch.random(x, case["batch"], 0.8, seed=case["bseed"])
        y_off, y_edg = oracle.ref_apply_batch(x, b)
        empty = Batch.empty()
        out = dict(x_off=x.out_off, x_edg=x.out_edg, del_u=b.del_u, del_v=b.del_v, ins_u=b.ins_u, ins_v=b.ins_v,
                   y_off=y_off, y_edg=y_edg)
        q, qit, _ = oracle.ref_pagerank(0, x, empty)
        out["q"], out["q_iterations"] = q, qit
        r, it, _ = oracle.ref_pagerank(0, x, b)
        out["static_ranks"], out["static_iterations"] = r, it
        for m, tag in ((1, "ND"), (2, "DT"), (3, "DF"), (4, "DFP")):
            r, it, _ = oracle.ref_pagerank(m, x, b, q)
            out[f"{tag}_ranks"], out[f"{tag}_iterations"] = r, it
        for m, tag in ((0, "static"), (1, "ND"), (2, "DT"), (3, "DF"), (4, "DFP")):
            r, it, _ = oracle.ref_pagerank(m, x, b, None if m == 0 else q, asynchronous=True, threads=1)
            out[f"async_{tag}_ranks"], out[f"async_{tag}_iterations"] = r, it
        out["affected_df"] = oracle.ref_affected(0, x, b)
        out["affected_dt"] = oracle.ref_affected(1, x, b)
        np.savez_compressed(HERE / f"{name}.npz", **out)
        print(name, x.N, x.M, "batch", b.deletions, b.insertions, "static it", out["static_iterations"])


if __name__ == "__main__":
    main()
