"""Autotuner command line.

Picks, on the current GPU, a batched tile shape (lanes per pair x rows per
lane) per query-length bin and the query length from which pairs take the long
kernel's pipelined strips, by coordinate descent over timed executions of one
or more workloads (msa_nw_tune in the library), and writes the tuning table
that msa_nw_options.tuning_file reads:

  python -m msa_amd.tune --workload uneven,dna --out msa_amd/tuning/gfx950.json

With several workloads each is tuned on its own and the tables are merged:
a bin takes the shape chosen by the workload that puts the most cells in it,
and the long threshold is the candidate with the least summed slowdown over
the workloads' threshold trials (within 1%, the one nearest the built-in 8192).
"""
import argparse
import json
import math
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from msa_amd import _native as N  # noqa: E402
from msa_amd import workloads as W  # noqa: E402

BUILTIN_LONG_THRESHOLD = 8192    # the library's built-in table


def workload(name, tables):
    if name == "uneven":
        return W.uneven_batch(tables)
    if name == "uniform":
        return W.uniform_batch(tables)
    if name == "dna":
        return W.long_dna()
    if name == "small":
        return W.uniform_small(tables)
    raise SystemExit("unknown workload %r" % name)


def merge(tables):
    """Merges tuning tables measured on different workloads (see module doc)."""
    if len(tables) == 1:
        return tables[0]
    edges = [b["min_query"] for b in tables[0]["bin_cells"]]
    shapes = []
    for i, lo in enumerate(edges):
        best = max(tables, key=lambda t: t["bin_cells"][i]["cells"])["bin_cells"][i]
        shape = (best["lanes"], best["rows"]) if best["cells"] > 0 else (shapes[-1] if shapes else None)
        shapes.append(shape)
    first = next((s for s in shapes if s), (16, 32))
    shapes = [s or first for s in shapes]
    bins = []
    for lo, shape in zip(edges, shapes):
        if not bins or (bins[-1]["lanes"], bins[-1]["rows"]) != shape:
            bins.append({"min_query": lo, "lanes": shape[0], "rows": shape[1]})
    bins[0]["min_query"] = 0
    # Long threshold: the candidate with the least summed slowdown over the
    # workloads; a threshold a workload did not try moves none of its pairs.
    curves = [{c["long_threshold"]: c["ms"] for c in t["threshold_trials"]} for t in tables]
    candidates = sorted({th for c in curves for th in c})
    never = max(candidates)

    def slowdown(th):
        return sum(c.get(th, c[never]) / min(c.values()) for c in curves)

    # Candidates within 1% of the best count as ties, settled towards the
    # built-in threshold so that noise does not move it.
    best = min(slowdown(th) for th in candidates)
    ties = [th for th in candidates if slowdown(th) <= best * 1.01]
    long_threshold = min(ties, key=lambda th: abs(math.log(th / BUILTIN_LONG_THRESHOLD)))
    merged = dict(tables[0])
    merged["bins"] = bins
    merged["long_threshold"] = long_threshold
    merged["workloads"] = [{"workload": t.get("workload", ""), "sample": t["sample"], "bins": t["bins"],
                            "long_threshold": t["long_threshold"], "bin_cells": t["bin_cells"],
                            "threshold_trials": t["threshold_trials"], "measurements": t["measurements"]}
                           for t in tables]
    for key in ("sample", "bin_cells", "threshold_trials", "measurements", "workload"):
        merged.pop(key, None)
    return merged


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--workload", default="uneven,dna", help="comma-separated: uneven, uniform, dna, small")
    ap.add_argument("--budget-cells", type=float, default=0, help="sample size in cells (0 = the library's 2e12)")
    ap.add_argument("--out", default=None, help="tuning table to write (JSON)")
    args = ap.parse_args()
    if N.device_count() < 1:
        raise SystemExit("no GPU visible")
    tables = {n: N.table(n) for n in N.table_names()}
    results = []
    for name in args.workload.split(","):
        w = workload(name, tables)
        print("tuning on %s (%d sequences, %d pairs)" % (w.name, w.nseq, w.npairs), file=sys.stderr, flush=True)
        t = N.tune(w.blocks, w.offsets, w.sizes, w.table, w.penalty, budget_cells=int(args.budget_cells))
        t["workload"] = w.name
        for m in t["measurements"]:
            print("  %-34s long from %10d: %9.3f ms %9.1f GCUPS" % (m["step"], m["long_threshold"], m["ms"],
                                                                    m["gcups"]), file=sys.stderr, flush=True)
        results.append(t)
    table = merge(results)
    text = json.dumps(table, indent=1)
    if args.out:
        with open(args.out, "w") as f:
            f.write(text + "\n")
    print(text)
    return 0


if __name__ == "__main__":
    sys.exit(main())
