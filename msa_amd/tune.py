"""Autotuner command line.

Times, on the current GPU, every batched tile shape (lanes per pair x rows per
lane) and, for long queries, the long kernel's pipelined strips, per
query-length bin of one or more workloads (msa_nw_tune in the library), and
writes the tuning table that msa_nw_options.tuning_file reads:

  python -m msa_amd.tune --workload uneven,dna --out msa_amd/tuning/gfx950.json

With several workloads each is tuned on its own and the tables are merged
bin by bin: a bin measured in several workloads takes the shape with the
highest rate summed over their measurements, and the long threshold is the
lowest one at which some workload measured the long scheme winning.
"""
import argparse
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from msa_amd import _native as N  # noqa: E402
from msa_amd import workloads as W  # noqa: E402

DEFAULT_LONG_THRESHOLD = 8192


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
    rates = {}
    for t in tables:
        for m in t["measurements"]:
            if m["scheme"] != "batched":
                continue
            key = (m["min_query"], m["lanes"], m["rows"])
            rates[key] = rates.get(key, 0.0) + m["gcups"]
    edges = sorted({k[0] for k in rates})
    bins = []
    for lo in edges:
        best = max((v, k) for k, v in rates.items() if k[0] == lo)[1]
        if not bins or (bins[-1]["lanes"], bins[-1]["rows"]) != best[1:]:
            bins.append({"min_query": lo, "lanes": best[1], "rows": best[2]})
    bins[0]["min_query"] = 0
    wins = [t["long_threshold"] for t in tables
            if any(m["scheme"] == "long" and m["min_query"] >= t["long_threshold"] for m in t["measurements"])]
    merged = dict(tables[0])
    merged["bins"] = bins
    merged["long_threshold"] = min(wins) if wins else max(DEFAULT_LONG_THRESHOLD,
                                                           min(t["long_threshold"] for t in tables))
    merged["measurements"] = [dict(m, workload=t.get("workload", "")) for t in tables for m in t["measurements"]]
    merged["workloads"] = [t.get("workload", "") for t in tables]
    return merged


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--workload", default="uneven,dna", help="comma-separated: uneven, uniform, dna, small")
    ap.add_argument("--budget-cells", type=float, default=4e10, help="sample size per bin and candidate")
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
            print("  %-7s q[%5d,%5d] %2d lanes x %2d rows: %9.1f GCUPS" % (
                m["scheme"], m["min_query"], m["max_query"], m["lanes"], m["rows"], m["gcups"]),
                file=sys.stderr, flush=True)
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
