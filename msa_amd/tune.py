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

The table also records the tile grid {32, 64, 128} rows per lane x {16, 32, 64}
lanes per pair ("grid"): every kernel shape the library has, timed on each
workload with every pair forced onto it (scores checked equal to the tuned
plan's), and each grid point either mapped to one of those shapes or marked
dominated, with the reason and the measurements and compiler resource counts
(kernel_resources.json, written by the build) that show it.
"""
import argparse
import ctypes
import json
import math
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import numpy as np  # noqa: E402

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


GRID_WIDTHS = (32, 64, 128)     # rows of the query each lane holds
GRID_LANES = (16, 32, 64)       # lanes scoring one pair
LONG_LANES = 64                 # the long kernel: one wave per strip
LONG_LDS_FIXED = 1472           # the long kernel's LDS besides the profile: slot ring, line ring, table, flags


def time_plan(w, opts, reps=3):
    """(best ms of `reps` executions after one warm-up, plan info, scores)."""
    L = N.lib()
    stream = ctypes.c_void_p()
    N.check(L.msa_stream_create(ctypes.byref(stream)))
    ev = [ctypes.c_void_p(), ctypes.c_void_p()]
    for e in ev:
        N.check(L.msa_event_create(ctypes.byref(e)))
    out = N.DeviceBuffer(4 * w.npairs)
    plan = N.Plan(w.blocks, w.offsets, w.sizes, w.table, w.penalty, opts=opts, stream=stream.value)
    try:
        plan.execute(out.ptr, stream.value)
        best = float("inf")
        for _ in range(reps):
            N.check(L.msa_event_record(ev[0], stream))
            plan.execute(out.ptr, stream.value)
            N.check(L.msa_event_record(ev[1], stream))
            N.check(L.msa_event_sync(ev[1]))
            ms = ctypes.c_float()
            N.check(L.msa_event_elapsed_ms(ev[0], ev[1], ctypes.byref(ms)))
            best = min(best, ms.value)
        plan.status(stream.value)
        info = plan.info()
        scores = out.download(np.int32, w.npairs)
    finally:
        plan.close()
        out.close()
        for e in ev:
            L.msa_event_destroy(e)
        L.msa_stream_destroy(stream)
    return best, info, scores


def long_rows(w):
    """Rows per lane the long kernel takes for this workload's alphabet: its
    profile keeps one slot per code present plus the neutral slot."""
    present = set()
    for o, s in zip(w.offsets, w.sizes):
        present.update(np.unique(W.unpack(w.blocks, int(o), int(s))).tolist())
    slots = len(present - {24}) + 1
    return (32 if slots <= 8 else 16 if slots <= 16 else 8), slots


def resources():
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kernel_resources.json")
    try:
        with open(path) as f:
            kernels = json.load(f)["kernels"]
    except (OSError, ValueError, KeyError):
        return {}
    found = {}
    for name, r in kernels.items():
        for key, pat in (("batched_16x32", "nw_batched_kernel<0>"), ("batched_16x16", "nw_batched_kernel<1>"),
                         ("batched_32x16", "nw_batched_kernel<2>"), ("long_8", "nw_long_kernel<8>"),
                         ("long_16", "nw_long_kernel<16>"), ("long_32", "nw_long_kernel<32>"),
                         ("plain", "nw_plain_kernel<")):
            if pat in name:
                found[key] = r
    return found


def measure_shapes(w, tuning=None):
    """Every kernel shape timed on `w` with all pairs forced onto it, each
    checked equal, pair for pair, to the tuned plan's scores."""
    shapes = [("auto", N.options(N.KERNEL_AUTO, tuning_file=tuning)),
              ("batched_16x32", N.options(N.KERNEL_BATCHED, lanes=16, rows=32)),
              ("batched_16x16", N.options(N.KERNEL_BATCHED, lanes=16, rows=16)),
              ("batched_32x16", N.options(N.KERNEL_BATCHED, lanes=32, rows=16)),
              ("long", N.options(N.KERNEL_LONG)),
              ("plain", N.options(N.KERNEL_PLAIN))]
    cells = w.cells()
    rows, slots = long_rows(w)
    out, ref = {}, None
    for name, opts in shapes:
        ms, info, scores = time_plan(w, opts)
        if ref is None:
            ref = scores
        rec = {"ms": round(ms, 4), "gcups": round(cells / ms / 1e6, 1), "equal_to_auto": bool(np.array_equal(scores, ref)),
               "routing": {"batched": int(info.batched_pairs), "long": int(info.long_pairs),
                           "plain": int(info.plain_pairs)}}
        if name == "long":
            rec["rows_per_lane"] = rows
            rec["profile_slots"] = slots
        out[name] = rec
        print("  grid %-14s %10.3f ms %9.1f GCUPS equal %s" % (name, ms, rec["gcups"], rec["equal_to_auto"]),
              file=sys.stderr, flush=True)
    return out


def grid_map(measured, res):
    """The tile grid point by point: the shape that realises it, or why it is
    dominated, with the measurements (per workload) and resource counts."""
    def vg(key):
        return res.get(key, {}).get("vgprs")

    def ratio(name_a, name_b):
        return {wl: round(m[name_a]["ms"] / m[name_b]["ms"], 4) for wl, m in measured.items()}

    lanes32_cost = ratio("batched_32x16", "batched_16x16")      # 32-lane groups vs 16 at 16 rows
    rows32_gain = ratio("batched_16x32", "batched_16x16")       # 32 rows vs 16 at 16 lanes
    per_row = None
    if vg("batched_16x32") and vg("batched_16x16"):
        per_row = (vg("batched_16x32") - vg("batched_16x16")) / 16.0
    long_per_row = None
    if vg("long_32") and vg("long_16"):
        long_per_row = (vg("long_32") - vg("long_16")) / 16.0
    lds_per_cu = 160 * 1024
    points = []
    for width in GRID_WIDTHS:
        for lanes in GRID_LANES:
            p = {"rows_per_lane": width, "lanes_per_pair": lanes}
            if lanes == 16 and width == 32:
                p.update(status="measured", shape="batched_16x32",
                         timings={wl: m["batched_16x32"] for wl, m in measured.items()})
            elif lanes == LONG_LANES and width == 32:
                p.update(status="measured where the alphabet allows it", shape="long",
                         reason="the long kernel is the 64-lane shape; it holds 32 rows per lane when the data has "
                                "at most 7 codes (its LDS profile is slots x 64 x R bytes per wave), 16 up to 15 "
                                "codes, else 8: the protein workloads run it at 8 rows per lane",
                         timings={wl: m["long"] for wl, m in measured.items()})
            else:
                p["status"] = "dominated"
                profile = lanes * width                       # bytes per code of a lane group's profile
                if lanes < LONG_LANES:
                    est = None if per_row is None else round(vg("batched_16x16") + per_row * (width - 16) +
                                                             (vg("batched_32x16") - vg("batched_16x16")) *
                                                             (lanes // 16 - 1))
                    p["reason"] = ("the batched kernel's LDS keeps one 512-byte profile block per code per wave; "
                                   "%d lanes x %d rows needs %d bytes per code (25 codes x 4 waves = %d KiB per "
                                   "workgroup against 160 KiB per CU)%s" % (
                                       lanes, width, profile, 100 * profile // 1024,
                                       "" if est is None else "; registers grow %.1f VGPRs per row of a lane "
                                       "(16x16 -> 16x32), so about %d VGPRs%s" % (
                                           per_row, est, ", beyond the 256 a wave can address" if est > 256
                                           else ", at most %d waves per SIMD" % max(1, 512 // est))))
                    p["evidence"] = {"ms_32x16_over_16x16": lanes32_cost, "ms_16x32_over_16x16": rows32_gain,
                                     "estimated_vgprs": est}
                else:
                    est = None if long_per_row is None else round(vg("long_32") + long_per_row * (width - 32))
                    per_wave_dna = 5 * 64 * width + LONG_LDS_FIXED
                    p["reason"] = ("the long kernel at %d rows per lane: its profile is slots x 64 x %d bytes per "
                                   "wave, %d KiB for DNA (5 slots: at most %d waves per CU, fewer than the 8 its "
                                   "pipelining keeps) and %d KiB for protein (21 slots)%s" % (
                                       width, width, per_wave_dna // 1024, lds_per_cu // per_wave_dna,
                                       21 * 64 * width // 1024,
                                       "" if est is None else "; registers grow %.1f VGPRs per row (16 -> 32), so "
                                       "about %d VGPRs%s" % (long_per_row, est,
                                                             ", beyond the 256 a wave can address" if est > 256
                                                             else "")))
                    p["evidence"] = {"long_timings": {wl: m["long"] for wl, m in measured.items()},
                                     "estimated_vgprs": est}
            points.append(p)
    return {"axes": {"rows_per_lane": list(GRID_WIDTHS), "lanes_per_pair": list(GRID_LANES)},
            "shapes_measured": measured, "resources": res, "points": points}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--workload", default="uneven,dna", help="comma-separated: uneven, uniform, dna, small")
    ap.add_argument("--budget-cells", type=float, default=0, help="sample size in cells (0 = the library's 2e12)")
    ap.add_argument("--out", default=None, help="tuning table to write (JSON)")
    ap.add_argument("--no-grid", dest="grid", action="store_false",
                    help="skip timing every kernel shape over the tile grid")
    args = ap.parse_args()
    if N.device_count() < 1:
        raise SystemExit("no GPU visible")
    tables = {n: N.table(n) for n in N.table_names()}
    results = []
    loaded = []
    for name in args.workload.split(","):
        w = workload(name, tables)
        print("tuning on %s (%d sequences, %d pairs)" % (w.name, w.nseq, w.npairs), file=sys.stderr, flush=True)
        t = N.tune(w.blocks, w.offsets, w.sizes, w.table, w.penalty, budget_cells=int(args.budget_cells))
        t["workload"] = w.name
        for m in t["measurements"]:
            print("  %-34s long from %10d: %9.3f ms %9.1f GCUPS" % (m["step"], m["long_threshold"], m["ms"],
                                                                    m["gcups"]), file=sys.stderr, flush=True)
        results.append(t)
        loaded.append(w)
    table = merge(results)
    if args.grid:
        # The shapes under the merged table, on every workload.
        import tempfile
        with tempfile.NamedTemporaryFile("w", suffix=".json", delete=False) as f:
            f.write(json.dumps(table))
        try:
            measured = {w.name: measure_shapes(w, tuning=f.name) for w in loaded}
        finally:
            os.unlink(f.name)
        table["grid"] = grid_map(measured, resources())
    text = json.dumps(table, indent=1)
    if args.out:
        with open(args.out, "w") as f:
            f.write(text + "\n")
    print(text)
    return 0


if __name__ == "__main__":
    sys.exit(main())
