#!/usr/bin/env python3
"""All-pairs Needleman-Wunsch throughput on one MI355X.

One step is one pass of the pairwise stage over one batch: every pair of the
uneven-length database (2000 protein sequences, lengths clip(round(
lognormal(ln 400, 1.0)), 50, 5000), BLOSUM62 background residues, blosum62
with the reference's gap penalty 4, seed 3) scored by the batched persistent
kernel from inputs already resident in HBM, int32 scores written in reference
pair order.  The same pairs are also scored by the plain one-pair-per-
workgroup kernel, both results are compared pair by pair, and a random sample
of pairs is checked against the CPU oracle.

Prints one JSON line (rank 0, one GPU).  GCUPS = sum over pairs of the
unpadded L1*L2 cells / seconds.  The roofline object prices the batched
execution against the gfx950 VALU peak (256 CU x 64 lanes x 2.4 GHz): a cell
of the shifted recurrence needs at least 2 int32 VALU operations (one add,
one max3), a packed 16-bit cell 1.5 (one byte permute, one packed add and
one packed max3 for two cells of two targets), a duo cell 1 (one packed add
and one packed max3 for two cells of two queries: no permute), so the floor
is 2 x int32 cells + 1.5 x packed cells + 1 x duo cells lane-operations per
execution.  `int32_equivalent_frac` prices every cell at 2.

Usage: python bench.py [--gpus 1] [--steps 5] [--warmup 2] [--workload uneven]
"""
import argparse
import ctypes
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "all-pairs NW GCUPS on 1 MI355X (batched vs plain, bit-exact) + serving p50/p99 latency"
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12     # int32 lane-ops per second, in T
OPS_PER_CELL = 2
PACKED_OPS_PER_CELL = 1.5
DUO_OPS_PER_CELL = 1.0
DEFAULT_TUNING = os.path.join(ROOT, "msa_amd", "tuning", "gfx950.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(name, tables):
    from msa_amd import workloads as W
    if name == "uneven":
        return W.uneven_batch(tables)
    if name == "uniform":
        return W.uniform_batch(tables)
    if name == "small":
        return W.uniform_small(tables)
    if name == "dna":
        return W.long_dna()
    if re.fullmatch(r"case[1-6]", name):
        return W.reference_case(int(name[4:]), tables)
    raise SystemExit("unknown workload %r" % name)


class Timer:
    """hipEvent pair on one stream."""

    def __init__(self, N, stream):
        self.N, self.L, self.stream = N, N.lib(), stream
        self.a, self.b = ctypes.c_void_p(), ctypes.c_void_p()
        N.check(self.L.msa_event_create(ctypes.byref(self.a)))
        N.check(self.L.msa_event_create(ctypes.byref(self.b)))

    def start(self):
        self.N.check(self.L.msa_event_record(self.a, self.stream))

    def stop(self):
        self.N.check(self.L.msa_event_record(self.b, self.stream))
        self.N.check(self.L.msa_event_sync(self.b))
        ms = ctypes.c_float()
        self.N.check(self.L.msa_event_elapsed_ms(self.a, self.b, ctypes.byref(ms)))
        return ms.value


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` on `workload` as the committed rocprofv3
    counter summary states them (a line `traffic <workload> <kernel> <bytes>`),
    or None when no summary covers the pair."""
    folder = os.path.join(ROOT, "profiles")
    if not os.path.isdir(folder):
        return None
    for name in sorted(os.listdir(folder)):
        if not name.endswith(".md"):
            continue
        with open(os.path.join(folder, name)) as f:
            for line in f:
                parts = line.strip().strip("`").split()
                if len(parts) == 4 and parts[0] == "traffic" and parts[1] == workload and parts[2] == kernel:
                    return int(parts[3])
    return None


def run_plan(N, w, kernel, steps, warmup, stream, out, tuning=None):
    L = N.lib()
    t0 = time.time()
    plan = N.Plan(w.blocks, w.offsets, w.sizes, w.table, w.penalty,
                  opts=N.options(kernel, use_graph=True, tuning_file=tuning), stream=stream)
    plan_s = time.time() - t0
    for _ in range(warmup):
        plan.execute(out.ptr, stream)
    N.check(L.msa_device_sync())
    timer = Timer(N, stream)
    wall0 = time.time()
    timer.start()
    for _ in range(steps):
        plan.execute(out.ptr, stream)
    ms = timer.stop()
    N.check(L.msa_device_sync())
    wall = time.time() - wall0
    info = plan.info()
    scores = out.download(np.int32, w.npairs)
    plan.close()
    return {"ms": ms / steps, "wall_ms": 1e3 * wall / steps, "plan_s": plan_s, "info": info, "scores": scores}


def pair_index(c):
    from msa_amd import workloads as W
    return W.pair_of(c)


def host_cpus():
    """What the host offers the CPU baseline: os.cpu_count(), the affinity
    mask, the cgroup CPU quota (cpu.max, when one is set), and the CPU model.
    The cores used are the affinity mask, capped by the quota."""
    nproc = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = nproc
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
        if q != "max":
            quota = max(1, int(-(-int(q) // int(period))))
    except (OSError, ValueError):
        pass
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    cores = min(affinity, quota) if quota else affinity
    return {"nproc": nproc, "affinity": affinity, "cgroup_quota": quota, "model": model, "cores": cores}


def oracle_sample(w, scores, count, seed=11, max_cells=2e10):
    """Checks a random sample of pairs (at most `count`, and about `max_cells`
    cells of CPU work) against the oracle."""
    from oracle import museqa_ref as M
    count = int(max(8, min(count, max_cells // max(1.0, w.cells() / w.npairs))))
    idx = np.random.default_rng(seed).choice(w.npairs, min(count, w.npairs), replace=False)
    pairs = np.array([pair_index(int(c)) for c in idx], dtype=np.int32)
    exp = M.CpuOracle().align(w.blocks, w.offsets, w.sizes, w.table.astype(np.float32), w.penalty,
                              pairs=pairs, workers=min(16, os.cpu_count() or 1))
    return bool(np.array_equal(scores[idx].astype(np.float32), exp)), len(idx)


def cpu_baseline(w, seconds, seed=12):
    """The reference's sequential path on the host's cores over a bounded
    random sample of pairs of the same workload, in the reference's topology:
    W = C workers and W = C - 1 (C MPI ranks with an idle master,
    sequential.cpp:116, needleman.cpp:36-37), each worker on a contiguous
    share of the pair list (utils::partition, src/utils.hpp:176-183)."""
    from oracle import museqa_ref as M
    rng = np.random.default_rng(seed)
    lengths = w.lengths
    mean_cells = w.cells() / w.npairs
    cpus = host_cpus()
    C = cpus["cores"]
    use_ref = M.ReferenceBuild.available()
    engine = M.ReferenceBuild() if use_ref else M.CpuOracle()

    def run(pairs, workers):
        t0 = time.time()
        if use_ref:
            engine.align_list(w.blocks, w.offsets, w.sizes, pairs, w.table.astype(np.float32), w.penalty,
                              workers=workers)
        else:
            engine.align(w.blocks, w.offsets, w.sizes, w.table.astype(np.float32), w.penalty, pairs=pairs,
                         workers=workers)
        return time.time() - t0

    def sample(n):
        # Small workloads are sampled with repetition to fill the time budget.
        idx = rng.choice(w.npairs, n, replace=n > w.npairs)
        return np.array([pair_index(int(c)) for c in idx], dtype=np.int32)

    # Calibrate on a small sample, then size the timed sample to ~`seconds` at W = C.
    probe = sample(max(C * 4, 64))
    probe_rate = (lengths[probe[:, 0]] * lengths[probe[:, 1]]).sum() / max(run(probe, C), 1e-6)
    n = int(max(C * 8, seconds * probe_rate / mean_cells))
    pairs = sample(n)
    cells = int((lengths[pairs[:, 0]] * lengths[pairs[:, 1]]).sum())
    legs = []
    for workers in (C, max(C - 1, 1)):
        dt = run(pairs, workers)
        legs.append({"workers": workers, "gcups": round(cells / dt / 1e9, 4), "wall_s": round(dt, 2)})
    best = max(legs, key=lambda leg: leg["gcups"])
    return {"value": best["gcups"], "unit": "GCUPS", "cores": best["workers"],
            "kind": "reference" if use_ref else "port", "legs": legs, "host": cpus,
            "sample": "%d random pairs of the same workload (%.3e cells), timed at W = C = %d workers and at "
                      "W = C - 1 (the reference's topology: an idle master rank); `value` is the faster leg; one "
                      "thread per worker on contiguous shares as the reference's MPI ranks take them%s" % (
                          n, cells, C, "; the reference's own align() (oracle/_ref)" if use_ref else "")}


def serving_run(N, tables, rates, seconds, pool_size=48, slots=4, seed=5, log_fn=log, jobs=0, resident=False):
    """Poisson job streams through the serving loop.

    Jobs: 8..256 protein sequences (count uniform), lengths uniform in
    [100, 2000], BLOSUM62 background residues, blosum62 (seed 5); a pool of
    `pool_size` jobs is cycled.  Arrivals: Poisson at each rate (seed 50).
    Latency = scores on the host - scheduled (Poisson) arrival: the job's
    completion on the device plus the score copy of msa_srv_fetch, and the
    time the submitter thread fell behind its schedule before handing a job
    over, so a backlog shows in p50/p99.  Every job's scores are compared
    with the same job scored alone through msa_nw_align.  `jobs` > 0 fixes
    the number of jobs per rate instead of `seconds`.  `resident`: the pool's
    sequences are put into the server's store once, before the clock starts,
    and every job names them by id (msa_srv_submit_resident).
    """
    import threading
    from msa_amd import workloads as W
    pool = W.serving_jobs(tables, pool_size, seed=seed)
    T, p = tables["blosum62"]
    alone = [N.align(j.blocks, j.offsets, j.sizes, T, p) for j in pool]
    out = []
    for rate in rates:
        count = jobs if jobs > 0 else max(8, int(rate * seconds))
        arrivals = W.poisson_arrivals(rate, count, seed=50)
        srv = N.Server(T, p, slots=slots, max_length=2048, device_bytes=4 << 30, host_bytes=512 << 20)
        stored = [np.arange(len(j.sizes)) + srv.store_add(j.blocks, j.offsets, j.sizes) for j in pool] \
            if resident else None
        ids = [None] * count
        errors = []
        t0 = time.time()

        late = [0.0] * count

        def submitter():
            try:
                for k in range(count):
                    delay = arrivals[k] - (time.time() - t0)
                    if delay > 0:
                        time.sleep(delay)
                    # How far behind its scheduled arrival this job is handed
                    # over (the submitter plans jobs on its own thread).
                    late[k] = max(0.0, (time.time() - t0) - arrivals[k])
                    if resident:
                        ids[k] = srv.submit_resident(stored[k % pool_size])
                    else:
                        j = pool[k % pool_size]
                        ids[k] = srv.submit(j.blocks, j.offsets, j.sizes)
            except Exception as e:  # surfaced below
                errors.append(e)

        th = threading.Thread(target=submitter)
        th.start()
        lat, cells, exact, first, last = [], 0, True, None, None
        uploaded = seq_uploaded = 0
        for k in range(count):
            while ids[k] is None and th.is_alive():
                time.sleep(0.0002)
            if ids[k] is None:
                break
            srv.wait(ids[k], 120000)
            scores, st = srv.fetch(ids[k])
            exact = exact and bool(np.array_equal(scores, alone[k % pool_size]))
            lat.append((st["finished_us"] - st["arrival_us"] + st["fetch_us"]) / 1e3 + 1e3 * late[k])
            uploaded += st["upload_bytes"]
            seq_uploaded += st["sequence_bytes"]
            cells += st["cells"]
            first = st["arrival_us"] if first is None else min(first, st["arrival_us"])
            last = st["finished_us"] if last is None else max(last, st["finished_us"])
        th.join()
        srv.close()
        if errors:
            raise errors[0]
        span_s = max((last - first) / 1e6, 1e-9)
        rec = {"rate": rate, "jobs": len(lat), "jobs_per_s": round(len(lat) / span_s, 1),
               "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
               "gcups": round(cells / span_s / 1e9, 1), "bit_exact_vs_alone": exact,
               "submit_lag_max_ms": round(1e3 * float(max(late[:len(lat)] or [0.0])), 3),
               "sequences": "resident" if resident else "uploaded per job",
               "upload_bytes_per_job": int(uploaded / max(len(lat), 1)),
               "sequence_bytes_per_job": int(seq_uploaded / max(len(lat), 1))}
        log_fn("serving rate %g: %s" % (rate, rec))
        out.append(rec)
    return out


def measure(N, w, steps, warmup, plain_steps, oracle_pairs, stream, out, tuning, oracle_cells=2e10):
    """The auto plan (timed as hipGraph replays) and the plain kernel on one
    workload, both outputs compared on every pair, plus an oracle sample."""
    cells = w.cells()
    batched = run_plan(N, w, N.KERNEL_AUTO, steps, warmup, stream, out, tuning=tuning)
    log("auto plan: %.2f ms/step, %.1f GCUPS" % (batched["ms"], cells / batched["ms"] / 1e6))
    plain = run_plan(N, w, N.KERNEL_PLAIN, plain_steps, 1, stream, out)
    log("plain:     %.2f ms/step, %.1f GCUPS" % (plain["ms"], cells / plain["ms"] / 1e6))
    same = bool(np.array_equal(batched["scores"], plain["scores"]))
    oracle_ok, oracle_n = oracle_sample(w, batched["scores"], oracle_pairs, max_cells=oracle_cells)
    log("auto == plain on all pairs:", same, "| oracle sample of", oracle_n, "pairs:", oracle_ok)
    return batched, plain, same, oracle_ok, oracle_n


def valu_roofline(w, info, ms):
    """Roofline of one batched execution: the floor of VALU lane-operations
    (2 per int32 cell, 1.5 per packed cell, 1 per duo cell) over its time,
    against the peak."""
    cells = w.cells()
    packed = int(getattr(info, "packed_cells", 0))
    duo = int(getattr(info, "duo_cells", 0))
    ops = OPS_PER_CELL * (cells - packed) + PACKED_OPS_PER_CELL * (packed - duo) + DUO_OPS_PER_CELL * duo
    achieved = ops / (ms * 1e-3) / 1e12
    kernel = dominant_kernel(info)
    return {"bound": "valu", "kernel": kernel, "achieved": round(achieved, 3), "peak": round(VALU_PEAK_TOPS, 2),
            "unit": "TOP/s", "frac": round(achieved / VALU_PEAK_TOPS, 4), "traffic": measured_traffic(w.name, kernel),
            "ops_per_cell": round(ops / cells, 4), "packed_cells": packed, "duo_cells": duo,
            "int32_equivalent_frac": round(OPS_PER_CELL * cells / (ms * 1e-3) / 1e12 / VALU_PEAK_TOPS, 4)}


def config_record(w, batched, plain, same, oracle_ok, oracle_n):
    cells = w.cells()
    info = batched["info"]
    return {"description": w.description, "pairs": w.npairs, "cells": cells, "input_checksum": w.checksum(),
            "gcups": round(cells / batched["ms"] / 1e6, 2), "ms_per_step": round(batched["ms"], 3),
            "plain_gcups": round(cells / plain["ms"] / 1e6, 2),
            "auto_vs_plain": round(plain["ms"] / batched["ms"], 3),
            "routing": {"batched_pairs": int(info.batched_pairs), "long_pairs": int(info.long_pairs),
                        "plain_pairs": int(info.plain_pairs)},
            "roofline": valu_roofline(w, info, batched["ms"]),
            "bit_exact": {"auto_equals_plain_all_pairs": same, "oracle_sample_pairs": oracle_n,
                          "oracle_sample_equal": oracle_ok}}


def tuning_summary(path):
    if not path:
        return "built-in (16 lanes x 32 rows, long from 8192)"
    with open(path) as f:
        t = json.load(f)
    return {"file": os.path.relpath(path, ROOT), "bins": t["bins"], "long_threshold": t["long_threshold"]}


def dominant_kernel(info):
    """The kernel family that scores most of the plan's pairs' cells (by pairs)."""
    counts = {"nw_batched_kernel": info.batched_pairs, "nw_long_kernel": info.long_pairs,
              "nw_plain_kernel": info.plain_pairs}
    return max(counts, key=lambda k: counts[k])


def phylogeny_run(N, w, steps, warmup, cpu=True, variant=0):
    """Neighbor joining over the workload's pairwise scores, both on the GPU:
    the scores stay in device memory and msa_nj_build_device builds the tree.
    The oracle builds the same tree on one host core (the CPU baseline, kind
    "port") and the two trees must be identical."""
    from oracle import museqa_ref as M
    L = N.lib()
    n = w.nseq
    plan = N.Plan(w.blocks, w.offsets, w.sizes, w.table, w.penalty)
    d_scores = N.DeviceBuffer(4 * w.npairs)
    plan.execute(d_scores.ptr)
    plan.status()
    d_nodes = N.DeviceBuffer(N.NJ_NODE.itemsize * (2 * n - 1))
    for _ in range(warmup):
        N.nj_build_device(d_scores.ptr, n, variant, d_nodes.ptr)
    N.check(L.msa_device_sync())
    t0 = time.time()
    for _ in range(steps):
        N.nj_build_device(d_scores.ptr, n, variant, d_nodes.ptr)
    ms = 1e3 * (time.time() - t0) / steps
    nodes = d_nodes.download(N.NJ_NODE, 2 * n - 1)
    q_evals = sum(c * (c - 1) // 2 for c in range(2, n + 1))
    result = {
        "metric": "neighbor-joining tree build on 1 MI355X (after the pairwise stage)", "value": round(ms, 3),
        "unit": "ms per tree", "n_gpus": 1, "steps": steps, "warmup": warmup, "higher_is_better": False,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": w.name, "otus": n, "variant": "symmetric" if variant == 0 else "linear",
                   "joins": n - 1, "q_evaluations": q_evals},
        "q_evaluations_per_s": round(q_evals / (ms * 1e-3), 1),
        "roofline": {"bound": "hbm", "achieved": round(4.0 * q_evals / (ms * 1e-3) / 1e9, 1), "peak": 8000.0,
                     "unit": "GB/s", "frac": round(4.0 * q_evals / (ms * 1e-3) / 8e12, 4), "traffic": None,
                     "bytes_per_q": 4},
    }
    if cpu:
        scores = d_scores.download(np.int32, w.npairs).astype(np.float32)
        t0 = time.time()
        exp = M.NjOracle().build(scores, n, variant)
        cpu_s = time.time() - t0
        same = all(np.array_equal(nodes[f], exp[f]) for f in ("parent", "child", "level")) and bool(np.all(
            (nodes["distance"].view(np.uint32) == exp["distance"].view(np.uint32)) |
            (np.isnan(nodes["distance"]) & np.isnan(exp["distance"]))))
        result["bit_exact_vs_oracle"] = same
        result["cpu_baseline"] = {"value": round(1e3 * cpu_s, 1), "unit": "ms per tree", "cores": 1, "kind": "port",
                                  "sample": "the whole tree over the same scores (oracle/nj_oracle.c, -O3)"}
        result["vs_cpu"] = round(1e3 * cpu_s / ms, 1)
    plan.close()
    return result


def fasta_run(N, steps, warmup, target_mib, cpu=True, cpu_mib=64):
    """FASTA loading on the GPU (msa_fasta_parse_device) over a protein
    database already in device memory: parse, encode and pack into the
    sequence store, every step.  The store is compared with the oracle's
    (oracle/fasta_oracle.c) over the whole text.  CPU baseline: the
    reference's own parser (oracle/_ref/libref_fasta.so: fasta.cpp +
    encoder.cpp + database.cpp, one thread, reading the file from the page
    cache) on the first `cpu_mib` MiB of the same text, cut at a record."""
    import tempfile
    from msa_amd import workloads as W
    from oracle import museqa_ref as M
    text, desc = W.fasta_database(target_bytes=int(target_mib) << 20)
    n = len(text)
    log("fasta workload:", desc, "bytes", n)
    d_text = N.DeviceBuffer(n)
    d_text.upload(text)
    ld = N.FastaLoader()
    for _ in range(warmup):
        ld.parse_device(d_text.ptr, n)
    stages = {"total_ms": 0.0, "lines_ms": 0.0, "records_ms": 0.0, "emit_ms": 0.0}
    t0 = time.time()
    for _ in range(steps):
        ld.parse_device(d_text.ptr, n)
        for k, v in ld.timing().items():
            stages[k] += v / steps
    wall_ms = 1e3 * (time.time() - t0) / steps
    info = ld.info()
    got = ld.store()
    exp = M.FastaOracle().parse(text.tobytes())
    same = all(np.array_equal(g, e) for g, e in zip(got, exp))
    out_bytes = 2 * info["blocks"] + 28 * info["sequences"]      # blocks + offset, size, description range
    emit_bytes = n + 2 * info["blocks"]
    result = {
        "metric": "FASTA load (parse + encode + pack into the sequence store) on 1 MI355X",
        "value": round(n / (wall_ms * 1e-3) / 1e9, 2), "unit": "GB/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": round(wall_ms, 3), "higher_is_better": True, "scaling": "strong",
        "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": "fasta-protein-%dMiB" % target_mib, "description": desc, "bytes": n,
                   "sequences": info["sequences"], "residues": info["residues"], "blocks": info["blocks"],
                   "lines": info["lines"]},
        "stages_ms": {k: round(v, 3) for k, v in stages.items()},
        "gpu_gbps": round(n / (stages["total_ms"] * 1e-3) / 1e9, 2),
        "roofline": {"bound": "hbm", "kernel": "emit_kernel",
                     "achieved": round(emit_bytes / (stages["emit_ms"] * 1e-3) / 1e9, 1), "peak": 8000.0,
                     "unit": "GB/s", "frac": round(emit_bytes / (stages["emit_ms"] * 1e-3) / 8e12, 4),
                     "traffic": measured_traffic("fasta-protein-%dMiB" % target_mib, "emit_kernel"),
                     "bytes_per_launch": emit_bytes},
        "parse_roofline": {"achieved": round((n + out_bytes) / (wall_ms * 1e-3) / 1e9, 1), "peak": 8000.0,
                           "unit": "GB/s", "frac": round((n + out_bytes) / (wall_ms * 1e-3) / 8e12, 4)},
        "bit_exact_vs_oracle": bool(same),
    }
    if cpu:
        # A prefix of the text that ends at a record boundary.
        cut = min(n, int(cpu_mib) << 20)
        nxt = np.nonzero(text[cut:] == ord(">"))[0]
        cut = cut + int(nxt[0]) if len(nxt) else n
        sample = text[:cut].tobytes()
        if M.ReferenceFasta.available():
            with tempfile.NamedTemporaryFile(suffix=".fasta", delete=False) as f:
                f.write(sample)
                path = f.name
            ref = M.ReferenceFasta()
            ref.load(path)              # page cache warm
            t0 = time.time()
            reps = 0
            while True:
                ref.load(path)
                reps += 1
                if time.time() - t0 > 5.0:
                    break
            cpu_s = (time.time() - t0) / reps
            os.unlink(path)
            kind, what = "reference", "the reference's fasta.cpp + encoder.cpp + database.cpp (oracle/_ref)"
        else:
            ora = M.FastaOracle()
            t0 = time.time()
            ora.parse(sample)
            cpu_s = time.time() - t0
            kind, what = "port", "oracle/fasta_oracle.c"
        gbps = len(sample) / cpu_s / 1e9
        result["cpu_baseline"] = {"value": round(gbps, 4), "unit": "GB/s", "cores": 1, "kind": kind,
                                  "sample": "the first %.1f MiB of the same text (%s), one thread" %
                                            (len(sample) / 2**20, what)}
        result["vs_cpu"] = round(result["value"] / gbps, 1)
    ld.close()
    d_text.close()
    return result


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="uneven")
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--oracle-pairs", type=int, default=400)
    ap.add_argument("--serving", action="store_true", help="only the serving sweep")
    ap.add_argument("--rates", default="10,20,50,100,200,500")
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--serve-jobs", type=int, default=0, help="jobs per rate in the --serving sweep (0: rate x seconds)")
    ap.add_argument("--serve-rate", default="100,500",
                    help="rates (jobs/s, comma separated) of the serving figures in the main line; '' = none")
    ap.add_argument("--no-other-configs", dest="other_configs", action="store_false",
                    help="skip the uniform-512 and dna-64x30000 figures of the default line")
    ap.add_argument("--no-reference-cases", dest="reference_cases", action="store_false",
                    help="skip the figures over the reference's db/case1..6.fasta")
    ap.add_argument("--phylogeny", action="store_true",
                    help="only the neighbor-joining stage over the workload's scores (its own JSON line)")
    ap.add_argument("--fasta", action="store_true",
                    help="only the FASTA loader over a synthetic protein database (its own JSON line)")
    ap.add_argument("--fasta-mib", type=int, default=512, help="size of the FASTA database for --fasta")
    ap.add_argument("--tuning", default=DEFAULT_TUNING,
                    help="tuning table for the batched plan (default: the committed gfx950 table; '' = built-in)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path here (the pair list would be sharded by "
              "cell count, one process per GPU); run with --gpus 1", file=sys.stderr)
        return 2

    from msa_amd import _native as N
    if N.device_count() < 1:
        raise SystemExit("no GPU visible")
    tables = {n: N.table(n) for n in N.table_names()}
    if args.serving:
        rates = [float(r) for r in args.rates.split(",")]
        sweep = serving_run(N, tables, rates, args.serve_seconds, jobs=args.serve_jobs)
        resident = serving_run(N, tables, rates, args.serve_seconds, jobs=args.serve_jobs, resident=True)
        print(json.dumps({
            "metric": METRIC, "value": max(r["jobs_per_s"] for r in sweep), "unit": "jobs/s",
            "n_gpus": 1, "higher_is_better": True, "data": "synthetic", "dtype": "int32",
            "config": {"workload": "serving: Poisson jobs of 8-256 protein sequences, L~U[100,2000], blosum62",
                       "seconds_per_rate": None if args.serve_jobs > 0 else args.serve_seconds,
                       "jobs_per_rate": args.serve_jobs if args.serve_jobs > 0 else None, "slots": 4},
            "serving": sweep, "serving_resident": resident}), flush=True)
        return 0 if all(r["bit_exact_vs_alone"] for r in sweep + resident) else 1
    if args.fasta:
        res = fasta_run(N, args.steps, args.warmup, args.fasta_mib, cpu=not args.no_cpu)
        print(json.dumps(res), flush=True)
        return 0 if res["bit_exact_vs_oracle"] else 1
    w = make_workload(args.workload, tables)
    if args.phylogeny:
        res = phylogeny_run(N, w, args.steps, args.warmup, cpu=not args.no_cpu)
        print(json.dumps(res), flush=True)
        return 0 if res.get("bit_exact_vs_oracle", True) else 1
    cells = w.cells()
    log("workload:", w.description, "pairs", w.npairs, "cells %.4e" % cells, "checksum", w.checksum())

    L = N.lib()
    stream = ctypes.c_void_p()
    N.check(L.msa_stream_create(ctypes.byref(stream)))
    stream = stream.value
    out = N.DeviceBuffer(4 * w.npairs)

    tuning = args.tuning if args.tuning and os.path.exists(args.tuning) else None
    batched, plain, same, oracle_ok, oracle_n = measure(N, w, args.steps, args.warmup, args.plain_steps,
                                                        args.oracle_pairs, stream, out, tuning)

    gcups = cells / batched["ms"] / 1e6
    info = batched["info"]
    result = {
        "metric": METRIC,
        "value": round(gcups, 2),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(batched["ms"], 3),
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        # cells as exact 16-bit integers, two to a register, where the plan
        # packs them (strip-boundary lines and scores stay int32)
        "dtype": "u16x2+int32" if int(getattr(info, "packed_cells", 0)) > 0 else "int32",
        "data": "synthetic",
        "config": {
            "workload": w.name, "description": w.description, "sequences": w.nseq, "pairs": w.npairs,
            "cells": cells, "input_checksum": w.checksum(), "kernel": "batched persistent (auto routing)",
            "launches_per_step": int(info.launches), "bins": int(info.bins),
            "batched_pairs": int(info.batched_pairs), "long_pairs": int(info.long_pairs),
            "plain_routed_pairs": int(info.plain_pairs),
            "stacked_pairs": int(info.stacked_pairs), "packed_pairs": int(info.packed_pairs),
            "duo_pairs": int(info.duo_pairs),
            "graph": True, "parallelism": "single GPU",
            "tuning": tuning_summary(tuning),
        },
        "plain_gcups": round(cells / plain["ms"] / 1e6, 2),
        "plain_ms_per_step": round(plain["ms"], 3),
        "batched_vs_plain": round(plain["ms"] / batched["ms"], 3),
        "bit_exact": {"batched_equals_plain_all_pairs": same, "oracle_sample_pairs": oracle_n,
                      "oracle_sample_equal": oracle_ok},
        "wall_ms_per_step": round(batched["wall_ms"], 3),
        "roofline": valu_roofline(w, info, batched["ms"]),
    }
    # Host in, host out (msa_nw_align): planning, one upload, the kernels and
    # the score download, once -- the PCIe-inclusive rate, never `value`.
    t0 = time.time()
    e2e_scores = N.align(w.blocks, w.offsets, w.sizes, w.table, w.penalty,
                         opts=N.options(N.KERNEL_AUTO, tuning_file=tuning))
    e2e_s = time.time() - t0
    result["host_to_host"] = {"ms": round(1e3 * e2e_s, 2), "gcups": round(cells / e2e_s / 1e9, 1),
                              "equal": bool(np.array_equal(e2e_scores, batched["scores"])),
                              "plan_ms": round(1e3 * batched["plan_s"], 1)}
    log("host to host (msa_nw_align): %.1f ms" % (1e3 * e2e_s))
    out.close()
    # The other throughput workloads, each timed the same way (auto plan as a
    # hipGraph, plain kernel, all-pairs equality, oracle sample); parity and
    # context figures, not `value`.
    ok_others = True
    if args.workload == "uneven" and args.other_configs:
        result["other_configs"] = {}
        for name in ("uniform", "dna"):
            wo = make_workload(name, tables)
            log("workload:", wo.description, "pairs", wo.npairs, "cells %.4e" % wo.cells())
            o_out = N.DeviceBuffer(4 * wo.npairs)
            b, pl, eq, ok_o, n_o = measure(N, wo, args.steps, min(args.warmup, 2), args.plain_steps,
                                           args.oracle_pairs, stream, o_out, tuning)
            o_out.close()
            rec = config_record(wo, b, pl, eq, ok_o, n_o)
            result["other_configs"][wo.name] = rec
            ok_others = ok_others and eq and ok_o
    # The reference's own benchmark databases (db/case1..6.fasta, real viral
    # genomes), parsed by the GPU FASTA loader and timed the same way.
    if args.workload == "uneven" and args.reference_cases:
        result["reference_cases"] = {}
        for k in range(1, 7):
            wo = make_workload("case%d" % k, tables)
            log("workload:", wo.description, "pairs", wo.npairs, "cells %.4e" % wo.cells())
            o_out = N.DeviceBuffer(4 * wo.npairs)
            b, pl, eq, ok_o, n_o = measure(N, wo, min(args.steps, 3), 1, 1, 24, stream, o_out, tuning,
                                           oracle_cells=4e9)
            o_out.close()
            result["reference_cases"][wo.name] = config_record(wo, b, pl, eq, ok_o, n_o)
            ok_others = ok_others and eq and ok_o
    if args.serve_rate:
        rates = [float(r) for r in str(args.serve_rate).split(",")]
        result["serving"] = serving_run(N, tables, rates, args.serve_seconds)
        result["serving_resident"] = serving_run(N, tables, rates, args.serve_seconds, resident=True)
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(w, args.cpu_seconds)
        result["vs_cpu"] = round(gcups / result["cpu_baseline"]["value"], 1)
    print(json.dumps(result), flush=True)
    ok = same and oracle_ok and ok_others and result["host_to_host"]["equal"] and \
        all(r["bit_exact_vs_alone"] for r in result.get("serving", []) + result.get("serving_resident", []))
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
