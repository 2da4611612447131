"""A short, fixed run of the pairwise kernels (or of the FASTA loader) for
hardware-counter profiling.

Runs one warm-up and `--executions` timed executions of the plan for a
workload (no hipGraph, so each kernel is its own dispatch), or with
`--fasta-mib M` one warm-up and `--executions` parses of an M MiB protein
FASTA database already in device memory; then a 1 GiB device-to-device copy
whose byte counts calibrate FETCH_SIZE / WRITE_SIZE.  Meant to run under
rocprofv3, e.g.

  rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT -o run -- \\
      python3 -m msa_amd.profile_run --workload uneven
"""
import argparse
import ctypes
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from msa_amd import _native as N  # noqa: E402
from msa_amd import workloads as W  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="uneven", choices=["uneven", "uniform", "dna", "small"])
    ap.add_argument("--executions", type=int, default=2)
    ap.add_argument("--kernel", default="auto", choices=["auto", "plain", "batched", "long"])
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--count", type=int, default=0, help="sequences (0 = the workload's own count)")
    ap.add_argument("--fasta-mib", type=int, default=0, help="profile the FASTA loader on a database this large")
    args = ap.parse_args()
    if args.fasta_mib:
        return fasta(args)
    tables = {n: N.table(n) for n in N.table_names()}
    kw = {"count": args.count} if args.count else {}
    w = {"uneven": lambda: W.uneven_batch(tables, **kw), "uniform": lambda: W.uniform_batch(tables, **kw),
         "dna": lambda: W.long_dna(**kw), "small": lambda: W.uniform_small(tables)}[args.workload]()
    kernel = {"auto": N.KERNEL_AUTO, "plain": N.KERNEL_PLAIN, "batched": N.KERNEL_BATCHED,
              "long": N.KERNEL_LONG}[args.kernel]
    plan = N.Plan(w.blocks, w.offsets, w.sizes, w.table, w.penalty, opts=N.options(kernel))
    out = N.DeviceBuffer(4 * w.npairs)
    L = N.lib()
    plan.execute(out.ptr)
    N.check(L.msa_device_sync())
    t0 = time.time()
    for _ in range(args.executions):
        plan.execute(out.ptr)
    N.check(L.msa_device_sync())
    dt = (time.time() - t0) / args.executions
    print("%s: %d pairs, %.4e cells, %.2f ms per execution, %.1f GCUPS" % (
        w.name, w.npairs, w.cells(), 1e3 * dt, w.cells() / dt / 1e9), flush=True)
    nbytes = args.copy_mib << 20
    src, dst = N.DeviceBuffer(nbytes), N.DeviceBuffer(nbytes)
    N.check(L.msa_memset(src.ptr, 1, nbytes))
    N.check(L.msa_memcpy(dst.ptr, src.ptr, nbytes))
    N.check(L.msa_device_sync())
    print("copy: %d bytes device to device" % nbytes, flush=True)
    plan.close()


def fasta(args):
    text, desc = W.fasta_database(target_bytes=args.fasta_mib << 20)
    d_text = N.DeviceBuffer(len(text))
    d_text.upload(text)
    ld = N.FastaLoader()
    ld.parse_device(d_text.ptr, len(text))
    t0 = time.time()
    for _ in range(args.executions):
        ld.parse_device(d_text.ptr, len(text))
    dt = (time.time() - t0) / args.executions
    print("fasta: %s; %d bytes, %.3f ms per parse, %.1f GB/s, %s" % (
        desc, len(text), 1e3 * dt, len(text) / dt / 1e9, ld.info()), flush=True)
    L = N.lib()
    nbytes = args.copy_mib << 20
    src, dst = N.DeviceBuffer(nbytes), N.DeviceBuffer(nbytes)
    N.check(L.msa_memset(src.ptr, 1, nbytes))
    N.check(L.msa_memcpy(dst.ptr, src.ptr, nbytes))
    N.check(L.msa_device_sync())
    print("copy: %d bytes device to device" % nbytes, flush=True)
    ld.close()


if __name__ == "__main__":
    main()
