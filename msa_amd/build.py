"""Builds the native parts of msa_amd in-tree.

  libmsa_hip.so   the C-ABI library: HIP kernels for gfx950 plus their host
                  planner and runtime layer (hipcc)
  _museqa*.so     the Python binding of the C++ host layer (museqa::pairwise
                  surface), linked against libmsa_hip.so

Every HIP translation unit is compiled with packed FP32 instructions turned
off (-Xclang -target-feature -Xclang -packed-fp32-ops); hipcc's host pass
warns that it does not know that feature, which is harmless.
"""
import json
import os
import subprocess
import sys
import sysconfig

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MSA_OFFLOAD_ARCH", "gfx950")

# Longest compiles first (the batched kernel's three tile configurations).
HIP_SOURCES = ["nw_batched.hip", "nw_batched_32x16.hip", "nw_batched_16x16.hip", "nw_batched_packed.hip", "nw_batched_duo.hip",
               "runtime.hip", "nw_plain.hip",
               "nw_long.hip", "nw_plan.hip", "nw_noalloc.hip", "nw_serve.hip", "nw_tune.hip", "nj.hip", "pgalign.hip",
               "fasta.hip"]
HIP_FLAGS = ["-std=c++17", "-O3", "-fPIC", "--offload-arch=" + ARCH,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-command-line-argument"]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


# The pairwise kernels' sources: their compiles also report each kernel's
# registers, LDS and occupancy (the compiler's kernel-resource-usage remarks),
# kept in kernel_resources.json next to the library for the tuner's record.
RESOURCE_SOURCES = ("nw_batched.hip", "nw_batched_32x16.hip", "nw_batched_16x16.hip", "nw_batched_packed.hip", "nw_batched_duo.hip",
                    "nw_plain.hip",
                    "nw_long.hip")
RESOURCE_FILE = os.path.join(PKG, "kernel_resources.json")
_RESOURCE_KEYS = {"TotalSGPRs": "sgprs", "VGPRs": "vgprs", "AGPRs": "agprs", "ScratchSize [bytes/lane]": "scratch",
                  "Occupancy [waves/SIMD]": "waves_per_simd", "VGPRs Spill": "vgpr_spill",
                  "LDS Size [bytes/block]": "static_lds"}


def parse_resource_remarks(text):
    """{kernel name (demangled): {vgprs, agprs, sgprs, waves_per_simd, ...}}
    from hipcc's -Rpass-analysis=kernel-resource-usage output."""
    kernels, name = {}, None
    for line in text.splitlines():
        if "remark:" not in line:
            continue
        body = line.split("remark:", 1)[1].split(" [-Rpass", 1)[0].strip()
        key, _, value = body.partition(": ")
        if key == "Function Name":
            name = value.strip()
            kernels[name] = {}
        elif name and key in _RESOURCE_KEYS:
            try:
                kernels[name][_RESOURCE_KEYS[key]] = int(value)
            except ValueError:
                pass
    if kernels:
        try:
            names = list(kernels)
            out = subprocess.run(["c++filt"], input="\n".join(names), capture_output=True, text=True, check=True)
            kernels = dict(zip(out.stdout.splitlines(), kernels.values()))
        except (OSError, subprocess.CalledProcessError):
            pass
    return kernels


def build_hip(jobs=None):
    jobs = jobs or max(1, min(8, os.cpu_count() or 1))
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    remarks = []

    def reap(proc, src, log):
        rc = proc.wait()
        if log is not None:
            log.close()
            with open(log.name) as f:
                text = f.read()
            other = [ln for ln in text.splitlines() if "remark:" not in ln and ln.strip() and
                     not ln.lstrip().startswith(("|", "^")) and not ln.strip()[0].isdigit()]
            if other:
                print("\n".join(other), file=sys.stderr, flush=True)
            remarks.append(text)
        if rc != 0:
            raise RuntimeError("hipcc failed on %s" % src)

    for src in HIP_SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [HIPCC] + HIP_FLAGS + ["-I", INCLUDE, "-I", CSRC, "-c", os.path.join(CSRC, src), "-o", obj]
        log = None
        if src in RESOURCE_SOURCES:
            cmd.append("-Rpass-analysis=kernel-resource-usage")
            log = open(obj + ".remarks", "w")
        print("+", " ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd, stderr=log), src, log))
        if len(procs) >= jobs:
            reap(*procs.pop(0))
    for entry in procs:
        reap(*entry)
    out = os.path.join(PKG, "libmsa_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-pthread"])
    os.replace(tmp, out)
    resources = {}
    for text in remarks:
        resources.update(parse_resource_remarks(text))
    with open(RESOURCE_FILE, "w") as f:
        json.dump({"arch": ARCH, "source": "hipcc -Rpass-analysis=kernel-resource-usage", "kernels": resources}, f,
                  indent=1)
    return out


def build_host():
    """libmuseqa.so (the C++ host layer) and the _museqa Python module."""
    host = os.path.join(CSRC, "host")
    common = ["-std=c++17", "-O2", "-fPIC", "-Wall", "-I", INCLUDE, "-I", host]
    rpath = "-Wl,-rpath,$ORIGIN"
    lib = os.path.join(PKG, "libmuseqa.so")
    _run(["g++"] + common + ["-shared", "-o", lib + ".tmp", os.path.join(host, "museqa.cpp"),
                              "-L", PKG, "-lmsa_hip", rpath])
    os.replace(lib + ".tmp", lib)
    import pybind11
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    mod = os.path.join(PKG, "_museqa" + suffix)
    _run(["g++"] + common + ["-shared", "-fvisibility=hidden", "-I", pybind11.get_include(),
                              "-I", sysconfig.get_paths()["include"], "-o", mod + ".tmp",
                              os.path.join(host, "pymuseqa.cpp"), "-L", PKG, "-lmuseqa", "-lmsa_hip", rpath])
    os.replace(mod + ".tmp", mod)
    return mod


def build_all():
    build_hip()
    build_host()


if __name__ == "__main__":
    build_all()
