"""Build liberod.so (HIP kernels + C ABI) for gfx950 in-tree with hipcc.

    python -m elasticapp_amd.build          # or: python elasticapp_amd/build.py

The op-by-op FP64 kernels and both fused instantiations are compiled with
-ffp-contract=off: the op-by-op and fused FP64 kernels round every product and
sum on their own, like the CPU restatement they are checked against, and the
FP32-bulk kernel's only fused multiply-adds are the explicit ones in its
source, so its results do not depend on the tile shape.  Packed FP32
instructions are disabled for all device code.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liberod.so")
ARCH = os.environ.get("EROD_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++20", "-fPIC", "--offload-arch=" + ARCH,
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-Wno-unused-result", "-Wno-unused-value"]
SOURCES = {
    "erod_ops.hip": ["-ffp-contract=off"],
    "erod_fused64.hip": ["-ffp-contract=off"],
    "erod_fused32.hip": ["-ffp-contract=off"],
    "erod_fused.hip": [],
    "erod_capi.hip": [],
    "erod_serve.hip": [],
    "erod_io.hip": [],
}


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    cc = hipcc()
    objs, cmds = [], []
    for src, extra in SOURCES.items():
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        cmds.append([cc] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    # the translation units compile independently: one hipcc per host core, at most 8
    workers = max(1, min(8, len(os.sched_getaffinity(0)), len(cmds)))
    with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as ex:
        def run(cmd):
            if verbose:
                print(" ".join(cmd), flush=True)
            subprocess.run(cmd, check=True)
        for f in [ex.submit(run, c) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    cmd = [cc, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    # the C++ host-layer driver (links liberod.so next to it)
    drv = os.path.join(HERE, "timoshenko")
    cmd = [os.environ.get("CXX", "g++"), "-std=c++20", "-O2", "-Wall", "-I", os.path.join(HERE, "..", "include"),
           "-o", drv, os.path.join(CSRC, "host", "timoshenko.cpp"), "-L", HERE, "-lerod", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-q" not in sys.argv))
