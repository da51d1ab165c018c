"""Build recipe for libgat.so (host C++ + gfx950 HIP) and the CPU oracle.

Run `python -m graphautotune_amd.build` (or __graft_entry__.build()). It always
recompiles. Host C++ is compiled with g++, the engine with hipcc for gfx950,
and everything is linked in-tree into graphautotune_amd/libgat.so.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libgat.so")

HOST_SOURCES = ["abi.cpp", "pattern.cpp", "graph.cpp", "plan.cpp"]
HIP_SOURCES = ["engine.hip", "mlp.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
# Packed FP32 instructions are switched off for all device code: some of
# their forms have given wrong results on this hardware without any error.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC", "-Wall", "-Wno-unused-function",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + INCLUDE, "-I" + CSRC]
HOST_FLAGS = ["-O2", "-std=c++20", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter",
              "-I" + INCLUDE, "-I" + CSRC]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def build_lib(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        jobs.append([CXX] + HOST_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
    for src in HIP_SOURCES:
        if not os.path.exists(os.path.join(CSRC, src)):
            continue
        obj = os.path.join(BUILD, src + ".o")
        jobs.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    objs = [j[-1] for j in jobs]
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, LIB)
    # pm command-line tool next to the library (links libgat.so by rpath)
    os.makedirs(os.path.join(PKG, "bin"), exist_ok=True)
    pm = os.path.join(PKG, "bin", "pm")
    _run([CXX, "-O2", "-std=c++20", "-I" + INCLUDE, os.path.join(CSRC, "pm_main.cpp"), "-o", pm + ".tmp",
          "-L" + PKG, "-lgat", "-Wl,-rpath,$ORIGIN/.."])
    os.replace(pm + ".tmp", pm)
    return LIB


def build_oracle(verbose=False):
    sys.path.insert(0, ROOT)
    from oracle import build as ob  # the oracle's own recipe
    return ob.build(verbose=verbose)


def build(verbose=False):
    lib = build_lib(verbose)
    orc = build_oracle(verbose)
    return lib, orc


if __name__ == "__main__":
    print(build(verbose=True))
