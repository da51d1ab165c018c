"""Compile oracle.c into oracle/liboracle.so with gcc (checker only)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


def build(verbose=False):
    # -mpopcnt: the clique counter's popcounts as one instruction (any
    # x86-64 host of an MI355X has it); no -march=native, since the library
    # is built here and runs on the GPU box's host
    cmd = ["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-pthread", "-mpopcnt", "-Wall", "-Wextra",
           "-Wno-unused-parameter", "-o", LIB + ".tmp", os.path.join(HERE, "oracle.c"),
           os.path.join(HERE, "clique_mt.c"), "-lm"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout)
    if verbose and r.stdout.strip():
        print(r.stdout)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
