"""Build recipe of the oracle's C++/OpenMP restatement (``oracle/cpu_lnl.cpp``) into
``oracle/_build/libcpu_lnl.so``.  Test infrastructure: only tests/, bench.py's cpu_baseline leg
and __graft_entry__.build() use it.  Built for x86-64-v3 (AVX2 + FMA), which the hosts of the GPU
boxes support."""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpu_lnl.cpp")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libcpu_lnl.so")


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = ["g++", "-O3", "-march=x86-64-v3", "-ffast-math", "-fopenmp", "-std=c++17", "-shared", "-fPIC", "-o", tmp,
           SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print("built", build(verbose=True))
    sys.exit(0)
