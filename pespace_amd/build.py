"""Build recipe for the gfx950 HIP library ``libpespace_amd.so`` (in-tree).

Generates ``csrc/qnm_table.inc`` from ``data/qnm_table.json`` and compiles
``csrc/pespace_amd.hip`` with hipcc for gfx950.
"""

from __future__ import annotations

import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpespace_amd.so")
ARCH = os.environ.get("PESPACE_AMD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
    # keep packed FP32 math off: the compiler must not emit v_pk_* FP32 ops
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fno-gpu-rdc", "-Wno-unused-result",
]


def gen_qnm_inc():
    with open(os.path.join(HERE, "data", "qnm_table.json")) as fh:
        doc = json.load(fh)
    order = ["22", "21", "33", "32", "44", "43"]
    lines = ["// generated from data/qnm_table.json by pespace_amd/build.py: Kerr n=0 QNM",
             "// frequencies (Leaver's method), fring/fdamp in units of the remnant mass.",
             "#pragma once",
             "namespace pes {",
             f"constexpr double kQnmAmin = {doc['amin']!r};",
             f"constexpr double kQnmStep = {doc['step']!r};",
             f"constexpr int kQnmN = {doc['n']};"]
    for name in ("fring", "fdamp"):
        key = "kQnmFring" if name == "fring" else "kQnmFdamp"
        lines.append(f"__constant__ static const double {key}[6][{doc['n']}] = {{")
        for m in order:
            vals = ", ".join(repr(float(v)) for v in doc["modes"][m][name])
            lines.append(f"  {{{vals}}},")
        lines.append("};")
    lines.append("}  // namespace pes")
    path = os.path.join(CSRC, "qnm_table.inc")
    text = "\n".join(lines) + "\n"
    if not os.path.exists(path) or open(path).read() != text:
        with open(path, "w") as fh:
            fh.write(text)
    return path


def build(verbose=False):
    gen_qnm_inc()
    src = os.path.join(CSRC, "pespace_amd.hip")
    tmp = LIB + ".tmp"
    cmd = ["hipcc", *HIPCC_FLAGS, "-o", tmp, src]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    print("built", LIB)
    sys.exit(0)
