"""Build liblst_amd.so (HIP for gfx950) in-tree.

Flags that matter for parity with the reference CPU results:
  -ffp-contract=off   no FMA contraction (the oracle is built the same way)
  no fast-math        IEEE division/sqrt, denormals kept
and, for this pool, packed FP32 instructions are disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "_lib")
LIB_PATH = os.path.join(LIB_DIR, "liblst_amd.so")
ARCH = os.environ.get("LST_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = [
    "es.cpp",
    "prepare.cpp",
    "stages_hits_md_ls_t3.hip",
    "stages_t5.hip",
    "stages_pixel.hip",
    "stages_clean.hip",
    "t5dnn_train.hip",
    "engine.hip",
]

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-variable",
    "-Wno-unused-function",
]


def build(verbose: bool = False, jobs: int = 8) -> str:
    """Compile every translation unit (in parallel) and link the shared library."""
    os.makedirs(LIB_DIR, exist_ok=True)
    objs, procs = [], []
    for src in SOURCES:
        obj = os.path.join(LIB_DIR, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else ["-x", "hip"]
        cmd = [HIPCC, *FLAGS, *lang, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
        while len([p for _, p in procs if p.poll() is None]) >= jobs:
            procs[0][1].wait()
    failed = []
    for src, p in procs:
        out = p.communicate()[0].decode(errors="replace")
        text = "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
        if p.returncode != 0:
            failed.append((src, out))
        elif verbose and text.strip():
            print(text)
    if failed:
        for src, out in failed:
            print(f"--- {src} failed ---\n{out}", file=sys.stderr)
        raise RuntimeError("HIP build failed: " + ", ".join(s for s, _ in failed))
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", LIB_PATH + ".tmp", *objs]
    subprocess.check_call(cmd)
    os.replace(LIB_PATH + ".tmp", LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
