"""Build libct2.so for gfx950 in-tree (ct-2_amd/ct2/libct2.so).

Every HIP source in csrc/ is compiled with hipcc, in parallel, then linked
into one shared library.  Flags:
  -ffp-contract=off   no implicit FMA contraction: the exact-arithmetic
                      kernels must keep the reference's FP32 operation order;
                      kernels that want FMAs call them explicitly.
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      keeps the compiler from emitting packed FP32 ops.
Usage: python ct-2_amd/build.py [-j N] [--arch gfx950]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "ct2", "libct2.so")
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def flags(arch):
    return ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={arch}", "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-result", "-I", os.path.join(HERE, "..", "include")]


def compile_one(src, arch):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *flags(arch), "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
    return obj


def build(arch="gfx950", jobs=None, verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    jobs = jobs or min(8, len(srcs))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: compile_one(s, arch), srcs))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={arch}", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT} from {len(srcs)} sources")
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("--arch", default="gfx950")
    a = ap.parse_args()
    try:
        build(a.arch, a.j)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
