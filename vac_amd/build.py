"""Builds libvacgpu.so (the HIP kernels and the C ABI) in-tree for gfx950.

    python -m vac_amd.build

FP64 parity needs the arithmetic exactly as written: FMA contraction is off
(-ffp-contract=off) and packed-FP32 instruction selection is disabled.  FP32
divide and square root (used only by the FP32-bulk tier) run at hardware
rate instead of the correctly rounded software sequence.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRCS = [os.path.join(HERE, "csrc", "vacgpu.hip"), os.path.join(HERE, "csrc", "batch.hip"),
        os.path.join(HERE, "csrc", "noalloc.hip")]
OUT = os.path.join(HERE, "libvacgpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("VACGPU_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    # FP32-bulk tier only: hardware-rate FP32 divide and sqrt (about 1 ulp);
    # FP64 arithmetic is IEEE correctly rounded either way
    "-fno-hip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
    "-Wall", "-Wno-unused-function",
]


def build(verbose: bool = False) -> str:
    """Compiles the translation units in parallel (one hipcc each), then links."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    cflags = [f for f in FLAGS if f != "-shared"]
    procs = []
    objs = []
    for src in SRCS:
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        cmd = [HIPCC, *cflags, "-c", src, "-o", obj + ".tmp"]
        if verbose:
            print(" ".join(cmd))
        procs.append((src, obj, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                                 stderr=subprocess.STDOUT, text=True)))
    failed = False
    for src, obj, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed = True
        else:
            os.replace(obj + ".tmp", obj)
    if failed:
        raise RuntimeError("hipcc failed for libvacgpu.so")
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", OUT + ".tmp"]
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("linking libvacgpu.so failed")
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
