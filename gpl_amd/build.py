"""Builds libgpl_hip.so in-tree: the HIP kernels and the host controller for gfx950 with hipcc,
the host-side design code with g++ (OpenMP), linked into one shared library next to this file.

Run directly (`python -m gpl_amd.build`) or through __graft_entry__.build().
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgpl_hip.so")
ARCH = os.environ.get("GPL_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -ffp-contract=off: no fused multiply-add, the reference's float arithmetic is unfused.
# The packed-FP32 feature is turned off for every device compile.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
CXX_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-fopenmp", "-ffp-contract=off"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")
    elif verbose and r.stdout.strip():
        # hipcc's host pass warns that it does not know the packed-FP32 feature; harmless
        lines = [l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l]
        if lines:
            print("\n".join(lines))


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    design_o = os.path.join(BUILD, "design.o")
    _run(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, "design.cpp"), "-o", design_o], verbose)
    objs.append(design_o)
    for src in ("kernels.hip", "nesterov.cpp", "noalloc.hip", "ip_gpu.hip", "route.hip"):
        o = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        _run([HIPCC, *HIP_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", o], verbose)
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp, "-lgomp"], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
