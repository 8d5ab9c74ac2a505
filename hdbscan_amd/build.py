"""Builds the gfx950 shared library libhdbscan_rocm.so in-tree with hipcc."""

from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCES = [os.path.join(HERE, "csrc", "hdbscan_rocm.hip")]
LIBRARY = os.path.join(HERE, "_lib", "libhdbscan_rocm.so")

# -ffp-contract=off: a*b+c stays two correctly rounded operations, as in the
# reference's kernels (compiled with --fmad=false).  The default HIP division
# and square root are correctly rounded; no fast-math flag is used.
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(verbose=False):
    os.makedirs(os.path.dirname(LIBRARY), exist_ok=True)
    tmp = LIBRARY + ".tmp"
    cmd = ["hipcc", *FLAGS, "-o", tmp, *SOURCES]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + proc.stdout + proc.stderr)
    if verbose:
        print(" ".join(cmd))
    os.replace(tmp, LIBRARY)
    return LIBRARY


if __name__ == "__main__":
    print(build(verbose=True))
