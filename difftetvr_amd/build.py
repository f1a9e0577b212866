"""Builds the gfx950 shared library ``difftetvr_amd/libdifftetvr_amd.so`` in-tree with hipcc.

Flags:
  -ffp-contract=off      identical float rounding of the projection/coverage maths in every kernel
  -munsafe-fp-atomics    native global float atomics for the gradient scatter (buffers are device memory)
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed FP32 instructions
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = ["api.hip", "binning.hip", "sort.hip", "render.hip", "proj.hip", "optim.hip", "regularizer.hip",
           "serve.hip", "split.hip"]
LIB = os.path.join(HERE, "libdifftetvr_amd.so")
ARCH = os.environ.get("DTV_OFFLOAD_ARCH", "gfx950")


def hipcc_command(out=LIB, extra=()):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    return [hipcc, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
            "-munsafe-fp-atomics", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I", os.path.join(ROOT, "include"), *extra,
            *[os.path.join(HERE, "csrc", s) for s in SOURCES], "-o", out]


def build(verbose=False):
    """Compiles every HIP source into the in-tree shared library (always rebuilds)."""
    cmd = hipcc_command(out=LIB + ".tmp")
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    print(LIB)
    sys.exit(0)
