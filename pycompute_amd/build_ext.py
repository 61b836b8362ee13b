"""Build libpycompute_amd.so in-tree with hipcc (host runtime + precompiled
gfx950 kernels). Run ``python -m pycompute_amd.build_ext``.

The generated LM modules are compiled at run time by hiprtc through this
library (and cached on disk); see pycompute_amd/program.py.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libpycompute_amd.so")
ARCH = "gfx950"

# Packed-FP32 VALU forms are kept out of every kernel built here.
HIP_DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                    # MFMA sources and results in VGPRs rather than AGPRs: the
                    # MLP kernels feed every accumulator straight into VALU work
                    # (bias-seeded accumulators, activations), which the AGPR
                    # form pays for with v_accvgpr_read copies and registers
                    "-mllvm", "-amdgpu-mfma-vgpr-form"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def host_sources():
    return [os.path.join(CSRC, "runtime.cpp")]


def device_sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith("_lib.hip"))


def build(verbose: bool = True) -> str:
    srcs = host_sources() + device_sources()
    cmd = [hipcc(), "-O3", "-fPIC", "-shared", "-std=c++17", f"--offload-arch={ARCH}",
           *HIP_DEVICE_FLAGS, "-I", os.path.join(ROOT, "include"), "-o", LIB + ".tmp",
           *srcs, "-lhiprtc"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
