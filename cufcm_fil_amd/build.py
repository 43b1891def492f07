"""Builds the HIP library libcufcm_fil.so in-tree with hipcc for gfx950.

    python -m cufcm_fil_amd.build          (or __graft_entry__.build())

All device code is compiled with packed-FP32 instruction generation disabled
(-target-feature -packed-fp32-ops).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcufcm_fil.so")
SOURCES = ["api.hip", "apply.hip", "gmres.hip", "fil.hip", "step.hip", "gnn.hip", "nearfield.hip",
           "densepc.hip", "batch.hip", "cilia.hip", "body.hip", "swim.hip", "probe.hip"]
ARCH = os.environ.get("CFM_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]
# rocSOLVER factorises the blob mobility once for the resident-inverse preconditioner
LIBS = ["-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]


def build(verbose: bool = False, extra=None) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc, *FLAGS, *(extra or []), "-o", tmp, *srcs, *LIBS]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True, extra=sys.argv[1:])
