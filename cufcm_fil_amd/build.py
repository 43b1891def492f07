"""Builds the HIP libraries in-tree with hipcc for gfx950.

    python -m cufcm_fil_amd.build          (or __graft_entry__.build())

* libcufcm_fil.so -- the mobility library (include/cufcm_fil.h): every source of
  csrc/, compiled one object per source in parallel under _build/, then linked.
* libcfm_probe.so -- the FP32 issue-rate probe (include/cfm_probe.h), a
  measurement helper kept out of the mobility library.

All device code is compiled with packed-FP32 instruction generation disabled
(-target-feature -packed-fp32-ops).  Every call compiles every source.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcufcm_fil.so")
PROBE_SRC = os.path.join(HERE, "probe", "valu_probe.hip")
PROBE_LIB = os.path.join(HERE, "libcfm_probe.so")
SOURCES = ["api.hip", "apply.hip", "gmres.hip", "fil.hip", "step.hip", "gnn.hip", "nearfield.hip",
           "densepc.hip", "batch.hip", "cilia.hip", "body.hip", "swim.hip", "noalloc.hip",
           "coarse.hip", "squirmer.hip", "apply_sym.hip"]
ARCH = os.environ.get("CFM_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]
# rocSOLVER factorises the blob mobility once for the resident-inverse preconditioner
LIBS = ["-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = (r.stdout + r.stderr).strip()
    # hipcc's host pass does not know the packed-FP32 feature; that warning is harmless
    lines = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln]
    if lines and (verbose or r.returncode):
        print("\n".join(lines), flush=True)
    if r.returncode:
        raise subprocess.CalledProcessError(r.returncode, cmd)


def build(verbose: bool = False, extra=None, jobs: int | None = None) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJ, exist_ok=True)
    extra = list(extra or [])
    jobs = jobs or min(len(SOURCES), max(1, min(16, os.cpu_count() or 1)))
    objs = [os.path.join(OBJ, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    cmds = [[hipcc, *FLAGS, *extra, "-c", "-o", o, os.path.join(CSRC, s)]
            for s, o in zip(SOURCES, objs)]
    with cf.ThreadPoolExecutor(jobs) as ex:
        for f in [ex.submit(_run, c, verbose) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs, *LIBS], verbose)
    os.replace(tmp, LIB)
    ptmp = PROBE_LIB + ".tmp"
    _run([hipcc, *FLAGS, *extra, "-shared", "-o", ptmp, PROBE_SRC], verbose)
    os.replace(ptmp, PROBE_LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True, extra=sys.argv[1:])
