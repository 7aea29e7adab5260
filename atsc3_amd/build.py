"""Build libatsc3fec.so (gfx950) in-tree with hipcc.

    python -m atsc3_amd.build          # or atsc3_amd.build.build()

Flags that matter for results:
  -ffp-contract=off      no FMA contraction: the plain min-sum must round
                         every FP32 operation once, in source order
                         (bit-exactness with lab/ldpc_kernel.c, which is
                         built with -ffp-contract=off, lab/build_ldpc_kernel.py:36)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                         no packed FP32 VALU forms in device code
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libatsc3fec.so")
SOURCES = ["ldpc_qc.hip", "ldpc_qc_phase.hip", "ldpc_batch.hip", "ldpc_csr.hip", "demap.hip",
           "ldm.hip", "unrolled.hip"]
ARCH = os.environ.get("ATSC3_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def build(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc, *HIPCC_FLAGS, "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed ({r.returncode}):\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
