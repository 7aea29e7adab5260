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
SOURCES = ["ldpc_qc.hip", "ldpc_qc_phase.hip", "ldpc_batch.hip", "ldpc_batch_f32.hip",
           "ldpc_batch_f16.hip", "ldpc_batch_i8.hip", "ldpc_csr.hip", "demap.hip", "ldm.hip",
           "unrolled.hip", "baseband.hip", "sum_product.hip", "cpe.hip", "ldpc_resident.hip", "ldpc_utc.hip"]
ARCH = os.environ.get("ATSC3_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]
OBJDIR = os.path.join(CSRC, "_obj")


def build(verbose=False, jobs=None):
    """Compile every source to an object (in parallel), then link the .so."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJDIR, exist_ok=True)
    jobs = jobs or max(1, min(len(SOURCES), len(os.sched_getaffinity(0)), 16))
    procs, objs = [], []
    pending = list(SOURCES)
    while pending or procs:
        while pending and len(procs) < jobs:
            s = pending.pop(0)
            obj = os.path.join(OBJDIR, s.replace(".hip", ".o"))
            objs.append(obj)
            cmd = [hipcc, *HIPCC_FLAGS, "-c", os.path.join(CSRC, s), "-o", obj]
            if verbose:
                print(" ".join(cmd))
            procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                              stderr=subprocess.PIPE, text=True)))
        s, p = procs.pop(0)
        _, err = p.communicate()
        if p.returncode != 0:
            for _, q in procs:
                q.kill()
            raise RuntimeError(f"hipcc failed on {s} ({p.returncode}):\n{err[-4000:]}")
    tmp = LIB + ".tmp"
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed ({r.returncode}):\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    # the resident engine's tables of every A/322 code, for the host-pointer
    # drop-in ldpc_minsum_f32 (read beside the library at its first call)
    from .resident import DROPIN_TABLES, write_dropin_tables
    write_dropin_tables(os.path.join(HERE, DROPIN_TABLES))
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
