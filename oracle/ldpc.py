"""FP32 min-sum oracle (ctypes over minsum_oracle.c) and the alpha ladder.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Pinning:
  * ``minsum_f32`` (C restatement of lab/ldpc_kernel.c:65-146) -- checked
    field-by-field against golden fixtures written by the upstream kernel
    compiled from its own source (tests/golden/, made by
    tests/golden/make_golden.py), and live against ``_ref/ldpc_kernel.so``
    when that library is present; against ``scalar_minsum`` (pure Python) on
    brute-force tiny codes; and against invariants (conv => H.bits = 0 and
    bad = 0; noiseless codeword => its 0; all-positive LLRs => zero codeword
    at its 0, the smoke case of lab/build_ldpc_kernel.py:101-120).
  * ``ladder`` (lab/m9_fast.py:843-891, lab/m6_bicm.py:341-355) -- checked
    against per-rung ``minsum_f32`` calls by construction in the tests
    (first converged rung wins; else the fewest unsatisfied wins).
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}


def _load(which):
    if which in _LIBS:
        return _LIBS[which]
    if which == "oracle":
        path, abi, fn = (os.path.join(HERE, "liboracle_minsum.so"),
                         "oracle_minsum_abi", "oracle_minsum_f32")
    elif which == "ref":
        path, abi, fn = (os.path.join(HERE, "_ref", "ldpc_kernel.so"),
                         "ldpc_kernel_abi", "ldpc_minsum_f32")
    elif which in ("plant1", "plant2"):          # negative controls (make plant)
        path, abi, fn = (os.path.join(HERE, "_plant", f"liboracle_{which}.so"),
                         "oracle_minsum_abi", "oracle_minsum_f32")
    else:
        raise ValueError(which)
    if not os.path.exists(path):
        _LIBS[which] = None
        return None
    lib = ctypes.CDLL(path)
    getattr(lib, abi).restype = ctypes.c_int32
    if getattr(lib, abi)() != 1:
        raise OSError(f"{path}: ABI mismatch")
    f = getattr(lib, fn)
    f.restype = ctypes.c_int32
    i32 = np.ctypeslib.ndpointer(np.int32, flags="C")
    f32 = np.ctypeslib.ndpointer(np.float32, flags="C")
    u8 = np.ctypeslib.ndpointer(np.uint8, flags="C")
    f.argtypes = [f32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                  ctypes.c_int32, i32, i32, i32, i32, ctypes.c_int32,
                  ctypes.c_float, u8, u8, i32, i32, f32]
    _LIBS[which] = f
    return f


def available(which="oracle"):
    return _load(which) is not None


def minsum_f32(llr2d, code, iters=50, alpha=0.75, which="oracle"):
    """(nblk, n) LLRs -> (bits u8, conv bool, its i32, bad i32)."""
    f = _load(which)
    if f is None:
        raise RuntimeError(f"{which} min-sum library not built (make -C oracle)")
    cptr, cidx, vptr, vedge = code.csr()
    llr = np.ascontiguousarray(llr2d, np.float32)
    if llr.ndim == 1:
        llr = llr[None]
    nblk, n = llr.shape
    assert n == code.n
    bits = np.empty((nblk, n), np.uint8)
    conv = np.zeros(nblk, np.uint8)
    its = np.empty(nblk, np.int32)
    bad = np.empty(nblk, np.int32)
    work = np.empty(n + len(cidx), np.float32)
    rc = f(llr, nblk, n, len(cptr) - 1, len(cidx), cptr, cidx, vptr, vedge,
           int(iters), np.float32(alpha), bits, conv, its, bad, work)
    if rc != 0:
        raise RuntimeError(f"min-sum oracle rejected the code (rc {rc})")
    return bits, conv.view(bool), its, bad


def minsum_f32_csr(llr2d, n, cptr, cidx, vptr, vedge, iters, alpha,
                   which="oracle"):
    """Same as ``minsum_f32`` on explicit CSR tables (tiny synthetic codes)."""
    class _C:
        pass
    c = _C()
    c.n = n
    c.csr = lambda: (cptr, cidx, vptr, vedge)
    return minsum_f32(llr2d, c, iters, alpha, which)


def ladder(llr2d, code, ladder=(1.0, 0.85, 0.75), iters=50, which="oracle"):
    """Alpha ladder: rungs tried on still-unconverged rows only; a rung's
    result is taken if it converged or has fewer unsatisfied checks than the
    best so far (lab/m9_fast.py:843-891)."""
    llr2d = np.ascontiguousarray(llr2d, np.float32)
    nb = llr2d.shape[0]
    bits = np.empty((nb, code.n), np.uint8)
    conv = np.zeros(nb, bool)
    its = np.zeros(nb, np.int32)
    bad = np.full(nb, 1 << 30, np.int32)
    todo = np.arange(nb)
    for a in ladder:
        if len(todo) == 0:
            break
        b, c, it, d = minsum_f32(llr2d[todo], code, iters, a, which)
        take = c | (d < bad[todo])
        sel = todo[take]
        bits[sel] = b[take]
        its[sel] = it[take]
        bad[sel] = d[take]
        conv[sel] = c[take]
        todo = todo[~c]
    return bits, conv, its, bad


def scalar_minsum(llr, checks, n, iters, alpha):
    """Pure-Python restatement on Python floats rounded to binary32 after
    every operation (struct round trip).  For tiny codes only; pins the C
    oracle by a third, independent route."""
    import struct

    def f32(x):
        return struct.unpack("f", struct.pack("f", x))[0]

    alpha = f32(alpha)
    lam = [f32(float(x)) for x in llr]
    tot = list(lam)
    msgs = [[0.0] * len(c) for c in checks]
    var_edges = [[] for _ in range(n)]
    for ci, c in enumerate(checks):
        for k, v in enumerate(c):
            var_edges[v].append((ci, k))
    for it in range(iters + 1):
        unsat = 0
        for c in checks:
            p = 0
            for v in c:
                p ^= tot[v] < 0.0
            unsat += p
        if unsat == 0 or it == iters:
            bits = [1 if t < 0.0 else 0 for t in tot]
            return (np.array(bits, np.uint8), unsat == 0,
                    it if unsat == 0 else iters, 0 if unsat == 0 else unsat)
        for ci, c in enumerate(checks):
            vals = [f32(tot[v] - msgs[ci][k]) for k, v in enumerate(c)]
            mags = [abs(x) for x in vals]
            lo1, lo2, first = math.inf, math.inf, 0
            for k, a in enumerate(mags):
                if a < lo1:
                    lo2, lo1, first = lo1, a, k
                elif a < lo2:
                    lo2 = a
            par = 0
            for x in vals:
                par ^= x < 0.0
            for k, x in enumerate(vals):
                out = f32((lo2 if k == first else lo1) * alpha)
                if par ^ (x < 0.0):
                    out = -out
                msgs[ci][k] = out
        for v in range(n):
            s = 0.0
            for ci, k in var_edges[v]:        # ascending check index
                s = f32(s + msgs[ci][k])
            tot[v] = f32(lam[v] + s)
    raise AssertionError("unreachable")
