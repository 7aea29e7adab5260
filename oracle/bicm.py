"""A/322 bit interleaver, restated by running the transmitter on labels.

TEST INFRASTRUCTURE (see oracle/__init__.py).  The product builds its
de-interleave map ``lam_of_q`` in closed form (atsc3_amd/bicm.py); this
module gets the same map a different way: it writes the codeword positions
0..N-1 through the transmitter's three interleavers exactly as A/322 6.2
describes them -- fill an array, read it out in the other order -- and
reads off which codeword position lands at each q-stream slot.

Follows lab/m6_bicm.py:76-190 (parity interleave = identity on this
codeword order, :76-107; group-wise Y_j = X_pi(j), :137-149; block
interleaver Type A column-write / row-read and Type B NQCB_IG-group blocks
plus an uninterleaved part 2, :152-182; composite lam_of_q, :185-190).

Pinned by tests/test_bicm.py: the 6.2.3.1 worked example (lab/m6_bicm.py:
380-406), the permutation property for every (N, mod, rate), and sha256
fixtures of the upstream's own composite map for every (N, mod, rate)
(tests/golden/ref_bicm.npz, written by tests/golden/make_ref_vectors.py).
"""
from __future__ import annotations

import json
import os

import numpy as np

_DATA = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                     "atsc3_amd", "data", "bitint_tables.json")
with open(_DATA) as _f:
    _T = json.load(_f)

MOD_BITS = dict(_T["mod_bits"])
_TYPE_A = {(n, m): v for n, m, *v in _T["type_a"]}
_TYPE_B = {(n, m): v for n, m, *v in _T["type_b"]}
_BI_TYPE = {(n, nb, r): t for n, nb, r, t in _T["bi_type"]}
_GW = {(n, m, r): p for n, m, r, p in _T["groupwise"]}
_DEMUX = {(n, m): v for n, m, *v in _T["demux"]}


def combos(mods=("QPSK", "16QAM", "64QAM", "256QAM")):
    """Every (N, mod, rate) the tables define a cell count for."""
    return sorted((n, m, r) for (n, m, r) in _GW
                  if m in mods and _DEMUX.get((n, m), [None, None])[1] is not None)


def groupwise_tx(u, n, mod, rate):
    """6.2.2: output group j carries input group pi(j)."""
    pi = _GW[(n, mod, rate)]
    groups = np.asarray(u).reshape(n // 360, 360)
    out = np.empty_like(groups)
    for j in range(n // 360):
        out[j] = groups[pi[j]]
    return out.ravel()


def blockint_tx(v, n, mod, rate):
    """6.2.3: Type A (column-write, row-read; part 1 Nr1 rows, part 2 Nr2)
    or Type B (groups of NQCB_IG rows of 360 written row-wise, read
    column-wise; part 2 passed through)."""
    v = np.asarray(v)
    nc = MOD_BITS[mod]
    out = []
    if _BI_TYPE[(n, nc, rate)] == "A":
        nr1, nr2, _ = _TYPE_A[(n, mod)]
        for rows, seg in ((nr1, v[:nr1 * nc]), (nr2, v[nr1 * nc:])):
            if rows == 0:
                continue
            arr = np.empty((rows, nc), v.dtype)
            for c in range(nc):                      # write column by column
                arr[:, c] = seg[c * rows:(c + 1) * rows]
            out.append(arr.reshape(-1))              # read row by row
    else:
        nq, np1, _ = _TYPE_B[(n, mod)]
        blk = nq * 360
        for b0 in range(0, np1, blk):
            arr = np.empty((nq, 360), v.dtype)
            for r in range(nq):                      # one bit group per row
                arr[r] = v[b0 + r * 360:b0 + (r + 1) * 360]
            out.append(arr.T.reshape(-1))            # read column by column
        out.append(v[np1:])
    return np.concatenate(out)


def lam_of_q(n, mod, rate):
    """q-stream slot j carries codeword bit lam_of_q[j]."""
    lam = np.arange(n, dtype=np.int64)               # parity interleave: identity
    return blockint_tx(groupwise_tx(lam, n, mod, rate), n, mod, rate)


def map_sha256(m):
    """The fixture key: sha256 of the map as little-endian int32."""
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(m, "<i4").tobytes()).hexdigest()
