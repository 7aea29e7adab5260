"""Update-then-check normalized min-sum in float32 or float64 (numpy).

TEST INFRASTRUCTURE (see oracle/__init__.py).

This is the other decode contract of the upstream: the torch decoder
``GpuMinSum.decode`` (lab/m9_gpu.py:138-211) and the scalar / batched
float64 decoders ``min_sum_decode`` (lab/m3_ldpc.py:245-286) and
``min_sum_decode_batch(dtype=float64)`` (lab/m3_ldpc.py:509-630).  It differs
from the compiled kernel's contract (lab/ldpc_kernel.c:82-143) in three ways:

* messages are updated first and the syndrome is checked after each update,
  so a block reports its >= 1 (a clean block stops after one update);
* the posterior is accumulated from the channel value:
  tot[v] = ((llr[v] + m_e1) + m_e2) + ... over v's edges in ascending
  raveled edge order, i.e. ascending check index (lab/m9_gpu.py:184-188,
  np.add.at in lab/m3_ldpc.py:276);
* each check row is padded to the code's largest degree with entries of
  magnitude 1e9 that take part in the two-minimum (lab/m9_gpu.py:157-158,
  lab/m3_ldpc.py:259-260); they change a message only when every real
  |x| of the check is >= 1e9.

Two rules for a check row holding NaN magnitudes (NaN LLRs, or inf - inf
sums), where the upstream's own implementations differ; on NaN-free rows
they agree:

* rule "torch" (GpuMinSum at float32 and float64, lab/m9_gpu.py:160-166):
  min1 = amin, which propagates NaN; the edge that takes min2 is the first
  with |x| == min1 (edge 0 when min1 is NaN, since no entry compares equal);
  min2 = amin over the other entries, NaN-propagating;
* rule "numpy" (min_sum_decode and min_sum_decode_batch(float64),
  lab/m3_ldpc.py:264-270, 565-569): min1, min2 = the two smallest entries
  in np.partition's order, which puts NaN last; the edge that takes min2 is
  np.argmin's, which is the first NaN edge when there is one.

iters == 0 follows min_sum_decode: bits = llr < 0, conv 0, its 0, bad -1.

Pinned by tests/golden/utc_*.npz: outputs of the upstream's own
``min_sum_decode`` (float64, rule "numpy") and ``GpuMinSum`` (float32 and
float64, rule "torch", run on torch's CPU device, whose elementwise IEEE
operations and exact min/argmax match its device path) on the same LLRs,
written by tests/golden/make_ref_vectors.py; the *_nonfinite and infnan sets
hold NaN and +-inf LLRs.
"""
from __future__ import annotations

import numpy as np

BIG = 1e9


def padded(checks, n):
    """(idx, mask) (nc, dmax); pad slots point at the dummy column n."""
    nc = len(checks)
    dmax = max(len(c) for c in checks)
    idx = np.full((nc, dmax), n, np.int64)
    mask = np.zeros((nc, dmax), bool)
    for i, c in enumerate(checks):
        idx[i, :len(c)] = c
        mask[i, :len(c)] = True
    return idx, mask


def var_columns(idx, mask, n):
    """vcol[v, k] = the k-th smallest raveled edge index of variable v
    (pad = nc*dmax, a slot whose message stays 0)."""
    ne = idx.size
    e = np.arange(ne)[mask.ravel()]
    var = idx.ravel()[mask.ravel()]
    order = np.lexsort((e, var))
    v_s, e_s = var[order], e[order]
    deg = np.bincount(v_s, minlength=n)
    start = np.concatenate([[0], np.cumsum(deg)])
    within = np.arange(len(e_s)) - start[v_s]
    vcol = np.full((n, max(1, int(deg.max()))), ne, np.int64)
    vcol[v_s, within] = e_s
    return vcol


RULES = ("torch", "numpy")


def utc_minsum(llr, checks, n, iters=50, alpha=0.75, dtype=np.float32, rule="torch",
               _plant=None):
    """(nblk, n) -> (bits u8, conv bool, its i32, bad i32); ``rule`` picks
    the NaN handling of the check update ("torch" or "numpy", above).

    _plant (negative controls for the tests only): "zero_first" sums
    (0 + m_1 + ...) then adds llr (the compiled kernel's association);
    "no_pad" drops the 1e9 row padding; "check_first" checks the channel
    values before the first update."""
    if rule not in RULES:
        raise ValueError(f"rule must be one of {RULES}")
    dt = np.dtype(dtype)
    lam = np.ascontiguousarray(llr, dt).reshape(-1, n)
    nb = lam.shape[0]
    bits_out = (lam < 0).astype(np.uint8)
    conv = np.zeros(nb, bool)
    its = np.full(nb, iters, np.int32)
    bad = np.full(nb, -1, np.int32)
    if nb == 0 or iters == 0:
        return bits_out, conv, its, bad
    if _plant == "check_first":
        clean = _syndrome(lam < 0, checks) == 0
        if clean.any():
            conv = clean.copy()
            its[clean] = 0
            bad[clean] = 0
    idx, mask = padded(checks, n)
    nc, dmax = idx.shape
    vcol = var_columns(idx, mask, n)
    a = dt.type(alpha)
    big = dt.type(BIG)
    tot = np.zeros((nb, n + 1), dt)
    tot[:, :n] = lam
    m = np.zeros((nb, nc * dmax + 1), dt)
    ar = np.arange(dmax)
    for it in range(iters):
        v = tot[:, idx] - m[:, :-1].reshape(nb, nc, dmax)
        v = np.where(mask, v, dt.type(np.inf) if _plant == "no_pad" else big)
        sg = v < 0
        parity = (np.sum(sg & mask, axis=2) & 1).astype(bool)
        mag = np.abs(v)
        if rule == "torch":
            # np.min propagates NaN as torch.amin does; argmax of an all-False
            # row is 0
            m1 = mag.min(axis=2, keepdims=True)
            first = np.argmax(mag == m1, axis=2)
            onehot = ar[None, None, :] == first[:, :, None]
            m2 = np.where(onehot, dt.type(np.inf), mag).min(axis=2, keepdims=True)
        else:
            # ascending with NaN last; the first NaN edge, else the first
            # minimal one, takes min2
            srt = np.sort(mag, axis=2)
            m1, m2 = srt[:, :, 0:1], srt[:, :, 1:2]
            isnan = np.isnan(mag)
            first = np.where(isnan.any(axis=2), np.argmax(isnan, axis=2),
                             np.argmax(mag == m1, axis=2))
            onehot = ar[None, None, :] == first[:, :, None]
        out = np.where(onehot, m2, m1)
        sgn = np.where(parity[:, :, None] ^ sg, dt.type(-1.0), dt.type(1.0))
        mm = np.where(mask, (a * sgn) * out, dt.type(0.0))
        m[:, :-1] = mm.reshape(nb, -1)
        if _plant == "zero_first":
            acc = np.zeros_like(lam)
            for k in range(vcol.shape[1]):
                acc = acc + m[:, vcol[:, k]]
            acc = lam + acc
        else:
            acc = lam.copy()
            for k in range(vcol.shape[1]):
                acc = acc + m[:, vcol[:, k]]
        tot[:, :n] = acc
        b = acc < 0
        bb = np.concatenate([b, np.zeros((nb, 1), bool)], axis=1)
        syn = np.sum(bb[:, idx] & mask, axis=2) & 1
        cnt = syn.sum(axis=1).astype(np.int32)
        live = ~conv
        newly = live & (cnt == 0)
        bits_out[live] = b[live]
        bad[live] = cnt[live]
        its[newly] = it + 1
        conv |= newly
        if conv.all():
            break
    return bits_out, conv, its, bad


def _syndrome(bits, checks):
    return np.array([[int(np.bitwise_xor.reduce(b[c].astype(np.uint8))) for c in checks]
                     for b in np.atleast_2d(bits)]).sum(axis=1)
