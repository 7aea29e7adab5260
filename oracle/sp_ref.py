"""CPU oracle of the exact (float64, log-tanh) sum-product decode and of the
near-miss rescue policy around it.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

* ``sum_product`` restates lab/m16_margin.py:325-390
  (sum_product_decode_batch; with no deadline it is
  lab/e58_weighted.py:134-187): flooding BP in float64 over the check rows
  padded to the widest check (lab/m3_ldpc.py:228-242 pack), per iteration
    - syndrome of the hard decisions of the totals; converged rows leave
      with their iteration count,
    - the last iteration (or a spent budget) leaves with the current
      decisions and unsatisfied count,
    - v = total - message, t = tanh(clip(v, +-38) / 2), parity of the signs,
      lt = log(clip(|t|, 1e-300, 1)) (0 on padding), lsum = the padded row's
      sum (numpy's pairwise order), message = sign * 2 atanh(clip(exp(
      clip(lsum - lt, -700, -1e-16)), 0, 1 - 1e-16)),
    - totals rebuilt as LLR + messages in ascending check order.
  The budget is an iteration count here (``stop_at``), standing for the
  wall-clock deadline of the upstream, so that the shed path is
  deterministic.
* ``select`` / ``rescue`` restate the candidate policy of
  lab/m44_ldm.py:2527-2555 and lab/m9_fast.py:1126-1151: rows that failed
  with at most ``near_max`` unsatisfied checks, closest first (stable),
  capped at ``max_blocks``; converged rescues replace bits, conv, its, bad.

Pinning (tests/test_sum_product.py): on tree-shaped codes BP is exact, so
the totals after enough iterations equal the brute-force a-posteriori LLRs
over every codeword; a degree-3 check's message equals the closed-form
box-plus ln((1 + e^(a+b)) / (e^a + e^b)); decoding a real A/322 codeword
converges and a shuffled control does not (the upstream's own gates,
lab/gate_e58.py:165-190, lab/gate_e60.py:283-300).
"""
from __future__ import annotations

import numpy as np

LOGMIN = -700.0


def pack(checks, n):
    """(idx, mask) of shape (n_checks, widest check); padding -> column n."""
    nc = len(checks)
    dmax = max(len(c) for c in checks)
    idx = np.full((nc, dmax), n, np.int64)
    mask = np.zeros((nc, dmax), bool)
    for i, c in enumerate(checks):
        idx[i, :len(c)] = c
        mask[i, :len(c)] = True
    return idx, mask


def unsat(tot, idx, mask):
    """Unsatisfied checks of the hard decisions of ``tot`` (rows, n + 1)."""
    hard = (tot[:, idx] < 0) & mask
    return np.bitwise_xor.reduce(hard, axis=2).sum(axis=1, dtype=np.int32)


def messages(tot, m, idx, mask):
    """New check-to-variable messages from totals and old messages."""
    v = tot[:, idx] - m
    t = np.tanh(np.clip(v, -38.0, 38.0) * 0.5)
    neg = t < 0
    par = np.bitwise_xor.reduce(neg & mask, axis=2)
    lt = np.log(np.clip(np.abs(t), 1e-300, 1.0))
    lt[:, ~mask] = 0.0
    lsum = lt.sum(axis=2, keepdims=True)           # pairwise, padded width
    le = np.clip(lsum - lt, LOGMIN, -1e-16)
    mag = 2.0 * np.arctanh(np.clip(np.exp(le), 0.0, 1.0 - 1e-16))
    flip = par[:, :, None] ^ neg
    return np.where(mask, np.where(flip, -mag, mag), 0.0)


def totals(lam, m, idx, n):
    """LLR + messages, each variable's terms added in ascending check order
    (np.add.at walks idx in C order)."""
    rows = lam.shape[0]
    tot = np.zeros((rows, n + 1))
    tot[:, :n] = lam
    np.add.at(tot, (np.arange(rows)[:, None, None], idx[None, :, :]), m)
    tot[:, n] = 0.0
    return tot


def iterate(lam2d, checks, n, k):
    """Totals after k message updates, with no stopping rule."""
    idx, mask = pack(checks, n)
    lam = np.asarray(lam2d, np.float64)
    tot = totals(lam, np.zeros((lam.shape[0],) + idx.shape), idx, n)
    m = np.zeros((lam.shape[0],) + idx.shape)
    for _ in range(k):
        m = messages(tot, m, idx, mask)
        tot = totals(lam, m, idx, n)
    return tot[:, :n]


def sum_product(lam2d, checks, n, iters=100, stop_at=None):
    """(bits u8, conv bool, its i32, bad i32, shed bool) per row.

    stop_at: no iteration after this one is started (the rows still failing
    then leave with their current decisions; ``shed`` marks them)."""
    idx, mask = pack(checks, n)
    lam = np.asarray(lam2d, np.float64)
    rows = lam.shape[0]
    bits = (lam < 0).astype(np.uint8)
    conv = np.zeros(rows, bool)
    its = np.full(rows, iters, np.int32)
    bad = np.full(rows, -1, np.int32)
    shed = np.zeros(rows, bool)
    act = np.arange(rows)
    m = np.zeros((rows,) + idx.shape)
    tot = totals(lam, m, idx, n)
    for it in range(iters + 1):
        u = unsat(tot, idx, mask)
        done = u == 0
        if done.any():
            sel = act[done]
            bits[sel] = tot[done, :n] < 0
            conv[sel] = True
            its[sel] = it
            bad[sel] = 0
            keep = ~done
            act, tot, m, u = act[keep], tot[keep], m[keep], u[keep]
            if len(act) == 0:
                break
        timed_out = stop_at is not None and it >= stop_at
        if it == iters or timed_out:
            bits[act] = tot[:, :n] < 0
            bad[act] = u
            its[act] = it
            shed[act] = it < iters
            break
        m = messages(tot, m, idx, mask)
        tot = totals(lam[act], m, idx, n)
    return bits, conv, its, bad, shed


def select(conv, bad, near_max=1500, max_blocks=8):
    """Rows to offer the rescue: failed, bad <= near_max, closest first."""
    conv = np.asarray(conv, bool)
    bad = np.asarray(bad)
    near = np.flatnonzero(~conv & (bad <= near_max))
    return near[np.argsort(bad[near], kind="stable")][:max_blocks]


def rescue(lam2d, checks, n, bits, conv, its, bad, near_max=1500, max_blocks=8,
           iters=100):
    """Ladder results updated by the sum-product rescue of the near misses."""
    bits, conv, its, bad = (np.array(x) for x in (bits, conv, its, bad))
    rows = select(conv, bad, near_max, max_blocks)
    if len(rows):
        b, c, i, _, _ = sum_product(np.asarray(lam2d)[rows], checks, n, iters)
        for k in np.flatnonzero(c):
            r = rows[k]
            bits[r] = b[k]
            conv[r] = True
            its[r] = i[k]
            bad[r] = 0
    return bits, conv, its, bad, rows
