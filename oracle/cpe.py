"""CPU oracle of the cell-side steps in front of the demapper: per-symbol
common-phase-error correction, the dummy-cell SNR and the per-cell noise
weights of the weighted-LLR path.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

* ``cpe_correct`` restates lab/m6_cells.py:176-209 (= lab/m9_gpu.py:283-299
  on the device): cells normalised by the pool's RMS, then per OFDM symbol
  `iters` rounds of decision-directed gain fitting -- each cell's hard
  decision is the nearest point of the alphabet its region carries (dummy
  cells: their known value), c = vdot(hard, z) / max(vdot(hard, hard), 1e-12),
  z /= c -- returning the corrected cells and each symbol's total gain.
* ``frame_snr_db`` restates lab/m16_margin.py:303-318 (channel SNR off the
  known +-1 dummy cells).
* ``weighted_llr`` restates the weighted-LLR path of lab/m9_fast.py:1096-1115
  with the raw max-log differences of lab/m9_fast.py:811-822: per cell the
  squared distance to the nearest point, per pool part (part 0 = preamble
  spare, part l + 1 = OFDM symbol l, lab/m9_fast.py:927-940) the mean of it
  over the gathered cells, floored at 1e-6, and each cell's LLRs divided by
  its part's value.

Pinning (tests/test_cpe.py): noiseless cells under a per-symbol complex gain
give back the points and the gain exactly; dummy-only symbols reduce to the
least-squares gain; hard decisions equal brute-force nearest points; the SNR
of dummies a s + n matches its definition computed by least squares; the
weighted LLRs equal the FP64 demapper with the per-cell noise variance.
"""
from __future__ import annotations

import numpy as np

from . import demap as OD


def nearest(z, pts):
    """Nearest alphabet point of every cell (first index on ties)."""
    z = np.asarray(z, np.complex128)
    pts = np.asarray(pts, np.complex128)
    out = np.empty_like(z)
    for lo in range(0, len(z), 4096):
        d = np.abs(z[lo:lo + 4096, None] - pts[None, :])
        out[lo:lo + 4096] = pts[d.argmin(1)]
    return out


def cpe_correct(pool, symbol_of, alphabets, region, dummy_values, iters=3, decim=1):
    """decim > 1: the gain of a symbol is estimated from its cells at pool
    indices divisible by decim (lab/m9_fast.py:667-678), then applied to all."""
    pool = np.asarray(pool, np.complex128)
    out = pool / np.sqrt(np.mean(np.abs(pool) ** 2))
    gains = {}
    for sym in np.unique(symbol_of):
        m = np.flatnonzero(symbol_of == sym)
        z = out[m]
        r = np.asarray(region)[m]
        est = m % decim == 0
        c_tot = 1.0 + 0j
        for _ in range(iters):
            zs, rs, ms = z[est], r[est], m[est]
            hard = np.empty_like(zs)
            for rid, pts in enumerate(alphabets):
                sel = rs == rid
                if sel.any():
                    hard[sel] = nearest(zs[sel], pts)
            sel = rs == 2
            if sel.any():
                hard[sel] = np.asarray(dummy_values)[ms[sel]]
            c = np.vdot(hard, zs) / max(np.vdot(hard, hard).real, 1e-12)
            z = z / c
            c_tot *= c
        out[m] = z
        gains[int(sym)] = complex(c_tot)
    return out, gains


def frame_snr_db(pool, dummy_start, scr_bits):
    """``scr_bits``: the scrambling sequence over the whole pool."""
    d = np.asarray(pool, np.complex128)[dummy_start:]
    if len(d) < 64:
        return None
    s = 1.0 - 2.0 * np.asarray(scr_bits, np.float64)[dummy_start:dummy_start + len(d)]
    a = float(np.dot(d.real, s) / np.dot(s, s))
    noise = float(np.mean(np.abs(d - a * s) ** 2))
    if noise <= 0 or a == 0.0:
        return None
    return float(10.0 * np.log10(a * a / noise))


def weighted_llr(cells2d, pts, pid, npart):
    """(nblk, ncell) cells, part id of each -> (q-order LLRs (nblk, ncell*nb),
    per-part noise variance)."""
    cells2d = np.asarray(cells2d, np.complex128)
    nb = int(round(np.log2(len(pts))))
    raw, _ = OD.demap_q(cells2d, pts, None, np.ones(len(cells2d)))
    d2 = np.empty(cells2d.shape)
    for b in range(len(cells2d)):
        d2[b] = np.abs(cells2d[b][:, None] - np.asarray(pts)[None, :]).min(1) ** 2
    cnt = np.maximum(np.bincount(pid.ravel(), minlength=npart), 1)
    s2p = np.bincount(pid.ravel(), weights=d2.ravel(), minlength=npart) / cnt
    s2p = np.maximum(s2p, 1e-6)
    q = raw.reshape(len(cells2d), -1, nb) / s2p[pid][:, :, None]
    return q.reshape(len(cells2d), -1), s2p
