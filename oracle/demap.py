"""FP64 max-log demap oracle.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

``demap_q`` restates lab/m6_bicm.py:213-225 (``demap_llr``) batched over
blocks: for each cell z and label bit i (MSB first),

    LLR_i = w * (min_{p: y_i = 1} |z - p|^2 - min_{p: y_i = 0} |z - p|^2) / sigma2

with w = 1 (no CSI, the upstream case) or the per-cell channel power |h|^2 of
equalised cells z = y / h.  ``sigma2`` defaults per block to
max(mean_cells w * min_p |z - p|^2, 1e-9); with w = 1 that is exactly the
upstream estimator (lab/m6_bicm.py:219-220, lab/m9_fast.py:821).  Distances
are |z - p| via hypot then squared, as ``np.abs`` does upstream.

``to_codeword`` is the de-interleave ``lam[lam_of_q] = q`` of
lab/m6_bicm.py:331-335.

``ldm_core_inputs`` restates lab/m15_ldm_llr.py:77-85 (``powers``,
``llr_gaussian``): the core layer is demapped from y / sqrt(P_core) with
sigma2 = (P_enh + N0) / P_core.

Pinning: tests/test_oracle.py checks ``demap_q`` against the QPSK closed form
(LLR_y0 = 2*sqrt(2)*Im z / sigma2, LLR_y1 = 2*sqrt(2)*Re z / sigma2), against
a brute-force per-cell loop on a few cells for every constellation, the sign
rule on noiseless cells (LLR sign = transmitted bit), scale invariance in w
and sigma2, and the estimator identity.
"""
from __future__ import annotations

import numpy as np


def _masks(nb):
    k = np.arange(1 << nb)
    return np.array([((k >> (nb - 1 - i)) & 1).astype(bool) for i in range(nb)])


def demap_q(cells, pts, csi=None, sigma2=None, chunk=2048):
    """cells (B, C) complex -> q-ordered FP64 LLRs (B, C*nb), sigma2 (B,)."""
    z = np.atleast_2d(np.asarray(cells, np.complex128))
    pts = np.asarray(pts, np.complex128)
    B, C = z.shape
    nb = int(np.log2(len(pts)))
    ones = _masks(nb)
    w = np.ones((B, C)) if csi is None else np.atleast_2d(np.asarray(csi, np.float64))
    diff = np.empty((B, C, nb))
    d2min = np.empty((B, C))
    for lo in range(0, C, chunk):
        zz = z[:, lo:lo + chunk]
        d2 = np.abs(zz[:, :, None] - pts[None, None, :]) ** 2
        d2min[:, lo:lo + chunk] = d2.min(2)
        for i in range(nb):
            diff[:, lo:lo + chunk, i] = (d2[:, :, ones[i]].min(2)
                                         - d2[:, :, ~ones[i]].min(2))
    if sigma2 is None:
        s2 = np.maximum((w * d2min).mean(axis=1), 1e-9)
    else:
        s2 = np.broadcast_to(np.asarray(sigma2, np.float64), (B,)).copy()
    llr = diff * (w / s2[:, None])[:, :, None]
    return llr.reshape(B, C * nb), s2


def to_codeword(q, lam_of_q):
    """q-ordered (B, N) -> codeword order: lam[:, lam_of_q] = q."""
    q = np.atleast_2d(q)
    out = np.empty_like(q)
    out[:, lam_of_q] = q
    return out


def demap(cells, pts, lam_of_q, csi=None, sigma2=None):
    q, s2 = demap_q(cells, pts, csi, sigma2)
    return to_codeword(q, lam_of_q), s2


def ldm_powers(inj_db):
    """(P_core, P_enh) for injection level inj_db (lab/m15_ldm_llr.py:77-79)."""
    r = 10 ** (-inj_db / 10.0)
    return 1 / (1 + r), r / (1 + r)


def ldm_core_inputs(y, inj_db, n0):
    """Core-layer demap inputs: (y / sqrt(P_core), (P_enh + N0) / P_core)."""
    pc, pe = ldm_powers(inj_db)
    return np.asarray(y) / np.sqrt(pc), (pe + n0) / pc


def normwise_rel_err(got, want):
    """max|got - want| / max|want| per block -- the relative criterion of
    lab/m16_demap_fast.py:136-139."""
    got = np.atleast_2d(got)
    want = np.atleast_2d(want)
    return (np.abs(got - want).max(axis=1)
            / np.maximum(np.abs(want).max(axis=1), 1e-300))
