"""FP64 oracle of the LDM cancellation step.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

``cancel`` restates lab/e69_enh.py:102-106: re-modulate the decoded core
codeword (s = map_cells(cw[lam_of_q]), lab/m6_bicm.py:228-236), gain =
vdot(s, y) / vdot(s, s), residual y - gain * s; blocks that did not decode
are left as they are.

Pinning: tests/test_ldm.py checks it on noiseless superpositions (the
residual equals sqrt(P_enh) * x_enh exactly up to rounding, gain equals
sqrt(P_core)), the gain estimator against a least-squares solve, and the
pass-through of non-converged blocks.
"""
from __future__ import annotations

import numpy as np


def remodulate(cw, lam_of_q, pts):
    cw = np.atleast_2d(np.asarray(cw, np.int64))
    nb = int(np.log2(len(pts)))
    q = cw[:, lam_of_q].reshape(cw.shape[0], -1, nb)
    k = np.zeros(q.shape[:2], np.int64)
    for i in range(nb):
        k |= q[:, :, i] << (nb - 1 - i)
    return np.asarray(pts)[k]


def cancel(y, cw, conv, lam_of_q, pts, out_scale=1.0):
    y = np.atleast_2d(np.asarray(y, np.complex128))
    s = remodulate(cw, lam_of_q, pts)
    num = np.sum(np.conj(s) * y, axis=1)
    den = np.sum(np.abs(s) ** 2, axis=1)
    gain = np.where(np.asarray(conv, bool), num / den, 0.0)
    r = np.where(np.asarray(conv, bool)[:, None], y - gain[:, None] * s, y)
    return r * out_scale, gain
