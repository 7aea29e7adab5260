"""A/322 BICM host tables: NUC constellations, label bit masks and the
composite bit (de)interleaver permutation ``lam_of_q``.

These are set-up tables the demap kernel consumes; none of this runs per cell.

* ``points(mod, rate)``      lab/m6_bicm.py:197-200, lab/spec_nuc.py:77-80
  (quadrant rule x = [w, -conj w, conj w, -w]; label k = decimal value of
  (y0..y_{nb-1}), y0 the MSB).  QPSK = [1+j, -1+j, 1-j, -1-j]/sqrt(2)
  (lab/m6_bicm.py:69).
* ``bit_masks(nb)``          lab/m6_bicm.py:203-210.
* ``bicm_map(n, mod, rate)`` lab/m6_bicm.py:76-190: ``lam_of_q[j]`` is the
  LDPC codeword bit carried by q_j.  The parity interleaver is the identity
  here (lab/m6_bicm.py:76-107: the encoder already emits 6.2.1's order), the
  group-wise interleaver moves 360-bit groups (:137-149) and the block
  interleaver is Type A column/row or Type B (:152-182).  The receiver's
  de-interleave is ``lam[lam_of_q] = q`` (lab/m6_bicm.py:331-335).
* ``map_cells``              lab/m6_bicm.py:228-236 (transmitter side, used
  for synthetic workloads and SIC re-modulation).
"""
from __future__ import annotations

import functools
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

with open(os.path.join(HERE, "data", "bitint_tables.json")) as _f:
    _BI = json.load(_f)
_NUC = np.load(os.path.join(HERE, "data", "nuc_positions.npz"), allow_pickle=False)

MOD_BITS = dict(_BI["mod_bits"])
MODS = ("QPSK", "16QAM", "64QAM", "256QAM")
TYPE_A = {(n, m): tuple(v) for n, m, *v in _BI["type_a"]}
TYPE_B = {(n, m): tuple(v) for n, m, *v in _BI["type_b"]}
DEMUX = {(n, m): tuple(v) for n, m, *v in _BI["demux"]}
BI_TYPE = {(n, nb, r): t for n, nb, r, t in _BI["bi_type"]}
GROUPWISE = {(n, m, r): tuple(p) for n, m, r, p in _BI["groupwise"]}
GROUPWISE_HDR = {(n, m): tuple(h) for n, m, h in _BI["groupwise_hdr"]}
BI_RATES = list(_BI["rates"])

QPSK_POINTS = np.array([1 + 1j, -1 + 1j, 1 - 1j, -1 - 1j]) / np.sqrt(2)


def nuc_positions(m, rate):
    """Annex C position vector w (length M/4) for an M-point 2D NUC."""
    return np.asarray(_NUC[f"w{m}_{rate.replace('/', '_')}"], np.complex128)


def constellation(w):
    w = np.asarray(w)
    return np.concatenate([w, -np.conj(w), np.conj(w), -w])


@functools.lru_cache(maxsize=None)
def points(mod, rate):
    """Full M-point alphabet (complex128), index = label."""
    if mod == "QPSK":
        return QPSK_POINTS.copy()
    m = {"16QAM": 16, "64QAM": 64, "256QAM": 256}[mod]
    return constellation(nuc_positions(m, rate))


def bit_masks(nb):
    """masks[i, k] = bit i (MSB first) of label k."""
    k = np.arange(1 << nb)
    return np.array([((k >> (nb - 1 - i)) & 1).astype(bool) for i in range(nb)])


def cells_per_block(n, mod):
    return DEMUX[(n, mod)][1]


def groupwise_map(n, mod, rate, invert=False):
    """u_of_v[b]: index into u carried by v_b (output group j holds input
    group pi(j))."""
    p = np.asarray(GROUPWISE[(n, mod, rate)], np.int64)
    assert len(p) == n // 360
    if invert:
        p = np.argsort(p)
    return (360 * p[:, None] + np.arange(360)[None, :]).ravel()


def blockint_map(n, mod, rate):
    """v_of_q[j]: index into v carried by q_j (Type A or Type B)."""
    nc = MOD_BITS[mod]
    typ = BI_TYPE[(n, nc, rate)]
    q_of_v = np.empty(n, np.int64)
    if typ == "A":
        nr1, nr2, nc2 = TYPE_A[(n, mod)]
        assert nc2 == nc and nr1 + nr2 == n // nc
        i = np.arange(nc * nr1)
        col, row = np.divmod(i, nr1)
        q_of_v[:nc * nr1] = row * nc + col
        if nr2:
            i2 = np.arange(n - nc * nr1)
            col2, row2 = np.divmod(i2, nr2)
            q_of_v[nc * nr1:] = (nr1 + row2) * nc + col2
    else:
        nq, np1, np2 = TYPE_B[(n, mod)]
        assert nq == nc and np1 + np2 == n
        blk = nq * 360
        row, col = np.divmod(np.arange(blk), 360)
        within = col * nq + row
        base = (np.arange(np1 // blk) * blk)[:, None]
        q_of_v[:np1] = (base + within[None, :]).ravel()
        q_of_v[np1:] = np.arange(np1, n)
    v_of_q = np.empty(n, np.int64)
    v_of_q[q_of_v] = np.arange(n)
    return v_of_q


@functools.lru_cache(maxsize=None)
def bicm_map(n, mod, rate, gw_invert=False):
    """lam_of_q[j] = LDPC codeword bit carried by q_j (parity interleave is
    the identity)."""
    u_of_v = groupwise_map(n, mod, rate, gw_invert)
    v_of_q = blockint_map(n, mod, rate)
    out = u_of_v[v_of_q]
    out.setflags(write=False)
    return out


def map_cells(bits_q, mod, rate):
    """Transmitter mapper: q-ordered bits (..., ncell*nb) -> cells."""
    pts = points(mod, rate)
    nb = MOD_BITS[mod]
    b = np.asarray(bits_q, np.uint8)
    shp = b.shape[:-1]
    b = b.reshape(*shp, -1, nb).astype(np.int64)
    k = np.zeros(b.shape[:-1], np.int64)
    for i in range(nb):
        k |= b[..., i] << (nb - 1 - i)
    return pts[k]
