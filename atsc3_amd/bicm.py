"""A/322 BICM set-up tables: constellations, label bits and the composite
bit de-interleave map ``lam_of_q``.

Set-up only: the demap kernel consumes these once per (N, mod, rate).

Labels (A/322 6.3): a cell carries nb bits y0..y_{nb-1}, y0 the most
significant bit of the label k.  Every alphabet here follows the quadrant
rule of A/322 Annex C: the M/4 base points w (first quadrant) are indexed by
the low nb-2 bits of k; y1 flips the real part, y0 the imaginary part.
QPSK is that rule with the single base point (1 + j)/sqrt(2) (Table C.1.1).
Interfaces mirrored: ``points`` = lab/m6_bicm.py:197-200 (points_for) over
lab/spec_nuc.py:77-80,162-166; ``bit_masks`` = lab/m6_bicm.py:203-210;
``map_cells`` = lab/m6_bicm.py:228-236; ``bicm_map`` =
lab/m6_bicm.py:185-190.

Bit interleaver (A/322 6.2), transmitter order u -> v -> q:
  * parity interleave 6.2.1: the identity on this library's codeword order,
    because its encoder (atsc3_amd/codes.py) already emits the parity bits
    in the 6.2.1 order (lab/m6_bicm.py:76-107 records the same fact;
    tests/test_bicm.py re-derives it from 6.1.3.2);
  * group-wise 6.2.2: output bit group g of v carries input group pi(g);
  * block interleave 6.2.3: Type A writes v column by column into a part-1
    array of Nr1 rows and a part-2 array of Nr2 rows (nb columns each) and
    reads them row by row, one row per cell; Type B takes the first Npart1
    bits in blocks of NQCB_IG groups, each block written as NQCB_IG rows of
    360 and read column by column, and passes part 2 through.
``lam_of_q[j]`` is computed in one pass per q slot j by inverting those
read-outs in closed form: slot j -> the v position read there -> the u
position group-wise -> the codeword bit.  The receiver's de-interleave is
``lam[lam_of_q] = q`` (lab/m6_bicm.py:331-335), fused into the demap kernel.
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
BI_RATES = list(_BI["rates"])

_ORDER = {"16QAM": 16, "64QAM": 64, "256QAM": 256}


def base_points(mod, rate):
    """The M/4 first-quadrant base points w (complex128)."""
    if mod == "QPSK":
        return np.array([(1 + 1j) / np.sqrt(2)])
    return np.asarray(_NUC[f"w{_ORDER[mod]}_{rate.replace('/', '_')}"], np.complex128)


def quadrant_alphabet(w):
    """Alphabet indexed by label: base point w[k mod M/4], real part negated
    when y1 = 1, imaginary part when y0 = 1."""
    w = np.asarray(w, np.complex128)
    m = 4 * len(w)
    nb = m.bit_length() - 1
    k = np.arange(m)
    y0 = (k >> (nb - 1)) & 1
    y1 = (k >> (nb - 2)) & 1
    b = w[k & (len(w) - 1)]
    return (1 - 2 * y1) * b.real + 1j * (1 - 2 * y0) * b.imag


@functools.lru_cache(maxsize=None)
def points(mod, rate):
    """Full M-point alphabet (complex128), index = label."""
    return quadrant_alphabet(base_points(mod, rate))


def bit_masks(nb):
    """masks[i, k] is True where label bit y_i of label k is 1 (y0 = MSB)."""
    shift = np.arange(nb - 1, -1, -1)[:, None]
    return ((np.arange(1 << nb)[None, :] >> shift) & 1).astype(bool)


def cells_per_block(n, mod):
    return DEMUX[(n, mod)][1]


def block_read_index(n, mod, rate):
    """v position the block interleaver reads out at q slot j (A/322 6.2.3)."""
    nb = MOD_BITS[mod]
    j = np.arange(n, dtype=np.int64)
    if BI_TYPE[(n, nb, rate)] == "A":
        nr1, nr2, _ = TYPE_A[(n, mod)]
        row, col = np.divmod(j, nb)             # cell row, bit column
        in2 = row >= nr1
        return np.where(in2, nb * nr1 + col * nr2 + (row - nr1), col * nr1 + row)
    nq, np1, _ = TYPE_B[(n, mod)]
    span = nq * 360
    blk, w = np.divmod(j, span)
    col, row = np.divmod(w, nq)                 # column read out, group row
    return np.where(j < np1, blk * span + row * 360 + col, j)


@functools.lru_cache(maxsize=None)
def bicm_map(n, mod, rate, invert_groupwise=False):
    """lam_of_q[j] = LDPC codeword bit carried by q-stream bit j.

    invert_groupwise reads 6.2.2 the other way round (group g of v carries
    input group pi^-1(g)): a negative control only."""
    pi = np.asarray(GROUPWISE[(n, mod, rate)], np.int64)
    if invert_groupwise:
        pi = np.argsort(pi)
    grp, pos = np.divmod(block_read_index(n, mod, rate), 360)
    out = pi[grp] * 360 + pos
    out.setflags(write=False)
    return out


def map_cells(bits_q, mod, rate):
    """Transmitter mapper: q-ordered bits (..., ncell*nb) -> cells."""
    nb = MOD_BITS[mod]
    b = np.asarray(bits_q, np.int64)
    b = b.reshape(*b.shape[:-1], -1, nb)
    label = b @ (1 << np.arange(nb - 1, -1, -1, dtype=np.int64))
    return points(mod, rate)[label]
