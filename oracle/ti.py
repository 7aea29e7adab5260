"""CPU oracle of the time de-interleavers in front of the demapper.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

* ``tbi_read_order`` / ``tbi_deinterleave`` / ``tbi_interleave`` restate the
  A/322 7.1.5.4 twisted block interleaver equations as lab/m6_tbi.py:47-71
  states them: R_i = i mod Nr, T_i = R_i mod Nc, C_i = (T_i + floor(i/Nr))
  mod Nc, theta_i = Nr C_i + R_i, skipping theta_i < N_virtual Nr.
* ``hti_block_positions`` / ``hti_frame_index``: the transmitted positions
  of each data FEC Block's cells, by inverting the read order as a
  permutation (the upstream's fec_block, lab/m6_tbi.py:74-96, and its
  batched gather lab/m9_fast.py:1087-1093).
* ``cti_*`` restate the convolutional de-interleaver's index map of
  lab/m10_cti.py:108-152: q(i) = i + Nrows ((start_row + i) mod Nrows).
* ``CtiStream`` restates the receive stream of lab/m44_ldm.py:1187-1304
  (frames appended, missing ones zero-filled, complete FEC Blocks taken once
  the CTI's reach Nrows (Nrows - 1) is buffered) on an unbounded array; with
  ``cap_frames`` it applies the upstream's capacity rule (lab/m44_ldm.py:
  1262-1283): cells before the next output are released first, and when the
  consumer has fallen further behind the oldest cells are dropped and the
  next output moves up to the first cell still held (an overrun).

Pinning (tests/test_ti.py): A/327 Figure 6.5's printed 4x3 example with one
virtual FEC Block (output b g a f d e c h, lab/m6_tbi.py:41-42); the read
order is a bijection over legal geometries; interleave -> de-interleave is
the identity; the CTI map is injective with delay k Nrows at commutator row
k, and the interleave -> de-interleave round trip restores every valid cell;
A/322 9.3.9.1's signalled fec_block_start inverts to C.
"""
from __future__ import annotations

import numpy as np

# A/327 Figure 6.5 (as transcribed in lab/m6_tbi.py:41-42).
GOLD = dict(nrows=4, ncols=3, n_virtual=1, expect="b g a f d e c h".split())


def tbi_read_order(nrows, ncols, n_virtual=0):
    i = np.arange(nrows * ncols, dtype=np.int64)
    r = i % nrows
    c = (r % ncols + i // nrows) % ncols
    theta = nrows * c + r
    keep = theta >= n_virtual * nrows
    return theta[keep] - n_virtual * nrows


def tbi_interleave(mem, nrows, ncols, n_virtual=0):
    return np.asarray(mem)[tbi_read_order(nrows, ncols, n_virtual)]


def tbi_deinterleave(cells, nrows, ncols, n_virtual=0):
    order = tbi_read_order(nrows, ncols, n_virtual)
    cells = np.asarray(cells)
    if len(order) != len(cells):
        raise ValueError("cell count does not match the geometry")
    out = np.empty_like(cells)
    out[order] = cells
    return out


def hti_block_positions(nrows, ncols, n_virtual, j):
    """Transmitted positions (within the TI block) of data FEC Block j."""
    order = tbi_read_order(nrows, ncols, n_virtual)
    inv = np.empty(len(order), np.int64)
    inv[order] = np.arange(len(order))
    return inv[j * nrows:(j + 1) * nrows]


def hti_frame_index(plp_size, nti, nrows, ncols, n_virtual=0):
    """(nti * (ncols - n_virtual), nrows) pool index of every block cell."""
    per_ti = plp_size // nti
    return np.stack([ti * per_ti + hti_block_positions(nrows, ncols, n_virtual, j)
                     for ti in range(nti) for j in range(ncols - n_virtual)])


def cti_out_index(i, nrows, start_row):
    i = np.asarray(i, np.int64)
    return i + nrows * ((start_row + i) % nrows)


def cti_in_index(q, nrows, start_row):
    q = np.asarray(q, np.int64)
    return q - nrows * ((start_row + q) % nrows)


def cti_interleave(cells, nrows, start_row, n_out=None):
    """Transmit side; outputs fed by no input cell carry 0."""
    cells = np.asarray(cells)
    n = len(cells) if n_out is None else n_out
    out = np.zeros(n, cells.dtype)
    q = np.arange(n, dtype=np.int64)
    i = cti_in_index(q, nrows, start_row)
    ok = (i >= 0) & (i < len(cells))
    out[q[ok]] = cells[i[ok]]
    return out


def cti_deinterleave(recv, nrows, start_row, n_out=None):
    recv = np.asarray(recv)
    n = len(recv) if n_out is None else n_out
    q = cti_out_index(np.arange(n), nrows, start_row)
    valid = q < len(recv)
    out = np.zeros(n, recv.dtype)
    out[valid] = recv[q[valid]]
    return out, valid


def solve_C(fec_block_start, start_row, nrows):
    return fec_block_start - nrows * ((start_row + fec_block_start) % nrows)


class CtiStream:
    """Frames in, complete FEC Blocks out (capacity-limited with cap_frames)."""

    def __init__(self, nrows, ncell, plp_size, cap_frames=None):
        self.nrows, self.ncell, self.plp_size = nrows, ncell, plp_size
        self.reach = nrows * (nrows - 1)
        self.cap = None if cap_frames is None else max(cap_frames, 3) * plp_size
        self.cells = np.zeros(0, np.complex64)
        self.next_frame = None
        self.overruns = 0

    def reset(self, origin_frame, start_row, fec_block_start):
        self.start_row = start_row % self.nrows
        self.C0 = solve_C(fec_block_start, self.start_row, self.nrows)
        if not 0 <= self.C0 < self.ncell:
            raise ValueError("C outside one FEC Block")
        self.cells = np.zeros(0, np.complex64)
        self.next_frame = origin_frame
        self.out_ptr = self.C0
        self.block_id = 0
        self.base = 0

    def _room(self):
        if self.cap is None:
            return
        end, n = len(self.cells), self.plp_size
        if end - self.base + n > self.cap:
            self.base = max(self.base, self.out_ptr)
            if end - self.base + n > self.cap:
                self.base += end - self.base + n - self.cap
                self.out_ptr = max(self.out_ptr, self.base)
                self.overruns += 1

    def push(self, frame_idx, cells):
        if self.next_frame is None or frame_idx < self.next_frame:
            return
        zero = np.zeros(self.plp_size, np.complex64)
        while self.next_frame < frame_idx:
            self._room()
            self.cells = np.concatenate([self.cells, zero])
            self.next_frame += 1
        self._room()
        self.cells = np.concatenate([self.cells, zero if cells is None
                                     else np.asarray(cells, np.complex64)])
        self.next_frame = frame_idx + 1

    def available_blocks(self):
        if self.next_frame is None:
            return 0
        return max(0, (len(self.cells) - self.reach - self.out_ptr) // self.ncell)

    def take_blocks(self, nmax=None):
        nb = self.available_blocks()
        if nmax is not None:
            nb = min(nb, nmax)
        if nb <= 0:
            return []
        i = np.arange(self.out_ptr, self.out_ptr + nb * self.ncell, dtype=np.int64)
        c = self.cells[cti_out_index(i, self.nrows, self.start_row)]
        out = [(self.block_id + m, c[m * self.ncell:(m + 1) * self.ncell]) for m in range(nb)]
        self.out_ptr += nb * self.ncell
        self.block_id += nb
        return out
