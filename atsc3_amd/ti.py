"""Time de-interleaving fused into the demapper's cell read.

The receive chain gathers each FEC Block's cells out of the time
interleaver's output before demapping it:

* hybrid time interleaver (twisted block interleaver, A/322 7.1.5.4): the
  upstream builds, once per PLP layout, the index of every block cell in the
  PLP's cells and gathers ``plp0[hidx]`` (lab/m9_fast.py:1087-1093, built
  from lab/m6_tbi.py:74-96 fec_block);
* convolutional time interleaver (A/322 7.1.4): a stream over frames in
  which output cell i is received cell q(i) = i + Nrows ((start_row + i) mod
  Nrows) (lab/m10_cti.py:108-146), buffered until the interleaver's reach is
  in (lab/m44_ldm.py:1187-1304 CtiStream).

Here the demap kernel reads through either map instead of a materialised
gather: ``HtiGather`` keeps the index table on the device; ``CtiStreamDev``
keeps the received frames in a device ring buffer and the kernel evaluates
q(i) itself.  Where a consumer needs the de-interleaved cells themselves
(the LDM cancellation subtracts the re-modulated core layer from them),
``gather`` / ``take_blocks`` materialise them with one gather kernel.
"""
from __future__ import annotations

import collections

import numpy as np
import torch

from .demap import GATHER_CTI, GATHER_TABLE, GatherSpec, gather_cells_dev


def _virtual_before(i, nrows, ncols, nv):
    """Virtual-block positions among read positions 0 .. i-1 (vectorised).

    A read position p = m Nr + R is virtual when (R mod Nc + m) mod Nc < nv.
    With F(t) = #{y < t : y mod Nc < nv} = floor(t / Nc) nv + min(t mod Nc, nv),
    the count over R in [0, L) at column m is
    G(L, m) = floor(L / Nc) nv + F(m + L mod Nc) - F(m)."""
    if nv == 0:
        return np.zeros_like(i)

    def F(t):
        return (t // ncols) * nv + np.minimum(t % ncols, nv)

    b = nrows % ncols
    m_all = np.arange(ncols + 1, dtype=np.int64)
    d = F(m_all[:-1] + b) - F(m_all[:-1])
    full = np.concatenate([[0], np.cumsum(d)])        # sum over columns m' < m
    m, r0 = i // nrows, i % nrows
    per_col = (nrows // ncols) * nv
    return m * per_col + full[m] + (r0 // ncols) * nv + F(m + r0 % ncols) - F(m)


def hti_block_index(nrows, ncols, nv, j):
    """Transmitted positions (within a TI block) of the cells of data FEC
    Block j: the read position i at which memory column C = j + nv, row r
    leaves (C_i = (r mod Nc + floor(i / Nr)) mod Nc solved for floor(i / Nr)),
    less the virtual positions read before it."""
    r = np.arange(nrows, dtype=np.int64)
    c = j + nv
    i = nrows * ((c - r % ncols) % ncols) + r
    return i - _virtual_before(i, nrows, ncols, nv)


def hti_index(plp_size, nti, nrows, ncols, n_virtual=0):
    """(nti * (ncols - n_virtual), nrows) int32: index in the PLP's cells of
    every cell of every data FEC Block, TI block by TI block."""
    if not 0 <= n_virtual < ncols:
        raise ValueError("need 0 <= n_virtual < ncols")
    per_ti = plp_size // nti
    if per_ti < nrows * (ncols - n_virtual):
        raise ValueError("TI block smaller than its FEC Blocks")
    rows = [ti * per_ti + hti_block_index(nrows, ncols, n_virtual, j)
            for ti in range(nti) for j in range(ncols - n_virtual)]
    out = np.stack(rows)
    if out.max() >= 2 ** 31:
        raise ValueError("PLP too large for 32-bit cell indices")
    return out.astype(np.int32)


class HtiGather:
    """Device index table of one PLP layout; demaps every FEC Block of the
    PLP's cells through it."""

    def __init__(self, plp_size, nti, nrows, ncols, n_virtual=0, device="cuda", frames=1,
                 stride=None):
        """frames > 1: the same layout in ``frames`` consecutive frames of
        ``stride`` cells (the PLP at the start of each), one table."""
        idx = hti_index(plp_size, nti, nrows, ncols, n_virtual).astype(np.int64)
        if frames > 1:
            stride = int(stride)
            if stride < plp_size:
                raise ValueError("frame stride shorter than the PLP")
            idx = (np.arange(frames)[:, None, None] * stride + idx[None]).reshape(-1, nrows)
            if idx.max() >= 2 ** 31:
                raise ValueError("frames too large for 32-bit cell indices")
        self.plp_size = int(plp_size) if frames == 1 else (frames - 1) * stride + int(plp_size)
        idx = idx.astype(np.int32)
        self.nblk, self.ncell = idx.shape
        self.idx = torch.from_numpy(np.ascontiguousarray(idx.reshape(-1))).to(device)
        self.spec = GatherSpec(GATHER_TABLE, 0, 0, 0, self.idx.data_ptr(), 0, 0)

    def demap(self, demapper, cells, csi=None, sigma2=None, out=None):
        """cells: the PLP's cells (1-D complex64 device tensor, transmitted
        order) -> ((nblk, n) codeword-order LLRs, sigma2 per block)."""
        if demapper.ncell != self.ncell:
            raise ValueError("demapper block size differs from the FEC Block size")
        if cells.numel() < self.plp_size:
            raise ValueError("fewer cells than the PLP holds")
        return demapper.demap_gather_dev(cells, self.spec, self.nblk, csi, sigma2, out)

    def gather(self, cells):
        """The gathered blocks themselves ((nblk, ncell) device tensor)."""
        return gather_cells_dev(cells.contiguous(), self.spec, self.nblk, self.ncell)


class CtiStreamDev:
    """Received frames in (device ring buffer), complete FEC Blocks out.

    Same bookkeeping as lab/m44_ldm.py:1187-1304: absolute cell indices count
    from the first cell of the frame whose L1 signalled ``start_row``; a
    missing frame is zero-filled so the commutator phase (a function of the
    cell index) survives it; a block can be taken once the received stream
    reaches Nrows (Nrows - 1) cells past its last cell.  Absolute index a
    lives in ring slot a mod capacity."""

    def __init__(self, nrows, ncell, plp_size, cap_frames=4, device="cuda", with_csi=False):
        self.nrows, self.ncell, self.plp_size = int(nrows), int(ncell), int(plp_size)
        self.reach = self.nrows * (self.nrows - 1)
        self.cap = max(int(cap_frames), 3) * self.plp_size
        if self.cap < self.reach + 2 * self.ncell + self.plp_size:
            raise ValueError("ring too small for the interleaver's reach")
        self.device = torch.device(device)
        self.ring = torch.zeros(self.cap, dtype=torch.complex64, device=self.device)
        self.csi = (torch.zeros(self.cap, dtype=torch.float32, device=self.device)
                    if with_csi else None)
        self.origin_frame = None
        self.stats = collections.Counter()

    def reset(self, origin_frame, start_row, fec_block_start):
        self.start_row = int(start_row) % self.nrows
        c0 = int(fec_block_start) - self.nrows * ((self.start_row + int(fec_block_start))
                                                  % self.nrows)
        if not 0 <= c0 < self.ncell:
            raise ValueError(f"C={c0} outside [0,{self.ncell}): start_row / "
                             f"fec_block_start disagree")
        self.C0 = c0
        self.origin_frame = self.next_frame = int(origin_frame)
        self.base = self.end = 0
        self.out_ptr = c0
        self.block_id = 0
        self.ring.zero_()
        if self.csi is not None:
            self.csi.zero_()

    def _write(self, a, w=None):
        n = self.plp_size
        if self.end - self.base + n > self.cap:
            self.base = max(self.base, self.out_ptr)
            if self.end - self.base + n > self.cap:
                self.base += self.end - self.base + n - self.cap
                self.out_ptr = max(self.out_ptr, self.base)
                self.stats["buffer_overrun"] += 1
        o = self.end % self.cap
        k = min(n, self.cap - o)
        for dst, src in ((self.ring, a), (self.csi, w)):
            if dst is None:
                continue
            if src is None:
                dst[o:o + k].zero_()
                if k < n:
                    dst[:n - k].zero_()
            else:
                dst[o:o + k].copy_(src[:k])
                if k < n:
                    dst[:n - k].copy_(src[k:])
        self.end += n

    def push(self, frame_idx, cells, csi=None):
        """Append one frame's cells (device or host, plp_size of them); None
        for a frame that never arrived."""
        if self.origin_frame is None:
            return
        while self.next_frame < frame_idx:
            self._write(None)
            self.next_frame += 1
            self.stats["zero_filled_frames"] += 1
        if frame_idx < self.next_frame:
            self.stats["late_frame_dropped"] += 1
            return
        if cells is None:
            self.stats["zero_filled_frames"] += 1
            self._write(None)
        else:
            c = torch.as_tensor(cells).to(self.device, torch.complex64).reshape(-1)
            if c.numel() != self.plp_size:
                raise ValueError(f"a frame holds {self.plp_size} cells")
            w = None
            if self.csi is not None:
                w = (torch.ones(self.plp_size, dtype=torch.float32, device=self.device)
                     if csi is None else
                     torch.as_tensor(csi).to(self.device, torch.float32).reshape(-1))
            self._write(c, w)
        self.next_frame += 1

    def available_blocks(self):
        if self.origin_frame is None:
            return 0
        return max(0, (self.end - self.reach - self.out_ptr) // self.ncell)

    def _spec(self):
        return GatherSpec(GATHER_CTI, self.nrows, self.start_row, 0, None, self.out_ptr,
                          self.cap)

    def take_demap(self, demapper, nmax=None, sigma2=None, out=None):
        """Demap every complete block (at most nmax) straight from the ring:
        (first block id, (nb, n) LLRs, sigma2 per block)."""
        if demapper.ncell != self.ncell:
            raise ValueError("demapper block size differs from the FEC Block size")
        nb = self.available_blocks()
        if nmax is not None:
            nb = min(nb, int(nmax))
        first = self.block_id
        if nb <= 0:
            return first, None, None
        llr, s2 = demapper.demap_gather_dev(self.ring, self._spec(), nb, self.csi, sigma2, out)
        self.out_ptr += nb * self.ncell
        self.block_id += nb
        return first, llr, s2

    def take_blocks(self, nmax=None):
        """The de-interleaved cells of every complete block (for checking)."""
        nb = self.available_blocks()
        if nmax is not None:
            nb = min(nb, int(nmax))
        if nb <= 0:
            return self.block_id, torch.zeros((0, self.ncell), dtype=torch.complex64,
                                              device=self.device)
        if self.device.type == "cuda":
            cells = gather_cells_dev(self.ring, self._spec(), nb, self.ncell)
        else:                                  # host ring (bookkeeping checks)
            i = torch.arange(self.out_ptr, self.out_ptr + nb * self.ncell, dtype=torch.int64)
            q = i + self.nrows * ((self.start_row + i) % self.nrows)
            cells = self.ring[q % self.cap].view(nb, self.ncell)
        first = self.block_id
        self.out_ptr += nb * self.ncell
        self.block_id += nb
        return first, cells
