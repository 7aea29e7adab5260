"""Cells in front of the demapper: common-phase-error correction per OFDM
symbol, the dummy-cell SNR and per-part noise weights, on the device.

* ``Cpe`` -- lab/m6_cells.py:176-209 ``cpe_correct`` (batched on the GPU in
  lab/m9_gpu.py:283-299, threaded in lab/m9_fast.py:619-634): the pool
  normalised by its RMS, then per symbol a few rounds of decision-directed
  complex-gain fitting against the alphabet each cell is known to carry.
* ``frame_snr_db`` -- lab/m16_margin.py:303-318: the channel SNR off the
  known +-1 dummy cells (the switch of the weighted-LLR path).
* ``WeightedDemap`` -- lab/m9_fast.py:1096-1115: LLRs divided per pool part
  (preamble spare, each OFDM symbol) by the mean squared distance of that
  part's cells to their nearest points, instead of one sigma^2 per block.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .baseband import scrambler_sequence
from .demap import GATHER_TABLE, GatherSpec, _check_quadrant_rule


def _stream(dev):
    return torch.cuda.current_stream(dev).cuda_stream


def _pts(p, dev):
    """Alphabet as (M, 2) float64 (the hard decisions use the exact points)."""
    p = np.asarray(p, np.complex128)
    return torch.from_numpy(np.ascontiguousarray(np.stack([p.real, p.imag], 1))).to(dev)


class Cpe:
    """Device layout of one pool geometry (symbol of every cell, region of
    every cell, the region alphabets and dummy values)."""

    MAX_ITERS = 8

    def __init__(self, symbol_of, region, alphabets, dummy_values=None, device="cuda"):
        dev = self.device = torch.device(device)
        symbol_of = np.asarray(symbol_of)
        region = np.asarray(region)
        self.npool = len(symbol_of)
        if region.shape != symbol_of.shape:
            raise ValueError("one region per cell")
        if len(alphabets) > 2:
            raise ValueError("at most two alphabets (regions 0 and 1)")
        if not np.isin(region, (0, 1, 2)).all():
            raise ValueError("regions are 0, 1 (alphabets) or 2 (dummy)")
        for rid in (0, 1):
            if (region == rid).any() and rid >= len(alphabets):
                raise ValueError(f"region {rid} has no alphabet")
        if (region == 2).any() and dummy_values is None:
            raise ValueError("dummy cells need their values")
        self.symbols, inv = np.unique(symbol_of, return_inverse=True)
        self._inv = inv
        self.cells, self.seg = self._lists(np.arange(self.npool))
        self._est = {1: (self.cells, self.seg)}
        self.region = torch.from_numpy(region.astype(np.uint8)).to(dev)
        # quadrant-rule alphabets are searched through their base points
        self.alph, self.msizes = [], []
        for a in alphabets:
            a = np.asarray(a, np.complex128)
            try:
                _check_quadrant_rule(a)
                self.alph.append(_pts(a[: len(a) // 4], dev))
                self.msizes.append(-(len(a) // 4))
            except ValueError:
                self.alph.append(_pts(a, dev))
                self.msizes.append(len(a))
        self.dummy = None
        if dummy_values is not None:
            self.dummy = torch.from_numpy(np.asarray(dummy_values, np.complex64)).to(dev)
        self.ws = torch.empty(_lib.lib().atsc3_cpe_ws_doubles(self.npool), dtype=torch.float64,
                              device=dev)

    def _lists(self, which):
        """Per-symbol CSR list (device) of the pool cells ``which``."""
        inv = self._inv[which]
        order = which[np.argsort(inv, kind="stable")]
        seg = np.concatenate([[0], np.cumsum(np.bincount(inv, minlength=len(self.symbols)))])
        return (torch.from_numpy(order.astype(np.int32)).to(self.device),
                torch.from_numpy(seg.astype(np.int32)).to(self.device))

    def run(self, pool, iters=3, out=None, decim=1):
        """pool (npool,) complex64 device tensor -> (corrected cells,
        gains (nsym,) complex128 in the order of ``self.symbols``).  With
        decim > 1 only cells at pool indices divisible by decim estimate the
        gains (the upstream's fast path, decim 4); every cell is corrected."""
        if not (pool.is_cuda and pool.dtype == torch.complex64 and pool.numel() == self.npool):
            raise ValueError(f"pool must be a ({self.npool},) complex64 device tensor")
        if not 0 <= iters <= self.MAX_ITERS:
            raise ValueError(f"iters must be 0..{self.MAX_ITERS}")
        pool = pool.contiguous()
        dev = self.device
        out = torch.empty_like(pool) if out is None else out
        gains = torch.empty(len(self.symbols), dtype=torch.complex128, device=dev)
        p0 = self.alph[0].data_ptr() if len(self.alph) > 0 else None
        p1 = self.alph[1].data_ptr() if len(self.alph) > 1 else None
        m0 = self.msizes[0] if len(self.alph) > 0 else 0
        m1 = self.msizes[1] if len(self.alph) > 1 else 0
        if decim not in self._est:
            self._est[decim] = self._lists(np.arange(0, self.npool, int(decim)))
        ec, es = self._est[decim]
        _lib.call("atsc3_cpe_dev", torch.view_as_real(pool).data_ptr(), self.npool,
                  self.cells.data_ptr(), self.seg.data_ptr(), ec.data_ptr(), es.data_ptr(),
                  len(self.symbols),
                  self.region.data_ptr(), p0, m0, p1, m1,
                  None if self.dummy is None else torch.view_as_real(self.dummy).data_ptr(),
                  int(iters), torch.view_as_real(out).data_ptr(),
                  torch.view_as_real(gains).data_ptr(), self.ws.data_ptr(), _stream(dev))
        return out, gains


def dummy_values(pool_len, dummy_start):
    """Known dummy-cell values (+-1 from the scrambling sequence over the
    pool, lab/m6_cells.py:159-173)."""
    dv = np.zeros(pool_len, np.complex64)
    if pool_len > dummy_start:
        seq = scrambler_sequence(pool_len)
        dv[dummy_start:] = 1.0 - 2.0 * seq[dummy_start:pool_len].astype(np.float32)
    return dv


def frame_snr_db(pool, dummy_start):
    """Channel SNR (dB) off the dummy cells pool[dummy_start:], or None."""
    n = pool.numel() - dummy_start
    if n < 64:
        return None
    dev = pool.device
    bits = torch.from_numpy(scrambler_sequence(pool.numel())[dummy_start:].copy()).to(dev)
    d = pool[dummy_start:].contiguous()
    out = torch.empty(2, dtype=torch.float64, device=dev)
    _lib.call("atsc3_dummy_snr_dev", torch.view_as_real(d).data_ptr(), n, bits.data_ptr(),
              out.data_ptr(), _stream(dev))
    v = out.cpu().numpy()
    return float(v[0]) if v[1] else None


class WeightedDemap:
    """Per-part noise weights for one block layout: ``part[b, s]`` is the part
    of block b's cell s (e.g. 1 + the OFDM symbol of the pool cell it is
    gathered from, 0 for the preamble spare, lab/m9_fast.py:927-940)."""

    def __init__(self, demapper, part, nparts=None, gather_idx=None, device="cuda"):
        dev = self.device = torch.device(device)
        part = np.asarray(part)
        self.nblk, self.ncell = part.shape
        if self.ncell != demapper.ncell:
            raise ValueError("part table does not match the demapper's block size")
        if nparts is None:
            nparts = int(part.max()) + 1 if part.size else 1
        self.nparts = int(nparts)
        if part.size and (part.min() < 0 or part.max() >= self.nparts):
            raise ValueError("part ids outside 0 .. nparts-1")
        flat = part.reshape(-1)
        order = np.argsort(flat, kind="stable")
        seg = np.concatenate([[0], np.cumsum(np.bincount(flat, minlength=self.nparts))])
        self.dm = demapper
        self.part = torch.from_numpy(flat.astype(np.int32)).to(dev)
        self.order = torch.from_numpy(order.astype(np.int32)).to(dev)
        self.seg = torch.from_numpy(seg.astype(np.int32)).to(dev)
        self.idx = None
        if gather_idx is not None:
            gi = np.asarray(gather_idx).reshape(-1)
            if gi.size != flat.size:
                raise ValueError("one gather index per block cell")
            self.idx = torch.from_numpy(gi.astype(np.int32)).to(dev)
            self.idx_end = int(gi.max()) + 1 if gi.size else 0
        self.d2 = torch.empty(flat.size, dtype=torch.float32, device=dev)
        self.ws = torch.empty(max(_lib.lib().atsc3_demap_ws_doubles(self.nblk, self.ncell), 1),
                              dtype=torch.float64, device=dev)

    def run(self, cells, csi=None):
        """cells: (nblk, ncell) block cells, or the pool when built with a
        gather table -> ((nblk, n) codeword-order LLRs, s2 per part)."""
        dev = self.device
        nb = self.dm.nb
        llr = torch.empty((self.nblk, self.ncell * nb), dtype=torch.float32, device=dev)
        s2 = torch.empty(self.nparts, dtype=torch.float32, device=dev)
        if self.nblk == 0:
            return llr, s2
        spec = None
        if self.idx is not None:
            spec = GatherSpec(GATHER_TABLE, 0, 0, 0, self.idx.data_ptr(), 0, 0)
            src = cells.contiguous().view(-1)
            if src.numel() < self.idx_end:
                raise ValueError("pool shorter than the gather table reaches")
        else:
            if tuple(cells.shape) != (self.nblk, self.ncell):
                raise ValueError("cells must be (nblk, ncell)")
            src = cells.contiguous()
        if csi is not None:
            csi = csi.contiguous()
        _lib.call("atsc3_demap_weighted_dev", torch.view_as_real(src).data_ptr(),
                  None if csi is None else csi.data_ptr(),
                  None if spec is None else ctypes.byref(spec), self.nblk, self.ncell,
                  self.dm.points.data_ptr(), nb, self.dm.lam_of_q.data_ptr(),
                  self.part.data_ptr(), self.order.data_ptr(), self.seg.data_ptr(), self.nparts,
                  llr.data_ptr(), s2.data_ptr(), self.d2.data_ptr(), self.ws.data_ptr(),
                  _stream(dev))
        return llr, s2
