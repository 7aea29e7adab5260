"""Exact sum-product rescue of near-miss codewords, on the device.

The upstream offers the codewords the min-sum ladder failed on, but only by
a little (at most ``near_max`` unsatisfied checks), closest first and at
most ``max_blocks`` of them, to a float64 log-tanh belief-propagation decode
with a time budget; converged rescues replace the ladder's results
(lab/m44_ldm.py:2527-2555, lab/m9_fast.py:1126-1151; the decoder is
lab/m16_margin.py:325-390).

Here the selection, the decode and the write-back are kernels on the
caller's stream (csrc/sum_product.hip): ``rescue_dev`` chains after the
batched ladder with no host synchronisation.  ``sum_product_decode_batch``
is the host-array drop-in with the upstream's signature.
"""
from __future__ import annotations

import time

import numpy as np
import torch

from . import _lib
from .ldpc import resolve_code


class SumProduct:
    """Device tables of one code for the float64 sum-product decoder.

    Checks are laid out in QC order (group, position) when the code has the
    360-lifted structure, so neighbouring threads read neighbouring
    variables; slot k of a check is its k-th smallest variable, as in the
    upstream's padded index array (lab/m3_ldpc.py:228-242)."""

    MAX_DMAX = 128

    def __init__(self, code, device="cuda"):
        self.device = torch.device(device)
        self.code = code
        self.n = int(code.n)
        cptr, cidx, _, _ = code.csr()                   # check-major, listed order
        nc = len(cptr) - 1
        deg = np.diff(cptr).astype(np.int64)
        c = np.repeat(np.arange(nc, dtype=np.int64), deg)
        v = cidx.astype(np.int64)
        self.nc, self.dmax = nc, int(deg.max())
        if self.dmax > self.MAX_DMAX:
            raise ValueError(f"check degree {self.dmax} above {self.MAX_DMAX}")
        perm = self._thread_order(code, nc)
        inv = np.empty(nc, np.int64)
        inv[perm] = np.arange(nc)
        slot = np.arange(c.size) - cptr[c].astype(np.int64)               # position in its check
        cvar = np.full((self.dmax, nc), self.n, np.int32)
        cvar[slot, inv[c]] = v
        pos = slot * nc + inv[c]                          # message slot of each edge
        order = np.argsort(v, kind="stable")              # per variable, ascending check
        vptr = np.concatenate([[0], np.cumsum(np.bincount(v, minlength=self.n))])
        dev = self.device
        self.cvar = torch.from_numpy(np.ascontiguousarray(cvar.reshape(-1))).to(dev)
        self.vptr = torch.from_numpy(vptr.astype(np.int32)).to(dev)
        self.vedge = torch.from_numpy(pos[order].astype(np.int32)).to(dev)
        self._ws = {}

    @staticmethod
    def _thread_order(code, nc):
        """Check index of each thread: QC order where the code has it."""
        if hasattr(code, "check_group_pos"):
            g, j = code.check_group_pos(np.arange(nc))
            key = np.asarray(g, np.int64) * 360 + np.asarray(j, np.int64)
            if np.unique(key).size == nc:
                return np.argsort(key, kind="stable")
        return np.arange(nc)

    def workspace(self, rows):
        ws = self._ws.get(rows)
        if ws is None:
            dev = self.device
            ws = (torch.empty(rows * self.n, dtype=torch.float64, device=dev),
                  torch.empty(rows * self.dmax * self.nc, dtype=torch.float64, device=dev),
                  torch.empty(_lib.lib().atsc3_sp_ctl_words(rows), dtype=torch.int32,
                              device=dev))
            self._ws = {rows: ws}
        return ws

    def decode_rows(self, llr, rows, sel=None, nsel=None, row_off=None, iters=100,
                    budget_s=None):
        """Enqueue the decode of ``rows`` call rows; returns device
        (bits (rows, n) u8, conv u8, its i32, bad i32, shed u8).  Calls with
        the same row count share this object's workspace: enqueue them on
        one stream (calls on different streams would overlap on it)."""
        dev = self.device
        bits = torch.empty((rows, self.n), dtype=torch.uint8, device=dev)
        conv = torch.empty(rows, dtype=torch.uint8, device=dev)
        its = torch.empty(rows, dtype=torch.int32, device=dev)
        bad = torch.empty(rows, dtype=torch.int32, device=dev)
        shed = torch.empty(rows, dtype=torch.uint8, device=dev)
        if rows == 0:
            return bits, conv, its, bad, shed
        tot, msg, ctl = self.workspace(rows)
        ptr = lambda t: None if t is None else t.data_ptr()
        _lib.call("atsc3_sum_product_dev", llr.data_ptr(), ptr(row_off), ptr(sel), ptr(nsel),
                  rows, self.n, self.nc, self.dmax, self.cvar.data_ptr(), self.vptr.data_ptr(),
                  self.vedge.data_ptr(), int(iters), float(budget_s or 0.0), tot.data_ptr(),
                  msg.data_ptr(), ctl.data_ptr(), bits.data_ptr(), conv.data_ptr(),
                  its.data_ptr(), bad.data_ptr(), shed.data_ptr(),
                  torch.cuda.current_stream(dev).cuda_stream)
        return bits, conv, its, bad, shed

    def decode(self, llr, iters=100, budget_s=None):
        """(rows, n) float32 device LLRs -> device (bits, conv, its, bad, shed)."""
        if not (llr.is_cuda and llr.dtype == torch.float32 and llr.dim() == 2
                and llr.shape[1] == self.n):
            raise ValueError(f"llr must be a (rows, {self.n}) float32 device tensor")
        llr = llr.contiguous()
        return self.decode_rows(llr, llr.shape[0], iters=iters, budget_s=budget_s)


_TABLES = {}


def tables_for(code, device):
    key = (id(code), str(torch.device(device)))
    hit = _TABLES.get(key)
    if hit is None or hit.code is not code:
        if len(_TABLES) > 32:
            _TABLES.clear()
        hit = _TABLES[key] = SumProduct(code, device)
    return hit


def rescue_dev(llr, bits, conv, its, bad, code, near_max=1500, max_blocks=8, iters=100,
               budget_s=0.12, row_off=None, blk_code=None, code_id=0, skip_if_idle=False):
    """Sum-product rescue of the near misses of a decoded batch, in place on
    the device tensors (bits (B, n) or flat with ``row_off``, conv u8, its
    i32, bad i32).  With ``blk_code`` (a mixed batch) only the rows of code
    ``code_id`` (= ``code``) are offered.  Returns device (rows [max_blocks]
    i32, count [1] i32, sp_conv [max_blocks] u8, sp_shed [max_blocks] u8):
    the rows offered, how many, and which converged / ran out of time.

    With ``skip_if_idle`` the offered count is read on the host (one small
    synchronisation, as the upstream's host-side selection implies) and the
    decode is not launched when nothing is offered; otherwise every step is
    enqueued without synchronisation (an idle decode's launches return at
    once)."""
    dev = llr.device
    nblk = conv.numel()
    rows = torch.zeros(max_blocks, dtype=torch.int32, device=dev)
    count = torch.zeros(1, dtype=torch.int32, device=dev)
    if nblk == 0:
        z = torch.zeros(max_blocks, dtype=torch.uint8, device=dev)
        return rows, count, z, z.clone()
    sp = tables_for(code, dev)
    st = torch.cuda.current_stream(dev).cuda_stream
    _lib.call("atsc3_sp_select_dev", conv.data_ptr(), bad.data_ptr(),
              None if blk_code is None else blk_code.data_ptr(), int(code_id), nblk,
              int(near_max), int(max_blocks), rows.data_ptr(), count.data_ptr(), st)
    if skip_if_idle and int(count.item()) == 0:
        z = torch.zeros(max_blocks, dtype=torch.uint8, device=dev)
        return rows, count, z, z.clone()
    ro = None if row_off is None else row_off
    b2, c2, i2, _, s2 = sp.decode_rows(llr, max_blocks, rows, count, ro, iters, budget_s)
    _lib.call("atsc3_sp_merge_dev", rows.data_ptr(), count.data_ptr(), max_blocks, sp.n,
              None if ro is None else ro.data_ptr(), b2.data_ptr(), c2.data_ptr(),
              i2.data_ptr(), bits.data_ptr(), conv.data_ptr(), its.data_ptr(), bad.data_ptr(),
              st)
    return rows, count, c2, s2


def sum_product_decode_batch(lam2d, checks, n, iters=100, deadline=None, stats=None,
                             device="cuda"):
    """Drop-in for lab/m16_margin.py:325 sum_product_decode_batch: host
    arrays in and out, (bits u8, conv bool, its i32, bad i32).  ``deadline``
    is a time.perf_counter() value; the remaining time becomes the device
    budget.  Shed rows are counted in stats["sp_timeout"]."""
    code = resolve_code(checks, n)
    lam = np.ascontiguousarray(np.asarray(lam2d), np.float32)
    budget = None
    if deadline is not None:
        budget = max(deadline - time.perf_counter(), 1e-9)
    sp = tables_for(code, device)
    t = torch.from_numpy(lam).to(sp.device)
    b, c, i, d, s = sp.decode(t, iters, budget)
    torch.cuda.synchronize(sp.device)
    shed = int(s.sum().item())
    if stats is not None and shed:
        stats["sp_timeout"] = stats.get("sp_timeout", 0) + shed
    return (b.cpu().numpy(), c.cpu().numpy().astype(bool), i.cpu().numpy(),
            d.cpu().numpy())
