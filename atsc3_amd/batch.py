"""Batched min-sum over a queue of codewords of mixed A/322 codes.

One persistent kernel (atsc3_minsum_batch_dev) decodes every codeword of a
frame's uneven mix -- 64800- and 16200-bit blocks at any rates -- with a
bounded number of decode slots; converged codewords retire and free their
slot for the next one in the device queue.  Per codeword the results are
those of ldpc_minsum_f32 (lab/ldpc_kernel.c:65-146) in FP32 mode.

It stands where the upstream splits a frame into per-code sub-batches and
decodes each with its own call (lab/m9_fast.py:843-925 ldpc_batch over
ThreadPoolExecutor chunks; lab/m44_ldm.py:2497-2576).

``BatchPlan`` owns every device buffer of one batch shape (codes per block,
slots, grid), so a decode enqueues kernels only and can be captured in a
HIP graph (``capture``) and replayed per frame.
"""
from __future__ import annotations

import collections
import ctypes

import numpy as np
import torch

from . import _lib
from .codes import LdpcCode
from .ldpc import MSG_BYTES, MSG_FORMATS, device_tables

ROW = LdpcCode.ROW
MAX_RUNGS = 4            # alpha ladder rungs run inside the kernel


class _CodeTab(ctypes.Structure):
    """struct atsc3_code_tab (include/atsc3_fec.h)."""
    _fields_ = ([(f, ctypes.c_int32) for f in
                 ("n", "ncg", "nvg", "nbe", "nsc", "nsv", "type_a", "max_vdeg")]
                + [(f, ctypes.c_void_p) for f in
                   ("cg_info", "be_off", "be_miss", "vg_info", "ve_off", "ve_miss",
                    "c_steps", "v_steps")])


assert ctypes.sizeof(_CodeTab) == 96


class CodeBank:
    """Device tables of several codes plus their descriptor array."""

    def __init__(self, codes, device):
        self.device = torch.device(device)
        self.codes = list(codes)
        self.tabs = [device_tables(c, self.device) for c in self.codes]
        arr = (_CodeTab * len(self.codes))()
        for i, t in enumerate(self.tabs):
            q = t.qc
            if q is None:
                raise ValueError(f"code {i} has no QC form with variable degree <= 32")
            if q["nvg"] > 255 or q["nbe"] >= 2048:
                raise ValueError(f"code {i} exceeds the kernel's table limits")
            d = arr[i]
            d.n, d.ncg, d.nvg, d.nbe = t.n, q["ncg"], q["nvg"], q["nbe"]
            d.nsc, d.nsv = len(q["c_steps"]), len(q["v_steps"])
            d.type_a, d.max_vdeg = int(q["type_a"]), int(q["max_vdeg"])
            for f in ("cg_info", "be_off", "be_miss", "vg_info", "ve_off", "ve_miss",
                      "c_steps", "v_steps"):
                setattr(d, f, q[f].data_ptr())
        raw = np.frombuffer(bytes(arr), np.uint8).copy()
        self.desc = torch.from_numpy(raw).to(self.device)
        self.max_nvg = max(t.qc["nvg"] for t in self.tabs)
        self.max_nbe = max(t.qc["nbe"] for t in self.tabs)
        self.max_vdeg = max(int(t.qc["max_vdeg"]) for t in self.tabs)

    def __len__(self):
        return len(self.codes)


def max_grid(msg_dtype="float32", xcd=False):
    """Workgroups of the persistent kernel (grid-wide or XCD-partitioned
    mode) that fit on the current device."""
    fn = "atsc3_minsum_batch_xcd_max_grid" if xcd else "atsc3_minsum_batch_max_grid"
    g = getattr(_lib.lib(), fn)(MSG_FORMATS[msg_dtype])
    if g <= 0:
        raise _lib.FecError(fn, g)
    return g


# parts of the queue in the XCD-partitioned mode: one per XCD of an MI355X
XCD_PARTS = 8


def xcd_parts(n_of, nparts=XCD_PARTS):
    """(order, part_off): codewords dealt to ``nparts`` parts longest first,
    each to the part with the fewest code bits so far (ties: lowest part), a
    part keeping the input order of its codewords."""
    n_of = np.asarray(n_of, np.int64)
    load = np.zeros(nparts, np.int64)
    part = np.empty(len(n_of), np.int64)
    for b in sorted(range(len(n_of)), key=lambda b: (-n_of[b], b)):
        p = int(np.argmin(load))
        part[b] = p
        load[p] += n_of[b]
    order = np.concatenate([np.flatnonzero(part == p) for p in range(nparts)]).astype(np.int32)
    off = np.concatenate([[0], np.cumsum(np.bincount(part, minlength=nparts))]).astype(np.int32)
    return order, off


class BatchPlan:
    """Buffers and launch parameters for decoding one batch composition.

    block_codes: one LdpcCode per codeword, in queue order.  The LLRs of
    codeword b occupy ``llr[off[b] : off[b] + n_b]`` of the flat ``llr``
    buffer; its bits the same range of ``bits``.
    """

    def __init__(self, block_codes, device="cuda", iters=50, alpha=0.75,
                 msg_dtype="float32", slots=None, grid=None, msg_scale=None, ladder=None,
                 xcd=False):
        """xcd: the XCD-partitioned mode (atsc3_minsum_batch_xcd_dev): each XCD
        decodes its own part of the queue with barriers among its own
        workgroups; ``slots`` then counts all XCDs' slots (default: every
        codeword of the largest part on each XCD)."""
        self.device = torch.device(device)
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        uniq, idx = [], []
        for c in block_codes:
            for i, u in enumerate(uniq):
                if u is c:
                    idx.append(i)
                    break
            else:
                uniq.append(c)
                idx.append(len(uniq) - 1)
        self.bank = CodeBank(uniq, self.device)
        self.nblk = len(idx)
        self.fmt = MSG_FORMATS[msg_dtype]
        self.msg_dtype = msg_dtype
        self.iters, self.alpha = int(iters), float(np.float32(alpha))
        self.ladder = _ladder_arg(ladder)
        ns = np.array([uniq[i].n for i in idx], np.int64)
        self.n_of = ns
        self.off = np.concatenate([[0], np.cumsum(ns)]).astype(np.int64)
        self.total = int(self.off[-1])
        self.code_idx = np.array(idx, np.int32)
        self.xcd = bool(xcd)
        cap = max_grid(msg_dtype, self.xcd)
        self.grid = cap if grid is None else int(grid)
        if not 1 <= self.grid <= cap:
            raise ValueError(f"grid {self.grid} outside 1..{cap} (co-resident capacity)")
        dev = self.device
        if self.xcd:
            order, part_off = xcd_parts(ns)
            widest = int(np.diff(part_off).max(initial=1))
            want = XCD_PARTS * widest if slots is None else int(slots)
            self.slots = max(16, min(want, 2048))
            self.order = torch.from_numpy(order).to(dev)
            self.part_off = torch.from_numpy(part_off).to(dev)
        else:
            self.slots = max(1, min(self.nblk, 1024 if slots is None else int(slots)))
        self.blk_code = torch.from_numpy(self.code_idx).to(dev)
        self.blk_off = torch.from_numpy(self.off[:-1].copy()).to(dev)
        self.llr = torch.zeros(self.total, dtype=torch.float32, device=dev)
        self.bits = torch.empty(self.total, dtype=torch.uint8, device=dev)
        self.conv = torch.empty(self.nblk, dtype=torch.uint8, device=dev)
        self.its = torch.empty(self.nblk, dtype=torch.int32, device=dev)
        self.bad = torch.empty(self.nblk, dtype=torch.int32, device=dev)
        self.tot = torch.empty(self.slots * self.bank.max_nvg * ROW, dtype=torch.float32,
                               device=dev)
        self.msg = torch.empty(self.slots * self.bank.max_nbe * ROW * MSG_BYTES[self.fmt],
                               dtype=torch.uint8, device=dev)
        words = (_lib.lib().atsc3_minsum_batch_xcd_ctl_words if self.xcd
                 else _lib.lib().atsc3_minsum_batch_ctl_words)
        nctl = words(max(self.nblk, 0), self.slots)
        self.ctl = torch.zeros(nctl, dtype=torch.int32, device=dev)
        self.status_word = _lib.lib().atsc3_minsum_batch_status_word()
        self.msg_scale = None
        if msg_scale is not None:
            self.msg_scale = torch.as_tensor(msg_scale, dtype=torch.float32,
                                             device=dev).contiguous()
            if self.msg_scale.numel() != self.nblk:
                raise ValueError("msg_scale needs one value per codeword")

    def bank_index(self, code):
        """Index of ``code`` in this plan's code bank (its blk_code value)."""
        for i, c in enumerate(self.bank.codes):
            if c is code:
                return i
        raise KeyError("code not in this plan")

    def llr_view(self, b):
        return self.llr[int(self.off[b]):int(self.off[b + 1])]

    def run(self, stream=None):
        """Enqueue the decode of ``self.llr`` on the current stream."""
        if self.nblk == 0:
            return
        st = (stream or torch.cuda.current_stream(self.device)).cuda_stream
        head = (self.llr.data_ptr(), self.nblk, self.blk_off.data_ptr(),
                self.blk_code.data_ptr(), self.bank.desc.data_ptr(), len(self.bank),
                self.bank.max_nvg, self.bank.max_nbe, self.bank.max_vdeg, self.iters,
                self.alpha, *self.ladder, self.fmt,
                None if self.msg_scale is None else self.msg_scale.data_ptr())
        tail = (self.slots, self.grid, self.tot.data_ptr(), self.msg.data_ptr(),
                self.ctl.data_ptr(), self.bits.data_ptr(), self.conv.data_ptr(),
                self.its.data_ptr(), self.bad.data_ptr(), st)
        if self.xcd:
            _lib.call("atsc3_minsum_batch_xcd_dev", *head, self.order.data_ptr(),
                      self.part_off.data_ptr(), int(self.part_off.numel()) - 1, *tail)
        else:
            _lib.call("atsc3_minsum_batch_dev", *head, *tail)

    def check(self):
        """Synchronize and raise if the kernel aborted (a barrier wait timed
        out, e.g. when the workgroups could not all be resident)."""
        if self.nblk and int(self.ctl[self.status_word].item()) != 0:
            raise RuntimeError("batched min-sum aborted: grid barrier wait timed out")

    def rounds(self):
        """Rounds of the last launch (XCD-partitioned: of the busiest XCD)."""
        return int(self.ctl[9].item())

    def capture(self, pre=None):
        """A torch.cuda.CUDAGraph replaying ``pre()`` (e.g. the demaps that fill
        ``self.llr``) followed by the decode."""
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream(self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):
            if pre is not None:
                pre()
            self.run()
        torch.cuda.current_stream(self.device).wait_stream(s)
        torch.cuda.synchronize(self.device)
        with torch.cuda.graph(g, stream=s):
            if pre is not None:
                pre()
            self.run()
        return g

    def block_results(self, b):
        """(bits u8 [n_b], conv, its, bad) of codeword b as device tensors."""
        lo, hi = int(self.off[b]), int(self.off[b + 1])
        return self.bits[lo:hi], self.conv[b], self.its[b], self.bad[b]


def decode_mixed(llrs, codes, iters=50, alpha=0.75, msg_dtype="float32", slots=None,
                 grid=None):
    """Decode a list of per-codeword LLR vectors (device tensors) with their
    codes in one batched launch.  Returns a list of device bit vectors and
    device tensors (conv u8, its i32, bad i32) over the list."""
    if len(llrs) != len(codes):
        raise ValueError("one code per LLR vector")
    if not llrs:
        raise ValueError("empty batch")
    dev = llrs[0].device
    plan = BatchPlan(codes, dev, iters, alpha, msg_dtype, slots, grid)
    for b, l in enumerate(llrs):
        if l.numel() != plan.n_of[b]:
            raise ValueError(f"codeword {b}: {l.numel()} LLRs for a length-{plan.n_of[b]} code")
        plan.llr_view(b).copy_(l.reshape(-1))
    plan.run()
    plan.check()
    bits = [plan.bits[int(plan.off[b]):int(plan.off[b + 1])] for b in range(plan.nblk)]
    return bits, plan.conv, plan.its, plan.bad


_SINGLE = collections.OrderedDict()


def _ladder_arg(ladder):
    """(host float array or None, nrungs) for the C call."""
    if ladder is None:
        return (None, 0)
    a = [float(np.float32(x)) for x in ladder]
    if not 1 <= len(a) <= MAX_RUNGS:
        raise ValueError(f"ladder needs 1..{MAX_RUNGS} alphas")
    return ((ctypes.c_float * len(a))(*a), len(a))


def decode_batch_dev(llr, code, iters=50, alpha=0.75, msg_dtype="float32", msg_scale=None,
                     slots=None, grid=None, ladder=None):
    """The batched engine on a (nblk, n) batch of ONE code (device tensors in
    and out, no host sync): same results as ldpc.decode_dev, or with
    ``ladder`` (1..4 alphas) those of ldpc.ladder_dev."""
    dev = llr.device
    nblk, n = llr.shape
    # the stream is part of the key: the cached tables were allocated on the
    # stream that built them, so another stream gets its own copy and no
    # entry is ever freed while a launch on a different stream reads it
    key = (id(code), nblk, str(dev), torch.cuda.current_stream(dev).cuda_stream)
    hit = _SINGLE.get(key)
    if hit is None or hit[0] is not code:
        bank = CodeBank([code], dev)
        blk_code = torch.zeros(nblk, dtype=torch.int32, device=dev)
        blk_off = torch.arange(nblk, dtype=torch.int64, device=dev) * n
        hit = (code, bank, blk_code, blk_off)
        while len(_SINGLE) >= 64:
            _SINGLE.popitem(last=False)          # evict the oldest entry only
        _SINGLE[key] = hit
    else:
        _SINGLE.move_to_end(key)
    _, bank, blk_code, blk_off = hit
    fmt = MSG_FORMATS[msg_dtype]
    bits = torch.empty((nblk, n), dtype=torch.uint8, device=dev)
    conv = torch.empty(nblk, dtype=torch.uint8, device=dev)
    its = torch.empty(nblk, dtype=torch.int32, device=dev)
    bad = torch.empty(nblk, dtype=torch.int32, device=dev)
    if nblk == 0:
        return bits, conv, its, bad
    cap = max_grid(msg_dtype)
    grid = cap if grid is None else min(int(grid), cap)
    slots = nblk if slots is None else max(1, min(int(slots), nblk))
    tot = torch.empty(slots * bank.max_nvg * ROW, dtype=torch.float32, device=dev)
    msg = torch.empty(slots * bank.max_nbe * ROW * MSG_BYTES[fmt], dtype=torch.uint8,
                      device=dev)
    ctl = torch.empty(_lib.lib().atsc3_minsum_batch_ctl_words(nblk, slots), dtype=torch.int32,
                      device=dev)
    sp = None
    if msg_scale is not None:
        msg_scale = torch.as_tensor(msg_scale, dtype=torch.float32, device=dev).contiguous()
        sp = msg_scale.data_ptr()
    llr = llr.contiguous()
    _lib.call("atsc3_minsum_batch_dev", llr.data_ptr(), nblk, blk_off.data_ptr(),
              blk_code.data_ptr(), bank.desc.data_ptr(), 1, bank.max_nvg, bank.max_nbe,
              bank.max_vdeg, int(iters), float(np.float32(alpha)), *_ladder_arg(ladder), fmt, sp,
              slots, grid, tot.data_ptr(),
              msg.data_ptr(), ctl.data_ptr(), bits.data_ptr(), conv.data_ptr(), its.data_ptr(),
              bad.data_ptr(), torch.cuda.current_stream(dev).cuda_stream)
    return bits, conv, its, bad
