"""Resident FP16 min-sum: one workgroup per codeword, posteriors in LDS.

``atsc3_minsum_resident_dev`` decodes a queue of codewords of mixed A/322
codes with FP16 check-to-variable messages (each check's messages kept as
one compressed record: FP16 alpha*min1 and alpha*min2, the edge taking min2,
the signs) and the codeword's variable posteriors held in the workgroup's
LDS as FP16.  The sums are FP32 in the order of each variable group's entry
list (FP32 mode: the reference's ascending check order); the posterior is
rounded to FP16 where it is stored.  It stands where the upstream decodes a
frame's blocks with per-code sub-batched min-sum calls (lab/m9_fast.py:
843-925, lab/m44_ldm.py:2497-2576 over lab/ldpc_kernel.c:65-146), as the
FP16 message form of that decode.

``ResidentPlan`` owns the device buffers of one batch composition, so a
decode enqueues two launches and can be captured in a HIP graph.
"""
from __future__ import annotations

import ctypes
import functools

import numpy as np
import torch

from . import _lib
from .ldpc import _stream_ptr

Z = 360
MAX_N = 64800
MAX_DC = 89
MAX_DV = 32


_TABLES = ("grp_c", "edge_c", "grp_v", "edge_v", "c_steps", "c_tasks", "v_steps", "v_tasks",
           "c_off", "v_soff", "v_sdata", "c_soff", "c_sdata")
PROW = 424               # padded FP16 posterior row in LDS: 0..359, then 0..63 again


class _ResCode(ctypes.Structure):
    """struct atsc3_res_code (include/atsc3_fec.h)."""
    _fields_ = ([(f, ctypes.c_int32) for f in
                 ("n", "ncg", "nvg", "nbe", "max_dc", "max_dv", "nsc", "nsv")]
                + [(f, ctypes.c_void_p) for f in _TABLES])


assert ctypes.sizeof(_ResCode) == 136

# tasks packed per wave step by degree: check steps of degree <= 16 take 2
# tasks; variable steps 4 (degree <= 4), 2 (<= 8) or 1 (<= 16)
CHECK_PACK = ((16, 2),)
VAR_PACK = ((3, 8), (4, 4), (16, 2))     # FP16 / INT8: more record loads in flight
VAR_PACK_F32 = ((4, 4), (8, 2), (16, 1))  # FP32: two record words, ordered sums
FAST_MAX = 16


def _pack_tasks(deg, generic, pack):
    """Tasks (group | c6 << 8: rows c6*64 .. c6*64+63 of a 360-row group) in
    the order groups sorted by (not generic, -degree), cut into steps
    D | log2(G) << 7 | start << 10 | generic << 22 of G equal-degree tasks
    (G in {1, 2, 4, 8}); a generic step holds one task.  The waves take the
    steps in this order, so the slowest ones (generic steps: runtime loops,
    one task each; then the highest degrees) come first and the phase ends
    on light steps."""
    deg = np.asarray(deg, np.int64)
    generic = np.asarray(generic, bool)
    order = np.lexsort((np.arange(len(deg)), -deg, ~generic))
    tasks = np.array([g | (c6 << 8) for g in order for c6 in range(6)], np.int64)
    steps, i = [], 0
    while i < len(tasks):
        g = tasks[i] & 255
        d = int(deg[g])
        if generic[g]:
            steps.append(min(d, 127) | (i << 10) | (1 << 22))
            i += 1
            continue
        gmax = next(gg for lim, gg in pack if d <= lim)
        run = 1
        while (run < gmax and i + run < len(tasks) and not generic[tasks[i + run] & 255]
               and deg[tasks[i + run] & 255] == d):
            run += 1
        k = 1 << (run.bit_length() - 1)
        steps.append(d | ((k.bit_length() - 1) << 7) | (i << 10))
        i += k
    return tasks, np.asarray(steps, np.int64)


F32_LDS_ROWS = 113       # FP32 posterior rows (360 floats) a workgroup holds in LDS
TEAM_MAX = 32            # workgroups per codeword in team mode
TEAM_DIRECT = 0x100      # ATSC3_TEAM_DIRECT: team mode without LDS copies
TEAM_GROW = 0x200        # ATSC3_TEAM_GROW: one workgroup per codeword, the last grow into teams
# grow mode's team size: the slow codewords of a launch's last round take 7
# waiting workgroups of their XCD (2566 blocks of 64800 r10/15, FP16: 9.25 ms
# one workgroup per codeword, 9.00 ms teams of 4, 8.97 ms teams of 8)
GROW_TEAM = 8
SIGN_LO = 24             # edge signs a record's word 1 holds (the rest: word 2)


@functools.lru_cache(maxsize=None)
def _host_tables(code, f32=False):
    """int32 tables of the resident kernel, from the code's QC form.

    grp_c[g]      dc | b0 << 7 | has_missing << 18  (check groups)
    edge_c[b]     s | row << 9 | (missing j + 1) << 18 | global << 28  (base
                  edges, check-major; row: where the variable group's
                  posteriors live)
    grp_v[t]      dv | e0 << 6 | n0 << 17 | row << 23 | global << 31
                  (variable groups)
    FP16 (f32 False): every group's row is its index, in LDS.  FP32: the
    F32_LDS_ROWS groups of highest degree (most edge reads) get LDS rows, the
    others rows of the workgroup's global slot.
    edge_v[2e]    u | g << 9 | k << 17 | sp << 24: entry e of the variable
                  list (summation order at p = 0) meets check (g, (p + u) %
                  360) as that check's edge k, whose sign the check's record
                  keeps at bit sp (dc - 1 - k for checks of the fast steps,
                  which build the sign word by shifting, else k)
    edge_v[2e+1]  missing variable position + 1 (0 = full circulant)
    c_steps, c_tasks, v_steps, v_tasks  the work of each phase (_pack_tasks);
                  generic: a missing entry, degree > 16, or (variables) an
                  entry naming a check edge >= SIGN_LO (its sign is in word 2)
    """
    q = code.qc()
    cg = q["cg_ptr"].astype(np.int64)
    be = q["be_word"].astype(np.int64)
    vg = q["vg_ptr"].astype(np.int64)
    ve = q["ve_word"].astype(np.int64)
    nbe = len(be)
    dc = np.diff(cg)
    dv = np.diff(vg)
    miss = (be >> 18) & 1023
    hasmiss = np.array([int((miss[cg[g]:cg[g + 1]] > 0).any()) for g in range(len(dc))],
                       np.int64)
    vb, vu, vreg, vmiss = ve & 2047, (ve >> 11) & 511, (ve >> 20) & 1, (ve >> 21) & 1023
    g_of_b = np.searchsorted(cg, np.arange(nbe), side="right") - 1
    k_of_b = np.arange(nbe) - cg[g_of_b]
    n0 = np.array([int((vreg[vg[t]:vg[t + 1]] == 0).sum()) for t in range(len(dv))], np.int64)
    if (dc.max() > MAX_DC or dv.max() > MAX_DV or code.n > MAX_N or len(dc) > 255
            or nbe >= 2048):
        raise ValueError(f"code {code.rate} {code.n} exceeds the resident kernel's tables")
    vkmax = np.array([int(k_of_b[vb[vg[t]:vg[t + 1]]].max()) for t in range(len(dv))])
    vmissg = np.array([int((vmiss[vg[t]:vg[t + 1]] > 0).any()) for t in range(len(dv))])
    c_generic = (hasmiss > 0) | (dc > FAST_MAX)
    c_tasks, c_steps = _pack_tasks(dc, c_generic, CHECK_PACK)
    v_tasks, v_steps = _pack_tasks(dv, (vmissg > 0) | (dv > FAST_MAX) | (vkmax >= SIGN_LO),
                                   VAR_PACK_F32 if f32 else VAR_PACK)
    # a variable step with any group that has a first region (Type A: n0 > 0
    # entries summed before the rest) carries bit 23
    ta = n0 > 0

    def _ta(w):
        first, cnt = (w >> 10) & 4095, 1 << ((w >> 7) & 7)
        return any(ta[v_tasks[first + i] & 255] for i in range(cnt))
    v_steps = np.array([w | (1 << 23 if _ta(w) else 0) for w in v_steps], np.int64)
    nvg = len(dv)
    row = np.arange(nvg, dtype=np.int64)
    glob = np.zeros(nvg, np.int64)
    if f32:
        order = np.lexsort((np.arange(nvg), -dv))
        nl = min(nvg, F32_LDS_ROWS)
        row[order[:nl]] = np.arange(nl)
        row[order[nl:]] = np.arange(nvg - nl)
        glob[order[nl:]] = 1
    t_of_b = (be >> 9) & 511
    edge_c = (be & 511) | (row[t_of_b] << 9) | (miss << 18) | (glob[t_of_b] << 28)
    grp_v = dv | (vg[:-1] << 6) | (n0 << 17) | (row << 23) | (glob << 31)
    # FP16 / INT8: byte offset of the first of a wave's 64 posteriors on each
    # (base edge, 64-row block) in the padded LDS rows
    c64 = np.arange(6, dtype=np.int64) * 64
    c_off = 2 * (row[t_of_b][:, None] * PROW + (c64[None, :] + (be & 511)[:, None]) % Z)
    gb, kb = g_of_b[vb], k_of_b[vb]
    sp = np.where(c_generic[gb], kb, dc[gb] - 1 - kb)
    # each fast variable step flattened: per task a head word t | c6 << 8 |
    # n0 << 11 | row << 17 | global << 25, then the D entry words
    ev0 = vu | (gb << 9) | (kb << 17) | (sp << 24)
    v_soff, v_sdata = [], []
    for w in v_steps:
        v_soff.append(len(v_sdata))
        if (w >> 22) & 1:
            continue
        d, cnt, first = w & 127, 1 << ((w >> 7) & 7), (w >> 10) & 4095
        for q in range(cnt):
            task = int(v_tasks[first + q])
            t, c6 = task & 255, task >> 8
            v_sdata.append(t | (c6 << 8) | (int(n0[t]) << 11) | (int(row[t]) << 17)
                           | (int(glob[t]) << 25))
            v_sdata.extend(int(x) for x in ev0[vg[t]:vg[t] + d])
    if not v_sdata:
        v_sdata.append(0)
    # check steps flattened the same way: per task g | c6 << 8, then the D
    # words the check step reads (FP32: edge_c; FP16 / INT8: c_off for c6)
    c_soff, c_sdata = [], []
    for w in c_steps:
        c_soff.append(len(c_sdata))
        if (w >> 22) & 1:
            continue
        d, cnt, first = w & 127, 1 << ((w >> 7) & 7), (w >> 10) & 4095
        for q in range(cnt):
            task = int(c_tasks[first + q])
            g, c6 = task & 255, task >> 8
            c_sdata.append(task)
            c_sdata.extend(int(x) for x in (edge_c[cg[g]:cg[g] + d] if f32
                                            else c_off[cg[g]:cg[g] + d, c6]))
    if not c_sdata:
        c_sdata.append(0)
    i32 = lambda a: np.ascontiguousarray(np.asarray(a, np.int64).astype(np.uint32).view(np.int32))
    return dict(n=code.n, ncg=len(dc), nvg=nvg, nbe=nbe, max_dc=int(dc.max()),
                global_rows=int(glob.sum()),
                max_dv=int(dv.max()), nsc=len(c_steps), nsv=len(v_steps),
                c_steps=i32(c_steps), c_tasks=i32(c_tasks), v_steps=i32(v_steps),
                c_off=i32(c_off.ravel()), v_soff=i32(v_soff), v_sdata=i32(v_sdata),
                c_soff=i32(c_soff), c_sdata=i32(c_sdata),
                v_tasks=i32(v_tasks),
                grp_c=i32(dc | (cg[:-1] << 7) | (hasmiss << 18)),
                edge_c=i32(edge_c),
                grp_v=i32(grp_v),
                edge_v=i32(np.stack([ev0, vmiss],
                                    1).ravel()))


# Team launches need every workgroup of their grid co-resident; two running
# at once on different streams could each hold compute units the other
# waits for.  Each team launch on a device is therefore ordered after the
# previous one (an event the next launch's stream waits on when it is
# another stream; nothing is waited on the host).  Not during graph
# capture, where an event recorded outside the capture cannot be waited on.
_TEAM_LAST = {}


def _team_turn_begin(device):
    if torch.cuda.is_current_stream_capturing():
        return
    ev = _TEAM_LAST.get(str(device))
    cur = torch.cuda.current_stream(device)
    if ev is not None and ev[1] != cur.cuda_stream:
        cur.wait_event(ev[0])


def _team_turn_end(device):
    if torch.cuda.is_current_stream_capturing():
        return
    cur = torch.cuda.current_stream(device)
    ev = torch.cuda.Event()
    ev.record(cur)
    _TEAM_LAST[str(device)] = (ev, cur.cuda_stream)


class ResidentBank:
    """Device tables of several codes plus their descriptor array, for one
    storage mode (msg_fmt 1 = FP16, 0 = FP32)."""

    def __init__(self, codes, device, msg_fmt=1):
        self.device = torch.device(device)
        self.codes = list(codes)
        self.fmt = int(msg_fmt)
        self.host = [_host_tables(c, self.fmt == 0) for c in self.codes]
        self.dev = []
        arr = (_ResCode * len(self.codes))()
        for i, h in enumerate(self.host):
            d = {k: torch.from_numpy(h[k]).to(self.device) for k in _TABLES}
            self.dev.append(d)
            r = arr[i]
            for f in ("n", "ncg", "nvg", "nbe", "max_dc", "max_dv", "nsc", "nsv"):
                setattr(r, f, h[f])
            for f in _TABLES:
                setattr(r, f, d[f].data_ptr())
        raw = np.frombuffer(bytes(arr), np.uint8).copy()
        self.desc = torch.from_numpy(raw).to(self.device)
        self.max_checks = max(h["ncg"] * Z for h in self.host)
        self.max_n = max(h["n"] for h in self.host)
        self.max_dc = max(h["max_dc"] for h in self.host)
        self.max_dv = max(h["max_dv"] for h in self.host)
        self.max_global_rows = max(h["global_rows"] for h in self.host)

    def __len__(self):
        return len(self.codes)


def default_grid():
    """One workgroup per compute unit of the current device."""
    g = _lib.lib().atsc3_minsum_resident_grid()
    if g <= 0:
        raise _lib.FecError("atsc3_minsum_resident_grid", g)
    return g


class ResidentPlan:
    """Buffers and launch parameters for decoding one batch composition.

    The workspace holds the device queue word and every workgroup's records:
    runs of one plan must not overlap (enqueue them on one stream).

    block_codes: one LdpcCode per codeword, in queue order; codeword b's
    LLRs occupy ``llr[off[b] : off[b] + n_b]`` of the flat ``llr`` buffer,
    its bits the same range of ``bits``."""

    def __init__(self, block_codes, device="cuda", iters=50, alpha=0.75, grid=None,
                 msg_dtype="float16", msg_scale=None, team=None, order=None, direct=False,
                 ladder=None, grow=False):
        """msg_dtype "float16": the FP16 form; "int8": INT8 record magnitudes
        with a per-codeword step (msg_scale, one per codeword, or None: 6 *
        median|llr| / 127 on the device); "float32": the plain decoder,
        bit-exact with the reference kernel.

        team: None (one workgroup per codeword), or 1..32: team mode, each
        codeword decoded by that many workgroups of one XCD (same results;
        grid at most one workgroup per compute unit).  order: team mode's
        queue order -- None (0, 1, ...), codeword indices, or "auto": ranked
        on the device before every decode, the codewords whose channel
        decisions leave the most checks unsatisfied first (they tend to need
        the most iterations, and a slow codeword that starts late sets the
        launch's time).  direct: team mode without the members' LDS copies
        of the posteriors (every member reads the team's rows in L2; same
        results).  ladder: team mode only, 1..4 alphas: the alpha ladder
        inside the launch (``alpha`` unused), the results of ldpc.ladder_dev.
        grow: with team = S, grow mode: every codeword starts on one
        workgroup (the resident engine's decode); once the queue is empty, a
        codeword still running takes S - 1 waiting workgroups of its XCD and
        finishes as a team, so the launch's slow last codewords no longer
        leave most compute units idle (same results; no ladder)."""
        fmts = {"float32": 0, "float16": 1, "int8": 2}
        if msg_dtype not in fmts:
            raise ValueError("the resident engine stores float32, float16 or int8 records")
        self.msg_dtype = msg_dtype
        self.fmt = fmts[msg_dtype]
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
        self.bank = ResidentBank(uniq, self.device, self.fmt)
        self.nblk = len(idx)
        self.iters, self.alpha = int(iters), float(np.float32(alpha))
        ns = np.array([uniq[i].n for i in idx], np.int64)
        self.off = np.concatenate([[0], np.cumsum(ns)]).astype(np.int64)
        self.total = int(self.off[-1])
        self.code_idx = np.array(idx, np.int32)
        g = default_grid() if grid is None else int(grid)
        if g < 1:
            raise ValueError("grid must be >= 1")
        self.team = None if team is None else int(team)
        if self.team is None:
            self.grid = max(1, min(g, self.nblk))
        else:
            if not 1 <= self.team <= TEAM_MAX:
                raise ValueError(f"team must be 1..{TEAM_MAX}")
            # every workgroup must be co-resident (one per compute unit)
            self.grid = min(g, default_grid())
        self.direct = bool(direct)
        if self.direct and self.team is None:
            raise ValueError("direct applies to team mode only")
        self.grow = bool(grow)
        if self.grow and (self.team is None or ladder is not None):
            raise ValueError("grow needs a team size and no ladder")
        self.ladder = None
        if ladder is not None:
            if self.team is None:
                raise ValueError("the ladder runs in team mode only")
            self.ladder = [float(np.float32(a)) for a in ladder]
            if not 1 <= len(self.ladder) <= 4 or min(self.ladder) < 0:
                raise ValueError("ladder needs 1..4 alphas >= 0")
        dev = self.device
        self.blk_code = torch.from_numpy(self.code_idx).to(dev)
        self.blk_off = torch.from_numpy(self.off[:-1].copy()).to(dev)
        self.llr = torch.zeros(self.total, dtype=torch.float32, device=dev)
        self.bits = torch.empty(self.total, dtype=torch.uint8, device=dev)
        self.conv = torch.empty(self.nblk, dtype=torch.uint8, device=dev)
        self.its = torch.empty(self.nblk, dtype=torch.int32, device=dev)
        self.bad = torch.empty(self.nblk, dtype=torch.int32, device=dev)
        if self.team is None:
            nbytes = _lib.lib().atsc3_minsum_resident_ws_bytes(
                self.grid, self.bank.max_checks, self.bank.max_global_rows, self.fmt)
        else:
            nbytes = _lib.lib().atsc3_minsum_resident_team_ws_bytes(
                self.grid, self.team | (TEAM_GROW if self.grow else 0), self.bank.max_checks,
                self.bank.max_global_rows, self.fmt)
        if nbytes <= 0:
            raise _lib.FecError("atsc3_minsum_resident_ws_bytes", nbytes)
        self.order = None
        self.order_auto = isinstance(order, str) and order == "auto"
        if order is not None and self.team is None:
            raise ValueError("order applies to team mode only")
        if self.order_auto:
            self.order = torch.empty(self.nblk, dtype=torch.int32, device=dev)
            self.order_key = torch.empty(self.nblk, dtype=torch.float32, device=dev)
        elif order is not None:
            o = np.asarray(order, np.int64)
            if o.shape != (self.nblk,) or not np.array_equal(np.sort(o), np.arange(self.nblk)):
                raise ValueError("order must be a permutation of the codeword indices")
            self.order = torch.from_numpy(o.astype(np.int32)).to(dev)
        self.ws = torch.empty(int(nbytes), dtype=torch.uint8, device=dev)
        self.ws_bytes = int(nbytes)
        self.msg_scale = None
        if msg_scale is not None:
            if self.fmt != 2:
                raise ValueError("msg_scale applies to INT8 records only")
            self.msg_scale = torch.as_tensor(msg_scale, dtype=torch.float32,
                                             device=dev).contiguous()
            if self.msg_scale.numel() != self.nblk:
                raise ValueError("msg_scale needs one value per codeword")

    def run(self, llr=None, out=None, post=None):
        """Enqueue the decode of ``llr`` (flat, or None: self.llr as filled)
        into ``out`` = (bits, conv, its, bad) device tensors of the plan's
        sizes (None: the plan's own buffers, overwritten by the next run).
        ``post``: a float32 device tensor of the plan's length that receives
        every bit's final posterior LLR (the soft output)."""
        src = self.llr if llr is None else llr
        if src.numel() != self.total or src.dtype != torch.float32 or not src.is_contiguous():
            raise ValueError("llr must be a contiguous float32 tensor of the plan's length")
        bits, conv, its, bad = (self.bits, self.conv, self.its, self.bad) if out is None else out
        if (bits.numel() != self.total or conv.numel() != self.nblk or its.numel() != self.nblk
                or bad.numel() != self.nblk):
            raise ValueError("output tensors do not match the plan")
        b = self.bank
        args = (src.data_ptr(), self.nblk,
                self.blk_off.data_ptr(), self.blk_code.data_ptr(), b.desc.data_ptr(),
                len(b), b.max_checks, b.max_global_rows, b.max_n, b.max_dc, b.max_dv,
                self.iters, self.alpha, self.fmt,
                None if self.msg_scale is None else self.msg_scale.data_ptr(),
                self.grid, self.ws.data_ptr(),
                self.ws_bytes, bits.data_ptr(),
                conv.data_ptr(), its.data_ptr(), bad.data_ptr())
        if self.team is not None:
            if post is not None:
                raise ValueError("team mode writes no posteriors")
            _team_turn_begin(self.device)
            if self.order_auto:
                _lib.call("atsc3_minsum_resident_team_order_dev", src.data_ptr(), self.nblk,
                          self.blk_off.data_ptr(), self.blk_code.data_ptr(), b.desc.data_ptr(),
                          self.fmt, self.order_key.data_ptr(), self.order.data_ptr(),
                          _stream_ptr(self.device))
            lad = None
            if self.ladder is not None:
                lad = (ctypes.c_float * len(self.ladder))(*self.ladder)
            _lib.call("atsc3_minsum_resident_team_dev", *args[:15],
                      None if self.order is None else self.order.data_ptr(), lad,
                      0 if self.ladder is None else len(self.ladder), self.grid,
                      self.team | (TEAM_DIRECT if self.direct else 0)
                      | (TEAM_GROW if self.grow else 0), *args[16:],
                      _stream_ptr(self.device))
            _team_turn_end(self.device)
        elif post is None:
            _lib.call("atsc3_minsum_resident_dev", *args, _stream_ptr(self.device))
        else:
            if (post.numel() != self.total or post.dtype != torch.float32
                    or not post.is_contiguous()):
                raise ValueError("post must be a contiguous float32 tensor of the plan's length")
            _lib.call("atsc3_minsum_resident_llr_dev", *args, post.data_ptr(),
                      _stream_ptr(self.device))
        return bits, conv, its, bad

    def check(self):
        """Interface of BatchPlan.check (synchronises).  One workgroup per
        codeword never waits for another, so nothing can abort; team mode
        raises if a bounded team wait ran out (codewords left undecoded hold
        its = -1)."""
        if self.team is None:
            return None
        w = _lib.lib().atsc3_minsum_resident_team_status_word()
        if int(self.ws[4 * w:4 * w + 4].view(torch.int32).item()) != 0:
            raise RuntimeError("resident team decode aborted: a bounded team wait ran out")
        return None

    def rounds(self):
        """Interface of BatchPlan.rounds: no grid-wide rounds here."""
        return 0

    def llr_view(self, b):
        return self.llr[int(self.off[b]):int(self.off[b + 1])]

    def block_results(self, b):
        """(bits u8 [n_b], conv, its, bad) of codeword b as device tensors."""
        lo, hi = int(self.off[b]), int(self.off[b + 1])
        return self.bits[lo:hi], self.conv[b], self.its[b], self.bad[b]

    def capture(self, pre=None):
        """A torch.cuda.CUDAGraph replaying ``pre()`` (e.g. the demaps that
        fill ``self.llr``) followed by the decode of ``self.llr`` into the
        plan's buffers (the queue reset is a kernel, not a memset)."""
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
        torch.cuda.synchronize(self.device)
        return g

    def block_bits(self, b):
        return self.bits[int(self.off[b]):int(self.off[b + 1])]


_PLANS = {}            # a few recent plans (LRU): callers alternate codes, e.g. LDM layers
_PLANS_MAX = 8


def decode_resident_dev(llr, code, iters=50, alpha=0.75, grid=None, msg_dtype="float16",
                        msg_scale=None, team=None, direct=False, ladder=None, grow=None):
    """(nblk, n) device LLRs of one code -> device (bits (nblk, n), conv, its, bad).
    msg_dtype "float32" is the plain decoder (bit-exact with the reference);
    "int8" takes msg_scale (one step per codeword) or derives it.  team:
    team mode (each codeword on that many workgroups of one XCD; same
    results); direct: team mode without LDS copies of the posteriors;
    ladder: team mode, 1..4 alphas run as the alpha ladder (alpha unused).
    grow: S (2..32): grow mode (one workgroup per codeword; once the queue
    is empty a running codeword finishes on a team of S; same results).

    Enqueues two launches and no host synchronisation; the plan (tables,
    queue, per-workgroup records) is kept for the next call of the same
    shape on the same stream."""
    if not (isinstance(llr, torch.Tensor) and llr.is_cuda):
        raise TypeError("decode_resident_dev needs a CUDA (HIP) tensor")
    llr = llr.contiguous()
    if llr.dtype != torch.float32:
        llr = llr.float()
    nblk, n = llr.shape
    if n != code.n:
        raise ValueError(f"LLR width {n} != code length {code.n}")
    dev = llr.device
    bits = torch.empty((nblk, n), dtype=torch.uint8, device=dev)
    conv = torch.empty(nblk, dtype=torch.uint8, device=dev)
    its = torch.empty(nblk, dtype=torch.int32, device=dev)
    bad = torch.empty(nblk, dtype=torch.int32, device=dev)
    if nblk == 0:
        return bits, conv, its, bad
    # a plan's workspace (queue word, records) serves one stream at a time:
    # the stream is part of the key
    lad = None if ladder is None else tuple(float(np.float32(a)) for a in ladder)
    key = (id(code), nblk, int(iters), float(np.float32(alpha)), grid, str(dev), msg_dtype,
           _stream_ptr(dev), team, bool(direct), lad, grow)
    if grow is not None:
        if team is not None or ladder is not None:
            raise ValueError("grow takes the team size itself and runs no ladder")
        team = int(grow)
    if msg_scale is not None:
        # caller-given steps: a plan of its own (not cached)
        plan = ResidentPlan([code] * nblk, dev, iters, alpha, grid, msg_dtype, msg_scale, team,
                            direct=direct, ladder=lad, grow=grow is not None)
        return plan.run(llr.reshape(-1), (bits, conv, its, bad))
    plan = _PLANS.pop(key, None)
    if plan is None or plan.bank.codes[0] is not code:
        plan = ResidentPlan([code] * nblk, dev, iters, alpha, grid, msg_dtype, team=team,
                            direct=direct, ladder=lad, grow=grow is not None)
    _PLANS[key] = plan
    while len(_PLANS) > _PLANS_MAX:
        _PLANS.pop(next(iter(_PLANS)))
    return plan.run(llr.reshape(-1), (bits, conv, its, bad))


# ---------------------------------------------------------------------------
# Tables for the host-pointer drop-in (ldpc_minsum_f32): the C side cannot
# build the resident engine's tables, so the build writes them, for every
# A/322 code, to a file beside the library; the drop-in recognises a code by
# csr_hash of the CSR arrays the caller passes.
# ---------------------------------------------------------------------------
DROPIN_TABLES = "libatsc3fec_tables.bin"
_M64 = (1 << 64) - 1


def csr_hash(n, cptr, cidx, vptr, vedge):
    """Order-sensitive 64-bit hash of (n, nc, nedges, cptr, cidx, vptr,
    vedge): the sum over the concatenated words w_i of mix(w_i, i), mix a
    64-bit finaliser (xor-shift-multiply).  The drop-in computes the same
    value (csrc/ldpc_csr.hip)."""
    w = np.concatenate([np.asarray(a, np.int64).astype(np.uint32).astype(np.uint64)
                        for a in (cptr, cidx, vptr, vedge)])
    i = np.arange(len(w), dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = w * np.uint64(0x9E3779B97F4A7C15) + i * np.uint64(0xC2B2AE3D27D4EB4F)
        x ^= x >> np.uint64(31)
        x *= np.uint64(0xBF58476D1CE4E5B9)
        x ^= x >> np.uint64(29)
        h = int(x.sum(dtype=np.uint64))
    nc, ne = len(cptr) - 1, len(cidx)
    return (h + n * 0x100000001B3 + nc * 0x1000193 + ne) & _M64


def write_dropin_tables(path, codes=None):
    """Serialise the FP32-mode resident tables of ``codes`` (default: all 24
    A/322 codes), little-endian: "A3RT", version 1, count; per code the
    csr_hash (u64), n, nc, nedges, ncg, nvg, nbe, max_dc, max_dv, nsc, nsv,
    global_rows (i32), then each table of _TABLES as (length, words)."""
    import struct
    from .codes import RATES, get_code
    if codes is None:
        codes = [get_code(r, n) for n in (16200, 64800) for r in RATES]
    out = [b"A3RT", struct.pack("<II", 1, len(codes))]
    for code in codes:
        cptr, cidx, vptr, vedge = code.csr()
        h = _host_tables(code, True)
        out.append(struct.pack("<Q", csr_hash(code.n, cptr, cidx, vptr, vedge)))
        out.append(struct.pack("<11i", code.n, len(cptr) - 1, len(cidx), h["ncg"], h["nvg"],
                               h["nbe"], h["max_dc"], h["max_dv"], h["nsc"], h["nsv"],
                               h["global_rows"]))
        for k in _TABLES:
            a = np.ascontiguousarray(h[k], "<i4")
            out.append(struct.pack("<i", len(a)))
            out.append(a.tobytes())
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        f.write(b"".join(out))
    import os
    os.replace(tmp, path)
    return path
