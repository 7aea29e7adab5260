"""GPU min-sum decoding behind the upstream call signatures.

Two decode contracts, as the upstream has them:

* the compiled kernel's (lab/ldpc_kernel.c:82-143): syndrome checked before
  each update (a clean block reports its = 0), tot = llr + (0 + m_1 + ...).
  Served by ``min_sum_decode_batch(..., dtype=float32)`` -- lab/m3_ldpc.py:509
  routes float32 to that kernel -- bit-exact on every field (bits, conv, its,
  bad), and by ``decode_dev`` / ``ladder_dev`` with the FP16 / INT8 message
  forms beside it.
* the update-then-check one of the torch decoder ``GpuMinSum.decode``
  (lab/m9_gpu.py:138-211) and of the float64 decoders ``min_sum_decode``
  (lab/m3_ldpc.py:245-286) and ``min_sum_decode_batch(dtype=float64)``:
  messages updated first (its >= 1), tot = ((llr + m_1) + m_2) + ..., rows
  padded with 1e9 entries.  Served by ``decode_utc_dev`` in FP32 or FP64,
  bit-exact with those functions (``GpuMinSum(dtype="float32")``,
  ``GpuMinSum(dtype="float64")`` -- the upstream default -- and the float64
  functions), also on NaN / inf LLRs, where GpuMinSum and the numpy
  functions follow different rules and each mirror follows its own.

Mirrors (same arguments, same return tuples):
  * ``min_sum_decode(llr, checks, n, iters=50, alpha=0.75)``
  * ``min_sum_decode_batch(lam2d, checks, n, iters=50, alpha=0.75, dtype)``
  * ``GpuMinSum(checks, n, iters, alpha, device, dtype).decode(llr2d, iters,
    alpha)``: a numpy array or a device tensor in, numpy
    ``(bits u8, conv bool, its i32, bad i32)`` out.
  * ``ladder_decode`` -- the alpha ladder of lab/m9_fast.py:843-891 /
    lab/m6_bicm.py:341-355 (rungs 1.0, 0.85, 0.75 on still-failing rows; a
    rung's result is taken when it converged or has fewer unsatisfied checks).

``checks`` may be an ``LdpcCode`` or the upstream list-of-lists; a list is
matched to its A/322 code (N, rate from N - len(checks)) and verified equal,
so the QC kernels can serve it; an unrecognised matrix runs on the CSR kernels.

Message storage (``msg_dtype``, kernel contract): "float32" (plain,
bit-exact), "float16", "int8" (per-block scale).  Variable sums accumulate
in FP32 in every mode.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .codes import LdpcCode, get_code, RATES

MSG_FORMATS = {"float32": 0, "fp32": 0, "float16": 1, "fp16": 1, "int8": 2}
# batches at least this large run on the per-iteration phase kernels; smaller
# ones on the fused one-workgroup-per-codeword kernel (lower latency)
PHASE_MIN_BLOCKS = 256
MSG_BYTES = {0: 4, 1: 2, 2: 1}
_DEV_TABLES = {}
_RESOLVED = {}


def _stream_ptr(device):
    return torch.cuda.current_stream(device).cuda_stream


def resolve_code(checks, n=None):
    """LdpcCode for ``checks`` (an LdpcCode or a list of column lists)."""
    if isinstance(checks, LdpcCode):
        return checks
    key = id(checks)
    hit = _RESOLVED.get(key)
    if hit is not None and hit[0] is checks:
        return hit[1]
    nc = len(checks)
    code = None
    if n in (16200, 64800):
        k = n - nc
        if (k * 15) % n == 0:
            rate = f"{k * 15 // n}/15"
            if rate in RATES:
                try:
                    cand = get_code(rate, n)
                except ValueError:
                    cand = None
                if cand is not None and cand.checks == [list(c) for c in checks]:
                    code = cand
    if code is None:
        code = _GenericCode(checks, n)
    _RESOLVED[key] = (checks, code)
    return code


class _GenericCode:
    """A parity-check matrix that is not an A/322 QC code: CSR only."""

    def __init__(self, checks, n):
        self.n = int(n)
        self.checks = checks
        self.nc = len(checks)
        cn = np.array([len(c) for c in checks], np.int32)
        cptr = np.concatenate([[0], np.cumsum(cn)]).astype(np.int32)
        cidx = np.fromiter((v for c in checks for v in c), np.int32, int(cn.sum()))
        vedge = np.argsort(cidx, kind="stable").astype(np.int32)
        vptr = np.concatenate([[0], np.cumsum(np.bincount(cidx, minlength=self.n))]).astype(np.int32)
        self._csr = (cptr, cidx, vptr, vedge)

    def csr(self):
        return self._csr

    def qc(self):
        raise ValueError("not a 360-lifted QC code")


class DeviceTables:
    """Code tables resident on one device (built once per code, like
    GpuMinSum.__init__, lab/m9_gpu.py:95-136)."""

    def __init__(self, code, device):
        dev = torch.device(device)
        self.code = code
        self.n = code.n
        cptr, cidx, vptr, vedge = code.csr()
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(dev)
        self.cptr, self.cidx, self.vptr, self.vedge = t(cptr), t(cidx), t(vptr), t(vedge)
        self.nc = len(cptr) - 1
        self.nedges = len(cidx)
        self.max_check_degree = int(np.diff(cptr).max(initial=0))
        self.dmax = max(1, self.max_check_degree)     # row padding of the upstream
        try:
            q = code.kernel_tables()
        except (ValueError, AttributeError):
            q = None
        self.qc = None
        if q is not None and q["max_vdeg"] <= 32:
            self.qc = dict(q)
            for k in ("cg_info", "be_off", "be_miss", "vg_info", "ve_off", "ve_miss",
                      "c_steps", "v_steps"):
                self.qc[k] = t(q[k])


def device_tables(code, device):
    dev = torch.device(device)
    if dev.index is None:
        dev = torch.device(dev.type, torch.cuda.current_device())
    key = (id(code), str(dev))
    hit = _DEV_TABLES.get(key)
    if hit is None or hit.code is not code:
        hit = DeviceTables(code, dev)
        _DEV_TABLES[key] = hit
    return hit


def decode_dev(llr, code, iters=50, alpha=0.75, msg_dtype="float32", engine="auto",
               msg_scale=None):
    """Decode device-resident LLRs (nblk, n) float32; no host sync.

    Returns device tensors (bits u8 (nblk,n), conv u8, its i32, bad i32).
    engine: "qc" (fused 360-lifted kernel, one workgroup per codeword),
    "phase" (the same decode as per-iteration phase kernels, for large
    batches), "batch" (the persistent batched engine, atsc3_amd.batch),
    "csr" (generic kernel; FP32 only), "resident" (any format: one
    workgroup per codeword with its posteriors in LDS, atsc3_amd.resident;
    FP32 is bit-exact with the plain decoder; it runs in grow mode, where
    the codewords still running once the queue is empty finish on teams of
    GROW_TEAM workgroups -- the same results), "team" (the resident engine
    with each codeword on 8 or 16 compute units of one XCD: the same
    results, far lower latency for a few tens of codewords)
    or "auto" (the QC engines when the code has a QC form, else csr: FP32
    below PHASE_MIN_BLOCKS blocks the team mode, above it the resident
    engine; FP16/INT8 the batched engine, which keeps the FP16/INT8 message
    form of the QC engines, whose posteriors are FP32, and equals the fused
    engine bit for bit at 2-3x its speed for a few tens of blocks).

    Failure on the device: the batched engine leaves its = -1 (bad = -1,
    conv 0) for every codeword it did not finish (a bounded grid-barrier
    wait ran out); this function does not read the outputs back, so a
    caller that stays on the device checks ``(its < 0).any()`` itself (the
    host conversions -- min_sum_decode_batch, GpuMinSum.decode,
    ladder_decode -- raise on it).
    """
    if not (isinstance(llr, torch.Tensor) and llr.is_cuda):
        raise TypeError("decode_dev needs a CUDA (HIP) tensor")
    llr = llr.contiguous()
    if llr.dtype != torch.float32:
        llr = llr.float()
    dev = llr.device
    tabs = device_tables(code, dev)
    nblk, n = llr.shape
    if n != tabs.n:
        raise ValueError(f"LLR width {n} != code length {tabs.n}")
    fmt = MSG_FORMATS[msg_dtype]
    bits = torch.empty((nblk, n), dtype=torch.uint8, device=dev)
    conv = torch.empty(nblk, dtype=torch.uint8, device=dev)
    its = torch.empty(nblk, dtype=torch.int32, device=dev)
    bad = torch.empty(nblk, dtype=torch.int32, device=dev)
    if nblk == 0:
        return bits, conv, its, bad
    if engine == "auto":
        if tabs.qc is None:
            engine = "csr"
        elif fmt != 0:
            # the batched engine: the QC engines' FP16/INT8 form (32 blocks
            # of 64800 r10/15: 0.74 ms against 2.35 ms fused)
            engine = "batch"
        elif nblk < PHASE_MIN_BLOCKS:
            # the team mode, bit-exact like every FP32 engine (32 blocks:
            # 0.35 ms against 2.4 ms fused)
            engine = "team"
        else:
            # the resident engine's FP32 mode (bit-exact, ~2.4x the phased
            # engine on 2566 blocks)
            engine = "resident"
    if engine in ("resident", "team"):
        if msg_scale is not None and fmt != 2:
            raise ValueError("msg_scale applies to INT8 messages only")
        from .resident import GROW_TEAM, decode_resident_dev
        team = None if engine == "resident" else (16 if nblk <= 16 else 8)
        # the resident engine runs grow mode (the launch's last codewords
        # finish on teams; same results; 2566 blocks of 64800 r10/15 at
        # 21 dB: FP32 11.82 -> 11.41 ms, FP16 8.16 -> 7.87 ms)
        grow = GROW_TEAM if engine == "resident" else None
        return decode_resident_dev(llr, code, iters, alpha,
                                   msg_dtype={0: "float32", 1: "float16", 2: "int8"}[fmt],
                                   msg_scale=msg_scale, team=team, direct=team is not None,
                                   grow=grow)
    if engine == "batch":
        if tabs.qc is None:
            raise ValueError("this code has no QC form with variable degree <= 32")
        from .batch import decode_batch_dev
        return decode_batch_dev(llr, code, iters, alpha, msg_dtype, msg_scale)
    use_qc = engine in ("qc", "phase")
    if use_qc and tabs.qc is None:
        raise ValueError("this code has no QC form with variable degree <= 32; "
                         "use engine='csr'")
    stream = _stream_ptr(dev)
    if use_qc:
        q = tabs.qc
        row = LdpcCode.ROW
        tot = torch.empty(nblk * q["nvg"] * row, dtype=torch.float32, device=dev)
        msg = torch.empty(nblk * q["nbe"] * row * MSG_BYTES[fmt], dtype=torch.uint8, device=dev)
        scale_ptr = None
        if msg_scale is not None:
            msg_scale = torch.as_tensor(msg_scale, dtype=torch.float32, device=dev).contiguous()
            if msg_scale.numel() != nblk:
                raise ValueError("msg_scale needs one value per block")
            scale_ptr = msg_scale.data_ptr()
        tables = (q["ncg"], q["nvg"], q["nbe"], q["type_a"], q["cg_info"].data_ptr(),
                  q["be_off"].data_ptr(), q["be_miss"].data_ptr(), q["vg_info"].data_ptr(),
                  q["ve_off"].data_ptr(), q["ve_miss"].data_ptr(), q["c_steps"].data_ptr(),
                  len(q["c_steps"]), q["v_steps"].data_ptr(), len(q["v_steps"]),
                  q["max_cdeg"], q["max_vdeg"], int(iters), float(np.float32(alpha)), fmt,
                  scale_ptr, tot.data_ptr(), msg.data_ptr())
        outs = (bits.data_ptr(), conv.data_ptr(), its.data_ptr(), bad.data_ptr(), stream)
        if engine == "phase":
            ctl = torch.empty(_lib.lib().atsc3_minsum_qc_ctl_words(nblk), dtype=torch.int32,
                              device=dev)
            _lib.call("atsc3_minsum_qc_phased_dev", llr.data_ptr(), nblk, n, *tables,
                      ctl.data_ptr(), *outs)
        else:
            _lib.call("atsc3_minsum_qc_dev", llr.data_ptr(), nblk, n, *tables, *outs)
    else:
        if fmt != 0:
            raise ValueError("the CSR engine stores FP32 messages only")
        tot = torch.empty((nblk, n), dtype=torch.float32, device=dev)
        msg = torch.empty((nblk, tabs.nedges), dtype=torch.float32, device=dev)
        _lib.call("atsc3_minsum_csr_f32_dev", llr.data_ptr(), nblk, n, tabs.nc,
                  tabs.nedges, tabs.cptr.data_ptr(), tabs.cidx.data_ptr(),
                  tabs.vptr.data_ptr(), tabs.vedge.data_ptr(), int(iters),
                  float(np.float32(alpha)), tot.data_ptr(), msg.data_ptr(),
                  bits.data_ptr(), conv.data_ptr(), its.data_ptr(), bad.data_ptr(),
                  stream)
    return bits, conv, its, bad


def _to_numpy(bits, conv, its, bad):
    its_h = its.cpu().numpy()
    if its_h.size and its_h.min() < 0:
        # the batched engine marks codewords it did not finish (a grid
        # barrier wait timed out) with its = -1
        raise RuntimeError(f"decode aborted: {int((its_h < 0).sum())} codewords not decoded")
    return (bits.cpu().numpy(), conv.cpu().numpy().astype(bool), its_h, bad.cpu().numpy())


UTC_DTYPES = {"float32": 0, "float64": 1}
# how a check holding NaN magnitudes updates: GpuMinSum's torch.amin rule or
# min_sum_decode's np.partition / np.argmin rule (identical without NaN)
NAN_RULES = {"torch": 0, "numpy": 1}


def decode_utc_dev(llr, code, iters=50, alpha=0.75, nan_rule="torch"):
    """Update-then-check decode of device LLRs (nblk, n), float32 or float64
    (the arithmetic follows the tensor's dtype); no host sync.

    Contract of lab/m9_gpu.py:138-211 (GpuMinSum, float32 / float64) and
    lab/m3_ldpc.py:245-286 (min_sum_decode, float64): its >= 1 for
    iters >= 1, tot = ((llr + m_1) + m_2) + ... in ascending check order,
    rows padded to the largest check degree with 1e9 entries; iters == 0
    gives bits = llr < 0, conv 0, its 0, bad -1.  ``nan_rule`` "torch"
    follows GpuMinSum where a check's magnitudes include NaN (torch.amin
    propagates it, lab/m9_gpu.py:160-166), "numpy" follows min_sum_decode
    (np.partition sorts NaN last and np.argmin picks the first NaN,
    lab/m3_ldpc.py:264-270).
    Returns device tensors (bits u8 (nblk, n), conv u8, its i32, bad i32).
    """
    if nan_rule not in NAN_RULES:
        raise ValueError(f"nan_rule must be one of {tuple(NAN_RULES)}")
    if not (isinstance(llr, torch.Tensor) and llr.is_cuda):
        raise TypeError("decode_utc_dev needs a CUDA (HIP) tensor")
    if llr.dtype not in (torch.float32, torch.float64):
        raise TypeError("update-then-check decoding runs in float32 or float64")
    if iters < 0:
        raise ValueError("iters must be >= 0")
    llr = llr.contiguous()
    dev = llr.device
    tabs = device_tables(code, dev)
    nblk, n = llr.shape
    if n != tabs.n:
        raise ValueError(f"LLR width {n} != code length {tabs.n}")
    dt = 1 if llr.dtype == torch.float64 else 0
    bits = torch.empty((nblk, n), dtype=torch.uint8, device=dev)
    conv = torch.empty(nblk, dtype=torch.uint8, device=dev)
    its = torch.empty(nblk, dtype=torch.int32, device=dev)
    bad = torch.empty(nblk, dtype=torch.int32, device=dev)
    for lo in range(0, nblk, 65535):
        hi = min(nblk, lo + 65535)
        nb = hi - lo
        wsb = int(_lib.lib().atsc3_minsum_utc_ws_bytes(nb, n, tabs.nedges, dt))
        ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
        _lib.call("atsc3_minsum_utc_dev", llr[lo:hi].data_ptr(), dt, NAN_RULES[nan_rule], nb,
                  n, tabs.nc,
                  tabs.nedges, tabs.cptr.data_ptr(), tabs.cidx.data_ptr(),
                  tabs.vptr.data_ptr(), tabs.vedge.data_ptr(), tabs.dmax, int(iters),
                  float(alpha), ws.data_ptr(), wsb, bits[lo:hi].data_ptr(),
                  conv[lo:hi].data_ptr(), its[lo:hi].data_ptr(), bad[lo:hi].data_ptr(),
                  _stream_ptr(dev))
    return bits, conv, its, bad


def min_sum_decode_batch(lam2d, checks, n, iters=50, alpha=0.75, dtype=np.float64,
                         msg_dtype="float32", device="cuda", engine="auto"):
    """lab/m3_ldpc.py:509: (nblk, n) LLRs -> (bits, conv, its, bad) numpy.

    dtype float32: the compiled-kernel contract (lab/ldpc_kernel.c:65-146,
    the upstream's float32 route), bit-exact, on the fast engines (msg_dtype
    selects the FP16 / INT8 message forms).  dtype float64 (the upstream's
    default): the update-then-check float64 decoder, row-exact with
    lab/m3_ldpc.py:245-286 (NaN LLRs included), run in FP64 on the GPU.
    """
    dt = np.dtype(dtype)
    if dt not in (np.float32, np.float64):
        raise ValueError("dtype must be float32 or float64")
    code = resolve_code(checks, n)
    lam = np.asarray(lam2d) if not isinstance(lam2d, torch.Tensor) else lam2d
    nblk = int(lam.shape[0])
    if nblk == 0:
        return (np.zeros((0, n), np.uint8), np.zeros(0, bool),
                np.full(0, iters, np.int32), np.full(0, -1, np.int32))
    tdt = torch.float32 if dt == np.float32 else torch.float64
    if isinstance(lam, torch.Tensor):
        t = lam.to(device=device, dtype=tdt)
    else:
        t = torch.from_numpy(np.ascontiguousarray(lam, dt)).to(device)
    if dt == np.float64:
        return _to_numpy(*decode_utc_dev(t, code, iters, alpha, "numpy"))
    return _to_numpy(*decode_dev(t, code, iters, alpha, msg_dtype, engine))


def min_sum_decode(llr, checks, n, iters=50, alpha=0.75, device="cuda"):
    """lab/m3_ldpc.py:245: one block, float64 update-then-check ->
    (bits u8 (n,), converged bool, n_iter int, n_unsatisfied int)."""
    b, c, i, d = min_sum_decode_batch(np.asarray(llr, np.float64).reshape(1, -1), checks, n,
                                      iters, alpha, dtype=np.float64, device=device)
    return b[0], bool(c[0]), int(i[0]), int(d[0])


class GpuMinSum:
    """lab/m9_gpu.py:84: batched min-sum for ONE code, built once, reused.

    dtype "float32" / "float64" (the upstream default): the upstream's
    update-then-check contract in that precision, bit-exact with
    GpuMinSum.decode (lab/m9_gpu.py:138-211), its NaN propagation included.  contract="kernel" instead
    serves the compiled kernel's contract (lab/ldpc_kernel.c:82-143) on the
    fast engines, where msg_dtype selects the FP16 / INT8 message forms.
    """

    def __init__(self, checks, n=None, iters=50, alpha=0.75, device="cuda",
                 dtype="float64", contract="upstream", msg_dtype="float32", engine="auto"):
        if contract not in ("upstream", "kernel"):
            raise ValueError("contract is 'upstream' or 'kernel'")
        if contract == "upstream" and dtype not in UTC_DTYPES:
            raise NotImplementedError(
                f"dtype {dtype!r}: the update-then-check decoder runs in float32 or "
                "float64 (reduced-precision messages: contract='kernel', msg_dtype)")
        if contract == "kernel" and dtype != "float32":
            raise ValueError("the kernel contract computes in float32")
        self.code = resolve_code(checks, n if n is not None else getattr(checks, "n", None))
        self.n = self.code.n
        self.iters, self.alpha = iters, alpha
        self.device = torch.device(device)
        self.dtype_name, self.contract = dtype, contract
        self.dt = torch.float64 if dtype == "float64" else torch.float32
        self.msg_dtype, self.engine = msg_dtype, engine
        self.tables = device_tables(self.code, self.device)
        self.stats = dict(calls=0, iters=0, blocks=0)

    def decode(self, llr2d, iters=None, alpha=None):
        """(nblk, n) LLRs (ndarray or device tensor) -> numpy
        (bits u8 (nblk, n), converged bool, iters i32, unsatisfied i32).
        iters None or 0 means the constructor's value (lab/m9_gpu.py:145)."""
        iters = iters or self.iters
        if iters < 1:
            raise ValueError("iters must be >= 1")
        alpha = self.alpha if alpha is None else alpha
        if isinstance(llr2d, torch.Tensor):
            t = llr2d.to(self.device, dtype=self.dt)
        else:
            a = np.ascontiguousarray(np.asarray(llr2d, np.float64))
            t = torch.from_numpy(a).to(self.device, dtype=self.dt)
        if t.dim() != 2 or t.shape[1] != self.n:
            raise ValueError(f"LLRs must be (nblk, {self.n})")
        if self.contract == "upstream":
            out = _to_numpy(*decode_utc_dev(t, self.code, iters, alpha, "torch"))
        else:
            out = _to_numpy(*decode_dev(t, self.code, iters, alpha, self.msg_dtype,
                                        self.engine))
        self.stats["calls"] += 1
        self.stats["blocks"] += int(t.shape[0])
        self.stats["iters"] += int(out[2].max(initial=0))
        return out


def ladder_dev(llr, code, ladder=(1.0, 0.85, 0.75), iters=50, msg_dtype="float32",
               engine="auto"):
    """Alpha ladder on device tensors; returns device (bits, conv, its, bad).

    Up to 4 rungs run inside one launch, with no host sync: engine "team"
    (the resident engine's team mode; "auto" picks it for FP32 batches
    below PHASE_MIN_BLOCKS codewords, where it is the faster) or "batch" /
    "auto" (the batched persistent kernel).  Otherwise one decode per rung
    and one host sync per rung (the number of rows still failing)."""
    nb = llr.shape[0]
    dev = llr.device
    if len(ladder) <= 4 and (engine == "team" or (engine == "auto" and msg_dtype == "float32"
                                                  and nb < PHASE_MIN_BLOCKS)):
        from .resident import decode_resident_dev
        llr = llr.contiguous()
        if llr.dtype != torch.float32:
            llr = llr.float()
        return decode_resident_dev(llr, code, iters, ladder[0], msg_dtype=msg_dtype,
                                   team=16 if nb <= 16 else 8, direct=True, ladder=ladder)
    if engine in ("auto", "batch") and len(ladder) <= 4:
        from .batch import decode_batch_dev
        return decode_batch_dev(llr, code, iters, ladder[0], msg_dtype, ladder=ladder)
    bits = torch.empty((nb, code.n), dtype=torch.uint8, device=dev)
    conv = torch.zeros(nb, dtype=torch.uint8, device=dev)
    its = torch.zeros(nb, dtype=torch.int32, device=dev)
    bad = torch.full((nb,), 1 << 30, dtype=torch.int32, device=dev)
    todo = torch.arange(nb, device=dev)
    for a in ladder:
        if todo.numel() == 0:
            break
        sub = llr if todo.numel() == nb else llr.index_select(0, todo)
        b, c, it, d = decode_dev(sub, code, iters, a, msg_dtype, engine)
        take = (c != 0) | (d < bad.index_select(0, todo))
        sel = todo[take]
        bits[sel] = b[take]
        its[sel] = it[take]
        bad[sel] = d[take]
        conv[sel] = c[take]
        todo = todo[c == 0]
    return bits, conv, its, bad


def ladder_decode(lam2d, checks, n, ladder=(1.0, 0.85, 0.75), iters=50,
                  msg_dtype="float32", device="cuda"):
    code = resolve_code(checks, n)
    t = torch.as_tensor(np.ascontiguousarray(lam2d, np.float32)).to(device)
    return _to_numpy(*ladder_dev(t, code, ladder, iters, msg_dtype))


def minsum_f32_host(lam2d, checks, n, iters=50, alpha=0.75):
    """The host-pointer drop-in ``ldpc_minsum_f32`` (lab/ldpc_kernel.c:65),
    driven exactly the way lab/m3_ldpc.py:485-506 drives the upstream one."""
    code = resolve_code(checks, n)
    cptr, cidx, vptr, vedge = code.csr()
    llr = np.ascontiguousarray(lam2d, np.float32)
    nblk = llr.shape[0]
    bits = np.empty((nblk, n), np.uint8)
    conv = np.zeros(nblk, np.uint8)
    its = np.empty(nblk, np.int32)
    bad = np.empty(nblk, np.int32)
    work = np.empty(n + len(cidx), np.float32)
    p = lambda a: a.ctypes.data
    _lib.call("ldpc_minsum_f32", p(llr), nblk, n, len(cptr) - 1, len(cidx), p(cptr),
              p(cidx), p(vptr), p(vedge), int(iters), float(np.float32(alpha)), p(bits),
              p(conv), p(its), p(bad), p(work))
    return bits, conv.view(bool), its, bad
