"""GPU max-log demapping behind the upstream demap call signatures.

Mirrors:
  * ``demap_llr(cells, mod, rate, sigma2=None)`` -- lab/m6_bicm.py:213-225:
    one block of cells -> q-ordered LLRs (numpy).
  * ``demap_batch(cells2d, pts, ones, chunk=8, d2min_out=None)`` --
    lab/m9_fast.py:756 FrameDecoder.demap_batch: (nblk, ncell) -> (nblk,
    ncell*nb) q-ordered LLRs, sigma2 estimated per block; with d2min_out the
    unscaled differences and each cell's nearest-point distance.
  * ``demap_batch_qam(cells2d, ch, ex=None, threads=4, chunk=4)`` --
    lab/m44_ldm.py:2454, the same for a chain object's alphabet.
  * ``demap_q_batch_dev(cells2d, pts, sigma2, csi)`` -- lab/m9_gpu.py:301
    (GpuBicm.demap): device tensors, sigma2 given or estimated per block.
  * ``Demapper(n, mod, rate).demap_dev(cells, csi, sigma2)`` -- the fused
    form: demap plus the caller's de-interleave ``lam[lam_of_q] = q``
    (lab/m6_bicm.py:331-335), device tensors in and out, codeword order.

LLR_i = w * (min_{y_i=1} |z-p|^2 - min_{y_i=0} |z-p|^2) / sigma2, llr > 0 means
bit 0.  w is the per-cell CSI weight |h|^2 of equalised cells (1 if none);
sigma2 None -> per block max(mean(w * min_p |z-p|^2), 1e-9), reduced in FP64
on the device.
"""
from __future__ import annotations

import ctypes
import functools

import numpy as np
import torch

from . import _lib, bicm


def _check_quadrant_rule(pts):
    """The kernel evaluates max-log through the quadrant rule; verify it."""
    pts = np.asarray(pts, np.complex128)
    m = len(pts)
    nb = int(round(np.log2(m)))
    if 1 << nb != m or nb not in (2, 4, 6, 8):
        raise ValueError(f"unsupported alphabet size {m}")
    w = pts[: m // 4]
    ok = (np.allclose(pts[m // 4: m // 2], -np.conj(w), rtol=0, atol=0)
          and np.allclose(pts[m // 2: 3 * m // 4], np.conj(w), rtol=0, atol=0)
          and np.allclose(pts[3 * m // 4:], -w, rtol=0, atol=0)
          and (w.real >= 0).all() and (w.imag >= 0).all())
    if not ok:
        raise ValueError("alphabet does not follow the A/322 quadrant rule "
                         "with first-quadrant base points")
    return nb


def _stream(dev):
    return torch.cuda.current_stream(dev).cuda_stream


def demap_q_dev(cells, pts_dev, nb, lam_of_q_dev, csi=None, sigma2=None, out=None,
                want_sigma2=True):
    """Core launcher.  cells (B, C) complex64 device tensor.  want_sigma2
    False: no sigma2 tensor is returned (None), which saves the copy of a
    given sigma2 (one small launch) in per-frame graphs."""
    if not (isinstance(cells, torch.Tensor) and cells.is_cuda):
        raise TypeError("cells must be a device tensor")
    cells = cells.to(torch.complex64).contiguous()
    B, C = cells.shape
    dev = cells.device
    N = C * nb
    llr = out if out is not None else torch.empty((B, N), dtype=torch.float32, device=dev)
    s2out = (torch.empty(B, dtype=torch.float64, device=dev)
             if want_sigma2 or sigma2 is None else None)
    if B == 0 or C == 0:
        return llr, s2out
    csi_p = None
    if csi is not None:
        csi = torch.as_tensor(csi, device=dev).to(torch.float32).contiguous()
        assert csi.shape == (B, C)
        csi_p = csi.data_ptr()
    s2_p, ws, ws_p = None, None, None
    if sigma2 is not None:
        sigma2 = torch.as_tensor(sigma2, dtype=torch.float64, device=dev).reshape(-1)
        if sigma2.numel() == 1 and B > 1:
            sigma2 = sigma2.expand(B)
        sigma2 = sigma2.contiguous()
        s2_p = sigma2.data_ptr()
    else:
        nws = _lib.lib().atsc3_demap_ws_doubles(B, C)
        ws = torch.empty(max(nws, 1), dtype=torch.float64, device=dev)
        ws_p = ws.data_ptr()
    _lib.call("atsc3_demap_dev", torch.view_as_real(cells).data_ptr(), csi_p, s2_p, B, C,
              pts_dev.data_ptr(), nb, lam_of_q_dev.data_ptr(), llr.data_ptr(),
              None if s2out is None else s2out.data_ptr(), ws_p, _stream(dev))
    return llr, s2out


class _DemapGroup(ctypes.Structure):
    """struct atsc3_demap_group (include/atsc3_fec.h)."""
    _fields_ = ([(f, ctypes.c_void_p) for f in ("cells", "csi", "sigma2", "points", "lam_of_q",
                                                 "llr")]
                + [(f, ctypes.c_int32) for f in ("ncell", "nbits", "nblk", "blk0")])


assert ctypes.sizeof(_DemapGroup) == 64


class GroupedDemap:
    """The demaps of several (Demapper, cells, sigma2, out[, csi]) groups --
    a frame's blocks of different (N, constellation, rate) -- as ONE launch
    (``atsc3_demap_groups_dev``), with Demapper.demap_dev's LLRs.  The
    tensors are referenced, not copied: refill them in place between runs
    (the descriptor table holds their addresses), so the launch can be
    captured in a HIP graph."""

    def __init__(self, groups, device):
        self.device = torch.device(device)
        self.keep = []
        arr = (_DemapGroup * len(groups))()
        blk0, max_ncell = 0, 0
        for i, g in enumerate(groups):
            dm, cells, sigma2, out = g[:4]
            csi = g[4] if len(g) > 4 else None
            nblk = cells.shape[0]
            if (cells.dtype != torch.complex64 or not cells.is_contiguous()
                    or cells.shape[1] != dm.ncell):
                raise ValueError(f"group {i}: cells must be contiguous complex64 (B, ncell)")
            if (sigma2.dtype != torch.float64 or sigma2.numel() != nblk
                    or not sigma2.is_contiguous()):
                raise ValueError(f"group {i}: sigma2 must be contiguous float64 (B,)")
            if (out.dtype != torch.float32 or not out.is_contiguous()
                    or out.numel() != nblk * dm.ncell * dm.nb):
                raise ValueError(f"group {i}: out must be contiguous float32 (B, n)")
            if csi is not None and (csi.dtype != torch.float32 or csi.shape != cells.shape
                                    or not csi.is_contiguous()):
                raise ValueError(f"group {i}: csi must be contiguous float32 like the cells")
            for t in (cells, sigma2, out) + ((csi,) if csi is not None else ()):
                if t.device != self.device:
                    raise ValueError(f"group {i}: tensors must be on {self.device}")
            r = arr[i]
            r.cells = torch.view_as_real(cells).data_ptr()
            r.csi = None if csi is None else csi.data_ptr()
            r.sigma2 = sigma2.data_ptr()
            r.points = dm.points.data_ptr()
            r.lam_of_q = dm.lam_of_q.data_ptr()
            r.llr = out.data_ptr()
            r.ncell, r.nbits, r.nblk, r.blk0 = dm.ncell, dm.nb, nblk, blk0
            blk0 += nblk
            max_ncell = max(max_ncell, dm.ncell)
            self.keep.append((dm, cells, sigma2, out, csi))
        if blk0 > 65535:
            raise ValueError("at most 65535 blocks per grouped demap")
        self.ngroups, self.nblk, self.max_ncell = len(groups), blk0, max_ncell
        raw = np.frombuffer(bytes(arr), np.uint8).copy()
        self.table = torch.from_numpy(raw).to(self.device)

    def run(self):
        _lib.call("atsc3_demap_groups_dev", self.table.data_ptr(), self.ngroups, self.nblk,
                  self.max_ncell, _stream(self.device))


class GatherSpec(ctypes.Structure):
    """struct atsc3_gather (include/atsc3_fec.h)."""
    _fields_ = [("mode", ctypes.c_int32), ("nrows", ctypes.c_int32),
                ("start_row", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("idx", ctypes.c_void_p), ("first", ctypes.c_int64), ("ring", ctypes.c_int64)]


GATHER_TABLE, GATHER_CTI = 1, 2


def demap_gather_q_dev(pool, pts_dev, nb, lam_of_q_dev, spec, nblk, ncell, csi=None,
                       sigma2=None, out=None):
    """Demap ``nblk`` blocks of ``ncell`` cells read from ``pool`` (1-D
    complex64 device tensor) through ``spec`` (a GatherSpec)."""
    if not (isinstance(pool, torch.Tensor) and pool.is_cuda and pool.dtype == torch.complex64
            and pool.dim() == 1 and pool.is_contiguous()):
        raise TypeError("pool must be a contiguous 1-D complex64 device tensor")
    dev = pool.device
    N = ncell * nb
    llr = out if out is not None else torch.empty((nblk, N), dtype=torch.float32, device=dev)
    s2out = torch.empty(nblk, dtype=torch.float64, device=dev)
    if nblk == 0 or ncell == 0:
        return llr, s2out
    csi_p = None
    if csi is not None:
        if not (csi.is_cuda and csi.dtype == torch.float32 and csi.shape == pool.shape
                and csi.is_contiguous()):
            raise ValueError("csi must be a float32 device tensor shaped like the pool")
        csi_p = csi.data_ptr()
    s2_p, ws, ws_p = None, None, None
    if sigma2 is not None:
        sigma2 = torch.as_tensor(sigma2, dtype=torch.float64, device=dev).reshape(-1)
        if sigma2.numel() == 1 and nblk > 1:
            sigma2 = sigma2.expand(nblk)
        sigma2 = sigma2.contiguous()
        s2_p = sigma2.data_ptr()
    else:
        nws = _lib.lib().atsc3_demap_ws_doubles(nblk, ncell)
        ws = torch.empty(max(nws, 1), dtype=torch.float64, device=dev)
        ws_p = ws.data_ptr()
    _lib.call("atsc3_demap_gather_dev", torch.view_as_real(pool).data_ptr(), csi_p,
              ctypes.byref(spec), s2_p, nblk, ncell, pts_dev.data_ptr(), nb,
              lam_of_q_dev.data_ptr(), llr.data_ptr(), s2out.data_ptr(), ws_p, _stream(dev))
    return llr, s2out


def gather_cells_dev(pool, spec, nblk, ncell, out=None):
    """(nblk, ncell) complex64 device tensor of the cells ``spec`` maps to."""
    if not (pool.is_cuda and pool.dtype == torch.complex64 and pool.is_contiguous()):
        raise TypeError("pool must be a contiguous complex64 device tensor")
    out = out if out is not None else torch.empty((nblk, ncell), dtype=torch.complex64,
                                                  device=pool.device)
    if nblk and ncell:
        _lib.call("atsc3_gather_cells_dev", torch.view_as_real(pool).data_ptr(),
                  ctypes.byref(spec), nblk, ncell, torch.view_as_real(out).data_ptr(),
                  _stream(pool.device))
    return out


@functools.lru_cache(maxsize=None)
def _points_dev(key, device):
    mod, rate, scale = key
    pts = bicm.points(mod, rate) * scale
    nb = _check_quadrant_rule(pts)
    t = torch.from_numpy(np.stack([pts.real, pts.imag], 1).astype(np.float32)).to(device)
    return t.contiguous(), nb


@functools.lru_cache(maxsize=None)
def _identity_dev(N, device):
    return torch.arange(N, dtype=torch.int32, device=device)


class Demapper:
    """Fused demap + de-interleave for one (N, modulation, rate).

    point_scale a > 0 demaps against the alphabet a*p: with sigma2 scaled by
    a^2 this is the same LLR as demapping y / a against p (how the LDM core
    layer is demapped with the enhanced layer folded into the noise,
    lab/m15_ldm_llr.py:84-85), without a pass over the cells.
    """

    def __init__(self, n, mod, rate, device="cuda", point_scale=1.0):
        self.n, self.mod, self.rate = int(n), mod, rate
        self.device = torch.device(device)
        if not point_scale > 0:
            raise ValueError("point_scale must be positive")
        self.points, self.nb = _points_dev((mod, rate, float(point_scale)), self.device)
        self.ncell = bicm.cells_per_block(self.n, mod)
        lq = bicm.bicm_map(self.n, mod, rate)
        self.lam_of_q = torch.from_numpy(lq.astype(np.int32)).to(self.device)

    def demap_dev(self, cells, csi=None, sigma2=None, out=None, want_sigma2=True):
        """(B, ncell) cells -> ((B, n) f32 codeword-order LLRs, sigma2 (B,) f64;
        None when sigma2 is given and want_sigma2 is False)."""
        if cells.shape[1] != self.ncell:
            raise ValueError(f"{cells.shape[1]} cells per block, expected {self.ncell}")
        return demap_q_dev(cells, self.points, self.nb, self.lam_of_q, csi, sigma2, out,
                           want_sigma2)

    def demap_gather_dev(self, pool, spec, nblk, csi=None, sigma2=None, out=None):
        """Blocks read from ``pool`` through a GatherSpec (the time
        de-interleaver's FEC-block gather) -> same as demap_dev on the
        gathered cells."""
        return demap_gather_q_dev(pool, self.points, self.nb, self.lam_of_q, spec, nblk,
                                  self.ncell, csi, sigma2, out)


def _alphabet_dev(pts, dev):
    pts = np.asarray(pts)
    nb = _check_quadrant_rule(pts)
    p = torch.from_numpy(np.stack([pts.real, pts.imag], 1).astype(np.float32)).to(dev)
    return p, nb


def demap_q_batch_dev(cells2d, pts, sigma2=None, csi=None, device="cuda"):
    """(nblk, ncell) cells -> ((nblk, ncell*nb) q-order float32 device tensor,
    sigma2 (nblk,) float64 device tensor); sigma2 None -> per-block estimate."""
    dev = torch.device(device)
    p, nb = _alphabet_dev(pts, dev)
    if isinstance(cells2d, torch.Tensor):
        z = cells2d.to(dev, torch.complex64)
    else:
        z = torch.from_numpy(np.ascontiguousarray(cells2d, np.complex64)).to(dev)
    z = z.reshape(-1, z.shape[-1])
    ident = _identity_dev(z.shape[1] * nb, dev)
    return demap_q_dev(z, p, nb, ident, csi, sigma2)


def demap_batch(cells2d, pts, ones, chunk=8, d2min_out=None, *, device="cuda"):
    """lab/m9_fast.py:756 FrameDecoder.demap_batch: (nblk, ncell) cells ->
    (nblk, ncell*nb) float32 max-log LLRs in q order, sigma^2 the mean of
    each row's nearest-point distances (floored at 1e-9).

    ones = the label bit masks (bit_masks(nb), y0 the MSB), as the upstream
    passes them; chunk is the upstream's CPU work split and has no effect
    here.  d2min_out: an (nblk, ncell) float32 array (or device tensor) that
    receives each cell's squared distance to the nearest point, and the
    LLRs come back unscaled (sigma^2 off), for per-cell noise weights
    (lab/m9_fast.py:767-771).  A device tensor in gives a device tensor out.
    """
    pts = np.asarray(pts)
    ones = np.asarray(ones)
    nb = _check_quadrant_rule(pts)
    if ones.shape != (nb, 1 << nb) or not np.array_equal(ones, bicm.bit_masks(nb)):
        raise ValueError("ones must be the MSB-first label bit masks of the alphabet")
    on_dev = isinstance(cells2d, torch.Tensor)
    dev = cells2d.device if on_dev else torch.device(device)
    if d2min_out is None:
        llr, _ = demap_q_batch_dev(cells2d, pts, device=dev)
        return llr if on_dev else llr.cpu().numpy()
    p, _ = _alphabet_dev(pts, dev)
    if on_dev:
        z = cells2d.to(dev, torch.complex64).contiguous()
    else:
        z = torch.from_numpy(np.ascontiguousarray(cells2d, np.complex64)).to(dev)
    B, C = z.shape
    if tuple(d2min_out.shape) != (B, C):
        raise ValueError(f"d2min_out must be ({B}, {C})")
    llr = torch.empty((B, C * nb), dtype=torch.float32, device=dev)
    d2 = torch.empty((B, C), dtype=torch.float32, device=dev)
    if B and C:
        ws = torch.empty(max(1, _lib.lib().atsc3_demap_ws_doubles(B, C)), dtype=torch.float64,
                         device=dev)
        _lib.call("atsc3_demap_raw_dev", torch.view_as_real(z).data_ptr(), None, B, C,
                  p.data_ptr(), nb, _identity_dev(C * nb, dev).data_ptr(), llr.data_ptr(),
                  d2.data_ptr(), ws.data_ptr(), _stream(dev))
    if isinstance(d2min_out, torch.Tensor):
        d2min_out.copy_(d2)
    else:
        d2min_out[...] = d2.cpu().numpy()
    return llr if on_dev else llr.cpu().numpy()


def demap_batch_qam(cells2d, ch, ex=None, threads=4, chunk=4, *, device="cuda"):
    """lab/m44_ldm.py:2454: (nblk, ncell) cells of the chain ``ch`` (any
    object with .mod, .rate, .mod_bits) -> (nblk, ncell*nb) float32 max-log
    LLRs in q order, per-row sigma^2 estimate.  ex / threads / chunk are the
    upstream's CPU thread pool and work split; unused here."""
    pts = bicm.points(ch.mod, ch.rate)
    if int(ch.mod_bits) != bicm.MOD_BITS[ch.mod]:
        raise ValueError("ch.mod_bits does not match ch.mod")
    return demap_batch(cells2d, pts, bicm.bit_masks(int(ch.mod_bits)), device=device)


def demap_llr(cells, mod, rate, sigma2=None, device="cuda"):
    """lab/m6_bicm.py:213: one block -> q-ordered LLRs (numpy float32)."""
    llr, _ = demap_q_batch_dev(np.asarray(cells)[None, :], bicm.points(mod, rate), sigma2,
                               device=device)
    return llr[0].cpu().numpy()
