"""CPU baseline harness: the upstream's multi-process decode topology.

TEST / CPU-BASELINE INFRASTRUCTURE ONLY (see oracle/__init__.py): bench.py's
``cpu_baseline`` leg times the CPU path with it; nothing in the product
imports it.

Topology (lab/m11_watch.py:67-113 ``_decode_proc_main``, :290-322
``_procs_start``, :343-428 FIFO re-order): ``multiprocessing`` spawn workers,
input cells in ``shared_memory``, work items on a Queue, results written back
into shared memory, the parent re-orders by item index.  Each worker runs
the CPU demap -- the float32 broadcast form of lab/m44_ldm.py:2431-2494
(``demap_f32_cpu``) -- then the bit de-interleave ``lam[:, lam_of_q] = q``,
then the FP32 min-sum C kernel: this repository's C restatement of the
upstream kernel (``liboracle_minsum.so``, equal to the upstream kernel built
from its own source on every golden fixture), so no binary built from the
upstream tree is needed where the baseline runs.
"""
from __future__ import annotations

import os
import time

import numpy as np


def demap_f32_cpu(cells, pts, csi=None, sigma2=None):
    """float32 max-log in the upstream cpu_fast form (lab/m44_ldm.py:2431-2494,
    lab/m9_fast.py:790-821): metric |p|^2 - 2 Re(z conj p) (the |z|^2 term
    cancels in the bitwise-minima difference), per-bit minima over the label
    bit axis, divided by sigma2; CSI weights multiply."""
    z = np.atleast_2d(cells)
    nblk, ncell = z.shape
    nb = int(np.log2(len(pts)))
    pr = np.asarray(pts.real, np.float32)
    pi = np.asarray(pts.imag, np.float32)
    p2 = (pts.real ** 2 + pts.imag ** 2).astype(np.float32)
    zr = np.asarray(z.real, np.float32).ravel()
    zi = np.asarray(z.imag, np.float32).ravel()
    q = zr[:, None] * pr[None, :]
    q += zi[:, None] * pi[None, :]
    q *= np.float32(-2.0)
    q += p2[None, :]
    q = q.reshape(nblk, ncell, len(pts))
    if sigma2 is None:
        z2 = zr.reshape(nblk, ncell) ** 2 + zi.reshape(nblk, ncell) ** 2
        w = 1.0 if csi is None else csi
        s2 = np.maximum((w * (z2 + q.min(2))).mean(1), 1e-9).astype(np.float32)
    else:
        s2 = np.broadcast_to(np.asarray(sigma2, np.float32), (nblk,))
    qr = q.reshape(nblk, ncell, *([2] * nb))
    sub = tuple(range(2, 2 + nb - 1))
    out = np.empty((nblk, ncell, nb), np.float32)
    head = (slice(None), slice(None))
    for i in range(nb):
        a = qr[head + (slice(None),) * i + (1,)].min(axis=sub)
        b = qr[head + (slice(None),) * i + (0,)].min(axis=sub)
        out[:, :, i] = a - b
    if csi is not None:
        out *= np.asarray(csi, np.float32)[:, :, None]
    out /= np.asarray(s2, np.float32)[:, None, None]
    return out.reshape(nblk, ncell * nb)


def _worker(shm_names, shapes, cfg, in_q, out_q):
    try:
        from multiprocessing import shared_memory
        import sys
        sys.path.insert(0, cfg["root"])
        from atsc3_amd import bicm
        from atsc3_amd.codes import get_code
        from oracle import bicm as OB
        from oracle import ldpc as OL
        shms = [shared_memory.SharedMemory(name=n) for n in shm_names]
        cells = np.ndarray(shapes["cells"], np.complex64, buffer=shms[0].buf)
        csi = np.ndarray(shapes["csi"], np.float32, buffer=shms[1].buf)
        bits = np.ndarray(shapes["bits"], np.uint8, buffer=shms[2].buf)
        meta = np.ndarray(shapes["meta"], np.int32, buffer=shms[3].buf)
        code = get_code(cfg["rate"], cfg["n"])
        pts = bicm.points(cfg["mod"], cfg["rate"])
        lq = OB.lam_of_q(cfg["n"], cfg["mod"], cfg["rate"])
        which = cfg["which"]

        def run(lo, hi):
            q = demap_f32_cpu(cells[lo:hi], pts, csi[lo:hi], cfg["sigma2"])
            lam = np.empty_like(q)
            lam[:, lq] = q
            b, c, it, d = OL.minsum_f32(lam, code, cfg["iters"], cfg["alpha"], which)
            bits[lo:hi] = b
            meta[lo:hi, 0] = c
            meta[lo:hi, 1] = it
            meta[lo:hi, 2] = d

        run(0, 1)                                  # warm-up (tables, caches)
        out_q.put(("ready", os.getpid()))
        while True:
            msg = in_q.get()
            if msg is None:
                break
            idx, lo, hi = msg
            t = time.perf_counter()
            run(lo, hi)
            out_q.put(("ok", (idx, time.perf_counter() - t)))
        for s in shms:
            s.close()
    except Exception as e:                         # noqa: BLE001
        import traceback
        out_q.put(("err", f"{type(e).__name__}: {e}\n{traceback.format_exc()}"))


def run_pool(cells, csi, sigma2, rate, n, mod, iters, alpha, procs, chunk=4,
             which=None, root=None):
    """Decode (demap + de-interleave + min-sum) the blocks with `procs`
    spawn workers; returns (seconds, bits, conv, its, bad, which)."""
    import multiprocessing as mp
    from multiprocessing import shared_memory
    from . import ldpc as OL
    if which is None:
        which = "oracle"
    root = root or os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    cells = np.ascontiguousarray(cells, np.complex64)
    csi = np.ascontiguousarray(csi, np.float32)
    nblk = cells.shape[0]
    shapes = dict(cells=cells.shape, csi=csi.shape, bits=(nblk, n), meta=(nblk, 3))
    sizes = [cells.nbytes, csi.nbytes, nblk * n, nblk * 12]
    shms = [shared_memory.SharedMemory(create=True, size=max(s, 1)) for s in sizes]
    try:
        np.ndarray(cells.shape, np.complex64, buffer=shms[0].buf)[:] = cells
        np.ndarray(csi.shape, np.float32, buffer=shms[1].buf)[:] = csi
        ctx = mp.get_context("spawn")
        in_q, out_q = ctx.Queue(), ctx.Queue()
        cfg = dict(rate=rate, n=n, mod=mod, iters=iters, alpha=alpha,
                   sigma2=sigma2, which=which, root=root)
        ps = [ctx.Process(target=_worker, args=([s.name for s in shms], shapes, cfg,
                                                 in_q, out_q), daemon=True)
              for _ in range(procs)]
        for p in ps:
            p.start()
        ready = 0
        while ready < procs:
            kind, payload = out_q.get(timeout=300)
            if kind == "err":
                raise RuntimeError(f"CPU worker failed: {payload}")
            ready += kind == "ready"
        items = [(i, lo, min(nblk, lo + chunk)) for i, lo in enumerate(range(0, nblk, chunk))]
        t0 = time.perf_counter()
        for it in items:
            in_q.put(it)
        done = {}
        while len(done) < len(items):
            kind, payload = out_q.get(timeout=600)
            if kind == "err":
                raise RuntimeError(f"CPU worker failed: {payload}")
            done[payload[0]] = payload[1]         # re-ordered by item index
        secs = time.perf_counter() - t0
        for _ in ps:
            in_q.put(None)
        for p in ps:
            p.join(timeout=30)
        bits = np.ndarray((nblk, n), np.uint8, buffer=shms[2].buf).copy()
        meta = np.ndarray((nblk, 3), np.int32, buffer=shms[3].buf).copy()
        return secs, bits, meta[:, 0].astype(bool), meta[:, 1], meta[:, 2], which
    finally:
        for s in shms:
            s.close()
            s.unlink()
