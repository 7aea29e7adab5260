"""Layered-division multiplexing: core + enhanced layer decode with SIC.

The two-layer chain of lab/m15_ldm_llr.py:77-85 (powers, Gaussian-folded core
demap) and lab/e69_enh.py:86-121 (core decode, re-modulation, per-block gain
``vdot(s, r) / vdot(s, s)``, subtraction, enhanced demap and decode), on
device tensors:

  1. core demap: the enhanced layer is folded into the noise --
     demap_llr(y / sqrt(P_core), sigma2 = (P_enh + N0) / P_core), run as a
     demap of y against sqrt(P_core) * points with sigma2 = P_enh + N0;
  2. core decode (min-sum, FP32 plain mode);
  3. cancellation (HIP kernel atsc3_ldm_cancel_dev): blocks that converged are
     re-modulated from their codeword and subtracted with the estimated gain;
  4. enhanced demap of the residual against sqrt(P_enh) * points with
     sigma2 = N0 (the known noise level of the synthetic configuration; the
     upstream live chain estimates it from the residual power), then decode.

Injection level: ``inj_db`` dB of the enhanced layer BELOW the core layer,
r = 10^(-inj_db/10), P_core = 1/(1+r), P_enh = r/(1+r) (lab/m15_ldm_llr.py:77-79).
Cell alignment: a 64800-bit QPSK core block spans 32400 cells and a 64800-bit
64-NUC enhanced block 10800, so core block b carries enhanced blocks 3b..3b+2
(the enhanced cell view is a reshape, no copy).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib, bicm, ldpc
from .codes import get_code
from .demap import Demapper


def powers(inj_db):
    r = 10.0 ** (-inj_db / 10.0)
    return 1.0 / (1.0 + r), r / (1.0 + r)


class LdmPipeline:
    def __init__(self, core=(64800, "QPSK", "4/15"), enh=(64800, "64QAM", "10/15"),
                 inj_db=5.0, iters=50, alpha=0.75, device="cuda", msg_dtype="float32",
                 engine="auto"):
        """engine: the decoder of both layers (ldpc.decode_dev's choice;
        "resident" needs msg_dtype float16)."""
        self.device = torch.device(device)
        self.core, self.enh = core, enh
        self.p_core, self.p_enh = powers(inj_db)
        self.iters, self.alpha, self.msg = iters, alpha, msg_dtype
        self.engine = engine
        n_c, mod_c, r_c = core
        n_e, mod_e, r_e = enh
        self.code_c, self.code_e = get_code(r_c, n_c), get_code(r_e, n_e)
        self.dm_c = Demapper(n_c, mod_c, r_c, self.device, point_scale=np.sqrt(self.p_core))
        self.dm_e = Demapper(n_e, mod_e, r_e, self.device, point_scale=np.sqrt(self.p_enh))
        self.ncell_c = self.dm_c.ncell
        self.ncell_e = self.dm_e.ncell
        if self.ncell_c % self.ncell_e:
            raise ValueError("core block must hold a whole number of enhanced blocks")
        self.ratio = self.ncell_c // self.ncell_e
        pts = bicm.points(mod_c, r_c)
        self.core_points = torch.from_numpy(
            np.stack([pts.real, pts.imag], 1).astype(np.float32)).to(self.device)
        self.core_nb = bicm.MOD_BITS[mod_c]

    def cancel(self, y, cw, conv, out_scale=1.0):
        """Residual y - g * s(cw) for converged blocks (device tensors)."""
        y = y.to(torch.complex64).contiguous()
        nblk, ncell = y.shape
        r = torch.empty_like(y)
        gain = torch.empty((nblk, 2), dtype=torch.float64, device=y.device)
        ws = torch.empty(max(1, _lib.lib().atsc3_ldm_ws_doubles(nblk, ncell)),
                         dtype=torch.float64, device=y.device)
        if nblk:
            _lib.call("atsc3_ldm_cancel_dev", torch.view_as_real(y).data_ptr(),
                      cw.contiguous().data_ptr(), conv.contiguous().data_ptr(), nblk, ncell,
                      self.code_c.n, self.dm_c.lam_of_q.data_ptr(),
                      self.core_points.data_ptr(), self.core_nb, float(out_scale),
                      torch.view_as_real(r).data_ptr(), gain.data_ptr(), ws.data_ptr(),
                      torch.cuda.current_stream(y.device).cuda_stream)
        return r, torch.view_as_complex(gain)

    def run(self, y, n0):
        """y: (ncore, ncell_core) complex64 device cells.  Returns a dict of
        device tensors: core (bits, conv, its, bad), enh (bits, conv, its, bad)
        with 3 enhanced blocks per core block, and the per-block gains."""
        llr_c, _ = self.dm_c.demap_dev(y, sigma2=self.p_enh + n0)
        core = ldpc.decode_dev(llr_c, self.code_c, self.iters, self.alpha, self.msg,
                               self.engine)
        r, gain = self.cancel(y, core[0], core[1])
        r_e = r.reshape(-1, self.ncell_e)
        llr_e, _ = self.dm_e.demap_dev(r_e, sigma2=n0)
        enh = ldpc.decode_dev(llr_e, self.code_e, self.iters, self.alpha, self.msg,
                              self.engine)
        return dict(core=core, enh=enh, gain=gain, llr_core=llr_c, llr_enh=llr_e, resid=r)


class LdmStream:
    """Seeded synthetic LDM cells on the device:
    y = sqrt(P_core) x_core + sqrt(P_enh) x_enh + n, n ~ CN(0, N0)."""

    def __init__(self, pipe, ncw, seed):
        from . import synth
        self.pipe = pipe
        nc, modc, rc = pipe.core
        ne, mode, re = pipe.enh
        self.cw_c = synth.codewords(rc, nc, ncw, seed)
        self.cw_e = synth.codewords(re, ne, ncw * pipe.ratio, seed + 1)
        dev = pipe.device
        self.xc = torch.from_numpy(synth.modulate(self.cw_c, nc, modc, rc).astype(np.complex64)).to(dev)
        self.xe = torch.from_numpy(synth.modulate(self.cw_e, ne, mode, re).astype(np.complex64)).to(dev)
        self.cw_c_dev = torch.from_numpy(self.cw_c).to(dev)
        self.cw_e_dev = torch.from_numpy(self.cw_e).to(dev)
        self.ncw = ncw
        self.gen = torch.Generator(device=dev)
        self.gen.manual_seed(seed + 4242)

    def batch(self, ncore, n0, start=0):
        p = self.pipe
        dev = p.device
        ic = (torch.arange(ncore, device=dev) + start) % self.ncw
        ie = (ic[:, None] * p.ratio + torch.arange(p.ratio, device=dev)[None, :]).reshape(-1)
        xc = self.xc.index_select(0, ic)
        xe = self.xe.index_select(0, ie).reshape(ncore, p.ncell_c)
        shp = xc.shape
        g = lambda: torch.randn(shp, generator=self.gen, device=dev, dtype=torch.float32)
        nz = torch.complex(g(), g()) * float(np.sqrt(n0 / 2.0))
        y = float(np.sqrt(p.p_core)) * xc + float(np.sqrt(p.p_enh)) * xe + nz
        return y.contiguous(), ic, ie
