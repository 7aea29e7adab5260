"""Seeded synthetic workloads for the demap + decode path.

Everything here is deterministic given (seed, shapes): numpy's PCG64
``default_rng``.  Used by the tests, the smoke check and bench.py; none of it
is on the timed path.

Transmit chain (the shape of lab/m6_bicm.py:367-373 with BCH and scrambler
bypassed): uniform info bits -> LDPC encode -> q = cw[lam_of_q] -> NUC map.

Channels:
  awgn      y = x + n,  n ~ CN(0, N0),  N0 = 10^(-Es/N0 / 10), E|x|^2 = 1
  rayleigh  y = h x + n, h = g1 + j g2 with g ~ N(0, 1) per cell, i.e. |h| is
            Rayleigh(sigma = 1) with uniform phase (E|h|^2 = 2); perfect CSI:
            the receiver gets z = y / h and w = |h|^2
  ldm       y = (x_core + 10^(inj/20) x_enh) / sqrt(1 + 10^(inj/10)) + n
"""
from __future__ import annotations

import numpy as np

from . import bicm
from .codes import get_code


def n0_of(esn0_db):
    return 10.0 ** (-esn0_db / 10.0)


def codewords(rate, n, nblk, seed):
    code = get_code(rate, n)
    rng = np.random.default_rng(seed)
    info = rng.integers(0, 2, (nblk, code.K), dtype=np.uint8)
    return code.encode(info)


def modulate(cw, n, mod, rate):
    lq = bicm.bicm_map(n, mod, rate)
    return bicm.map_cells(np.asarray(cw)[:, lq], mod, rate)


def complex_noise(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2)


def awgn_blocks(rate, n, mod, nblk, esn0_db, seed):
    """(cw uint8 (B,n), cells complex128 (B,C), N0)."""
    cw = codewords(rate, n, nblk, seed)
    x = modulate(cw, n, mod, rate)
    rng = np.random.default_rng(seed + 7919)
    n0 = n0_of(esn0_db)
    y = x + complex_noise(rng, x.shape) * np.sqrt(n0)
    return cw, y, n0


def rayleigh_blocks(rate, n, mod, nblk, esn0_db, seed):
    """(cw, equalised cells z, csi w = |h|^2, N0) under per-cell Rayleigh."""
    cw = codewords(rate, n, nblk, seed)
    x = modulate(cw, n, mod, rate)
    rng = np.random.default_rng(seed + 104729)
    h = rng.standard_normal(x.shape) + 1j * rng.standard_normal(x.shape)
    n0 = n0_of(esn0_db)
    y = h * x + complex_noise(rng, x.shape) * np.sqrt(n0)
    return cw, y / h, np.abs(h) ** 2, n0


def llr_bpsk_like(cw, snr_db, seed, scale=None):
    """Quick float32 LLRs for decoder-only tests: (1-2b)*a + sigma*N(0,1)."""
    cw = np.asarray(cw)
    rng = np.random.default_rng(seed)
    sigma = np.float32(10 ** (-snr_db / 20.0))
    a = np.float32(1.0)
    noise = rng.standard_normal(cw.shape, dtype=np.float32)
    y = (np.float32(1) - np.float32(2) * cw.astype(np.float32)) * a + sigma * noise
    s = np.float32(2.0) / (sigma * sigma) if scale is None else np.float32(scale)
    return (y * s).astype(np.float32)


class FadingStream:
    """Device-side generator for the live-frame workload: a stream of 256-NUC
    (or any NUC) blocks under per-cell Rayleigh fading with perfect CSI.

    Codewords: ``ncw`` distinct seeded codewords, block i carries codeword
    i % ncw; every block gets its own fading and noise, drawn on the device
    from a seeded torch generator, so FER statistics are those of independent
    channel realisations.  Produces equalised cells z = y / h (complex64) and
    CSI weights w = |h|^2 (float32) resident in device memory.
    """

    def __init__(self, rate, n, mod, ncw, seed, device):
        import torch
        self.rate, self.n, self.mod = rate, n, mod
        self.device = torch.device(device)
        self.cw = codewords(rate, n, ncw, seed)
        x = modulate(self.cw, n, mod, rate).astype(np.complex64)
        self.x = torch.from_numpy(x).to(self.device)
        self.cw_dev = torch.from_numpy(self.cw).to(self.device)
        self.ncw = ncw
        self.gen = torch.Generator(device=self.device)
        self.gen.manual_seed(seed + 31337)

    def batch(self, nblk, esn0_db, start=0):
        import torch
        idx = (torch.arange(nblk, device=self.device) + start) % self.ncw
        x = self.x.index_select(0, idx)
        shp = x.shape
        g = lambda: torch.randn(shp, generator=self.gen, device=self.device,
                                dtype=torch.float32)
        h = torch.complex(g(), g())
        n0 = n0_of(esn0_db)
        nz = torch.complex(g(), g()) * float(np.sqrt(n0 / 2.0))
        z = x + nz / h
        w = (h.real * h.real + h.imag * h.imag).contiguous()
        return z.contiguous(), w, idx, n0


class AwgnStream:
    """Device-side generator of blocks over complex AWGN (unit channel):
    ``ncw`` seeded codewords, block i carries codeword (start + i) % ncw,
    noise from a seeded torch generator.  ``esn0_db`` is a scalar or one
    value per block.  Returns cells (nblk, ncell) complex64, the codeword
    indices, and N0 per block (float64 device tensor)."""

    def __init__(self, rate, n, mod, ncw, seed, device):
        import torch
        self.rate, self.n, self.mod = rate, n, mod
        self.device = torch.device(device)
        self.cw = codewords(rate, n, ncw, seed)
        x = modulate(self.cw, n, mod, rate).astype(np.complex64)
        self.x = torch.from_numpy(x).to(self.device)
        self.cw_dev = torch.from_numpy(self.cw).to(self.device)
        self.ncw = ncw
        self.gen = torch.Generator(device=self.device)
        self.gen.manual_seed(seed + 2718)

    def batch(self, nblk, esn0_db, start=0):
        import torch
        idx = (torch.arange(nblk, device=self.device) + start) % self.ncw
        x = self.x.index_select(0, idx)
        es = torch.as_tensor(esn0_db, dtype=torch.float64, device=self.device).reshape(-1)
        n0 = torch.pow(10.0, -es / 10.0).expand(nblk).contiguous()
        g = lambda: torch.randn(x.shape, generator=self.gen, device=self.device,
                                dtype=torch.float32)
        amp = torch.sqrt(n0 / 2.0).to(torch.float32)[:, None]
        z = x + torch.complex(g() * amp, g() * amp)
        return z.contiguous(), idx, n0
