"""Fitting per-iteration, per-check-degree normalization factors.

The unrolled normalized min-sum (atsc3_amd.diff) is trained on synthetic
codewords: QPSK over AWGN with Eb/N0 uniform in a range, cells demapped by
the differentiable soft demapper, T unrolled iterations, binary
cross-entropy of the output LLRs against the sent bits, Adam on the (T,
classes) factors initialised at the plain value.  The learned table is then
compared with the plain decoder (one factor for every iteration and
degree) on fresh blocks; the plain bit-exact min-sum is the fallback for
blocks the learned decoder does not bring to a zero syndrome.

The upstream tunes a single alpha per rung of its ladder
(lab/m6_bicm.py:313, lab/m9_fast.py:843-891); this generalises that one
number to a table learned by gradient descent.
"""
from __future__ import annotations

import math
import time

import numpy as np
import torch

from . import ldpc, synth
from .codes import get_code
from .diff import SoftDemapper, UnrolledGraph, bce_on_llr, soft_demap, unrolled_minsum


def esn0_of_ebn0(ebn0_db, rate, mod):
    from .bicm import MOD_BITS
    r = [int(x) for x in rate.split("/")]
    return ebn0_db + 10.0 * math.log10(MOD_BITS[mod] * r[0] / r[1])


class Trainer:
    def __init__(self, n=16200, rate="6/15", mod="QPSK", iters=20, batch=512,
                 ebn0=(1.0, 3.0), alpha0=0.75, lr=0.01, ckpt_every=5, seed=0, device="cuda",
                 ncw=512):
        self.device = torch.device(device)
        self.n, self.rate, self.mod = n, rate, mod
        self.code = get_code(rate, n)
        self.graph = UnrolledGraph(self.code, self.device)
        self.dm = SoftDemapper(n, mod, rate, self.device)
        self.points = self.dm.points64.float()
        self.iters, self.batch, self.ebn0 = iters, batch, ebn0
        self.alpha0 = alpha0
        self.alphas = torch.full((iters, self.graph.ncls), alpha0, dtype=torch.float32,
                                 device=self.device, requires_grad=True)
        self.opt = torch.optim.Adam([self.alphas], lr=lr)
        self.ckpt_every = ckpt_every
        self.stream = synth.AwgnStream(rate, n, mod, ncw, seed, self.device)
        self.gen = torch.Generator(device=self.device)
        self.gen.manual_seed(seed + 99)
        self.step_no = 0
        self.history = []

    def sample(self, nblk, start):
        lo, hi = self.ebn0
        eb = lo + (hi - lo) * torch.rand(nblk, generator=self.gen, device=self.device,
                                         dtype=torch.float64)
        es = eb + (esn0_of_ebn0(0.0, self.rate, self.mod))
        z, idx, n0 = self.stream.batch(nblk, es, start=start)
        return torch.view_as_real(z).contiguous(), idx, n0

    def step(self):
        z, idx, n0 = self.sample(self.batch, self.step_no * self.batch)
        llr = soft_demap(z, self.points, n0.float(), self.dm)
        tot = unrolled_minsum(llr, self.alphas, self.graph, self.ckpt_every)
        bits = self.stream.cw_dev.index_select(0, idx)
        loss = bce_on_llr(tot, bits)
        self.opt.zero_grad(set_to_none=True)
        loss.backward()
        self.opt.step()
        self.step_no += 1
        v = float(loss.item())
        self.history.append(v)
        return v

    def train(self, steps, log=None):
        t0 = time.perf_counter()
        for k in range(steps):
            v = self.step()
            if log and (k % max(1, steps // 10) == 0 or k == steps - 1):
                log(f"[learn] step {k}: loss {v:.5f}")
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    @torch.no_grad()
    def evaluate(self, ebn0_db, nblk, chunk=2048, fallback_iters=50, seed_offset=10 ** 6):
        """FER at one Eb/N0 on the same fresh blocks: learned table, plain
        table (every factor alpha0), the plain bit-exact min-sum with early
        stop at iters = T and at the upstream's 50 iterations, the upstream's
        alpha ladder (1.0, 0.85, 0.75 at 50 iterations, lab/m9_fast.py:
        843-891), and the learned decoder with the plain 50-iteration
        min-sum as fallback."""
        es = esn0_of_ebn0(ebn0_db, self.rate, self.mod)
        plain = torch.full_like(self.alphas, self.alpha0)
        learned = self.alphas.detach()
        err = dict(learned=0, plain_unrolled=0, plain_minsum=0, plain_minsum_50=0,
                   plain_ladder_50=0, learned_fallback=0)
        done = 0
        while done < nblk:
            b = min(chunk, nblk - done)
            z, idx, n0 = self.stream.batch(b, es, start=seed_offset + done)
            llr = soft_demap(torch.view_as_real(z).contiguous(), self.points, n0.float(), self.dm)
            ref = self.stream.cw_dev.index_select(0, idx)
            for name, tab in (("learned", learned), ("plain_unrolled", plain)):
                tot, _ = self.graph.forward(llr, tab)
                hard = (tot < 0).to(torch.uint8)
                bad = (hard != ref).any(1)
                err[name] += int(bad.sum().item())
                if name == "learned":
                    syn_bad = self._syndrome_nonzero(hard)
                    fb = torch.nonzero(syn_bad).flatten()
                    bad_fb = bad.clone()
                    if fb.numel():
                        o = ldpc.decode_dev(llr.index_select(0, fb).contiguous(), self.code,
                                            fallback_iters, self.alpha0)
                        bad_fb[fb] = (o[0] != ref.index_select(0, fb)).any(1)
                    err["learned_fallback"] += int(bad_fb.sum().item())
            o = ldpc.decode_dev(llr, self.code, self.iters, self.alpha0)
            err["plain_minsum"] += int((o[0] != ref).any(1).sum().item())
            o = ldpc.decode_dev(llr, self.code, fallback_iters, self.alpha0)
            err["plain_minsum_50"] += int((o[0] != ref).any(1).sum().item())
            o = ldpc.ladder_dev(llr, self.code, (1.0, 0.85, 0.75), fallback_iters)
            err["plain_ladder_50"] += int((o[0] != ref).any(1).sum().item())
            done += b
        return {k: v / nblk for k, v in err.items()}, err

    def _syndrome_nonzero(self, hard):
        g = self.graph
        if not hasattr(self, "_chk"):
            cptr = g.cptr.cpu().numpy()
            deg = np.diff(cptr)
            rows = np.repeat(np.arange(len(deg)), deg)
            self._chk = (torch.from_numpy(rows).to(self.device), g.cidx.long())
        rows, cols = self._chk
        par = torch.zeros((hard.shape[0], g.nc), dtype=torch.int32, device=self.device)
        par.index_add_(1, rows, hard[:, cols].to(torch.int32))
        return (par & 1).any(1)
