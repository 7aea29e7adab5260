"""Uneven frames: demap + decode of a mix of codes and constellations.

A frame carries ``nblk`` FEC blocks, each with its own (N, rate,
constellation) drawn in seeded random proportions from 64800-bit rates
6/15, 8/15, 10/15, 12/15 and 16200-bit rates 4/15, 9/15 on QPSK / 16-, 64-,
256-NUC, and its own AWGN Es/N0.  The receive chain is the upstream's per-PLP
FEC stage (lab/m9_fast.py:1087-1166: demap, de-interleave, LDPC decode of
every block), here for all of a frame's blocks at once:

  * the demaps of every (N, constellation, rate) group present in the frame
    as ONE launch (atsc3_amd.demap.GroupedDemap), writing codeword-order
    LLRs straight into the decoder's flat LLR buffer (the groups' blocks are
    laid out contiguously);
  * ONE persistent batched decode (atsc3_amd.batch) over the whole mix.

``MixedPipeline`` owns every buffer, so a frame's whole chain is captured in
one HIP graph and replayed per frame.

Per-block Es/N0 is drawn around each (code, constellation)'s decoding
threshold (``THRESHOLD_DB``: the lowest Es/N0, on a 0.5 dB grid, at which 64
AWGN blocks all decoded with FP32 min-sum, 50 iterations, alpha 0.75), from
``thr - 1`` dB (blocks that fail after 50 iterations) to ``thr + 3`` dB
(3-5 iterations), so iteration counts span 3..50.
"""
from __future__ import annotations

import numpy as np
import torch

from . import bicm, synth
from .batch import BatchPlan
from .codes import get_code
from .demap import Demapper, GroupedDemap

CODES = ((64800, "6/15"), (64800, "8/15"), (64800, "10/15"), (64800, "12/15"),
         (16200, "4/15"), (16200, "9/15"))
MODS = ("QPSK", "16QAM", "64QAM", "256QAM")

THRESHOLD_DB = {
    (64800, "6/15"): (2.5, 8.5, 14.0, 18.0),
    (64800, "8/15"): (3.0, 9.0, 14.5, 19.0),
    (64800, "10/15"): (3.5, 9.5, 14.5, 19.5),
    (64800, "12/15"): (5.0, 11.0, 16.0, 21.5),
    (16200, "4/15"): (-2.0, 2.5, 5.5, 7.5),
    (16200, "9/15"): (3.5, 8.5, 13.0, 18.5),
}


def frame_mix(nblk=74, seed=7):
    """Seeded random proportions: a Dirichlet draw over the 24 (code,
    constellation) pairs, then nblk blocks sampled from it.  Returns the
    per-block specs [(n, rate, mod)] sorted into groups (the frame layout)."""
    rng = np.random.default_rng(seed)
    pairs = [(n, r, m) for (n, r) in CODES for m in MODS]
    p = rng.dirichlet(np.full(len(pairs), 0.8))
    pick = rng.choice(len(pairs), size=nblk, p=p)
    return [pairs[i] for i in sorted(pick)]


def block_esn0(specs, seed, lo=-1.0, hi=3.0):
    rng = np.random.default_rng(seed)
    thr = np.array([THRESHOLD_DB[(n, r)][MODS.index(m)] for n, r, m in specs])
    return thr + rng.uniform(lo, hi, len(specs))


class MixedPipeline:
    """Demap + batched decode of frames with a fixed block layout."""

    def __init__(self, specs, frames=1, device="cuda", iters=50, alpha=0.75,
                 msg_dtype="float32", slots=None, grid=None, engine="batch", team=None,
                 direct=False):
        """engine "batch": the persistent batched engine (slots, grid);
        "xcd": the same engine partitioned per XCD (slots over all XCDs);
        "resident": one workgroup per codeword (grid = workgroups, slots
        unused); "team": the resident engine with each codeword on a team
        of ``team`` workgroups of one XCD (one workgroup per compute
        unit), the queue ranked on the device before each decode (the
        codewords likely to need the most iterations first); ``direct``:
        the team mode without LDS copies of the posteriors."""
        self.device = torch.device(device)
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.specs = list(specs) * frames
        self.groups = []            # (spec, first block, count)
        for b, s in enumerate(self.specs):
            if self.groups and self.groups[-1][0] == s and self.groups[-1][1] + self.groups[-1][2] == b:
                g = self.groups[-1]
                self.groups[-1] = (g[0], g[1], g[2] + 1)
            else:
                self.groups.append((s, b, 1))
        codes = [get_code(r, n) for n, r, m in self.specs]
        if engine == "resident":
            from .resident import ResidentPlan
            self.plan = ResidentPlan(codes, self.device, iters, alpha, grid, msg_dtype)
        elif engine == "team":
            from .resident import ResidentPlan
            self.plan = ResidentPlan(codes, self.device, iters, alpha, None, msg_dtype,
                                     team=8 if team is None else team, order="auto",
                                     direct=direct)
        elif engine in ("batch", "xcd"):
            self.plan = BatchPlan(codes, self.device, iters, alpha, msg_dtype, slots, grid,
                                  xcd=engine == "xcd")
        else:
            raise ValueError(f"engine {engine!r}: 'batch', 'xcd', 'resident' or 'team'")
        self.engine = engine
        self.demappers = {}
        self.cells, self.sigma2 = [], []
        for (n, r, m), b0, cnt in self.groups:
            key = (n, m, r)
            if key not in self.demappers:
                self.demappers[key] = Demapper(n, m, r, self.device)
            nc = bicm.cells_per_block(n, m)
            self.cells.append(torch.zeros((cnt, nc), dtype=torch.complex64, device=self.device))
            self.sigma2.append(torch.ones(cnt, dtype=torch.float64, device=self.device))
        self.nblk = len(self.specs)
        self.grouped = GroupedDemap(
            [(self.demappers[(n, m, r)], z, s2, self._llr_view(b0, cnt, n))
             for ((n, r, m), b0, cnt), z, s2 in zip(self.groups, self.cells, self.sigma2)],
            self.device)

    def _llr_view(self, b0, cnt, n):
        lo, hi = int(self.plan.off[b0]), int(self.plan.off[b0 + cnt])
        return self.plan.llr[lo:hi].view(cnt, n)

    def stage(self, grouped=True):
        """The demaps: fill the decoder's LLR buffer (enqueue only); one
        launch for every group, or (grouped False) one per group."""
        if grouped:
            self.grouped.run()
            return
        for ((n, r, m), b0, cnt), z, s2 in zip(self.groups, self.cells, self.sigma2):
            self.demappers[(n, m, r)].demap_dev(z, sigma2=s2, out=self._llr_view(b0, cnt, n),
                                                want_sigma2=False)

    def run(self):
        self.stage()
        self.plan.run()

    def capture(self):
        return self.plan.capture(pre=self.stage)

    def load(self, cells_per_group, sigma2_per_group):
        for dst, src in zip(self.cells, cells_per_group):
            dst.copy_(src)
        for dst, src in zip(self.sigma2, sigma2_per_group):
            dst.copy_(src)


class MixedSource:
    """Seeded frames for a MixedPipeline: per group ``ncw`` codewords on the
    device, AWGN at each block's Es/N0, noise from a seeded generator."""

    def __init__(self, pipe, ncw=8, seed=11):
        self.pipe = pipe
        self.streams = []
        for i, ((n, r, m), b0, cnt) in enumerate(pipe.groups):
            self.streams.append(synth.AwgnStream(r, n, m, ncw, seed + 101 * i, pipe.device))

    def frame(self, esn0_db, start=0):
        """esn0_db: per block of the pipeline.  Returns (cells per group,
        sigma2 per group, codeword index per group)."""
        cells, s2, idx = [], [], []
        for st, ((n, r, m), b0, cnt) in zip(self.streams, self.pipe.groups):
            z, i, n0 = st.batch(cnt, esn0_db[b0:b0 + cnt], start=start)
            cells.append(z)
            s2.append(n0)
            idx.append(i)
        return cells, s2, idx

    def errors(self, idx):
        """Block errors of the pipeline's last decode against the codewords."""
        plan = self.pipe.plan
        bits = plan.bits
        errs = torch.zeros(plan.nblk, dtype=torch.bool, device=bits.device)
        for st, i, ((n, r, m), b0, cnt) in zip(self.streams, idx, self.pipe.groups):
            lo, hi = int(plan.off[b0]), int(plan.off[b0 + cnt])
            ref = st.cw_dev.index_select(0, i)
            errs[b0:b0 + cnt] = (bits[lo:hi].view(cnt, n) != ref).any(1)
        return errs
