"""Float64 torch-autograd reference of the unrolled normalized min-sum and of
the exhaustive max-log demapper.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): the checker for the hand-
written backward of atsc3_amd.diff.  It is written with plain tensor ops
(gather, topk, index_add) so torch's autograd supplies the reference
gradient; the min1/min2 selection is differentiated through ``topk``.

Update rule (lab/ldpc_kernel.c:82-143 without the syndrome stop):
x = tot[v(e)] - m[e]; m' = alpha[t][cls(c)] * prod_{e'!=e} sign(x_e') *
min_{e'!=e} |x_e'|; tot' = lam + sum_{e in v} m'.

``pattern`` returns the discrete choices of a forward (signs of every x,
argmin indices), so a finite-difference check can tell whether its +-h
evaluations stayed on one smooth piece.

Pinning: tests/test_unrolled.py checks the forward against the reference
min-sum (oracle/minsum_oracle.c): with every alpha equal to the plain
alpha, hard decisions after T iterations equal the plain decoder's bits
with iters = T on blocks that did not converge earlier; a degree-2 check
passes x_other * alpha exactly; alpha = 0 gives tot = lam.  The gradients
are checked against central finite differences.
"""
from __future__ import annotations

import numpy as np
import torch


class RefGraph:
    def __init__(self, code, device="cpu"):
        cptr, cidx, vptr, vedge = code.csr()
        deg = np.diff(cptr)
        dmax = int(deg.max())
        nc = len(deg)
        self.classes = sorted(set(deg.tolist()))
        pad = np.zeros((nc, dmax), np.int64)
        mask = np.zeros((nc, dmax), bool)
        eid = np.full((nc, dmax), -1, np.int64)
        for c in range(nc):
            d = deg[c]
            pad[c, :d] = cidx[cptr[c]:cptr[c + 1]]
            mask[c, :d] = True
            eid[c, :d] = np.arange(cptr[c], cptr[c + 1])
        dev = torch.device(device)
        self.pad = torch.from_numpy(pad).to(dev)
        self.mask = torch.from_numpy(mask).to(dev)
        self.cls = torch.from_numpy(np.searchsorted(self.classes, deg)).to(dev)
        self.n, self.nc, self.dmax = int(code.n), nc, dmax
        self.device = dev
        # variable-major order of the flattened (check, slot) entries,
        # ascending check index per variable: the summation order
        flat_v = pad[mask]
        flat_pos = np.flatnonzero(mask.ravel())
        order = np.lexsort((np.arange(len(flat_v)), flat_v))
        self.sum_pos = torch.from_numpy(flat_pos[order]).to(dev)
        self.sum_var = torch.from_numpy(flat_v[order]).to(dev)


def forward(lam, alphas, g, record=False):
    """lam (B, n) float64, alphas (T, ncls) -> tot_T (and the pattern)."""
    B = lam.shape[0]
    T = alphas.shape[0]
    tot = lam
    m = torch.zeros((B, g.nc, g.dmax), dtype=lam.dtype, device=lam.device)
    mask = g.mask[None]
    pats = []
    ar = torch.arange(g.dmax, device=lam.device)
    for t in range(T):
        x = tot[:, g.pad] - m
        x = torch.where(mask, x, torch.ones_like(x))
        neg = (x < 0) & mask
        sgn = torch.where(neg, -torch.ones_like(x), torch.ones_like(x))
        mag = torch.where(mask, x.abs(), torch.full_like(x, float("inf")))
        vals, idx = torch.topk(mag, 2, dim=-1, largest=False, sorted=True)
        m1, m2, i1 = vals[..., 0:1], vals[..., 1:2], idx[..., 0:1]
        mu = torch.where(ar[None, None, :] == i1, m2, m1)
        prod = sgn.prod(-1, keepdim=True)
        a = alphas[t][g.cls][None, :, None]
        m = torch.where(mask, a * (prod * sgn) * mu, torch.zeros_like(x))
        flat = m.reshape(B, -1)[:, g.sum_pos]
        s = torch.zeros_like(lam)
        # sequential-order sum is not needed for the gradient reference; an
        # index_add gives the same value up to rounding
        s = s.index_add(1, g.sum_var, flat)
        tot = lam + s
        if record:
            pats.append((neg.detach().clone(), idx.detach().clone()))
    return (tot, pats) if record else tot


def pattern_equal(p, q):
    return all(torch.equal(a[0], b[0]) and torch.equal(a[1], b[1]) for a, b in zip(p, q))


def soft_demap(z, pts, sigma2, lam_of_q, nb):
    """z (B, C, 2), pts (M, 2), sigma2 (B,) -> codeword-order LLRs (B, C*nb)."""
    B, C, _ = z.shape
    M = pts.shape[0]
    d = ((z[:, :, None, :] - pts[None, None, :, :]) ** 2).sum(-1)        # (B, C, M)
    lab = torch.arange(M, device=z.device)
    out = []
    for i in range(nb):
        one = ((lab >> (nb - 1 - i)) & 1).bool()
        d1 = d[..., one].min(-1).values
        d0 = d[..., ~one].min(-1).values
        out.append((d1 - d0) / sigma2[:, None])
    q = torch.stack(out, -1).reshape(B, C * nb)
    res = torch.empty_like(q)
    res[:, torch.as_tensor(lam_of_q, device=z.device).long()] = q
    return res
