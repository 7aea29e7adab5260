"""Differentiable demap + decode: hand-written backward behind
torch.autograd.Function (HIP kernels in csrc/unrolled.hip).

  * ``unrolled_minsum(lam, alphas, graph, ckpt_every)`` -- T = len(alphas)
    flooding iterations of normalized min-sum with a learnable factor per
    (iteration, check-degree class); returns the posterior LLRs tot_T.  The
    check and variable updates are those of lab/ldpc_kernel.c:82-143 (x = tot
    - m, two-minimum with the first-minimum rule, sign product, ordered
    variable sums) without the syndrome stop.  The backward uses the
    subgradient of the min1/min2 selection (the gradient reaches the edge
    attaining min1 through every other edge's message and the edge attaining
    min2 through the min1 edge's message; signs are piecewise constant).
    Checkpointing: the forward keeps (tot_t, m_t) every ``ckpt_every``
    iterations; the backward recomputes each segment's per-check states
    (min1, min2, indices, sign bits) from its checkpoint and backpropagates
    through it.
  * ``soft_demap(z, points, sigma2, demapper)`` -- exhaustive max-log
    demapping (lab/m6_bicm.py:213-225 semantics for any constellation) with
    gradients for the cells and the constellation points.

Both run in float32 or float64 (the dtype of the inputs).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib, bicm


def _sfx(dtype):
    if dtype == torch.float32:
        return "f32"
    if dtype == torch.float64:
        return "f64"
    raise TypeError("float32 or float64 tensors only")


def _stream(dev):
    return torch.cuda.current_stream(dev).cuda_stream


class UnrolledGraph:
    """CSR tables of one parity-check matrix on a device, with the check
    degree classes that index the normalization factors."""

    def __init__(self, code, device="cuda"):
        self.device = torch.device(device)
        cptr, cidx, vptr, vedge = code.csr()
        deg = np.diff(cptr)
        if deg.max() > 32:
            raise ValueError("check degree above 32")
        self.classes = [int(d) for d in sorted(set(deg.tolist()))]
        cls = np.searchsorted(self.classes, deg).astype(np.int32)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(self.device)
        self.cptr, self.cidx, self.vptr, self.vedge, self.ccls = (
            t(cptr), t(cidx), t(vptr), t(vedge), t(cls))
        self.n, self.nc, self.E = int(code.n), len(cptr) - 1, len(cidx)
        onehot = np.zeros((self.nc, len(self.classes)))
        onehot[np.arange(self.nc), cls] = 1.0
        self.onehot = torch.from_numpy(onehot).to(self.device)
        self.code = code

    @property
    def ncls(self):
        return len(self.classes)

    # -- raw launches ---------------------------------------------------------
    def check_fwd(self, tot, m_in, m_out, alpha_t, state=None):
        B = tot.shape[0]
        st = (None, None, None) if state is None else (s.data_ptr() for s in state)
        _lib.call(f"atsc3_unrolled_check_fwd_{_sfx(tot.dtype)}", tot.data_ptr(),
                  None if m_in is None else m_in.data_ptr(), m_out.data_ptr(),
                  self.cptr.data_ptr(), self.cidx.data_ptr(), self.ccls.data_ptr(),
                  alpha_t.data_ptr(), B, self.n, self.nc, self.E, *st, _stream(tot.device))

    def var_fwd(self, lam, m, tot):
        _lib.call(f"atsc3_unrolled_var_fwd_{_sfx(lam.dtype)}", lam.data_ptr(), m.data_ptr(),
                  self.vptr.data_ptr(), self.vedge.data_ptr(), tot.data_ptr(), lam.shape[0],
                  self.n, self.E, _stream(lam.device))

    def new_state(self, B, dtype):
        d = self.device
        return (torch.empty((B, self.nc, 2), dtype=dtype, device=d),
                torch.empty((B, self.nc), dtype=torch.int32, device=d),
                torch.empty((B, self.nc), dtype=torch.int32, device=d))

    def forward(self, lam, alphas, record_from=None, keep_every=None):
        """Plain forward (no autograd).  Returns tot_T and, when keep_every is
        given, the checkpoints {t: (tot_t, m_t)}."""
        T = alphas.shape[0]
        B = lam.shape[0]
        tot, m = lam, None
        ck = {}
        for t in range(T):
            if keep_every and t % keep_every == 0:
                ck[t] = (tot, m)
            m_new = torch.empty((B, self.E), dtype=lam.dtype, device=lam.device)
            self.check_fwd(tot, m, m_new, alphas[t].contiguous())
            tot_new = torch.empty_like(lam)
            self.var_fwd(lam, m_new, tot_new)
            tot, m = tot_new, m_new
        return tot, ck


class _UnrolledMinSum(torch.autograd.Function):
    @staticmethod
    def forward(ctx, lam, alphas, graph, ckpt_every):
        lam = lam.contiguous()
        alphas = alphas.detach().to(lam.dtype).contiguous()
        tot, ck = graph.forward(lam, alphas, keep_every=ckpt_every)
        ctx.graph, ctx.ck, ctx.K = graph, ck, ckpt_every
        ctx.save_for_backward(lam, alphas)
        ctx.alpha_dtype = alphas.dtype
        return tot

    @staticmethod
    def backward(ctx, g_out):
        lam, alphas = ctx.saved_tensors
        g, ck, K = ctx.graph, ctx.ck, ctx.K
        T, B = alphas.shape[0], lam.shape[0]
        dt, dev = lam.dtype, lam.device
        sfx = _sfx(dt)
        G = g_out.contiguous().to(dt)
        carry = None
        g_lam = torch.zeros_like(lam)
        g_alpha = torch.zeros((T, g.ncls), dtype=torch.float64, device=dev)
        part = torch.empty((B, g.nc), dtype=dt, device=dev)
        for s0 in reversed(range(0, T, K)):
            s1 = min(s0 + K, T)
            tot, m = ck[s0]
            states = []
            for t in range(s0, s1):
                st = g.new_state(B, dt)
                m_new = torch.empty((B, g.E), dtype=dt, device=dev)
                g.check_fwd(tot, m, m_new, alphas[t], st)
                states.append(st)
                if t + 1 < s1:
                    tot_new = torch.empty_like(lam)
                    g.var_fwd(lam, m_new, tot_new)
                    tot, m = tot_new, m_new
            for t in reversed(range(s0, s1)):
                st = states[t - s0]
                gx = torch.empty((B, g.E), dtype=dt, device=dev)
                _lib.call(f"atsc3_unrolled_check_bwd_{sfx}", G.data_ptr(),
                          None if carry is None else carry.data_ptr(), g.cptr.data_ptr(),
                          g.cidx.data_ptr(), g.ccls.data_ptr(), alphas[t].data_ptr(),
                          st[0].data_ptr(), st[1].data_ptr(), st[2].data_ptr(), B, g.n, g.nc,
                          g.E, gx.data_ptr(), part.data_ptr(), _stream(dev))
                g_alpha[t] = part.to(torch.float64).sum(0) @ g.onehot
                G_new = torch.empty_like(lam)
                _lib.call(f"atsc3_unrolled_var_bwd_{sfx}", gx.data_ptr(), G.data_ptr(),
                          g.vptr.data_ptr(), g.vedge.data_ptr(), G_new.data_ptr(),
                          g_lam.data_ptr(), B, g.n, g.E, _stream(dev))
                G, carry = G_new, gx
        g_lam += G
        return g_lam, g_alpha.to(ctx.alpha_dtype), None, None


def unrolled_minsum(lam, alphas, graph, ckpt_every=5):
    """lam (B, n) device LLRs (float32/float64), alphas (T, graph.ncls) ->
    tot_T (B, n).  Differentiable in lam and alphas."""
    if not (lam.is_cuda and lam.dim() == 2 and lam.shape[1] == graph.n):
        raise ValueError("lam must be a (B, n) device tensor")
    if alphas.dim() != 2 or alphas.shape[1] != graph.ncls:
        raise ValueError(f"alphas must be (T, {graph.ncls})")
    return _UnrolledMinSum.apply(lam, alphas, graph, max(1, int(ckpt_every)))


class SoftDemapper:
    """Exhaustive max-log demap of one (N, constellation, rate) with the
    de-interleave to codeword order."""

    def __init__(self, n, mod, rate, device="cuda"):
        self.device = torch.device(device)
        self.n, self.mod, self.rate = n, mod, rate
        self.nb = bicm.MOD_BITS[mod]
        self.M = 1 << self.nb
        self.ncell = bicm.cells_per_block(n, mod)
        lq = bicm.bicm_map(n, mod, rate)
        self.lam_of_q = torch.from_numpy(lq.astype(np.int32)).to(self.device)
        p = bicm.points(mod, rate)
        self.points64 = torch.from_numpy(np.stack([p.real, p.imag], 1)).to(self.device)


class _SoftDemap(torch.autograd.Function):
    @staticmethod
    def forward(ctx, z, pts, sigma2, dm):
        z, pts, sigma2 = z.contiguous(), pts.contiguous(), sigma2.contiguous()
        B, C = z.shape[0], z.shape[1]
        out = torch.empty((B, C * dm.nb), dtype=z.dtype, device=z.device)
        _lib.call(f"atsc3_soft_demap_fwd_{_sfx(z.dtype)}", z.data_ptr(), None, sigma2.data_ptr(),
                  pts.data_ptr(), dm.M, dm.nb, dm.lam_of_q.data_ptr(), B, C, out.data_ptr(),
                  _stream(z.device))
        ctx.save_for_backward(z, pts, sigma2)
        ctx.dm = dm
        return out

    @staticmethod
    def backward(ctx, g_llr):
        z, pts, sigma2 = ctx.saved_tensors
        dm = ctx.dm
        B, C = z.shape[0], z.shape[1]
        grid = _lib.lib().atsc3_soft_demap_bwd_grid(B, C)
        gz = torch.empty_like(z)
        gp = torch.empty((grid, dm.M, 2), dtype=z.dtype, device=z.device)
        gs = torch.zeros((B, grid), dtype=z.dtype, device=z.device)
        _lib.call(f"atsc3_soft_demap_bwd_{_sfx(z.dtype)}", z.data_ptr(), None, sigma2.data_ptr(),
                  pts.data_ptr(), dm.M, dm.nb, dm.lam_of_q.data_ptr(), B, C,
                  g_llr.contiguous().to(z.dtype).data_ptr(), gz.data_ptr(), gp.data_ptr(),
                  gs.data_ptr(), _stream(z.device))
        return gz, gp.sum(0), gs.sum(1), None


def soft_demap(z, points, sigma2, dm):
    """z (B, ncell, 2) real view of the cells, points (M, 2), sigma2 (B,) ->
    codeword-order LLRs (B, n).  Differentiable in z, points and sigma2."""
    if z.dim() != 3 or z.shape[1] != dm.ncell or z.shape[2] != 2:
        raise ValueError(f"z must be (B, {dm.ncell}, 2)")
    if points.shape != (dm.M, 2):
        raise ValueError(f"points must be ({dm.M}, 2)")
    return _SoftDemap.apply(z, points.to(z.dtype), sigma2.to(z.dtype), dm)


def bce_on_llr(tot, bits):
    """Mean binary cross-entropy of the bits under LLRs (llr > 0 -> bit 0):
    softplus(-(1 - 2 b) tot)."""
    sgn = 1.0 - 2.0 * bits.to(tot.dtype)
    return torch.nn.functional.softplus(-sgn * tot).mean()
