"""Differentiable unrolled min-sum and soft demapper.

CPU: the float64 autograd reference (oracle/unrolled_ref.py) is pinned by
closed forms (alpha = 0 passes the channel LLRs through; a degree-2 check
forwards alpha * the other LLR), by the reference min-sum (hard decisions
after T iterations equal the plain decoder's bits with iters = T on blocks
that did not converge earlier), and its gradients by central finite
differences (step 1e-4): exact to O(h^2) where the +-h evaluations keep
every sign and argmin, close where only argmins switch (kinks of min); the
loss jumps where a sign of x flips, and those parameters are not compared.

GPU: the HIP forward in FP32 is bit-exact with the reference C min-sum in
the same sense; the HIP FP64 forward/backward equal the float64 autograd
reference to 1e-9 relative and its alpha gradient matches central finite
differences; the FP32 backward agrees with the float64 reference; the
soft demapper's gradients (cells, points, sigma2) equal autograd's.
"""
import numpy as np
import pytest
import torch

from tests import _fixtures as F

F.ensure_oracle_built()

from oracle import ldpc as OL             # noqa: E402
from oracle import unrolled_ref as R      # noqa: E402
from atsc3_amd import bicm, synth         # noqa: E402
from atsc3_amd.codes import get_code      # noqa: E402

RATE, N = "6/15", 16200


class _Csr:
    def __init__(self, checks, n):
        self.n = n
        cn = np.array([len(c) for c in checks])
        cptr = np.concatenate([[0], np.cumsum(cn)]).astype(np.int32)
        cidx = np.concatenate([np.sort(c) for c in checks]).astype(np.int32)
        vedge = np.argsort(cidx, kind="stable").astype(np.int32)
        vptr = np.concatenate([[0], np.cumsum(np.bincount(cidx, minlength=n))]).astype(np.int32)
        self._csr = (cptr, cidx, vptr, vedge)

    def csr(self):
        return self._csr


def _llr(nblk, snr, seed, rate=RATE, n=N):
    cw = synth.codewords(rate, n, nblk, seed)
    return cw, synth.llr_bpsk_like(cw, snr, seed + 1)


def test_reference_closed_forms():
    g = R.RefGraph(_Csr([[0, 1]], 2))
    lam = torch.tensor([[1.5, -0.25]], dtype=torch.float64)
    a = torch.tensor([[0.5], [0.5], [0.5]], dtype=torch.float64)
    tot = R.forward(lam, a, g)
    assert torch.allclose(tot, torch.tensor([[1.5 - 0.125, -0.25 + 0.75]], dtype=torch.float64))
    code = get_code(RATE, N)
    gr = R.RefGraph(code)
    _, llr = _llr(2, 1.0, 3)
    lam = torch.from_numpy(llr.astype(np.float64))
    z = torch.zeros((4, len(gr.classes)), dtype=torch.float64)
    assert torch.equal(R.forward(lam, z, gr), lam)


@pytest.mark.parametrize("T", [1, 2, 4])
def test_reference_matches_plain_minsum(T):
    """Below threshold nothing converges early: the plain decoder with
    iters = T returns sign(tot_T) of the same update rule."""
    code = get_code(RATE, N)
    gr = R.RefGraph(code)
    _, llr = _llr(6, -0.5, 11)
    bits, conv, its, bad = OL.minsum_f32(llr, code, T, 0.75)
    keep = its == T
    assert keep.sum() >= 4
    tot = R.forward(torch.from_numpy(llr.astype(np.float64)),
                    torch.full((T, len(gr.classes)), 0.75, dtype=torch.float64), gr)
    hard = (tot < 0).numpy().astype(np.uint8)
    mism = (hard[keep] != bits[keep]).mean()
    assert mism < 1e-4, mism          # FP64 vs FP32 rounding only


def _fd_alpha(fn, alphas, h, pattern_fn):
    """Central differences per parameter, plus for each whether the +-h
    evaluations kept every sign of x (no jump of the loss) and every argmin
    (no kink) of the unperturbed forward."""
    base = pattern_fn(alphas)
    fd = np.empty(alphas.numel())
    sign_ok = np.zeros(alphas.numel(), bool)
    smooth = np.zeros(alphas.numel(), bool)
    flat = alphas.reshape(-1)
    for i in range(flat.numel()):
        ap, am = flat.clone(), flat.clone()
        ap[i] += h
        am[i] -= h
        ap, am = ap.reshape(alphas.shape), am.reshape(alphas.shape)
        pp, pm = pattern_fn(ap), pattern_fn(am)
        sign_ok[i] = all(torch.equal(x[0], y[0]) for q in (pp, pm) for x, y in zip(q, base))
        smooth[i] = sign_ok[i] and R.pattern_equal(pp, base) and R.pattern_equal(pm, base)
        fd[i] = (fn(ap) - fn(am)) / (2 * h)
    return fd, sign_ok, smooth


def _check_fd(fd, sign_ok, smooth, grad):
    """Where no discrete choice changes within +-h the central difference
    equals the gradient (to the O(h^2) truncation); where only argmins
    change (kinks of min, the loss stays continuous) it stays close."""
    rel = np.abs(fd - grad) / np.maximum(np.abs(grad), 1e-12)
    assert smooth.sum() >= 2, "too few smooth parameters for the check"
    assert rel[smooth].max() < 1e-6, rel[smooth]
    assert sign_ok.sum() >= len(fd) // 2
    assert np.median(rel[sign_ok]) < 1e-5, rel[sign_ok]
    assert np.quantile(rel[sign_ok], 0.9) < 1e-2, rel[sign_ok]


def test_reference_gradients_vs_finite_differences():
    code = get_code(RATE, N)
    gr = R.RefGraph(code)
    cw, llr = _llr(1, 1.2, 21)
    lam = torch.from_numpy(llr.astype(np.float64))
    bits = torch.from_numpy(cw)
    T = 20
    alphas = torch.full((T, len(gr.classes)), 0.75, dtype=torch.float64)
    alphas += torch.linspace(-0.05, 0.05, alphas.numel(), dtype=torch.float64).reshape(alphas.shape)

    def loss(a, l=lam):
        tot = R.forward(l, a, gr)
        return torch.nn.functional.softplus(-(1 - 2 * bits.double()) * tot).mean()

    a = alphas.clone().requires_grad_(True)
    lt = lam.clone().requires_grad_(True)
    L = loss(a, lt)
    ga, gl = torch.autograd.grad(L, (a, lt))
    _check_fd(*_fd_alpha(lambda x: float(loss(x)), alphas, 1e-4,
                         lambda x: R.forward(lam, x, gr, record=True)[1]),
              ga.reshape(-1).numpy())
    # directional derivative in lam
    d = torch.randn(lam.shape, generator=torch.Generator().manual_seed(5), dtype=torch.float64)
    h = 1e-4
    pp = R.forward(lam + h * d, alphas, gr, record=True)[1]
    pm = R.forward(lam - h * d, alphas, gr, record=True)[1]
    p0 = R.forward(lam, alphas, gr, record=True)[1]
    if R.pattern_equal(pp, p0) and R.pattern_equal(pm, p0):
        num = (float(loss(alphas, lam + h * d)) - float(loss(alphas, lam - h * d))) / (2 * h)
        assert abs(num - float((gl * d).sum())) <= 1e-6 * max(1.0, abs(num))


def test_reference_soft_demap_gradients():
    rng = np.random.default_rng(4)
    for mod, rate in (("QPSK", "6/15"), ("16QAM", "8/15")):
        nb = bicm.MOD_BITS[mod]
        p = bicm.points(mod, rate)
        pts = torch.tensor(np.stack([p.real, p.imag], 1), dtype=torch.float64)
        z = torch.tensor(rng.normal(0, 0.7, (2, 5, 2)), dtype=torch.float64)
        s2 = torch.tensor([0.3, 0.5], dtype=torch.float64)
        lq = np.arange(5 * nb)
        w = torch.tensor(rng.normal(size=(2, 5 * nb)), dtype=torch.float64)
        f = lambda zz, pp, ss: float((R.soft_demap(zz, pp, ss, lq, nb) * w).sum())
        zz, pp, ss = (t.clone().requires_grad_(True) for t in (z, pts, s2))
        gz, gp, gs = torch.autograd.grad((R.soft_demap(zz, pp, ss, lq, nb) * w).sum(), (zz, pp, ss))
        h = 1e-6
        for t, g in ((z, gz), (pts, gp), (s2, gs)):
            for i in range(0, t.numel(), max(1, t.numel() // 7)):
                tp, tm = t.clone().reshape(-1), t.clone().reshape(-1)
                tp[i] += h
                tm[i] -= h
                args_p = [z, pts, s2]
                args_m = [z, pts, s2]
                k = [id(z), id(pts), id(s2)].index(id(t))
                args_p[k], args_m[k] = tp.reshape(t.shape), tm.reshape(t.shape)
                num = (f(*args_p) - f(*args_m)) / (2 * h)
                assert abs(num - float(g.reshape(-1)[i])) <= 1e-5 * max(1.0, abs(num))


# ---------------------------------------------------------------------------- GPU
@pytest.mark.gpu
@pytest.mark.parametrize("T", [1, 2, 3, 5])
def test_hip_forward_fp32_bit_exact_with_plain(device, T):
    from atsc3_amd.diff import UnrolledGraph
    code = get_code(RATE, N)
    g = UnrolledGraph(code, device)
    _, llr = _llr(24, 0.2, 31)
    print(f"[gpu] unrolled fp32 forward vs plain T={T}", flush=True)
    bits, conv, its, bad = OL.minsum_f32(llr, code, T, 0.75)
    keep = its == T
    assert keep.sum() >= 8
    tot, _ = g.forward(torch.from_numpy(llr).to(device),
                       torch.full((T, g.ncls), 0.75, dtype=torch.float32, device=device))
    hard = (tot < 0).to(torch.uint8).cpu().numpy()
    np.testing.assert_array_equal(hard[keep], bits[keep])


def _ref_grads(code, llr, bits, alphas64, dev):
    gr = R.RefGraph(code, dev)
    lam = torch.from_numpy(llr.astype(np.float64)).to(dev).requires_grad_(True)
    a = alphas64.clone().to(dev).requires_grad_(True)
    tot = R.forward(lam, a, gr)
    b = torch.from_numpy(bits).to(dev).double()
    L = torch.nn.functional.softplus(-(1 - 2 * b) * tot).mean()
    gl, ga = torch.autograd.grad(L, (lam, a))
    return tot.detach(), gl, ga


@pytest.mark.gpu
@pytest.mark.parametrize("K", [1, 3, 20])
def test_hip_fp64_matches_autograd_reference(device, K):
    from atsc3_amd.diff import UnrolledGraph, bce_on_llr, unrolled_minsum
    code = get_code(RATE, N)
    g = UnrolledGraph(code, device)
    cw, llr = _llr(16, 1.0, 41)
    T = 20
    a64 = 0.75 + 0.1 * torch.rand((T, g.ncls), generator=torch.Generator().manual_seed(2),
                                  dtype=torch.float64) - 0.05
    print(f"[gpu] unrolled fp64 vs autograd K={K}", flush=True)
    tot_r, gl_r, ga_r = _ref_grads(code, llr, cw, a64, device)
    lam = torch.from_numpy(llr.astype(np.float64)).to(device).requires_grad_(True)
    a = a64.clone().to(device).requires_grad_(True)
    tot = unrolled_minsum(lam, a, g, K)
    L = bce_on_llr(tot, torch.from_numpy(cw).to(device))
    L.backward()
    torch.testing.assert_close(tot.detach(), tot_r, rtol=1e-10, atol=1e-9)
    torch.testing.assert_close(a.grad, ga_r, rtol=1e-9, atol=1e-12)
    torch.testing.assert_close(lam.grad, gl_r, rtol=1e-9, atol=1e-12)


@pytest.mark.gpu
def test_hip_fp64_alpha_grad_vs_finite_differences(device):
    from atsc3_amd.diff import UnrolledGraph, bce_on_llr, unrolled_minsum
    code = get_code(RATE, N)
    g = UnrolledGraph(code, device)
    gr = R.RefGraph(code, device)
    cw, llr = _llr(1, 1.2, 51)
    T = 20
    lam = torch.from_numpy(llr.astype(np.float64)).to(device)
    bits = torch.from_numpy(cw).to(device)
    a0 = torch.full((T, g.ncls), 0.75, dtype=torch.float64, device=device)
    a0 += torch.linspace(-0.05, 0.05, a0.numel(), dtype=torch.float64, device=device).reshape(a0.shape)
    print("[gpu] unrolled fp64 alpha FD", flush=True)
    a = a0.clone().requires_grad_(True)
    bce_on_llr(unrolled_minsum(lam, a, g, 2), bits).backward()
    _check_fd(*_fd_alpha(lambda x: float(bce_on_llr(g.forward(lam, x)[0], bits)), a0, 1e-4,
                         lambda x: R.forward(lam, x, gr, record=True)[1]),
              a.grad.reshape(-1).cpu().numpy())


@pytest.mark.gpu
def test_hip_fp32_backward_vs_fp64_reference(device):
    from atsc3_amd.diff import UnrolledGraph, bce_on_llr, unrolled_minsum
    code = get_code(RATE, N)
    g = UnrolledGraph(code, device)
    cw, llr = _llr(64, 1.5, 61)
    T = 20
    a64 = torch.full((T, g.ncls), 0.75, dtype=torch.float64)
    print("[gpu] unrolled fp32 backward vs fp64 reference", flush=True)
    tot_r, gl_r, ga_r = _ref_grads(code, llr, cw, a64, device)
    lam = torch.from_numpy(llr).to(device).requires_grad_(True)
    a = a64.float().to(device).requires_grad_(True)
    tot = unrolled_minsum(lam, a, g, 5)
    bce_on_llr(tot, torch.from_numpy(cw).to(device)).backward()
    rel_a = (a.grad.double() - ga_r).abs().max() / ga_r.abs().max()
    rel_l = (lam.grad.double() - gl_r).norm() / gl_r.norm()
    assert rel_a < 5e-3, float(rel_a)
    assert rel_l < 5e-3, float(rel_l)


@pytest.mark.gpu
@pytest.mark.parametrize("mod,rate", [("QPSK", "6/15"), ("16QAM", "8/15"), ("256QAM", "10/15")])
def test_soft_demap_vs_autograd(device, mod, rate):
    from atsc3_amd.diff import SoftDemapper, soft_demap
    n = 16200
    dm = SoftDemapper(n, mod, rate, device)
    _, y, n0 = synth.awgn_blocks(rate, n, mod, 2, 8.0, 3)
    z = torch.view_as_real(torch.from_numpy(y)).to(device).contiguous()
    s2 = torch.tensor([n0, 1.3 * n0], dtype=torch.float64, device=device)
    pts = dm.points64.clone()
    w = torch.randn((2, n), generator=torch.Generator(device=device).manual_seed(1),
                    dtype=torch.float64, device=device)
    print(f"[gpu] soft demap grads {mod}", flush=True)
    zz, pp, ss = (t.clone().requires_grad_(True) for t in (z, pts, s2))
    (soft_demap(zz, pp, ss, dm) * w).sum().backward()
    zr, pr, sr = (t.clone().requires_grad_(True) for t in (z, pts, s2))
    ref = R.soft_demap(zr, pr, sr, dm.lam_of_q, dm.nb)
    (ref * w).sum().backward()
    torch.testing.assert_close(soft_demap(z, pts, s2, dm), ref.detach(), rtol=1e-12, atol=1e-9)
    torch.testing.assert_close(zz.grad, zr.grad, rtol=1e-9, atol=1e-9)
    torch.testing.assert_close(pp.grad, pr.grad, rtol=1e-9, atol=1e-7)
    torch.testing.assert_close(ss.grad, sr.grad, rtol=1e-9, atol=1e-7)
    # FP32 forward within 1e-5 (normwise) of the FP64 one
    l32 = soft_demap(z.float(), pts.float(), s2.float(), dm).double()
    l64 = ref.detach()
    assert ((l32 - l64).abs().amax(1) / l64.abs().amax(1)).max() < 1e-5


@pytest.mark.gpu
def test_training_reduces_loss(device):
    from atsc3_amd.learn import Trainer
    print("[gpu] alpha training smoke", flush=True)
    tr = Trainer(batch=64, iters=10, lr=0.02, seed=3, device=device, ncw=64)
    l0 = np.mean([tr.step() for _ in range(3)])
    tr.train(40)
    l1 = np.mean(tr.history[-3:])
    assert l1 < l0
    assert not torch.allclose(tr.alphas.detach(), torch.full_like(tr.alphas, 0.75))
    fer, err = tr.evaluate(2.0, 256, chunk=128)
    assert set(fer) == {"learned", "plain_unrolled", "plain_minsum", "plain_minsum_50",
                        "plain_ladder_50", "learned_fallback"}
    assert fer["learned_fallback"] <= fer["learned"]
    # the same blocks: more iterations of the plain decoder never add errors
    # here (a block converged by iters is decoded the same at 50)
    assert fer["plain_minsum_50"] <= fer["plain_minsum"]
