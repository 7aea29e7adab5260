"""GPU demapper parity against the FP64 oracle: normwise relative error
max|got - want| / max|want| <= 1e-5 per block (the criterion of
lab/m16_demap_fast.py:136-139), sigma2 estimate within 1e-6 relative."""
import numpy as np
import pytest

from oracle import demap as OD
from atsc3_amd import bicm, synth
from tests._fixtures import ref_lam_of_q

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.mark.parametrize("n,mod,rate,esn0", [
    (64800, "256QAM", "10/15", 20.0), (64800, "256QAM", "6/15", 14.0),
    (64800, "256QAM", "8/15", 17.0), (64800, "64QAM", "10/15", 15.0),
    (64800, "16QAM", "8/15", 10.0), (64800, "QPSK", "4/15", 0.0),
    (16200, "64QAM", "9/15", 12.0), (16200, "QPSK", "9/15", 2.0),
    (16200, "256QAM", "4/15", 8.0)])
@pytest.mark.parametrize("given_sigma", [True, False])
def test_demap_vs_fp64_oracle(device, n, mod, rate, esn0, given_sigma):
    import torch
    from atsc3_amd.demap import Demapper
    print(f"[gpu] demap {n} {mod} {rate} sigma_given={given_sigma}", flush=True)
    cw, y, n0 = synth.awgn_blocks(rate, n, mod, 3, esn0, 31)
    y = y.astype(np.complex64).astype(np.complex128)      # the kernel's input precision
    want, s2w = OD.demap(y, bicm.points(mod, rate), ref_lam_of_q(n, mod, rate),
                         sigma2=n0 if given_sigma else None)
    dm = Demapper(n, mod, rate, device)
    z = torch.from_numpy(y.astype(np.complex64)).to(device)
    llr, s2 = dm.demap_dev(z, sigma2=n0 if given_sigma else None)
    torch.cuda.synchronize()
    got = llr.cpu().numpy().astype(np.float64)
    err = OD.normwise_rel_err(got, want)
    assert (err <= TOL).all(), err
    assert np.allclose(s2.cpu().numpy(), s2w, rtol=1e-6, atol=0)


def test_demap_with_csi_rayleigh(device):
    import torch
    from atsc3_amd.demap import Demapper
    n, mod, rate = 64800, "256QAM", "10/15"
    cw, z, w, n0 = synth.rayleigh_blocks(rate, n, mod, 2, 20.0, 5)
    z = z.astype(np.complex64).astype(np.complex128)
    w = w.astype(np.float32).astype(np.float64)
    for s2 in (n0, None):
        want, s2w = OD.demap(z, bicm.points(mod, rate), ref_lam_of_q(n, mod, rate), csi=w, sigma2=s2)
        dm = Demapper(n, mod, rate, device)
        llr, s2g = dm.demap_dev(torch.from_numpy(z.astype(np.complex64)).to(device),
                                csi=torch.from_numpy(w.astype(np.float32)).to(device), sigma2=s2)
        torch.cuda.synchronize()
        err = OD.normwise_rel_err(llr.cpu().numpy().astype(np.float64), want)
        assert (err <= TOL).all(), err
        assert np.allclose(s2g.cpu().numpy(), s2w, rtol=1e-6, atol=0)


def test_demap_reference_signatures(device):
    """The upstream's demap calls, made exactly as the upstream makes them:
    lab/m9_fast.py:1117 ``q = self.demap_batch(cells, self.pts0, self.ones0)``,
    lab/m9_fast.py:1105 with ``d2min_out`` (sigma^2 off, nearest distances
    out), lab/m44_ldm.py:2359 ``demap_batch_qam(cells, ch, self.ex,
    self.fec_threads)``, lab/m6_bicm.py:213 ``demap_llr``."""
    from atsc3_amd import demap as D
    n, mod, rate = 16200, "64QAM", "11/15"
    _, y, n0 = synth.awgn_blocks(rate, n, mod, 3, 16.0, 8)
    y = y.astype(np.complex64).astype(np.complex128)
    pts0, ones0 = bicm.points(mod, rate), bicm.bit_masks(6)
    want, _ = OD.demap_q(y, pts0)
    q = D.demap_batch(y, pts0, ones0)
    assert q.shape == (3, n) and q.dtype == np.float32
    assert (OD.normwise_rel_err(q.astype(np.float64), want) <= TOL).all()
    d2 = np.empty((3, n // 6), np.float32)
    raw = D.demap_batch(y, pts0, ones0, d2min_out=d2)
    want_raw, _ = OD.demap_q(y, pts0, sigma2=np.ones(3))
    assert (OD.normwise_rel_err(raw.astype(np.float64), want_raw) <= TOL).all()
    d2w = (np.abs(y[:, :, None] - pts0[None, None, :]) ** 2).min(axis=2)
    assert np.allclose(d2, d2w, rtol=1e-5, atol=1e-6)

    class Chain:                       # what the upstream's PlpChain exposes
        def __init__(self, mod, rate):
            self.mod, self.rate, self.mod_bits = mod, rate, bicm.MOD_BITS[mod]
    for m2, r2, nn in (("QPSK", "4/15", 64800), ("64QAM", "10/15", 64800)):
        _, y2, _ = synth.awgn_blocks(r2, nn, m2, 2, 12.0, 9)
        y2 = y2.astype(np.complex64).astype(np.complex128)
        q2 = D.demap_batch_qam(y2, Chain(m2, r2), None, 4)
        w2, _ = OD.demap_q(y2, bicm.points(m2, r2))
        assert (OD.normwise_rel_err(q2.astype(np.float64), w2) <= TOL).all()
    ql = D.demap_llr(y[0], mod, rate)
    assert OD.normwise_rel_err(ql[None].astype(np.float64), want[:1])[0] <= TOL
    with pytest.raises(ValueError):
        D.demap_batch(y, pts0, ~ones0)


def test_demap_noiseless_decodes(device):
    """Noiseless cells -> LLR signs are the codeword bits in codeword order."""
    import torch
    from atsc3_amd.demap import Demapper
    n, mod, rate = 64800, "256QAM", "10/15"
    cw = synth.codewords(rate, n, 2, 3)
    x = synth.modulate(cw, n, mod, rate)
    dm = Demapper(n, mod, rate, device)
    llr, _ = dm.demap_dev(torch.from_numpy(x.astype(np.complex64)).to(device), sigma2=0.01)
    got = (llr.cpu().numpy() < 0).astype(np.uint8)
    assert np.array_equal(got, cw)


def test_grouped_demap_equals_per_group(device):
    """GroupedDemap (every (N, constellation, rate) group of a frame in one
    launch) writes bit-for-bit the LLRs of one Demapper.demap_dev per group
    -- the same arithmetic -- with and without CSI weights, over the 24
    (code, constellation) pairs of the uneven frames; and within 1e-5
    normwise of the FP64 oracle."""
    import torch
    from atsc3_amd.demap import Demapper, GroupedDemap
    from atsc3_amd.mixed import CODES, MODS
    rng = np.random.default_rng(4)
    groups, ref, want = [], [], []
    for i, ((n, rate), mod) in enumerate([(c, m) for c in CODES for m in MODS]):
        nblk = 1 + i % 3
        cw, y, n0 = synth.awgn_blocks(rate, n, mod, nblk, 12.0, 100 + i)
        y = y.astype(np.complex64)
        dm = Demapper(n, mod, rate, device)
        z = torch.from_numpy(y).to(device)
        s2 = torch.from_numpy(np.full(nblk, n0, np.float64)).to(device)
        csi = None
        if i % 2:
            csi = torch.from_numpy(rng.uniform(0.2, 2.0, y.shape).astype(np.float32)).to(device)
        out = torch.full((nblk, n), np.nan, dtype=torch.float32, device=device)
        groups.append((dm, z, s2, out) + ((csi,) if csi is not None else ()))
        r, _ = dm.demap_dev(z, csi=csi, sigma2=s2)
        ref.append(r)
        w, _ = OD.demap(y.astype(np.complex128), bicm.points(mod, rate),
                        ref_lam_of_q(n, mod, rate), sigma2=np.full(nblk, n0),
                        csi=None if csi is None else csi.cpu().numpy().astype(np.float64))
        want.append(w)
    gd = GroupedDemap(groups, device)
    print(f"[gpu] grouped demap {gd.ngroups} groups {gd.nblk} blocks", flush=True)
    gd.run()
    torch.cuda.synchronize()
    for g, r, w in zip(groups, ref, want):
        assert torch.equal(g[3], r)
        err = OD.normwise_rel_err(g[3].cpu().numpy().astype(np.float64), w)
        assert (err <= TOL).all(), err
