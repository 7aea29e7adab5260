"""Cells in front of the demapper: common-phase-error correction, the
dummy-cell SNR and the per-part noise weights.

CPU: the oracle against closed forms -- noiseless cells under a per-symbol
complex gain come back as the sent points with the gain recovered; a
dummy-only symbol's gain is the least-squares coefficient; the SNR equals
its definition evaluated by least squares; the weighted LLRs equal the FP64
demapper with the part's noise variance as a per-cell weight, the variances
recomputed by explicit loops.

GPU: against the oracle on a synthetic frame (a 64-NUC PLP, a QPSK PLP and
dummy cells; per-symbol rotation and gain; AWGN): corrected cells within
2e-6 of the FP64 values (complex64 output), gains within 1e-9 relative,
identical hard decisions; SNR to 1e-9 dB; weighted LLRs within 1e-5
normwise and part variances within 1e-5 relative (FP32 nearest-point
distances against FP64)."""
import numpy as np
import pytest

from oracle import cpe as OC
from oracle import demap as OD
from oracle import bch as OB
from atsc3_amd import bicm

P64 = bicm.points("64QAM", "11/15")
PQ = bicm.points("QPSK", "2/15")


def _frame(seed, esn0_db=None, nsym=12, per=1500, spare=700, n0_plp=None, nq=600, ndummy=400,
           rot=0.15):
    """A small pool: spare cells (symbol -1), then symbols; region 0 cells
    first, then region 1, then dummies; each symbol rotated and scaled."""
    rng = np.random.default_rng(seed)
    npool = spare + nsym * per
    symbol_of = np.concatenate([np.full(spare, -1), np.repeat(np.arange(nsym), per)])
    n0c = npool - nq - ndummy if n0_plp is None else n0_plp
    region = np.full(npool, 2, np.int8)
    region[:n0c] = 0
    region[n0c:n0c + nq] = 1
    dv = np.zeros(npool, np.complex128)
    dv[n0c + nq:] = 1.0 - 2.0 * OB.sequence(npool)[n0c + nq:]
    x = np.empty(npool, np.complex128)
    x[:n0c] = P64[rng.integers(64, size=n0c)]
    x[n0c:n0c + nq] = PQ[rng.integers(4, size=nq)]
    x[n0c + nq:] = dv[n0c + nq:]
    syms = np.unique(symbol_of)
    g = (1.0 + rng.uniform(-0.1, 0.1, len(syms))) * np.exp(1j * rng.uniform(-rot, rot, len(syms)))
    gain = g[np.searchsorted(syms, symbol_of)]
    y = gain * x
    if esn0_db is not None:
        n0 = 10 ** (-esn0_db / 10)
        y = y + (rng.standard_normal(npool) + 1j * rng.standard_normal(npool)) * np.sqrt(n0 / 2)
    return dict(pool=y.astype(np.complex64), x=x, symbol_of=symbol_of, region=region, dv=dv,
                g=g, syms=syms, dummy_start=n0c + nq)


def test_noiseless_gains_come_back():
    f = _frame(1)
    pool = f["pool"].astype(np.complex128)
    out, gains = OC.cpe_correct(pool, f["symbol_of"], [P64, PQ], f["region"], f["dv"])
    rms = np.sqrt(np.mean(np.abs(pool) ** 2))
    # float32 cells: the recovered points are exact to the input rounding
    assert np.abs(out - f["x"]).max() < 1e-6
    for s, g in zip(f["syms"], f["g"]):
        assert abs(gains[int(s)] - g / rms) < 1e-6 * abs(g / rms)


def test_decimated_estimate_uses_every_kth_pool_cell():
    """decim = k: the gain comes from the pool cells at indices divisible by
    k only -- perturbing every other cell leaves the gains unchanged."""
    f = _frame(6, 14.0)
    pool = f["pool"].astype(np.complex128)
    keep = np.arange(len(pool)) % 4 == 0
    noisy = pool.copy()
    noisy[~keep] *= np.exp(0.7j)
    s = np.sqrt(np.mean(np.abs(pool) ** 2) / np.mean(np.abs(noisy) ** 2))
    _, g1 = OC.cpe_correct(pool, f["symbol_of"], [P64, PQ], f["region"], f["dv"], decim=4)
    _, g2 = OC.cpe_correct(noisy * s, f["symbol_of"], [P64, PQ], f["region"], f["dv"], decim=4)
    for k in g1:
        assert abs(g1[k] - g2[k]) < 1e-9 * abs(g1[k])


def test_dummy_only_symbol_is_least_squares():
    rng = np.random.default_rng(2)
    n = 900
    dv = 1.0 - 2.0 * rng.integers(0, 2, n)
    z = (0.8 + 0.3j) * dv + 0.2 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    out, gains = OC.cpe_correct(z, np.zeros(n, int), [P64], np.full(n, 2), dv, iters=1)
    zn = z / np.sqrt(np.mean(np.abs(z) ** 2))
    c, *_ = np.linalg.lstsq(dv[:, None].astype(complex), zn, rcond=None)
    assert abs(gains[0] - c[0]) < 1e-12 and np.allclose(out, zn / c[0], rtol=0, atol=1e-12)


def test_nearest_is_brute_force():
    rng = np.random.default_rng(3)
    z = rng.standard_normal(500) + 1j * rng.standard_normal(500)
    want = np.array([P64[np.argmin([abs(v - p) for p in P64])] for v in z])
    assert np.array_equal(OC.nearest(z, P64), want)


def test_frame_snr_definition():
    rng = np.random.default_rng(4)
    n, a, s2 = 5000, 0.9, 0.02
    bits = rng.integers(0, 2, n + 100).astype(np.uint8)
    s = 1.0 - 2.0 * bits[100:]
    d = a * s + (rng.standard_normal(n) + 1j * rng.standard_normal(n)) * np.sqrt(s2 / 2)
    pool = np.concatenate([np.zeros(100, complex), d])
    got = OC.frame_snr_db(pool, 100, bits)
    ah, *_ = np.linalg.lstsq(s[:, None], d.real, rcond=None)
    noise = np.mean(np.abs(d - ah[0] * s) ** 2)
    assert abs(got - 10 * np.log10(ah[0] ** 2 / noise)) < 1e-9
    assert abs(got - 10 * np.log10(a * a / s2)) < 0.2
    assert OC.frame_snr_db(pool[:150], 100, bits) is None


def test_weighted_llr_is_demap_with_part_weights():
    rng = np.random.default_rng(5)
    cells = (P64[rng.integers(64, size=(3, 200))]
             + 0.15 * (rng.standard_normal((3, 200)) + 1j * rng.standard_normal((3, 200))))
    pid = rng.integers(0, 4, (3, 200))
    q, s2 = OC.weighted_llr(cells, P64, pid, 5)
    for p in range(5):
        vals = [min(abs(c - P64) ** 2) for c, k in zip(cells.ravel(), pid.ravel()) if k == p]
        want = max(sum(vals) / max(len(vals), 1), 1e-6)
        assert abs(s2[p] - want) < 1e-12 * max(want, 1.0)
    ref, _ = OD.demap_q(cells, P64, 1.0 / s2[pid], np.ones(3))
    assert np.allclose(q, ref, rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------------------
# device
# ---------------------------------------------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("esn0,rot,decim", [(None, 0.15, 1), (16.0, 0.15, 1), (12.0, 0.3, 1),
                                            (12.0, 0.3, 4)])
def test_cpe_matches_oracle(device, esn0, rot, decim):
    import torch
    from atsc3_amd.cpe import Cpe
    f = _frame(10 + int(esn0 or 0), esn0, rot=rot)
    want, wg = OC.cpe_correct(f["pool"], f["symbol_of"], [P64, PQ], f["region"], f["dv"],
                              decim=decim)
    cpe = Cpe(f["symbol_of"], f["region"], [P64, PQ], f["dv"], device)
    print(f"[gpu] cpe esn0={esn0} rot={rot} decim={decim}", flush=True)
    out, gains = cpe.run(torch.from_numpy(f["pool"]).to(device), decim=decim)
    torch.cuda.synchronize()
    out = out.cpu().numpy()
    gains = gains.cpu().numpy()
    assert np.abs(out - want).max() < 2e-6
    for k, s in enumerate(cpe.symbols):
        assert abs(gains[k] - wg[int(s)]) < 1e-9 * abs(wg[int(s)])
    # identical hard decisions after correction
    r0 = f["region"] == 0
    assert np.array_equal(OC.nearest(out[r0], P64), OC.nearest(want[r0], P64))


@pytest.mark.gpu
def test_cpe_exhaustive_and_folded_searches(device):
    """A 256-NUC region (searched through its 64 base points) and an
    alphabet without the quadrant symmetry (searched exhaustively)."""
    import torch
    from atsc3_amd.cpe import Cpe
    p256 = bicm.points("256QAM", "10/15")
    skew = P64 * np.exp(0.2j)                   # breaks the quadrant rule
    rng = np.random.default_rng(40)
    nsym, per = 8, 1200
    npool = nsym * per
    symbol_of = np.repeat(np.arange(nsym), per)
    region = np.where(rng.random(npool) < 0.5, 0, 1).astype(np.int8)
    x = np.where(region == 0, p256[rng.integers(256, size=npool)],
                 skew[rng.integers(64, size=npool)])
    g = np.exp(1j * rng.uniform(-0.05, 0.05, nsym))[symbol_of]
    y = (g * x + 0.02 * (rng.standard_normal(npool) + 1j * rng.standard_normal(npool)))
    y = y.astype(np.complex64)
    want, wg = OC.cpe_correct(y, symbol_of, [p256, skew], region, None)
    cpe = Cpe(symbol_of, region, [p256, skew], None, device)
    assert cpe.msizes == [-64, 64]
    print("[gpu] cpe folded/exhaustive", flush=True)
    out, gains = cpe.run(torch.from_numpy(y).to(device))
    torch.cuda.synchronize()
    assert np.abs(out.cpu().numpy() - want).max() < 2e-6
    gains = gains.cpu().numpy()
    for k, s_ in enumerate(cpe.symbols):
        assert abs(gains[k] - wg[int(s_)]) < 1e-9 * abs(wg[int(s_)])


@pytest.mark.gpu
def test_dummy_snr_matches_oracle(device):
    import torch
    from atsc3_amd.cpe import frame_snr_db
    f = _frame(20, 14.0)
    want = OC.frame_snr_db(f["pool"], f["dummy_start"], OB.sequence(len(f["pool"])))
    print("[gpu] dummy snr", flush=True)
    got = frame_snr_db(torch.from_numpy(f["pool"]).to(device), f["dummy_start"])
    assert abs(got - want) < 1e-9
    assert frame_snr_db(torch.from_numpy(f["pool"][:f["dummy_start"] + 10]).to(device),
                        f["dummy_start"]) is None


@pytest.mark.gpu
@pytest.mark.parametrize("gathered", [False, True])
def test_weighted_demap_matches_oracle(device, gathered):
    import torch
    from atsc3_amd.demap import Demapper
    from atsc3_amd.cpe import WeightedDemap
    from atsc3_amd.ti import hti_index
    dm = Demapper(16200, "64QAM", "11/15", device)
    nr = dm.ncell                          # 2700 cells per block
    nblk = 8
    f = _frame(30, 13.0, nsym=10, per=2200, spare=800, n0_plp=nblk * nr, nq=100, ndummy=100)
    pool = f["pool"]
    idx = hti_index(nblk * nr, 1, nr, nblk) if gathered else np.arange(nblk * nr).reshape(nblk, nr)
    cells = pool[idx]
    pid = f["symbol_of"][idx] + 1
    npart = int(f["symbol_of"].max()) + 2
    q, s2 = OC.weighted_llr(cells, P64, pid, npart)
    want = OD.to_codeword(q, dm.lam_of_q.cpu().numpy())
    wd = WeightedDemap(dm, pid, npart, idx if gathered else None, device)
    src = torch.from_numpy(pool if gathered else cells).to(device)
    print(f"[gpu] weighted demap gathered={gathered}", flush=True)
    llr, s2g = wd.run(src)
    torch.cuda.synchronize()
    assert np.abs(s2g.cpu().numpy() - s2).max() <= 1e-5 * s2.max()
    got = llr.cpu().numpy().astype(np.float64)
    assert OD.normwise_rel_err(got, want).max() < 1e-5


@pytest.mark.gpu
def test_empty_inputs(device):
    import torch
    from atsc3_amd.cpe import Cpe, frame_snr_db
    cpe = Cpe(np.zeros(0, int), np.zeros(0, np.int8), [P64], None, device)
    out, gains = cpe.run(torch.zeros(0, dtype=torch.complex64, device=device))
    assert out.numel() == 0 and gains.numel() == 0
    assert frame_snr_db(torch.zeros(10, dtype=torch.complex64, device=device), 5) is None
    from atsc3_amd.demap import Demapper
    from atsc3_amd.cpe import WeightedDemap
    dm = Demapper(16200, "64QAM", "11/15", device)
    wd = WeightedDemap(dm, np.zeros((0, dm.ncell), np.int32), None, None, device)
    llr, s2 = wd.run(torch.zeros((0, dm.ncell), dtype=torch.complex64, device=device))
    assert llr.shape == (0, 16200)
