"""Pinning the CPU oracle (runs without a GPU)."""
import os

import numpy as np
import pytest

from tests import _fixtures as F

F.ensure_oracle_built()

from oracle import demap as OD        # noqa: E402
from oracle import ldpc as OL         # noqa: E402
from atsc3_amd import bicm, synth     # noqa: E402
from atsc3_amd.codes import get_code  # noqa: E402


@pytest.mark.parametrize("name", F.fixture_names())
def test_oracle_matches_reference_fixtures(name):
    """Fixtures were written by the upstream kernel compiled from its own
    source (tests/golden/make_golden.py)."""
    code, llr, iters, alpha, want = F.load(name)
    F.assert_same(OL.minsum_f32(llr, code, iters, alpha), want, name)


@pytest.mark.skipif(not OL.available("ref"), reason="upstream kernel not built here")
@pytest.mark.parametrize("rate,n,snr,alpha,iters", [
    ("6/15", 16200, 0.3, 0.75, 40), ("13/15", 16200, 5.0, 1.0, 20),
    ("5/15", 16200, -1.0, 0.85, 30), ("7/15", 64800, 0.9, 0.75, 12)])
def test_oracle_matches_reference_live(rate, n, snr, alpha, iters):
    code = get_code(rate, n)
    cw = synth.codewords(rate, n, 2, 77)
    llr = synth.llr_bpsk_like(cw, snr, 78)
    a = OL.minsum_f32(llr, code, iters, alpha, "oracle")
    b = OL.minsum_f32(llr, code, iters, alpha, "ref")
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def _tiny_code(rng, n=30, nc=15, dmin=2, dmax=6):
    checks = []
    for _ in range(nc):
        d = int(rng.integers(dmin, dmax + 1))
        checks.append(sorted(rng.choice(n, d, replace=False).tolist()))
    cover = set(v for c in checks for v in c)
    for v in range(n):                      # every column covered
        if v not in cover:
            checks[int(rng.integers(0, nc))].append(v)
    checks = [sorted(set(c)) for c in checks]
    cn = np.array([len(c) for c in checks], np.int32)
    cptr = np.concatenate([[0], np.cumsum(cn)]).astype(np.int32)
    cidx = np.array([v for c in checks for v in c], np.int32)
    vedge = np.argsort(cidx, kind="stable").astype(np.int32)
    vptr = np.concatenate([[0], np.cumsum(np.bincount(cidx, minlength=n))]).astype(np.int32)
    return checks, (cptr, cidx, vptr, vedge)


@pytest.mark.parametrize("seed", range(12))
def test_oracle_vs_scalar_restatement_tiny(seed):
    """Third route: pure-Python binary32 restatement, tiny random codes,
    coarse LLR grids (many magnitude ties) and tiny ones."""
    rng = np.random.default_rng(seed)
    n = 30
    checks, csr = _tiny_code(rng, n=n)
    llr = (rng.integers(-6, 7, (4, n)) * 0.5).astype(np.float32)
    llr[1] = rng.standard_normal(n).astype(np.float32)
    for alpha in (0.75, 1.0):
        got = OL.minsum_f32_csr(llr, n, *csr, iters=10, alpha=alpha)
        for b in range(len(llr)):
            bits, conv, its, bad = OL.scalar_minsum(llr[b], checks, n, 10, alpha)
            assert np.array_equal(got[0][b], bits)
            assert bool(got[1][b]) == conv and got[2][b] == its and got[3][b] == bad


@pytest.fixture(scope="module")
def planted():
    import subprocess
    subprocess.run(["make", "-s", "-C", os.path.join(F.ROOT, "oracle"), "plant"],
                   check=True, capture_output=True)
    return True


@pytest.mark.parametrize("which", ["plant1", "plant2"])
def test_planted_mistakes_are_caught(planted, which):
    """Negative controls: an oracle built with the variable sum in DESCENDING
    check order (plant1), or with one update before the first syndrome check
    (plant2), must disagree with the upstream fixtures.  The first-minimum
    rule itself needs no control: when two edges tie at min1, min2 == min1,
    so which of them receives min2 never changes a message value."""
    differs = 0
    for name in ("s16k_r9_near", "s64k_r10_nuc", "g16k_r9", "g16k_r4", "s16k_r3_smoke"):
        code, llr, iters, alpha, want = F.load(name)
        got = OL.minsum_f32(llr, code, iters, alpha, which)
        try:
            F.assert_same(got, want)
        except AssertionError:
            differs += 1
    assert differs >= 2


def test_smoke_case_all_positive():
    """lab/build_ldpc_kernel.py:101-120: all +5.0 on 16200 r3/15, alpha 1.0
    -> conv, its 0, bad 0, zero bits."""
    code = get_code("3/15", 16200)
    bits, conv, its, bad = OL.minsum_f32(np.full((2, 16200), 5.0, np.float32), code, 50, 1.0)
    assert conv.all() and (its == 0).all() and (bad == 0).all() and not bits.any()


def test_invariants_noiseless_and_converged():
    code = get_code("10/15", 64800)
    cw = synth.codewords("10/15", 64800, 2, 5)
    llr = ((1 - 2 * cw.astype(np.float32)) * 3.0).astype(np.float32)
    bits, conv, its, bad = OL.minsum_f32(llr, code, 50, 0.75)
    assert conv.all() and (its == 0).all() and np.array_equal(bits, cw)
    noisy = synth.llr_bpsk_like(cw, 2.0, 6)
    bits, conv, its, bad = OL.minsum_f32(noisy, code, 50, 0.75)
    syn = code.syndrome_count(bits)
    assert np.array_equal(syn == 0, conv) and (bad[conv] == 0).all()
    assert (bad[~conv] == syn[~conv]).all()


def test_ladder_semantics():
    code = get_code("6/15", 64800)
    cw = synth.codewords("6/15", 64800, 3, 204)
    llr = synth.llr_bpsk_like(cw, 0.8, 205)
    bits, conv, its, bad = OL.ladder(llr, code, (1.0, 0.85, 0.75), 50)
    r1 = OL.minsum_f32(llr, code, 50, 1.0)
    for b in range(3):
        if r1[1][b]:
            assert conv[b] and its[b] == r1[2][b]
    assert (bad[conv] == 0).all()


# ---------------------------------------------------------------------------
# demap oracle
# ---------------------------------------------------------------------------

def test_demap_qpsk_closed_form():
    rng = np.random.default_rng(1)
    z = rng.standard_normal((2, 500)) + 1j * rng.standard_normal((2, 500))
    s2 = np.array([0.3, 1.7])
    q, _ = OD.demap_q(z, bicm.points("QPSK", "2/15"), sigma2=s2)
    q = q.reshape(2, 500, 2)
    r2 = 2 * np.sqrt(2)
    assert np.allclose(q[:, :, 0], r2 * z.imag / s2[:, None], rtol=1e-12, atol=1e-12)
    assert np.allclose(q[:, :, 1], r2 * z.real / s2[:, None], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("mod,rate", [("16QAM", "8/15"), ("64QAM", "10/15"),
                                      ("256QAM", "10/15"), ("256QAM", "8/15")])
def test_demap_bruteforce_cells(mod, rate):
    pts = bicm.points(mod, rate)
    nb = int(np.log2(len(pts)))
    rng = np.random.default_rng(2)
    z = pts[rng.integers(0, len(pts), 20)] + 0.1 * (rng.standard_normal(20) + 1j * rng.standard_normal(20))
    w = rng.uniform(0.2, 3.0, 20)
    q, s2 = OD.demap_q(z[None], pts, csi=w[None])
    d2min = []
    for c in range(20):
        d = [abs(z[c] - p) ** 2 for p in pts]
        d2min.append(min(d))
        for i in range(nb):
            one = min(d[k] for k in range(len(pts)) if (k >> (nb - 1 - i)) & 1)
            zero = min(d[k] for k in range(len(pts)) if not (k >> (nb - 1 - i)) & 1)
            assert np.isclose(q[0, c * nb + i], w[c] * (one - zero) / s2[0], rtol=1e-12)
    assert np.isclose(s2[0], np.mean(w * np.array(d2min)), rtol=1e-12)


def test_demap_noiseless_sign_rule():
    for mod, rate in (("QPSK", "4/15"), ("64QAM", "10/15"), ("256QAM", "10/15")):
        pts = bicm.points(mod, rate)
        nb = int(np.log2(len(pts)))
        k = np.arange(len(pts))
        q, _ = OD.demap_q(pts[None], pts, sigma2=1.0)
        bits = ((k[:, None] >> (nb - 1 - np.arange(nb))[None, :]) & 1).ravel()
        # llr > 0 means bit 0 (points of 256-NUC 8/15 coincide: skip zeros)
        nz = q[0] != 0
        assert np.array_equal((q[0][nz] < 0).astype(int), bits[nz])
        assert nz.mean() > 0.9
