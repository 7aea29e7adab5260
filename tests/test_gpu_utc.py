"""GPU update-then-check decoder (csrc/ldpc_utc.hip) against the upstream's
own decoders: GpuMinSum.decode at float32 and float64 (lab/m9_gpu.py:84-211)
and min_sum_decode / min_sum_decode_batch at float64 (lab/m3_ldpc.py:245-286,
509-630), bit-exact on bits, conv, its and bad.

Fixtures: tests/golden/utc_*.npz (the upstream's outputs on golden LLR sets,
scaled sets past the 1e9 row padding, float64-native sets, hand-made graphs
where the association of the variable sum decides a bit, and sets holding
NaN and +-inf LLRs, on which GpuMinSum and min_sum_decode differ); live cases
against the oracle that those fixtures pin (oracle/utc_ref.py), under both
NaN rules.
"""
import numpy as np
import pytest

from atsc3_amd import synth
from atsc3_amd.codes import get_code
from oracle.utc_ref import utc_minsum
from tests.test_utc_oracle import NAMES, load_utc

pytestmark = pytest.mark.gpu


def _eq(got, want, tag):
    for g, w, nm in zip(got, want, ("bits", "conv", "its", "bad")):
        assert np.array_equal(np.asarray(g), np.asarray(w)), f"{tag}: {nm} differs"


@pytest.mark.parametrize("name", NAMES)
def test_utc_fixtures(device, name):
    from atsc3_amd import ldpc
    code, llr, iters, alpha, want = load_utc(name)
    print(f"[gpu] utc fixture {name} n={code.n} iters={iters}", flush=True)
    checks = [list(c) for c in code.checks]
    g32 = ldpc.GpuMinSum(checks, code.n, iters=iters, alpha=alpha, device=device,
                         dtype="float32")
    _eq(g32.decode(llr), want["f32"], f"{name} GpuMinSum float32")
    g64 = ldpc.GpuMinSum(checks, code.n, iters=iters, alpha=alpha, device=device)
    assert g64.dtype_name == "float64"                 # the upstream default
    _eq(g64.decode(llr), want["t64"], f"{name} GpuMinSum float64")
    _eq(ldpc.min_sum_decode_batch(llr, checks, code.n, iters, alpha), want["f64"],
        f"{name} min_sum_decode_batch float64")
    b, c, i, d = ldpc.min_sum_decode(llr[0], checks, code.n, iters, alpha)
    assert isinstance(c, bool) and isinstance(i, int) and isinstance(d, int)
    _eq((b[None], np.array([c]), np.array([i]), np.array([d])),
        tuple(x[:1] for x in want["f64"]), f"{name} min_sum_decode")


@pytest.mark.parametrize("rate,n,snrs,iters,alpha", [
    ("10/15", 64800, (2.6, 2.9, 3.4), 50, 0.75),
    ("4/15", 64800, (0.4, 0.8), 30, 0.85),          # Type A
    ("12/15", 16200, (3.0, 3.6), 20, 1.0),
    ("6/15", 16200, (1.2,), 3, 0.75),
    ("2/15", 16200, (-0.5,), 1, 0.75),
])
def test_utc_live_vs_oracle(device, rate, n, snrs, iters, alpha):
    """Device tensors in (no bus crossing, as lab/m9_gpu.py:148-152 allows),
    mixed convergence, both precisions, against the pinned oracle."""
    import torch
    from atsc3_amd import ldpc
    print(f"[gpu] utc live {n} {rate} iters={iters}", flush=True)
    code = get_code(rate, n)
    cw = synth.codewords(rate, n, 6, 91)
    llr = np.concatenate([synth.llr_bpsk_like(cw, s, 92 + k) for k, s in enumerate(snrs)])
    for dt, tdt in ((np.float32, torch.float32), (np.float64, torch.float64)):
        want = utc_minsum(llr, code.checks, n, iters, alpha, dt)
        t = torch.from_numpy(llr.astype(dt)).to(device)
        for rule in ("torch", "numpy"):     # identical on finite LLRs
            got = ldpc.decode_utc_dev(t, code, iters, alpha, rule)
            torch.cuda.synchronize()
            _eq(_host(got), want, f"{rate} {np.dtype(dt).name} {rule}")


def _host(got):
    return ([x.cpu().numpy() for x in got[:1]] + [got[1].cpu().numpy().astype(bool)]
            + [x.cpu().numpy() for x in got[2:]])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_utc_nonfinite_live_vs_oracle(device, dtype):
    """NaN and +-inf LLRs at seeded positions, scattered over several blocks
    in different numbers (a check then holds 0, 1 or several NaN entries,
    or inf - inf sums), both NaN rules, against the pinned oracle."""
    import torch
    from atsc3_amd import ldpc
    code = get_code("8/15", 16200)
    cw = synth.codewords("8/15", 16200, 6, 41)
    llr = synth.llr_bpsk_like(cw, 2.2, 42).astype(np.float64)
    rng = np.random.default_rng(43)
    for b, cnt in enumerate((0, 1, 3, 12, 60, 400)):
        pos = rng.choice(16200, 3 * cnt, replace=False)
        llr[b, pos[:cnt]] = np.nan
        llr[b, pos[cnt:2 * cnt]] = np.inf
        llr[b, pos[2 * cnt:]] = -np.inf
    t = torch.from_numpy(llr.astype(dtype)).to(device)
    for rule in ("torch", "numpy"):
        print(f"[gpu] utc nonfinite {np.dtype(dtype).name} {rule}", flush=True)
        want = utc_minsum(llr, code.checks, 16200, 20, 0.75, dtype, rule)
        got = ldpc.decode_utc_dev(t, code, 20, 0.75, rule)
        torch.cuda.synchronize()
        _eq(_host(got), want, f"nonfinite {np.dtype(dtype).name} {rule}")


def test_utc_iters_zero_and_properties(device):
    """iters 0 (float64): bits = llr < 0, conv 0, its 0, bad -1
    (lab/m3_ldpc.py:540-552); 256 blocks: conv <=> zero syndrome of the
    returned bits, its in 1..iters, bad = syndrome weight otherwise."""
    from atsc3_amd import ldpc
    code = get_code("6/15", 16200)
    cw = synth.codewords("6/15", 16200, 64, 17)
    llr = np.concatenate([synth.llr_bpsk_like(cw, s, 18 + k)
                          for k, s in enumerate((0.9, 1.1, 1.3, 1.8))])
    b, c, i, d = ldpc.min_sum_decode_batch(llr, code.checks, 16200, 0, 0.75)
    assert np.array_equal(b, (llr < 0).astype(np.uint8))
    assert not c.any() and (i == 0).all() and (d == -1).all()
    b, c, i, d = ldpc.GpuMinSum(code, iters=40, dtype="float32", device=device).decode(llr)
    syn = code.syndrome_count(b)
    assert np.array_equal(syn == 0, c)
    assert np.array_equal(d[~c], syn[~c]) and (d[c] == 0).all()
    assert ((i >= 1) & (i <= 40)).all() and 0 < c.sum() < len(c)
    idx = np.arange(0, 256, 37)
    _eq((b[idx], c[idx], i[idx], d[idx]),
        utc_minsum(llr[idx], code.checks, 16200, 40, 0.75, np.float32), "sample")
