"""The update-then-check oracle (oracle/utc_ref.py) against the upstream's
own decoders (CPU).

tests/golden/utc_*.npz hold what lab/m9_gpu.py:84-211 GpuMinSum.decode
(float32 "f32" and float64 "t64") and lab/m3_ldpc.py:245-286 min_sum_decode
(float64 "f64") return on the LLRs of the min-sum golden fixtures, on those
LLRs scaled past the 1e9 row padding, on float64-native LLRs, and on LLRs
holding NaN and +-inf (where GpuMinSum's NaN-propagating amin and numpy's
NaN-last partition / first-NaN argmin part ways).  Planted mistakes -- the
compiled kernel's association, no row padding, a syndrome check before the
first update, the other function's NaN rule -- must each fail at least one
fixture.
"""
import glob
import hashlib
import json
import os

import numpy as np
import pytest

from oracle.utc_ref import utc_minsum
from tests import _fixtures as F

GOLDEN = F.GOLDEN
NAMES = sorted(os.path.basename(p)[4:-4] for p in glob.glob(os.path.join(GOLDEN, "utc_*.npz")))


def load_utc(name):
    """-> (code, llr, iters, alpha, {"f32": (bits, conv, its, bad), "t64": ...,
    "f64": ...});
    code is an LdpcCode, or for the hand-made graphs a Tiny with .n/.checks."""
    from atsc3_amd.codes import get_code
    d = np.load(os.path.join(GOLDEN, f"utc_{name}.npz"), allow_pickle=False)
    meta = json.loads(str(d["meta"]))
    if "tiny" in meta:
        t = meta["tiny"]
        code = Tiny(meta["n"], [t["cidx"][a:b] for a, b in zip(t["cptr"][:-1], t["cptr"][1:])])
        llr = np.array(t["llr"], np.float64)
    else:
        code = get_code(meta["rate"], meta["n"])
    if "tiny" in meta:
        pass
    elif "f64" in meta:
        s = meta["f64"]
        llr = f64_llr(s["n"], s["rate"], s["mod"], s["nblk"], s["esn0"], s["seed"])
        assert hashlib.sha256(llr.tobytes()).hexdigest() == meta["llr_sha256"]
    else:
        _, llr, _, _, _ = F.load(meta["source"])
        if "scale" in meta:
            llr = (llr * np.float32(meta["scale"])).astype(np.float32)
        if "nonfinite" in meta:
            llr = nonfinite(llr, meta["nonfinite"]["count"], meta["nonfinite"]["seed"])
    want = {}
    for p in ("f32", "t64", "f64"):
        bits = np.unpackbits(d[f"{p}_bits"], axis=1)[:, :code.n]
        want[p] = (bits, d[f"{p}_conv"].astype(bool), d[f"{p}_its"], d[f"{p}_bad"])
    return code, llr, meta["iters"], meta["alpha"], want


class Tiny:
    def __init__(self, n, checks):
        self.n, self.checks = n, checks


def nonfinite(llr, count, seed):
    """``count`` NaN, +inf and -inf each at seeded distinct positions of
    every block, as tests/golden/make_ref_vectors.py planted them."""
    out = np.array(llr, np.float32)
    rng = np.random.default_rng(seed)
    for row in out:
        pos = rng.choice(row.size, 3 * count, replace=False)
        row[pos[:count]] = np.nan
        row[pos[count:2 * count]] = np.inf
        row[pos[2 * count:]] = -np.inf
    return out


# (dtype, NaN rule) each fixture output was produced under
PRECS = {"f32": (np.float32, "torch"), "t64": (np.float64, "torch"),
         "f64": (np.float64, "numpy")}
NONFINITE = ["s16k_r3_infnan", "g16k_r4_nonfinite", "s16k_r6_ties_nonfinite"]


def f64_llr(n, rate, mod, nblk, esn0, seed):
    from atsc3_amd import bicm, synth
    from oracle import demap as OD
    _, y, n0 = synth.awgn_blocks(rate, n, mod, nblk, esn0, seed)
    lam, _ = OD.demap(y, bicm.points(mod, rate), F.ref_lam_of_q(n, mod, rate), sigma2=n0)
    return np.ascontiguousarray(lam, np.float64)


def same(got, want):
    return all(np.array_equal(np.asarray(g), np.asarray(w)) for g, w in zip(got, want))


def test_fixture_set():
    assert len(NAMES) == 19
    assert {"s16k_r3_zeros", "s16k_r3_smoke", "g64k_r10_nuc20"} <= set(NAMES)
    assert set(NONFINITE) <= set(NAMES)


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("prec", sorted(PRECS))
def test_oracle_equals_upstream(name, prec):
    code, llr, iters, alpha, want = load_utc(name)
    dt, rule = PRECS[prec]
    assert same(utc_minsum(llr, code.checks, code.n, iters, alpha, dt, rule), want[prec])


def test_nan_rules_differ_on_nonfinite():
    """The two float64 functions agree on finite LLRs and part ways on NaN:
    swapping the NaN rules fails every non-finite fixture."""
    for name in NONFINITE:
        code, llr, iters, alpha, want = load_utc(name)
        assert llr.dtype == np.float32 and np.isnan(llr).any() and np.isinf(llr).any()
        assert not same(want["t64"], want["f64"]), name
        for prec, other in (("t64", "numpy"), ("f64", "torch")):
            got = utc_minsum(llr, code.checks, code.n, iters, alpha, np.float64, other)
            assert not same(got, want[prec]), (name, prec)
        got = utc_minsum(llr, code.checks, code.n, iters, alpha, np.float32, "numpy")
        assert not same(got, want["f32"]), name
    for name in sorted(set(NAMES) - set(NONFINITE)):
        _, _, _, _, want = load_utc(name)
        assert same(want["t64"], want["f64"]), name


def test_clean_block_reports_one_iteration():
    """The contract difference that matters to callers: a block whose channel
    decisions already satisfy H reports its = 1, not 0."""
    code, llr, iters, alpha, want = load_utc("s16k_r3_zeros")
    assert want["f32"][2].tolist() == [1] and want["f64"][2].tolist() == [1]


@pytest.mark.parametrize("plant,names", [
    ("zero_first", ["tiny_cancel_f64", "tiny_cancel_f32", "tiny_wide"]),
    ("no_pad", ["g16k_r4_x4e09", "s16k_r6_ties_x2e09"]),
    ("check_first", ["s16k_r3_zeros", "s16k_r3_smoke"]),
])
def test_planted_mistakes_fail(plant, names):
    """Each planted mistake fails in every precision and rule."""
    failed = dict.fromkeys(PRECS, 0)
    for name in names:
        code, llr, iters, alpha, want = load_utc(name)
        for prec, (dt, rule) in PRECS.items():
            got = utc_minsum(llr, code.checks, code.n, iters, alpha, dt, rule, _plant=plant)
            failed[prec] += not same(got, want[prec])
    assert all(failed.values()), (plant, failed)


def test_iters_zero_and_empty():
    from atsc3_amd.codes import get_code
    code = get_code("3/15", 16200)
    llr = np.linspace(-1, 1, 2 * 16200).reshape(2, 16200)
    b, c, i, d = utc_minsum(llr, code.checks, 16200, 0, 0.75, np.float64)
    assert np.array_equal(b, (llr < 0).astype(np.uint8))
    assert not c.any() and (i == 0).all() and (d == -1).all()
    b, c, i, d = utc_minsum(np.zeros((0, 16200)), code.checks, 16200, 5, 0.75)
    assert b.shape == (0, 16200)
