"""The update-then-check oracle (oracle/utc_ref.py) against the upstream's
own decoders (CPU).

tests/golden/utc_*.npz hold what lab/m9_gpu.py:84-211 GpuMinSum.decode
(float32) and lab/m3_ldpc.py:245-286 min_sum_decode (float64) return on the
LLRs of the min-sum golden fixtures, on those LLRs scaled past the 1e9 row
padding, and on float64-native LLRs.  Planted mistakes -- the compiled
kernel's association, no row padding, a syndrome check before the first
update -- must each fail at least one fixture.
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
    """-> (code, llr, iters, alpha, {"f32": (bits, conv, its, bad), "f64": ...});
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
    want = {}
    for p in ("f32", "f64"):
        bits = np.unpackbits(d[f"{p}_bits"], axis=1)[:, :code.n]
        want[p] = (bits, d[f"{p}_conv"].astype(bool), d[f"{p}_its"], d[f"{p}_bad"])
    return code, llr, meta["iters"], meta["alpha"], want


class Tiny:
    def __init__(self, n, checks):
        self.n, self.checks = n, checks


def f64_llr(n, rate, mod, nblk, esn0, seed):
    from atsc3_amd import bicm, synth
    from oracle import demap as OD
    _, y, n0 = synth.awgn_blocks(rate, n, mod, nblk, esn0, seed)
    lam, _ = OD.demap(y, bicm.points(mod, rate), F.ref_lam_of_q(n, mod, rate), sigma2=n0)
    return np.ascontiguousarray(lam, np.float64)


def same(got, want):
    return all(np.array_equal(np.asarray(g), np.asarray(w)) for g, w in zip(got, want))


def test_fixture_set():
    assert len(NAMES) == 16
    assert {"s16k_r3_zeros", "s16k_r3_smoke", "g64k_r10_nuc20"} <= set(NAMES)


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_oracle_equals_upstream(name, prec):
    code, llr, iters, alpha, want = load_utc(name)
    dt = np.float32 if prec == "f32" else np.float64
    assert same(utc_minsum(llr, code.checks, code.n, iters, alpha, dt), want[prec])


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
    """Each planted mistake fails in both precisions."""
    failed = {"f32": 0, "f64": 0}
    for name in names:
        code, llr, iters, alpha, want = load_utc(name)
        for prec, dt in (("f32", np.float32), ("f64", np.float64)):
            got = utc_minsum(llr, code.checks, code.n, iters, alpha, dt, _plant=plant)
            failed[prec] += not same(got, want[prec])
    assert failed["f32"] > 0 and failed["f64"] > 0, (plant, failed)


def test_iters_zero_and_empty():
    from atsc3_amd.codes import get_code
    code = get_code("3/15", 16200)
    llr = np.linspace(-1, 1, 2 * 16200).reshape(2, 16200)
    b, c, i, d = utc_minsum(llr, code.checks, 16200, 0, 0.75, np.float64)
    assert np.array_equal(b, (llr < 0).astype(np.uint8))
    assert not c.any() and (i == 0).all() and (d == -1).all()
    b, c, i, d = utc_minsum(np.zeros((0, 16200)), code.checks, 16200, 5, 0.75)
    assert b.shape == (0, 16200)
