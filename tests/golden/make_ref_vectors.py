#!/usr/bin/env python3
"""Write golden vectors from the upstream's own Python functions.

Run only in a container that holds the upstream tree (default
/root/reference); it imports the upstream modules from ``lab/`` to produce
DATA fixtures for functions nothing else pins byte for byte:

  ref_bicm.npz    the composite bit-interleaver map lam_of_q
                  (lab/m6_bicm.py:185-190 bicm_map) for every (N, mod, rate)
                  of QPSK..256QAM: sha256 of the map as little-endian int32,
                  plus the maps themselves (uint16) at the operating points
                  the demap parity tests use.
  ref_tables.json sha256 of the upstream kernel tables (cptr, cidx, vptr,
                  vedge as int32, lab/m3_ldpc.py:459-482 over
                  lab/m3_ldpc.py:166-212 parity_check) for every A/322 code.
  utc_<name>.npz  outputs (bits packed, conv, its, bad) of the update-then-
                  check decoders on the LLRs of golden fixture <name>:
                  GpuMinSum.decode at float32 and at float64 (lab/m9_gpu.py:
                  84-211, run on torch's CPU device) and min_sum_decode at
                  float64 (lab/m3_ldpc.py:245-286, one call per block); the
                  two float64 functions differ only where LLRs are NaN.  Sets marked
                  "f64" regenerate float64 LLRs (FP64 demap of seeded AWGN
                  cells) that are not float32 values, checked by sha256.

Nothing of the upstream is stored: only inputs' identities and outputs.
Usage:  python tests/golden/make_ref_vectors.py [/path/to/upstream]
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import bicm as OB                       # noqa: E402
from tests import _fixtures as F                    # noqa: E402

# the demap parity tests' operating points (N, mod, rate)
FULL_MAPS = [(64800, "256QAM", "10/15"), (64800, "256QAM", "6/15"),
             (64800, "256QAM", "8/15"), (64800, "64QAM", "10/15"),
             (64800, "16QAM", "8/15"), (64800, "QPSK", "4/15"),
             (16200, "64QAM", "9/15"), (16200, "QPSK", "9/15"),
             (16200, "256QAM", "4/15"), (16200, "64QAM", "11/15"),
             (16200, "QPSK", "6/15")]

# golden min-sum fixtures whose LLRs feed the update-then-check vectors
UTC_FROM = ["s16k_r3_smoke", "s16k_r3_zeros", "g64k_r10_nuc20", "s16k_r6_ties",
            "s16k_r6_ties_a85", "g16k_r4", "g64k_r4", "s16k_r9_near", "g16k_r11",
            "s16k_r3_infnan"]
# the LLRs of a fixture with NaN and +-inf planted at seeded positions:
# (fixture, count of each, seed)
UTC_NONFINITE = [("g16k_r4", 40, 5), ("s16k_r6_ties", 25, 6)]
# the LLRs of a fixture scaled so that most |x| exceed the 1e9 row padding
# (the padding then decides messages): (fixture, factor)
UTC_SCALED = [("g16k_r4", 4e9), ("s16k_r6_ties", 2e9)]
# float64-native LLR sets: (name, n, rate, mod, nblk, esn0 dB, seed, iters, alpha)
UTC_F64 = [("f64_16k_r6_qpsk", 16200, "6/15", "QPSK", 3, 0.6, 71, 50, 0.75),
           ("f64_64k_r10_nuc", 64800, "10/15", "256QAM", 2, 17.8, 72, 50, 0.85)]


def tiny_sets():
    """Small hand-made graphs whose LLRs make huge messages of opposite sign
    cancel on one variable, so the association of its sum decides its bit:
    llr-first ((l + m1) + m2) and zero-first (l + (m1 + m2)) differ.  The
    second is random with LLR magnitudes 2^-20 .. 2^70."""
    out = []
    for prec, e in (("f64", 55), ("f32", 26)):
        # every check of degree 3: no row padding enters the two-minimum
        checks = [[0, 1, 4], [0, 2, 5], [1, 2, 3]]
        big = 2.0 ** e
        llr = np.array([[-1.0, -big, big, 3.0, 2 * big, 2 * big],
                        [1.0, big, -big, -3.0, 2 * big, 2 * big]])
        out.append((f"tiny_cancel_{prec}", checks, 6, llr, 1, 0.75))
    rng = np.random.default_rng(77)
    n, nc = 48, 30
    checks = [sorted(rng.choice(n, rng.integers(3, 7), replace=False).tolist())
              for _ in range(nc)]
    llr = np.sign(rng.standard_normal((4, n))) * 2.0 ** rng.uniform(-20, 70, (4, n))
    out.append(("tiny_wide", checks, n, llr, 12, 0.75))
    return out


def nonfinite(llr, count, seed):
    """``count`` NaN, +inf and -inf each at seeded distinct positions of
    every block (float32)."""
    out = np.array(llr, np.float32)
    rng = np.random.default_rng(seed)
    for row in out:
        pos = rng.choice(row.size, 3 * count, replace=False)
        row[pos[:count]] = np.nan
        row[pos[count:2 * count]] = np.inf
        row[pos[2 * count:]] = -np.inf
    return out


def f64_llr(n, rate, mod, nblk, esn0, seed):
    """FP64 max-log LLRs of seeded AWGN cells, codeword order (not rounded
    to float32)."""
    from atsc3_amd import synth, bicm
    from oracle import demap as OD
    _, y, n0 = synth.awgn_blocks(rate, n, mod, nblk, esn0, seed)
    lam, _ = OD.demap(y, bicm.points(mod, rate), OB.lam_of_q(n, mod, rate), sigma2=n0)
    return np.ascontiguousarray(lam, np.float64)


def main(upstream):
    sys.path.insert(0, os.path.join(upstream, "lab"))
    import m3_ldpc as LD                            # noqa: E402  (upstream)
    import m6_bicm as UB                            # noqa: E402  (upstream)
    import m9_gpu as UG                             # noqa: E402  (upstream)
    from atsc3_amd.codes import RATES

    # -- bit-interleaver maps ------------------------------------------------
    keys, shas, full = [], [], {}
    for (n, mod, rate) in OB.combos():
        m = np.asarray(UB.bicm_map(n, mod, rate), np.int64)
        keys.append(f"{n}|{mod}|{rate}")
        shas.append(OB.map_sha256(m))
        if (n, mod, rate) in FULL_MAPS:
            full[f"map_{n}_{mod}_{rate.replace('/', '_')}"] = m.astype(np.uint16)
    np.savez_compressed(os.path.join(HERE, "ref_bicm.npz"), keys=np.array(keys),
                        sha256=np.array(shas), **full)
    print(f"ref_bicm.npz: {len(keys)} maps, {len(full)} stored in full")

    # -- kernel tables -------------------------------------------------------
    tables, checks_of = {}, {}
    for n in (16200, 64800):
        for rate in RATES:
            checks = LD.parity_check(rate, n)
            checks_of[(n, rate)] = checks
            cptr, cidx, vptr, vedge, ne = LD._kernel_tables(checks, n)
            h = hashlib.sha256()
            for a in (cptr, cidx, vptr, vedge):
                h.update(np.ascontiguousarray(a, "<i4").tobytes())
            tables[f"{n}|{rate}"] = dict(sha256=h.hexdigest(), nedges=int(ne),
                                         nc=len(checks))
    with open(os.path.join(HERE, "ref_tables.json"), "w") as f:
        json.dump(tables, f, indent=1, sort_keys=True)
    print(f"ref_tables.json: {len(tables)} codes")

    # -- update-then-check decoders -----------------------------------------
    def run(name, n, rate, llr, iters, alpha, meta, checks=None):
        checks = checks or checks_of.get((n, rate)) or LD.parity_check(rate, n)
        g = UG.GpuMinSum(checks, n, iters=iters, alpha=alpha, device="cpu",
                         dtype="float32")
        b32, c32, i32, d32 = g.decode(np.asarray(llr, np.float64))
        g64 = UG.GpuMinSum(checks, n, iters=iters, alpha=alpha, device="cpu",
                           dtype="float64")
        t64 = g64.decode(np.asarray(llr, np.float64))
        rows = [LD.min_sum_decode(np.asarray(r, np.float64), checks, n, iters, alpha)
                for r in llr]
        b64 = np.stack([r[0] for r in rows]).astype(np.uint8)
        c64 = np.array([r[1] for r in rows], bool)
        i64 = np.array([r[2] for r in rows], np.int32)
        d64 = np.array([r[3] for r in rows], np.int32)
        # the batched float64 decoder is row-exact with the scalar one
        bb = LD.min_sum_decode_batch(np.asarray(llr, np.float64), checks, n, iters, alpha,
                                     dtype=np.float64)
        assert np.array_equal(bb[0], b64) and np.array_equal(bb[2], i64), name
        assert np.array_equal(bb[1], c64) and np.array_equal(bb[3], d64), name
        np.savez_compressed(
            os.path.join(HERE, f"utc_{name}.npz"), meta=json.dumps(meta),
            f32_bits=np.packbits(np.asarray(b32, np.uint8), axis=1),
            f32_conv=np.asarray(c32, bool), f32_its=np.asarray(i32, np.int32),
            f32_bad=np.asarray(d32, np.int32),
            f64_bits=np.packbits(b64, axis=1), f64_conv=c64, f64_its=i64, f64_bad=d64,
            t64_bits=np.packbits(np.asarray(t64[0], np.uint8), axis=1),
            t64_conv=np.asarray(t64[1], bool), t64_its=np.asarray(t64[2], np.int32),
            t64_bad=np.asarray(t64[3], np.int32))
        print(f"utc_{name}: f32 its {np.asarray(i32).tolist()} conv "
              f"{np.asarray(c32).astype(int).tolist()}; f64 its {i64.tolist()} conv "
              f"{c64.astype(int).tolist()}; torch f64 its {np.asarray(t64[2]).tolist()}")

    for name in UTC_FROM:
        code, llr, iters, alpha, _ = F.load(name)
        run(name, code.n, code.rate, llr, iters, alpha,
            dict(source=name, n=code.n, rate=code.rate, iters=iters, alpha=alpha))
    for name, factor in UTC_SCALED:
        code, llr, iters, alpha, _ = F.load(name)
        llr = (llr * np.float32(factor)).astype(np.float32)
        run(f"{name}_x{factor:.0e}".replace("+", ""), code.n, code.rate, llr, iters, alpha,
            dict(source=name, scale=factor, n=code.n, rate=code.rate, iters=iters,
                 alpha=alpha))
    for name, cnt, seed in UTC_NONFINITE:
        code, llr, iters, alpha, _ = F.load(name)
        llr = nonfinite(llr, cnt, seed)
        run(f"{name}_nonfinite", code.n, code.rate, llr, iters, alpha,
            dict(source=name, nonfinite=dict(count=cnt, seed=seed), n=code.n, rate=code.rate,
                 iters=iters, alpha=alpha))
    for name, checks, n, llr, iters, alpha in tiny_sets():
        cn = [len(c) for c in checks]
        run(name, n, None, llr, iters, alpha,
            dict(tiny=dict(cptr=np.concatenate([[0], np.cumsum(cn)]).tolist(),
                           cidx=[v for c in checks for v in c], llr=llr.tolist()),
                 n=n, rate=None, iters=iters, alpha=alpha), checks=checks)
    for (name, n, rate, mod, nblk, esn0, seed, iters, alpha) in UTC_F64:
        llr = f64_llr(n, rate, mod, nblk, esn0, seed)
        sha = hashlib.sha256(llr.tobytes()).hexdigest()
        run(name, n, rate, llr, iters, alpha,
            dict(f64=dict(n=n, rate=rate, mod=mod, nblk=nblk, esn0=esn0, seed=seed),
                 llr_sha256=sha, n=n, rate=rate, iters=iters, alpha=alpha))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
