#!/usr/bin/env python3
"""Write the min-sum golden fixtures from the upstream compiled kernel.

Run in a container that has the upstream tree: ``make -C oracle`` first
builds ``oracle/_ref/ldpc_kernel.so`` from the upstream's own
lab/ldpc_kernel.c with the upstream's flags; this script drives that library
through ctypes on this repository's code tables and LLRs and records every
output field (bits, conv, its, bad).

Fixture kinds (all in tests/golden/minsum_*.npz):
  stored   the LLR block itself is in the file (float32, exact)
  seeded   the LLRs are regenerated from (rate, n, nblk, seed, snr) with
           atsc3_amd.synth (codewords + llr_bpsk_like, or the 256-NUC AWGN
           chain + FP64 demap); the file holds the sha256 of the regenerated
           float32 bytes so drift in the generator fails loudly instead of
           comparing different inputs.
Outputs are stored as packed bits plus conv/its/bad per block.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from atsc3_amd import bicm, synth                      # noqa: E402
from atsc3_amd.codes import get_code                   # noqa: E402
from oracle import bicm as OB                          # noqa: E402
from oracle import demap as OD                         # noqa: E402
from oracle import ldpc as OL                          # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def seeded_llr(spec):
    kind = spec["gen"]
    if kind == "bpsk":
        cw = synth.codewords(spec["rate"], spec["n"], spec["nblk"], spec["seed"])
        return synth.llr_bpsk_like(cw, spec["snr"], spec["seed"] + 1)
    if kind == "nuc_awgn":
        n, rate, mod = spec["n"], spec["rate"], spec["mod"]
        _, y, n0 = synth.awgn_blocks(rate, n, mod, spec["nblk"], spec["snr"],
                                     spec["seed"])
        lam, _ = OD.demap(y, bicm.points(mod, rate), OB.lam_of_q(n, mod, rate),
                          sigma2=n0)
        return lam.astype(np.float32)
    raise ValueError(kind)


def quantised(cw, seed, step=0.25):
    """LLRs on a coarse grid: many exact magnitude ties (first-min rule)."""
    llr = synth.llr_bpsk_like(cw, 0.5, seed, scale=1.5)
    return (np.round(llr / np.float32(step)) * np.float32(step)).astype(np.float32)


def run(name, code, llr, iters, alpha, meta, store_llr):
    bits, conv, its, bad = OL.minsum_f32(llr, code, iters, alpha, "ref")
    payload = dict(
        meta=json.dumps(dict(meta, rate=code.rate, n=code.n, iters=iters,
                             alpha=float(alpha))),
        llr_sha256=hashlib.sha256(np.ascontiguousarray(llr, np.float32).tobytes()).hexdigest(),
        bits_packed=np.packbits(bits, axis=1), conv=conv.astype(np.uint8),
        its=its, bad=bad)
    if store_llr:
        payload["llr"] = llr
    np.savez_compressed(os.path.join(OUT, f"minsum_{name}.npz"), **payload)
    print(f"{name:28s} its={its.tolist()} conv={conv.astype(int).tolist()} "
          f"bad={bad.tolist()}")


def main():
    if not OL.available("ref"):
        raise SystemExit("oracle/_ref/ldpc_kernel.so missing: run make -C oracle")
    # -- stored LLR vectors (16200 and one 64800 pair) -------------------------
    c = get_code("4/15", 16200)
    cw = synth.codewords("4/15", 16200, 3, 11)
    llr = synth.llr_bpsk_like(cw, -1.5, 12)
    for a in (1.0, 0.85, 0.75):
        run(f"s16k_r4_a{int(a * 100)}", c, llr, 50, a, dict(kind="stored"), True)
    c = get_code("9/15", 16200)
    cw = synth.codewords("9/15", 16200, 3, 21)
    run("s16k_r9_near", c, synth.llr_bpsk_like(cw, 1.4, 22), 50, 0.75,
        dict(kind="stored"), True)
    c = get_code("6/15", 16200)
    cw = synth.codewords("6/15", 16200, 3, 31)
    run("s16k_r6_ties", c, quantised(cw, 32), 30, 1.0, dict(kind="stored"), True)
    run("s16k_r6_ties_a85", c, quantised(cw, 33), 30, 0.85, dict(kind="stored"), True)
    # special vectors: the upstream build smoke (all +5.0, 16200 r3/15,
    # alpha 1.0 -> zero codeword at its 0), all zeros, and +-inf / NaN
    c = get_code("3/15", 16200)
    run("s16k_r3_smoke", c, np.full((2, 16200), 5.0, np.float32), 50, 1.0,
        dict(kind="stored"), True)
    run("s16k_r3_zeros", c, np.zeros((1, 16200), np.float32), 5, 0.75,
        dict(kind="stored"), True)
    cw = synth.codewords("3/15", 16200, 1, 41)
    sp = synth.llr_bpsk_like(cw, 0.0, 42)
    rng = np.random.default_rng(43)
    sp[0, rng.integers(0, 16200, 40)] = np.inf
    sp[0, rng.integers(0, 16200, 40)] = -np.inf
    sp[0, rng.integers(0, 16200, 5)] = np.nan
    run("s16k_r3_infnan", c, sp, 8, 0.75, dict(kind="stored"), True)
    c = get_code("10/15", 64800)
    spec = dict(gen="nuc_awgn", rate="10/15", n=64800, mod="256QAM", nblk=2,
                seed=101, snr=17.0)
    run("s64k_r10_nuc", c, seeded_llr(spec), 50, 0.75, dict(kind="stored"), True)
    # -- seeded 64800 / 16200 cases ------------------------------------------
    cases = [
        ("g64k_r10_nuc20", dict(gen="nuc_awgn", rate="10/15", n=64800,
                                mod="256QAM", nblk=4, seed=201, snr=20.0), 50, 0.75),
        ("g64k_r10_nuc16", dict(gen="nuc_awgn", rate="10/15", n=64800,
                                mod="256QAM", nblk=3, seed=202, snr=16.0), 50, 1.0),
        ("g64k_r4", dict(gen="bpsk", rate="4/15", n=64800, nblk=3, seed=203,
                         snr=-1.8), 50, 0.85),
        ("g64k_r6", dict(gen="bpsk", rate="6/15", n=64800, nblk=3, seed=204,
                         snr=0.8), 50, 0.75),
        ("g64k_r8", dict(gen="bpsk", rate="8/15", n=64800, nblk=3, seed=205,
                         snr=1.8), 50, 0.75),
        ("g64k_r12", dict(gen="bpsk", rate="12/15", n=64800, nblk=3, seed=206,
                          snr=4.6), 50, 0.75),
        ("g16k_r6", dict(gen="bpsk", rate="6/15", n=16200, nblk=4, seed=207,
                         snr=0.2), 50, 1.0),
        ("g16k_r9", dict(gen="bpsk", rate="9/15", n=16200, nblk=4, seed=208,
                         snr=2.2), 50, 0.75),
        ("g16k_r4", dict(gen="bpsk", rate="4/15", n=16200, nblk=4, seed=209,
                         snr=-2.0), 50, 0.75),
        ("g16k_r11", dict(gen="bpsk", rate="11/15", n=16200, nblk=4, seed=210,
                          snr=3.8), 50, 0.75),
        ("g16k_r2", dict(gen="bpsk", rate="2/15", n=16200, nblk=4, seed=211,
                         snr=-3.5), 50, 0.75),
    ]
    for name, spec, iters, alpha in cases:
        run(name, get_code(spec["rate"], spec["n"]), seeded_llr(spec), iters,
            alpha, dict(kind="seeded", spec=spec), False)


if __name__ == "__main__":
    main()
