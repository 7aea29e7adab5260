"""Bit-interleaver map, constellations and the A/322 table gates (CPU).

The composite map lam_of_q is pinned three ways:
  * the upstream's own map for every (N, mod, rate) of QPSK..256QAM, as
    sha256 fixtures (tests/golden/ref_bicm.npz, written by
    tests/golden/make_ref_vectors.py from lab/m6_bicm.py:185-190);
  * the A/322 6.2.3.1 worked example (lab/m6_bicm.py:380-406);
  * an independent restatement that runs the transmitter's interleavers on
    labels (oracle/bicm.py), equal to both.
Table gates restate lab/spec_bitint.py:394-456 (H1..H8) and
lab/spec_nuc.py:131-142 (unit power); the parity-interleave identity
restates lab/m6_bicm.py:110-134.
"""
import json
import os

import numpy as np
import pytest

from atsc3_amd import bicm, codes
from atsc3_amd.codes import get_code
from oracle import bicm as OB

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF = np.load(os.path.join(GOLDEN, "ref_bicm.npz"), allow_pickle=False)
REF_SHA = {tuple(str(k).split("|")): str(s) for k, s in zip(REF["keys"], REF["sha256"])}
COMBOS = OB.combos()


def _key(n, mod, rate):
    return (str(n), mod, rate)


def test_fixture_covers_every_combination():
    assert len(COMBOS) == 96
    assert set(REF_SHA) == {_key(*c) for c in COMBOS}


@pytest.mark.parametrize("n,mod,rate", COMBOS)
def test_map_equals_upstream_and_oracle(n, mod, rate):
    got = bicm.bicm_map(n, mod, rate)
    want = REF_SHA[_key(n, mod, rate)]
    assert OB.map_sha256(got) == want
    assert np.array_equal(OB.lam_of_q(n, mod, rate), got)
    # a permutation of 0..N-1 (lab/m6_bicm.py:409-425)
    assert np.array_equal(np.sort(got), np.arange(n))


def test_full_map_fixtures():
    """The maps the demap parity tests read are the upstream's, whole."""
    names = [f for f in REF.files if f.startswith("map_")]
    assert len(names) == 11
    for f in names:
        _, n, mod, a, b = f.split("_")
        m = REF[f].astype(np.int64)
        assert OB.map_sha256(m) == REF_SHA[(n, mod, f"{a}/{b}")]
        assert np.array_equal(m, bicm.bicm_map(int(n), mod, f"{a}/{b}"))


def test_worked_example_6231():
    """A/322 6.2.3.1 prints the block interleaver's output for 256QAM,
    N = 64800 (a Type A configuration; rate 2/15 is Type A there): first 16,
    the Part 1 / Part 2 seam, last 8 (lab/m6_bicm.py:391-398)."""
    head = [0, 7920, 15840, 23760, 31680, 39600, 47520, 55440,
            1, 7921, 15841, 23761, 31681, 39601, 47521, 55441]
    seam = [63359, 63360, 63540, 63720, 63900, 64080, 64260, 64440, 64620]
    tail = [63539, 63719, 63899, 64079, 64259, 64439, 64619, 64799]
    for v_of_q in (bicm.block_read_index(64800, "256QAM", "2/15"),
                   OB.blockint_tx(np.arange(64800), 64800, "256QAM", "2/15")):
        assert list(v_of_q[:16]) == head
        assert list(v_of_q[8 * 7920 - 1:8 * 7920 + 8]) == seam
        assert list(v_of_q[-8:]) == tail
    # negative control: the inverse permutation (q_of_v where v_of_q is
    # meant, the likeliest slip) fails the example
    wrong = np.argsort(bicm.block_read_index(64800, "256QAM", "2/15"))
    assert list(wrong[:16]) != head and list(wrong[-8:]) != tail


def test_inverted_groupwise_is_not_the_upstream_map():
    """6.2.2 read the other way round (the direction the upstream calls
    ASSUMPTION X2, lab/m6_bicm.py:140) gives a different map wherever the
    permutation is not an involution -- so the fixtures pin the direction."""
    differ = 0
    for (n, mod, rate) in COMBOS:
        inv = bicm.bicm_map(n, mod, rate, invert_groupwise=True)
        differ += OB.map_sha256(inv) != REF_SHA[_key(n, mod, rate)]
    assert differ >= 90


def test_inverted_groupwise_does_not_decode():
    """The upstream's control (lab/m6_bicm.py:498-506): a noiseless block
    de-interleaved with the inverted group-wise map fails the min-sum; the
    right map gives a clean codeword at its = 0."""
    from atsc3_amd import synth
    from oracle import ldpc as OL
    n, mod, rate = 16200, "64QAM", "11/15"
    cw = synth.codewords(rate, n, 1, 9)
    x = bicm.map_cells(cw[:, bicm.bicm_map(n, mod, rate)], mod, rate)
    from oracle import demap as OD
    code = get_code(rate, n)
    for inv, want_conv in ((False, True), (True, False)):
        lq = bicm.bicm_map(n, mod, rate, invert_groupwise=inv)
        llr, _ = OD.demap(x, bicm.points(mod, rate), lq, sigma2=0.05)
        bits, conv, its, bad = OL.minsum_f32(llr.astype(np.float32), code, 25, 0.75)
        assert bool(conv[0]) == want_conv
        if want_conv:
            assert its[0] == 0 and np.array_equal(bits[0], cw[0])


@pytest.mark.parametrize("n,rate", [(16200, "11/15"), (64800, "10/15"), (16200, "6/15")])
def test_parity_interleave_is_in_the_encoder(n, rate):
    """lab/m6_bicm.py:110-134: run the 6.1.3.2 Type B encoder verbatim (no
    permutation), apply the 6.2.1 parity interleaver
    u[K + 360 t + s] = lam[K + Q s + t]; the result is this library's
    codeword, so the BICM map must not apply 6.2.1 again."""
    code = get_code(rate, n)
    assert code.type == "B"
    K, q, m = code.K, code.Q2, n - code.K
    rows = codes.table_rows(rate, n)
    rng = np.random.default_rng(2)
    s = rng.integers(0, 2, K).astype(np.uint8)
    p = np.zeros(m, np.uint8)
    for kk in np.flatnonzero(s):
        for x in rows[kk // 360]:
            p[(x + q * (kk % 360)) % m] ^= 1
    np.bitwise_xor.accumulate(p, out=p)
    spec = np.concatenate([s, p])
    t, ss = np.meshgrid(np.arange(q), np.arange(360), indexing="ij")
    u = spec.copy()
    u[K + 360 * t + ss] = spec[K + q * ss + t]
    assert np.array_equal(code.encode(s[None])[0], u)
    assert not np.array_equal(spec, u)          # 6.2.1 is not the identity


def _bitint_json():
    with open(os.path.join(os.path.dirname(bicm.__file__), "data", "bitint_tables.json")) as f:
        return json.load(f)


def test_table_gates_h1_h8():
    """lab/spec_bitint.py:394-456 on the transcribed tables."""
    t = _bitint_json()
    mb = dict(t["mod_bits"])
    ta = {(n, m): v for n, m, *v in t["type_a"]}
    tb = {(n, m): v for n, m, *v in t["type_b"]}
    counts = dict.fromkeys(("H1", "H2", "H3", "H4", "H5", "H6", "H7", "H8"), 0)
    for n, m, r, perm in t["groupwise"]:                         # H1
        assert sorted(perm) == list(range(n // 360)), (n, m, r)
        counts["H1"] += 1
    for n, m, hdr in t["groupwise_hdr"]:                         # H2
        assert hdr == list(range(n // 360)), (n, m)
        counts["H2"] += 1
    for (n, m), (nr1, nr2, nc) in ta.items():                    # H3, H5
        if nr1 is None:
            continue
        assert nc == mb[m] and nr1 + nr2 == n // nc and nr1 == (n // nc) // 360 * 360
        assert nr1 * nc == tb[(n, m)][1]
        counts["H3"] += 1
        counts["H5"] += 1
    for (n, m), (nq, np1, np2) in tb.items():                    # H4
        if np1 is None:
            continue
        assert nq == mb[m] and np1 + np2 == n and np1 == n // (360 * nq) * (360 * nq)
        counts["H4"] += 1
    for n, m, bits, cells in t["demux"]:                         # H6
        if cells is None:
            continue
        assert bits == mb[m] and cells == n // bits
        counts["H6"] += 1
    for n, nb, r, typ in t["bi_type"]:                           # H7
        assert typ in ("A", "B")
        counts["H7"] += 1
    for n, m, labs in t["groupwise_labels"]:                     # H8
        assert [x[0] for x in labs] == t["rates"]
        assert [x[1] for x in labs] == list(range(1, len(t["rates"]) + 1))
        counts["H8"] += 1
    assert t["rates"] == codes.RATES
    assert all(v > 0 for v in counts.values()), counts
    assert counts["H1"] == 120 and counts["H7"] == 120


@pytest.mark.parametrize("mod", ["16QAM", "64QAM", "256QAM"])
def test_nuc_unit_power(mod):
    """lab/spec_nuc.py:131-142: mean |x|^2 = 1 within the tolerance the
    printed 4-decimal tables allow (1.5e-3 / 2e-4 / 2e-4)."""
    tol = {"16QAM": 1.5e-3, "64QAM": 2e-4, "256QAM": 2e-4}[mod]
    for rate in codes.RATES:
        p = bicm.points(mod, rate)
        assert len(p) == {"16QAM": 16, "64QAM": 64, "256QAM": 256}[mod]
        assert abs(np.mean(np.abs(p) ** 2) - 1) < tol, rate


def test_qpsk_and_quadrant_rule():
    """Table C.1.1: k = 2 y0 + y1 -> (1+j, -1+j, 1-j, -1-j)/sqrt(2); every
    alphabet: label bit y0 is the sign of Im, y1 the sign of Re."""
    assert np.allclose(bicm.points("QPSK", "2/15"),
                       np.array([1 + 1j, -1 + 1j, 1 - 1j, -1 - 1j]) / np.sqrt(2),
                       rtol=0, atol=1e-16)
    for mod in bicm.MODS:
        p = bicm.points(mod, "10/15")
        nb = bicm.MOD_BITS[mod]
        masks = bicm.bit_masks(nb)
        assert (p[masks[0]].imag <= 0).all() and (p[~masks[0]].imag >= 0).all()
        assert (p[masks[1]].real <= 0).all() and (p[~masks[1]].real >= 0).all()


def test_bit_masks_and_mapper():
    m = bicm.bit_masks(4)
    assert m.shape == (4, 16)
    assert list(np.flatnonzero(m[0])) == list(range(8, 16))      # y0 = MSB
    assert list(np.flatnonzero(m[3])) == list(range(1, 16, 2))
    rng = np.random.default_rng(1)
    for mod in bicm.MODS:
        nb = bicm.MOD_BITS[mod]
        bits = rng.integers(0, 2, (3, 10 * nb))
        x = bicm.map_cells(bits, mod, "8/15")
        lab = (bits.reshape(3, 10, nb) * (1 << np.arange(nb - 1, -1, -1))).sum(-1)
        assert np.array_equal(x, bicm.points(mod, "8/15")[lab])


def test_kernel_tables_equal_upstream():
    """sha256 of (cptr, cidx, vptr, vedge) against the upstream's own
    _kernel_tables(parity_check(rate, n)) (lab/m3_ldpc.py:166-212,459-482),
    every A/322 code: the byte-level pin of the check structure."""
    import hashlib
    with open(os.path.join(GOLDEN, "ref_tables.json")) as f:
        ref = json.load(f)
    assert len(ref) == 24
    for key, want in ref.items():
        n, rate = key.split("|")
        c = get_code(rate, int(n))
        h = hashlib.sha256()
        for a in c.csr():
            h.update(np.ascontiguousarray(a, "<i4").tobytes())
        assert h.hexdigest() == want["sha256"], key
        assert c.nedges == want["nedges"] and c.nc == want["nc"]
