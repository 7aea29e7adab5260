"""Baseband Packet step: BCH syndrome and descrambling (A/322 6.1.2.1, 5.2.3).

CPU: the oracle (oracle/bch.py) is pinned by the spec's printed scrambler
bytes, the minimal-polynomial property of every generator factor, the
generator degrees, encode -> zero syndrome and detection of every single
error; the product's syndrome matrix and scrambler equal the oracle.
GPU: the kernel's syndromes, ok flags and payload bytes equal the oracle on
valid BCH codewords, corrupted ones, and after a full demap + decode."""
import numpy as np
import pytest

from oracle import bch as OB
from atsc3_amd import baseband as BB
from atsc3_amd.codes import get_code


def test_scrambler_printed_vector():
    """lab/spec_bicm.py:88 SCRAMBLER_FIRST_BYTES = (0xC0, 0x6D, 0x3F)."""
    assert list(np.packbits(OB.sequence(24))) == [0xC0, 0x6D, 0x3F]
    np.testing.assert_array_equal(BB.scrambler_sequence(40000), OB.sequence(40000))


@pytest.mark.parametrize("n,m", [(16200, 14), (64800, 16)])
def test_generator_factors_are_minimal_polynomials(n, m):
    """Each factor divides x^(2^m) - x (its roots lie in GF(2^m)); a
    transcription slip would break it."""
    for p in BB.BCH_GEN_POLYS[n]:
        g = OB.poly_from_exps(p)
        assert g.bit_length() - 1 == m
        assert OB.gf2_mod((1 << (1 << m)) ^ 2, g) == 0, p
    g = OB.generator(BB.BCH_GEN_POLYS[n])
    assert g == BB.generator(n)[0]
    assert g.bit_length() - 1 == {16200: 168, 64800: 192}[n]


@pytest.mark.parametrize("n,rate", [(16200, "6/15"), (64800, "10/15")])
def test_encode_syndrome_and_single_errors(n, rate):
    g = OB.generator(BB.BCH_GEN_POLYS[n])
    k = get_code(rate, n).K
    mouter = g.bit_length() - 1
    rng = np.random.default_rng(5)
    cw = OB.encode(rng.integers(0, 2, k - mouter).astype(np.uint8), g)
    assert len(cw) == k and OB.syndrome(cw, g) == 0
    for pos in rng.choice(k, 12, replace=False):
        e = cw.copy()
        e[pos] ^= 1
        assert OB.syndrome(e, g) != 0


@pytest.mark.parametrize("n,rate", [(16200, "4/15"), (64800, "12/15")])
def test_syndrome_matrix_equals_division(n, rate):
    k = get_code(rate, n).K
    g, deg = BB.generator(n)
    m = BB.syndrome_matrix(k, n)
    rng = np.random.default_rng(9)
    for _ in range(3):
        bits = rng.integers(0, 2, k).astype(np.uint8)
        words = np.packbits(np.concatenate([bits, np.zeros((-k) % 32, np.uint8)])).view(">u4")
        par = np.zeros(deg, np.uint64)
        for w in range(len(words)):
            x = np.bitwise_and(np.uint32(words[w]), m[w])
            par += np.array([bin(int(v)).count("1") for v in x], np.uint64)
        s = sum(int(b & 1) << j for j, b in enumerate(par))
        assert s == OB.syndrome(bits, g)


def _bch_codewords(n, rate, nblk, seed):
    """Full transmit side: payload -> scramble -> BCH -> LDPC."""
    code = get_code(rate, n)
    g = OB.generator(BB.BCH_GEN_POLYS[n])
    mouter = g.bit_length() - 1
    rng = np.random.default_rng(seed)
    pay = rng.integers(0, 2, (nblk, code.K - mouter)).astype(np.uint8)
    seq = OB.sequence(code.K - mouter)
    info = np.stack([OB.encode(p ^ seq, g) for p in pay])
    return pay, code.encode(info)


@pytest.mark.gpu
@pytest.mark.parametrize("n,rate", [(16200, "4/15"), (16200, "9/15"), (64800, "10/15"),
                                    (64800, "13/15")])
def test_gpu_baseband_vs_oracle(device, n, rate):
    import torch
    print(f"[gpu] baseband {n} {rate}", flush=True)
    pay, cw = _bch_codewords(n, rate, 4, 17)
    cw[2, 5] ^= 1                      # a BCH error in block 2
    cw[3, :cw.shape[1] // 3] ^= 1      # many errors in block 3
    bb = BB.Baseband(rate, n, device)
    syn, ok, out = bb.run(torch.from_numpy(cw).to(device))
    torch.cuda.synchronize()
    g = OB.generator(BB.BCH_GEN_POLYS[n])
    s_ints = BB.syndrome_ints(syn)
    for b in range(4):
        s, by = OB.baseband_packet(cw[b], bb.kldpc, g)
        assert s_ints[b] == s, b
        assert bool(ok[b]) == (s == 0)
        assert out[b].cpu().numpy().tobytes() == by
    assert bool(ok[0]) and bool(ok[1]) and not bool(ok[2]) and not bool(ok[3])
    assert out[0].cpu().numpy().tobytes() == np.packbits(pay[0]).tobytes()


@pytest.mark.gpu
def test_gpu_chain_demap_decode_baseband(device):
    import torch
    from atsc3_amd import ldpc, synth
    from atsc3_amd.demap import Demapper
    n, rate, mod = 64800, "10/15", "256QAM"
    pay, cw = _bch_codewords(n, rate, 6, 23)
    x = synth.modulate(cw, n, mod, rate)
    rng = np.random.default_rng(3)
    n0 = synth.n0_of(21.0)
    y = (x + synth.complex_noise(rng, x.shape) * np.sqrt(n0)).astype(np.complex64)
    print("[gpu] demap + decode + baseband chain", flush=True)
    llr, _ = Demapper(n, mod, rate, device).demap_dev(torch.from_numpy(y).to(device), sigma2=n0)
    bits, conv, its, bad = ldpc.decode_dev(llr, get_code(rate, n), 50, 0.75)
    syn, ok, out = BB.Baseband(rate, n, device).run(bits)
    torch.cuda.synchronize()
    assert conv.all() and ok.all()
    for b in range(6):
        assert out[b].cpu().numpy().tobytes() == np.packbits(pay[b]).tobytes()
