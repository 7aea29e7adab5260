"""BCH syndrome and descrambling of decoded codewords: Baseband Packet bytes.

The step after the LDPC decode in the upstream receive chain
(lab/m6_bicm.py:357-364 PlpChain.baseband_packet; batched in
lab/m9_fast.py:942-949 via lab/m9_accel.py:150-236): of a decoded
codeword, the first Kldpc bits are the BCH codeword (systematic, MSB first);

  * BCH syndrome = (those Kldpc bits as a polynomial, first bit = highest
    degree) mod g(x), g = g1 * ... * g12 of A/322 6.1.2.1 (degree Mouter =
    168 for N = 16200, 192 for N = 64800); zero means no detected error
    (the upstream computes the syndrome only, it does not correct);
  * payload = the first Kpayload = Kldpc - Mouter bits XOR the A/322 5.2.3
    scrambling sequence, packed MSB-first into bytes.

On the device (csrc/baseband.hip) the syndrome is a GF(2) matrix-vector
product: bit j of the syndrome is the parity of (message AND column j of the
code's syndrome matrix), column j holding coefficient j of x^(Kldpc-1-i) mod
g for every message position i.
"""
from __future__ import annotations

import functools

import numpy as np
import torch

from . import _lib
from .codes import get_code

# A/322 Table 6.2 / 6.3 generator polynomial factors (exponents of x), as
# listed in lab/spec_bicm.py:125-169.  Data of the standard.
BCH_GEN_POLYS = {
    16200: ((14, 5, 3, 1, 0), (14, 11, 8, 6, 0), (14, 10, 9, 6, 2, 1, 0),
            (14, 12, 10, 8, 7, 4, 0), (14, 13, 11, 9, 8, 6, 4, 2, 0), (14, 13, 9, 8, 7, 3, 0),
            (14, 13, 11, 10, 7, 6, 5, 2, 0), (14, 11, 10, 9, 8, 5, 0), (14, 10, 9, 3, 2, 1, 0),
            (14, 12, 11, 9, 6, 3, 0), (14, 12, 11, 4, 0),
            (14, 13, 10, 8, 7, 6, 5, 3, 2, 1, 0)),
    64800: ((16, 5, 3, 2, 0), (16, 8, 6, 5, 4, 1, 0), (16, 11, 10, 9, 8, 7, 5, 4, 3, 2, 0),
            (16, 14, 12, 11, 9, 6, 4, 2, 0), (16, 12, 11, 10, 9, 8, 5, 3, 2, 1, 0),
            (16, 15, 14, 13, 12, 10, 9, 8, 7, 5, 4, 2, 0),
            (16, 15, 13, 11, 10, 9, 8, 6, 5, 2, 0), (16, 14, 13, 12, 9, 8, 6, 5, 2, 1, 0),
            (16, 11, 10, 9, 7, 5, 0), (16, 14, 13, 12, 10, 8, 7, 5, 2, 1, 0),
            (16, 13, 12, 11, 9, 5, 3, 2, 0), (16, 12, 11, 9, 7, 6, 5, 1, 0)),
}

# A/322 5.2.3 scrambler (the register as solved in lab/m4_scrambler.py:28-60):
# 16 stages X1..X16, the initial word 0xF180 with X1 = its MSB; per byte the
# output is (X3, X4, X5, X6, X10, X13, X14, X16) MSB first, then one Galois
# step toward X1: every stage takes the content of the stage above, X16
# takes the bit leaving X1, and that bit is also XORed into the new X3, X4,
# X5, X9, X10, X13 and X15.
SCRAMBLER_SEED = 0xF180
_SCR_OUT = (3, 4, 5, 6, 10, 13, 14, 16)
_SCR_FEED = (3, 4, 5, 9, 10, 13, 15)


def clmul(a, b):
    """Carry-less product of two GF(2) polynomials held as Python ints."""
    r = 0
    while b:
        if b & 1:
            r ^= a
        a <<= 1
        b >>= 1
    return r


@functools.lru_cache(maxsize=None)
def generator(n):
    """(g as an int, Mouter) for the 16200- or 64800-bit codes."""
    g = 1
    for exps in BCH_GEN_POLYS[n]:
        g = clmul(g, sum(1 << e for e in exps))
    return g, g.bit_length() - 1


def scrambler_sequence(nbits):
    """The scrambling sequence, MSB-first per byte (uint8 0/1)."""
    stage = [0] * 17                     # stage[1..16]
    for k in range(1, 17):
        stage[k] = (SCRAMBLER_SEED >> (16 - k)) & 1
    nby = (nbits + 7) // 8
    out = np.empty(nby * 8, np.uint8)
    for b in range(nby):
        for q, s in enumerate(_SCR_OUT):
            out[8 * b + q] = stage[s]
        fb = stage[1]
        nxt = [0] * 17
        for k in range(1, 16):
            nxt[k] = stage[k + 1] ^ (fb if k in _SCR_FEED else 0)
        nxt[16] = fb
        stage = nxt
    return out[:nbits]


@functools.lru_cache(maxsize=None)
def syndrome_matrix(kldpc, n):
    """uint32 [nwords][Mouter]: bit (31 - t) of [w][j] = coefficient j of
    x^(kldpc-1-(32w+t)) mod g."""
    g, deg = generator(n)
    nwords = (kldpc + 31) // 32
    nby = (deg + 7) // 8
    rows = np.zeros((nwords * 32, deg), np.uint8)
    r = 1
    for k in range(kldpc):
        i = kldpc - 1 - k
        rows[i] = np.unpackbits(np.frombuffer(r.to_bytes(nby, "little"), np.uint8),
                                bitorder="little")[:deg]
        r <<= 1
        if (r >> deg) & 1:
            r ^= g
    # pack 32 message positions per word, position t at bit 31 - t
    rows = rows.reshape(nwords, 32, deg)
    weights = (1 << (31 - np.arange(32, dtype=np.uint64))).astype(np.uint64)
    m = (rows.astype(np.uint64) * weights[None, :, None]).sum(1).astype(np.uint32)
    return np.ascontiguousarray(m)


class Baseband:
    """Device tables of one (rate, N) and the launcher."""

    def __init__(self, rate, n, device="cuda"):
        self.device = torch.device(device)
        self.code = get_code(rate, n)
        self.n = int(n)
        self.kldpc = int(self.code.K)
        _, self.mouter = generator(self.n)
        self.kpayload = self.kldpc - self.mouter
        if self.kpayload % 8:
            raise ValueError("payload is not a whole number of bytes")
        self.mask = torch.from_numpy(syndrome_matrix(self.kldpc, self.n).view(np.int32)).to(
            self.device)
        seq = np.packbits(scrambler_sequence(self.kpayload))
        self.seq = torch.from_numpy(seq).to(self.device)

    def run(self, bits):
        """bits (B, N) uint8 decoded codewords (device) -> (syndrome (B, 8)
        int32 little-endian 32-bit limbs, bch_ok (B,) uint8, payload bytes
        (B, Kpayload/8) uint8)."""
        if not (bits.is_cuda and bits.dtype == torch.uint8 and bits.dim() == 2
                and bits.shape[1] == self.n):
            raise ValueError(f"bits must be a (B, {self.n}) uint8 device tensor")
        bits = bits.contiguous()
        B = bits.shape[0]
        syn = torch.empty((B, 8), dtype=torch.int32, device=bits.device)
        ok = torch.empty(B, dtype=torch.uint8, device=bits.device)
        out = torch.empty((B, self.kpayload // 8), dtype=torch.uint8, device=bits.device)
        if B:
            _lib.call("atsc3_baseband_dev", bits.data_ptr(), B, self.n, self.kldpc,
                      self.mouter, self.mask.data_ptr(), self.seq.data_ptr(), syn.data_ptr(),
                      ok.data_ptr(), out.data_ptr(),
                      torch.cuda.current_stream(bits.device).cuda_stream)
        return syn, ok, out


def syndrome_ints(syn):
    """(B, 8) int32 limbs -> list of Python ints."""
    s = np.asarray(syn.cpu().numpy() if isinstance(syn, torch.Tensor) else syn).view(np.uint32)
    return [sum(int(w) << (32 * q) for q, w in enumerate(row)) for row in s]
