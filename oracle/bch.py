"""CPU oracle of the Baseband Packet step: BCH syndrome, BCH encoding and
the A/322 5.2.3 descrambler.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

* ``syndrome`` restates lab/m6_bicm.py:243-278 (_gen_poly, _poly_mod,
  bch_syndrome): the message bits as a polynomial (first bit = highest
  degree), reduced bit-serially mod g = g1 * ... * g12.
* ``encode`` restates lab/m6_bicm.py:281-289 (systematic: message, then the
  Mouter parity bits of m(x) x^Mouter mod g).
* ``sequence`` restates the scrambler of lab/m4_scrambler.py:128-160 with the
  register held as a 16-bit integer (bit k-1 = stage Xk).
* ``baseband_packet`` restates lab/m6_bicm.py:357-364.

Pinning (tests/test_baseband.py): the sequence reproduces the spec's printed
first bytes 0xC0 0x6D 0x3F (lab/spec_bicm.py:88); every generator factor
divides x^(2^m) - x (m = 14 / 16: the factors are the minimal polynomials the
tables transcribe); deg g = 168 / 192; encode -> syndrome 0; every single-bit
error is detected.
"""
from __future__ import annotations

import numpy as np


def poly_from_exps(exps):
    v = 0
    for e in exps:
        v ^= 1 << e
    return v


def gf2_mul(a, b):
    r = 0
    while a:
        if a & 1:
            r ^= b
        a >>= 1
        b <<= 1
    return r


def gf2_mod(v, g):
    dg = g.bit_length() - 1
    while v and v.bit_length() - 1 >= dg:
        v ^= g << (v.bit_length() - 1 - dg)
    return v


def generator(polys):
    g = 1
    for p in polys:
        g = gf2_mul(g, poly_from_exps(p))
    return g


def bits_to_int(bits):
    v = 0
    for b in np.asarray(bits, np.uint8):
        v = (v << 1) | int(b)
    return v


def syndrome(bits, g):
    return gf2_mod(bits_to_int(bits), g)


def encode(info_bits, g):
    deg = g.bit_length() - 1
    p = gf2_mod(bits_to_int(info_bits) << deg, g)
    par = np.array([(p >> (deg - 1 - i)) & 1 for i in range(deg)], np.uint8)
    return np.concatenate([np.asarray(info_bits, np.uint8), par])


def sequence(nbits, seed=0xF180):
    """Register word r: bit k-1 = stage Xk; seed loaded with X1 = its MSB."""
    r = 0
    for k in range(1, 17):
        r |= ((seed >> (16 - k)) & 1) << (k - 1)
    out_stages = (3, 4, 5, 6, 10, 13, 14, 16)          # MSB .. LSB of each byte
    feed_mask = 0
    for k in (3, 4, 5, 9, 10, 13, 15):
        feed_mask |= 1 << (k - 1)
    nby = (nbits + 7) // 8
    res = np.empty(nby * 8, np.uint8)
    for b in range(nby):
        for q, s in enumerate(out_stages):
            res[8 * b + q] = (r >> (s - 1)) & 1
        fb = r & 1
        r = (r >> 1) | (fb << 15)
        if fb:
            r ^= feed_mask
    return res[:nbits]


def baseband_packet(cw_bits, kldpc, g):
    mouter = g.bit_length() - 1
    nouter = np.asarray(cw_bits, np.uint8)[:kldpc]
    syn = syndrome(nouter, g)
    kpay = kldpc - mouter
    payload = nouter[:kpay] ^ sequence(kpay)
    return syn, np.packbits(payload).tobytes()
