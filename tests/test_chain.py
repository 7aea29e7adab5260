"""The whole receive chain on the device, frame cells to Baseband Packet
bytes, in the order of the upstream's frame decode (lab/m9_fast.py:1040-1160):
common-phase-error correction of the cell pool, the hybrid time
de-interleaver's gather fused into the demap, the alpha ladder inside the
batched decoder, the sum-product rescue of near misses, then BCH syndrome and
descrambling.  The frame is synthetic but built the upstream's transmit way
(lab/m6_bicm.py:367-373: payload XOR scrambling sequence, BCH, LDPC, BICM
map): a 16200-bit 64QAM 11/15 PLP of 74 FEC Blocks through the twisted block
interleaver, a QPSK 2/15 PLP, +-1 dummy cells, a per-symbol complex gain and
AWGN.  The check is the upstream's own verdict: every payload comes back
byte-identical with a zero BCH syndrome."""
import numpy as np
import pytest

from oracle import bch as OB
from oracle import ti as OT
from atsc3_amd import bicm, synth
from atsc3_amd.codes import get_code

pytestmark = pytest.mark.gpu

LADDER = (1.0, 0.85, 0.75)


def _plp(rate, mod, nblk, rng):
    from atsc3_amd.baseband import generator, scrambler_sequence
    code = get_code(rate, 16200)
    g, mouter = generator(16200)
    kpay = code.K - mouter
    pay = rng.integers(0, 2, (nblk, kpay), dtype=np.uint8)
    seq = scrambler_sequence(kpay)
    nouter = np.stack([OB.encode(p ^ seq, g) for p in pay])
    cw = code.encode(nouter)
    return code, pay, synth.modulate(cw, 16200, mod, rate)


def test_frame_to_baseband_packets(device):
    import torch
    from atsc3_amd import ldpc
    from atsc3_amd.baseband import Baseband
    from atsc3_amd.cpe import Cpe, dummy_values, frame_snr_db
    from atsc3_amd.demap import Demapper
    from atsc3_amd.rescue import rescue_dev
    from atsc3_amd.ti import HtiGather
    rng = np.random.default_rng(2024)
    code0, pay0, x0 = _plp("11/15", "64QAM", 74, rng)          # 74 x 2700 cells
    code16, pay16, x16 = _plp("2/15", "QPSK", 1, rng)          # 1 x 8100 cells
    n0c, n16, dstart = 74 * 2700, 8100, 74 * 2700 + 8100
    npool = dstart + 3572
    dv = dummy_values(npool, dstart)
    pool = np.concatenate([OT.tbi_interleave(x0.reshape(-1), 2700, 74, 0), x16[0],
                           dv[dstart:]]).astype(np.complex128)
    region = np.full(npool, 2, np.int8)
    region[:n0c] = 0
    region[n0c:dstart] = 1
    symbol_of = np.concatenate([np.full(2700, -1), np.arange(npool - 2700) // 5800])
    syms = np.unique(symbol_of)
    gs = (1 + rng.uniform(-0.1, 0.1, len(syms))) * np.exp(1j * rng.uniform(-0.1, 0.1, len(syms)))
    esn0 = 18.0
    noise = (rng.standard_normal(npool) + 1j * rng.standard_normal(npool)) * np.sqrt(
        synth.n0_of(esn0) / 2)
    y = (gs[np.searchsorted(syms, symbol_of)] * pool + noise).astype(np.complex64)
    print("[gpu] frame chain: CPE, HTI-gathered demap, ladder, rescue, baseband", flush=True)
    cpe = Cpe(symbol_of, region, [bicm.points("64QAM", "11/15"), bicm.points("QPSK", "2/15")],
              dv, device)
    z, _ = cpe.run(torch.from_numpy(y).to(device))
    snr = frame_snr_db(z, dstart)
    assert snr is not None and abs(snr - esn0) < 1.5
    dm0 = Demapper(16200, "64QAM", "11/15", device)
    llr0, _ = HtiGather(n0c, 1, 2700, 74, 0, device).demap(dm0, z)
    bits, conv, its, bad = ldpc.ladder_dev(llr0, code0, LADDER, 50, engine="batch")
    rescue_dev(llr0, bits, conv, its, bad, code0, budget_s=None)
    syn, ok, payload = Baseband("11/15", 16200, device).run(bits)
    dm16 = Demapper(16200, "QPSK", "2/15", device)
    llr16, _ = dm16.demap_dev(z[n0c:dstart].reshape(1, n16))
    b16, c16, _, _ = ldpc.decode_dev(llr16, code16, 50, 0.75)
    _, ok16, payload16 = Baseband("2/15", 16200, device).run(b16)
    torch.cuda.synchronize()
    assert bool(conv.all()) and bool(ok.all())
    assert np.array_equal(payload.cpu().numpy(), np.packbits(pay0, axis=1))
    assert bool(c16.all()) and bool(ok16.all())
    assert np.array_equal(payload16.cpu().numpy(), np.packbits(pay16, axis=1))
