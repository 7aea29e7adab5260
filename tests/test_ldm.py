"""LDM two-layer chain (core QPSK 4/15 + enhanced 64-NUC 10/15, -5 dB injection).

CPU: the FP64 cancellation oracle is pinned by closed forms -- re-modulation
equals the transmitter's cell mapping, the gain is the least-squares
projection coefficient (the residual is orthogonal to the re-modulated core
symbols), an exact scaled copy cancels to zero, and non-decoded blocks pass
through.  The oracle chain decodes both layers at 20 dB Es/N0.  (At the
configuration's 18 dB the enhanced layer sits at Pe/N0 = 11.8 dB, below what
normalized min-sum needs for 64-NUC 10/15: the core layer decodes and is
cancelled, the enhanced blocks run all 50 iterations and fail; 20 dB clears
it.)

GPU: every stage of LdmPipeline against the oracle on the same inputs --
core demap and enhanced demap within 1e-5 normwise relative of the FP64
oracle, both decodes bit-exact against the reference min-sum on the GPU's own
LLRs, the cancellation residual within FP32 rounding of the FP64 oracle.
"""
import numpy as np
import pytest

from oracle import demap as OD
from oracle import ldm as OL
from oracle import ldpc as OP
from atsc3_amd import bicm, synth
from atsc3_amd.codes import get_code

from tests._fixtures import ensure_oracle_built, ref_lam_of_q

CORE = (64800, "QPSK", "4/15")
ENH = (64800, "64QAM", "10/15")
INJ_DB = 5.0
ESN0 = 18.0
ESN0_DEC = 20.0


def _lq(cfg):
    n, mod, rate = cfg
    return ref_lam_of_q(n, mod, rate)


def _superpose(ncore, seed, n0):
    pc, pe = OD.ldm_powers(INJ_DB)
    cw_c = synth.codewords(CORE[2], CORE[0], ncore, seed)
    cw_e = synth.codewords(ENH[2], ENH[0], 3 * ncore, seed + 1)
    xc = synth.modulate(cw_c, *CORE)
    xe = synth.modulate(cw_e, *ENH).reshape(ncore, -1)
    rng = np.random.default_rng(seed + 99)
    y = np.sqrt(pc) * xc + np.sqrt(pe) * xe + synth.complex_noise(rng, xc.shape) * np.sqrt(n0)
    return cw_c, cw_e, y


@pytest.mark.parametrize("cfg", [CORE, ENH, (16200, "256QAM", "9/15")])
def test_remodulate_equals_transmit_mapping(cfg):
    n, mod, rate = cfg
    cw = synth.codewords(rate, n, 2, 3)
    got = OL.remodulate(cw, _lq(cfg), bicm.points(mod, rate))
    np.testing.assert_array_equal(got, synth.modulate(cw, n, mod, rate))


def test_cancel_closed_forms():
    rng = np.random.default_rng(1)
    cw = synth.codewords(CORE[2], CORE[0], 3, 5)
    pts = bicm.points(CORE[1], CORE[2])
    s = OL.remodulate(cw, _lq(CORE), pts)
    g = np.array([0.8 - 0.1j, 1.3 + 0.4j, 0.5j])
    # an exact scaled copy cancels completely, gain recovered
    r, gain = OL.cancel(g[:, None] * s, cw, np.ones(3, bool), _lq(CORE), pts)
    np.testing.assert_allclose(gain, g, rtol=1e-13)
    assert np.abs(r).max() < 1e-12
    # with an arbitrary second signal: least-squares coefficient, orthogonal residual
    y = g[:, None] * s + synth.complex_noise(rng, s.shape) * 0.7
    r, gain = OL.cancel(y, cw, np.array([1, 0, 1], bool), _lq(CORE), pts, out_scale=2.0)
    for b in (0, 2):
        ls = np.linalg.lstsq(s[b][:, None], y[b], rcond=None)[0][0]
        assert abs(gain[b] - ls) < 1e-12
        assert abs(np.vdot(s[b], r[b] / 2.0)) < 1e-8 * np.vdot(s[b], s[b]).real
    assert gain[1] == 0
    np.testing.assert_array_equal(r[1], 2.0 * y[1])


def test_oracle_chain_decodes_both_layers():
    ensure_oracle_built()
    n0 = synth.n0_of(ESN0_DEC)
    pc, pe = OD.ldm_powers(INJ_DB)
    cw_c, cw_e, y = _superpose(1, 11, n0)
    z, s2 = OD.ldm_core_inputs(y, INJ_DB, n0)
    llr_c, _ = OD.demap(z, bicm.points(CORE[1], CORE[2]), _lq(CORE), sigma2=s2)
    bits_c, conv_c, _, _ = OP.minsum_f32(llr_c, get_code(CORE[2], CORE[0]), 50, 0.75)
    assert conv_c.all() and (bits_c == cw_c).all()
    r, gain = OL.cancel(y, bits_c, conv_c, _lq(CORE), bicm.points(CORE[1], CORE[2]))
    assert abs(gain[0] - np.sqrt(pc)) < 0.01
    r = r.reshape(3, -1)
    llr_e, _ = OD.demap(r, np.sqrt(pe) * bicm.points(ENH[1], ENH[2]), _lq(ENH), sigma2=n0)
    bits_e, conv_e, _, _ = OP.minsum_f32(llr_e, get_code(ENH[2], ENH[0]), 50, 0.75)
    assert conv_e.all() and (bits_e == cw_e).all()


@pytest.mark.gpu
def test_cancel_kernel_vs_oracle(device):
    import torch
    from atsc3_amd.ldm import LdmPipeline
    print("[gpu] ldm cancel kernel", flush=True)
    pipe = LdmPipeline(device=device)
    n0 = synth.n0_of(ESN0)
    cw_c, _, y = _superpose(4, 21, n0)
    y = y.astype(np.complex64)
    conv = np.array([1, 0, 1, 1], np.uint8)
    cw = cw_c.copy()
    cw[3, ::7] ^= 1                                   # a wrong codeword still cancels as given
    for scale in (1.0, 1.7):
        r, gain = pipe.cancel(torch.from_numpy(y).to(device), torch.from_numpy(cw).to(device),
                              torch.from_numpy(conv).to(device), out_scale=scale)
        torch.cuda.synchronize()
        rw, gw = OL.cancel(y.astype(np.complex128), cw, conv.astype(bool), _lq(CORE),
                           bicm.points(CORE[1], CORE[2]).astype(np.complex64).astype(np.complex128),
                           out_scale=scale)
        np.testing.assert_allclose(gain.cpu().numpy(), gw, rtol=1e-9, atol=1e-12)
        err = np.abs(r.cpu().numpy() - rw).max(axis=1) / np.abs(rw).max(axis=1)
        assert (err < 4e-7).all(), err


@pytest.mark.gpu
def test_cancel_empty(device):
    import torch
    from atsc3_amd.ldm import LdmPipeline
    pipe = LdmPipeline(device=device)
    y = torch.zeros((0, pipe.ncell_c), dtype=torch.complex64, device=device)
    cw = torch.zeros((0, 64800), dtype=torch.uint8, device=device)
    conv = torch.zeros(0, dtype=torch.uint8, device=device)
    r, g = pipe.cancel(y, cw, conv)
    assert r.shape == y.shape and g.shape == (0,)


@pytest.mark.gpu
@pytest.mark.parametrize("esn0,seed", [(ESN0, 7), (ESN0_DEC, 9), (15.5, 8)])
def test_pipeline_stages_vs_oracle(device, esn0, seed):
    import torch
    from atsc3_amd.ldm import LdmPipeline
    print(f"[gpu] ldm pipeline stages esn0={esn0} seed={seed}", flush=True)
    ensure_oracle_built()
    pipe = LdmPipeline(device=device)
    n0 = synth.n0_of(esn0)
    pc, pe = OD.ldm_powers(INJ_DB)
    cw_c, cw_e, y = _superpose(3, seed, n0)
    y = y.astype(np.complex64)
    out = pipe.run(torch.from_numpy(y).to(device), n0)
    torch.cuda.synchronize()
    y64 = y.astype(np.complex128)
    # core demap: same LLR as demapping y/sqrt(Pc) with sigma2 (Pe+N0)/Pc
    z, s2 = OD.ldm_core_inputs(y64, INJ_DB, n0)
    want, _ = OD.demap(z, bicm.points(CORE[1], CORE[2]), _lq(CORE), sigma2=s2)
    llr_c = out["llr_core"].cpu().numpy()
    assert (OD.normwise_rel_err(llr_c.astype(np.float64), want) <= 1e-5).all()
    # core decode: bit-exact against the reference min-sum on the GPU's LLRs
    bits, conv, its, bad = OP.minsum_f32(llr_c, get_code(CORE[2], CORE[0]), 50, 0.75)
    g = [t.cpu().numpy() for t in out["core"]]
    np.testing.assert_array_equal(g[0], bits)
    np.testing.assert_array_equal(g[1].astype(bool), conv)
    np.testing.assert_array_equal(g[2], its)
    np.testing.assert_array_equal(g[3], bad)
    # cancellation
    pts_c = bicm.points(CORE[1], CORE[2]).astype(np.complex64).astype(np.complex128)
    rw, gw = OL.cancel(y64, g[0], g[1].astype(bool), _lq(CORE), pts_c)
    np.testing.assert_allclose(out["gain"].cpu().numpy(), gw, rtol=1e-9, atol=1e-12)
    r = out["resid"].cpu().numpy()
    err = np.abs(r - rw).max(axis=1) / np.abs(rw).max(axis=1)
    assert (err < 4e-7).all(), err
    # enhanced demap of the GPU residual, then bit-exact enhanced decode
    r_e = r.reshape(-1, pipe.ncell_e).astype(np.complex128)
    want_e, _ = OD.demap(r_e, np.sqrt(pe) * bicm.points(ENH[1], ENH[2]), _lq(ENH), sigma2=n0)
    llr_e = out["llr_enh"].cpu().numpy()
    assert (OD.normwise_rel_err(llr_e.astype(np.float64), want_e) <= 1e-5).all()
    bits, conv, its, bad = OP.minsum_f32(llr_e, get_code(ENH[2], ENH[0]), 50, 0.75)
    g = [t.cpu().numpy() for t in out["enh"]]
    np.testing.assert_array_equal(g[0], bits)
    np.testing.assert_array_equal(g[1].astype(bool), conv)
    np.testing.assert_array_equal(g[2], its)
    np.testing.assert_array_equal(g[3], bad)
    if esn0 == ESN0_DEC:
        assert out["core"][1].all() and out["enh"][1].all()
        assert (out["enh"][0].cpu().numpy() == cw_e).all()
