"""Exact sum-product rescue: the float64 oracle pinned by mathematics
(tree-code marginals, the box-plus closed form, numpy's summation order)
and by the upstream's own gates (a real codeword decodes, a shuffled
control does not, a deadline sheds), then the device decoder, selection
and write-back against it.

Tolerances: the device evaluates tanh/log/exp/atanh with its own
double-precision library, so totals are compared to 1e-9 (absolute plus
relative) only while no message saturates (small LLRs, few iterations);
decoded bits, convergence flags and iteration counts are compared exactly
on codewords that converge."""
import numpy as np
import pytest

from tests import _fixtures as F

F.ensure_oracle_built()

from oracle import ldpc as OL          # noqa: E402
from oracle import sp_ref as SP        # noqa: E402
from atsc3_amd import synth            # noqa: E402
from atsc3_amd.codes import get_code   # noqa: E402

LADDER = (1.0, 0.85, 0.75)


def _pairwise8(a):
    """numpy's float64 pairwise order for a row of up to 128 terms (the
    order the device kernel uses)."""
    n = len(a)
    if n < 8:
        r = 0.0
        for x in a:
            r += x
        return r
    m = n - n % 8
    p = [0.0] * 8
    for k in range(m):
        p[k % 8] += a[k]
    r = ((p[0] + p[1]) + (p[2] + p[3])) + ((p[4] + p[5]) + (p[6] + p[7]))
    for k in range(m, n):
        r += a[k]
    return r


@pytest.mark.parametrize("width", [3, 7, 8, 10, 17, 34, 86])
def test_numpy_row_sum_is_pairwise8(width):
    rng = np.random.default_rng(width)
    x = np.log(rng.random((3, 400, width)))
    x[:, :, width // 2:] *= rng.random((3, 400, width - width // 2)) < 0.7   # padding zeros
    got = x.sum(axis=2, keepdims=True)[..., 0]
    want = np.array([_pairwise8(list(r)) for r in x.reshape(-1, width)]).reshape(got.shape)
    assert np.array_equal(got, want)


def test_degree3_message_is_boxplus():
    rng = np.random.default_rng(3)
    for _ in range(20):
        a, b, c = rng.uniform(-6, 6, 3)
        tot = SP.iterate(np.array([[a, b, c]]), [[0, 1, 2]], 3, 1)[0]
        bp = np.log((1 + np.exp(a + b)) / (np.exp(a) + np.exp(b)))
        assert abs((tot[2] - c) - bp) < 1e-12


def _posterior_llr(lam, checks, n):
    """Brute-force a-posteriori LLRs over every codeword of a tiny code."""
    words = ((np.arange(1 << n)[:, None] >> np.arange(n)[None, :]) & 1).astype(np.uint8)
    ok = np.ones(len(words), bool)
    for c in checks:
        ok &= words[:, c].sum(1) % 2 == 0
    words = words[ok]
    logp = ((1 - 2.0 * words) * lam[None, :] / 2).sum(1)
    out = np.empty(n)
    for i in range(n):
        l0 = np.logaddexp.reduce(logp[words[:, i] == 0])
        l1 = np.logaddexp.reduce(logp[words[:, i] == 1])
        out[i] = l0 - l1
    return out


@pytest.mark.parametrize("seed", range(4))
def test_tree_code_totals_are_exact_marginals(seed):
    """On a cycle-free Tanner graph BP is exact after the graph's diameter."""
    checks = [[0, 1, 2], [2, 3, 4], [4, 5], [1, 6, 7], [7, 8, 9, 10]]
    n = 11
    rng = np.random.default_rng(seed)
    lam = rng.uniform(-3, 3, n)
    want = _posterior_llr(lam, checks, n)
    got = SP.iterate(lam[None, :], checks, n, 8)[0]
    assert np.abs(got - want).max() < 1e-9


def test_tree_check_catches_a_dropped_extrinsic_subtraction(monkeypatch):
    """The same check with a planted mistake in the oracle: messages formed
    from the full totals instead of total - message are not exact."""
    checks = [[0, 1, 2], [2, 3, 4], [4, 5], [1, 6, 7], [7, 8, 9, 10]]
    n = 11
    lam = np.random.default_rng(0).uniform(-3, 3, n)
    want = _posterior_llr(lam, checks, n)
    good = SP.messages
    monkeypatch.setattr(SP, "messages", lambda tot, m, idx, mask: good(tot, 0 * m, idx, mask))
    assert np.abs(SP.iterate(lam[None, :], checks, n, 8)[0] - want).max() > 1e-3


def test_oracle_decodes_and_shuffled_control_fails():
    """lab/gate_e58.py:165-190 / lab/gate_e60.py:283-300, on the 16200-bit
    8/15 code: near-threshold codewords converge to the sent bits where the
    min-sum ladder fails; the same LLRs shuffled do not converge."""
    code = get_code("8/15", 16200)
    cw = synth.codewords("8/15", 16200, 3, 11)
    llr = synth.llr_bpsk_like(cw, 1.4, 12)
    b, c, i, d, s = SP.sum_product(llr, code.checks, 16200, 100)
    assert c.all() and not s.any() and (d == 0).all()
    assert np.array_equal(b, cw)
    lad = OL.ladder(llr, code, LADDER, 50)
    assert not lad[1].all()
    shuf = llr[:, np.random.default_rng(5).permutation(16200)]
    b2, c2, i2, d2, s2 = SP.sum_product(shuf[:1], code.checks, 16200, 30)
    assert not c2.any() and i2[0] == 30 and d2[0] > 0


def test_oracle_stop_sheds_rows():
    code = get_code("11/15", 16200)
    lam = np.random.default_rng(6).standard_normal((3, 16200)) * 3.0
    b, c, i, d, s = SP.sum_product(lam, code.checks, 16200, 100, stop_at=2)
    assert not c.any() and s.all() and (i == 2).all() and (d > 0).all()


def test_select_policy():
    conv = np.array([0, 1, 0, 0, 0, 0, 0], bool)
    bad = np.array([9, 0, 3, 1600, 3, 1500, 1], np.int32)
    assert SP.select(conv, bad, 1500, 8).tolist() == [6, 2, 4, 0, 5]
    assert SP.select(conv, bad, 1500, 2).tolist() == [6, 2]
    assert SP.select(conv, bad, 5, 8).tolist() == [6, 2, 4]


# ---------------------------------------------------------------------------
# device
# ---------------------------------------------------------------------------

def _near_threshold(rate, snr, nblk, seed):
    code = get_code(rate, 16200)
    cw = synth.codewords(rate, 16200, nblk, seed)
    return code, cw, synth.llr_bpsk_like(cw, snr, seed + 1)


@pytest.mark.gpu
@pytest.mark.parametrize("k", [1, 2, 3])
def test_totals_match_oracle_before_saturation(device, k):
    """Small LLRs, k message updates (iters = k on a non-converging input:
    the totals of iteration k are what the rows leave with)."""
    import torch
    from atsc3_amd.rescue import SumProduct
    code = get_code("11/15", 16200)
    lam = (np.random.default_rng(60).standard_normal((2, 16200)) * 0.5).astype(np.float32)
    sp = SumProduct(code, device)
    print(f"[gpu] sum-product totals k={k}", flush=True)
    b, c, i, d, s = sp.decode(torch.from_numpy(lam).to(device), iters=k)
    torch.cuda.synchronize()
    tot = sp.workspace(2)[0].view(2, 16200).cpu().numpy()
    want = SP.iterate(lam, code.checks, 16200, k)
    err = np.abs(tot - want) / (1.0 + np.abs(want))
    assert err.max() < 1e-9, err.max()
    wb, wc, wi, wd, ws = SP.sum_product(lam, code.checks, 16200, k)
    assert np.array_equal(b.cpu().numpy(), wb) and np.array_equal(c.cpu().numpy(), wc)
    assert np.array_equal(i.cpu().numpy(), wi) and np.array_equal(d.cpu().numpy(), wd)


@pytest.mark.gpu
@pytest.mark.parametrize("rate,snr", [("8/15", 1.4), ("11/15", 3.6)])
def test_decode_matches_oracle(device, rate, snr):
    import torch
    from atsc3_amd.rescue import SumProduct
    code, cw, llr = _near_threshold(rate, snr, 6, 11)
    want = SP.sum_product(llr, code.checks, 16200, 100)
    sp = SumProduct(code, device)
    print(f"[gpu] sum-product decode {rate} {snr} dB", flush=True)
    got = [x.cpu().numpy() for x in sp.decode(torch.from_numpy(llr).to(device), iters=100)]
    torch.cuda.synchronize()
    assert want[1].all()
    for name, g, w in zip(("bits", "conv", "its", "bad", "shed"), got, want):
        assert np.array_equal(g.astype(w.dtype), w), name


@pytest.mark.gpu
def test_host_dropin_and_shuffled_control(device):
    from atsc3_amd.rescue import sum_product_decode_batch
    code, cw, llr = _near_threshold("8/15", 1.4, 2, 11)
    b, c, i, d = sum_product_decode_batch(llr, code.checks, 16200, iters=100)
    assert c.all() and np.array_equal(b, cw) and b.dtype == np.uint8 and c.dtype == bool
    shuf = llr[:, np.random.default_rng(5).permutation(16200)]
    b2, c2, i2, d2 = sum_product_decode_batch(shuf, code.checks, 16200, iters=30)
    assert not c2.any() and (i2 == 30).all() and (d2 > 0).all()


@pytest.mark.gpu
def test_deadline_sheds(device):
    """lab/gate_e60.py:262-278: never-converging input under a spent budget
    leaves shed and counted; a generous budget changes nothing.  A budget
    shorter than one iteration stops every row at iteration 1 (the stop is
    decided at the end of iteration 0)."""
    import time
    from atsc3_amd.rescue import sum_product_decode_batch
    code = get_code("11/15", 16200)
    lam = np.random.default_rng(6).standard_normal((8, 16200)) * 3.0
    print("[gpu] sum-product deadline", flush=True)
    st = {}
    b, c, i, d = sum_product_decode_batch(lam, code.checks, 16200, iters=100,
                                          deadline=time.perf_counter() + 1e-7, stats=st)
    assert not c.any() and st.get("sp_timeout", 0) == 8 and (i == 1).all() and (d > 0).all()
    ref = sum_product_decode_batch(lam, code.checks, 16200, iters=100)
    st = {}
    got = sum_product_decode_batch(lam, code.checks, 16200, iters=100,
                                   deadline=time.perf_counter() + 60.0, stats=st)
    assert st.get("sp_timeout", 0) == 0 and (got[2] == 100).all()
    for x, y in zip(got, ref):
        assert np.array_equal(x, y)


@pytest.mark.gpu
def test_rescue_after_ladder_matches_oracle(device):
    """Ladder in the batched kernel, then select / decode / merge on the
    device: equal to the oracle ladder followed by the oracle rescue."""
    import torch
    from atsc3_amd import ldpc
    from atsc3_amd.rescue import rescue_dev
    code, cw, llr = _near_threshold("8/15", 1.4, 12, 11)
    lad = OL.ladder(llr, code, LADDER, 50)
    want = SP.rescue(llr, code.checks, 16200, *lad, near_max=40, max_blocks=4)
    assert 0 < len(want[4]) == 4 and want[1].sum() > lad[1].sum()
    t = torch.from_numpy(llr).to(device)
    print("[gpu] rescue after ladder", flush=True)
    bits, conv, its, bad = ldpc.ladder_dev(t, code, LADDER, 50, engine="batch")
    rows, count, c2, s2 = rescue_dev(t, bits, conv, its, bad, code, near_max=40, max_blocks=4,
                                     budget_s=None)
    torch.cuda.synchronize()
    assert int(count.item()) == len(want[4])
    assert rows.cpu().numpy()[:len(want[4])].tolist() == want[4].tolist()
    for name, g, w in zip(("bits", "conv", "its", "bad"), (bits, conv, its, bad), want[:4]):
        assert np.array_equal(g.cpu().numpy().astype(w.dtype), w), name


@pytest.mark.gpu
def test_rescue_in_mixed_batch_only_touches_its_code(device):
    import torch
    from atsc3_amd.batch import BatchPlan
    from atsc3_amd.rescue import rescue_dev
    code_a, cw_a, llr_a = _near_threshold("8/15", 1.4, 4, 11)
    code_b, cw_b, llr_b = _near_threshold("11/15", 3.6, 4, 21)
    codes = [code_a, code_b] * 4
    llrs = [x for pair in zip(llr_a, llr_b) for x in pair]
    plan = BatchPlan(codes, device, 50, 0.75, ladder=LADDER)
    for b, l in enumerate(llrs):
        plan.llr_view(b).copy_(torch.from_numpy(l))
    plan.run()
    before = (plan.conv.clone(), plan.bad.clone())
    ia = plan.bank_index(code_a)
    rows, count, c2, _ = rescue_dev(plan.llr, plan.bits, plan.conv, plan.its, plan.bad, code_a,
                                    near_max=1500, max_blocks=8, budget_s=None,
                                    row_off=plan.blk_off, blk_code=plan.blk_code, code_id=ia)
    torch.cuda.synchronize()
    plan.check()
    nsel = int(count.item())
    sel = rows.cpu().numpy()[:nsel]
    assert all(plan.code_idx[s] == ia for s in sel)
    conv0, bad0 = (x.cpu().numpy() for x in before)
    want_sel = SP.select(conv0.astype(bool) | (plan.code_idx != ia), bad0, 1500, 8)
    assert sel.tolist() == want_sel.tolist()
    conv = plan.conv.cpu().numpy()
    bits = plan.bits.cpu().numpy()
    for b in range(len(codes)):
        if plan.code_idx[b] != ia:
            assert conv[b] == conv0[b]
        elif conv[b]:
            assert np.array_equal(bits[plan.off[b]:plan.off[b + 1]], cw_a[b // 2])


@pytest.mark.gpu
def test_widest_checks_degree_86(device):
    """64800 r7/15 has checks of degree 86: the second sign word and an
    80 + 6 pairwise row sum over the padded width."""
    import torch
    from atsc3_amd.rescue import SumProduct
    code = get_code("7/15", 64800)
    sp = SumProduct(code, device)
    assert sp.dmax == 86
    lam = (np.random.default_rng(70).standard_normal((2, 64800)) * 0.5).astype(np.float32)
    print("[gpu] sum-product degree 86 totals", flush=True)
    sp.decode(torch.from_numpy(lam).to(device), iters=2)
    torch.cuda.synchronize()
    tot = sp.workspace(2)[0].view(2, 64800).cpu().numpy()
    want = SP.iterate(lam, code.checks, 64800, 2)
    assert (np.abs(tot - want) / (1.0 + np.abs(want))).max() < 1e-9
    cw = synth.codewords("7/15", 64800, 2, 71)
    llr = synth.llr_bpsk_like(cw, 1.6, 72)
    want = SP.sum_product(llr, code.checks, 64800, 30)
    got = [x.cpu().numpy() for x in sp.decode(torch.from_numpy(llr).to(device), iters=30)]
    assert want[1].all() and np.array_equal(want[0], cw)
    for name, g, w in zip(("bits", "conv", "its", "bad", "shed"), got, want):
        assert np.array_equal(g.astype(w.dtype), w), name


@pytest.mark.gpu
def test_rescue_with_nothing_to_offer(device):
    """An all-converged batch and an empty batch: nothing is offered and the
    batch results are untouched."""
    import torch
    from atsc3_amd import ldpc
    from atsc3_amd.rescue import rescue_dev
    code, cw, llr = _near_threshold("8/15", 4.0, 3, 81)
    t = torch.from_numpy(llr).to(device)
    bits, conv, its, bad = ldpc.decode_dev(t, code, 50, 0.75)
    before = [x.clone() for x in (bits, conv, its, bad)]
    rows, count, _, _ = rescue_dev(t, bits, conv, its, bad, code, budget_s=None)
    torch.cuda.synchronize()
    assert int(count.item()) == 0 and bool(conv.all())
    rows, count, c2, _ = rescue_dev(t, bits, conv, its, bad, code, budget_s=None,
                                    skip_if_idle=True)
    assert int(count.item()) == 0 and not bool(c2.any())
    for x, y in zip(before, (bits, conv, its, bad)):
        assert torch.equal(x, y)
    e = torch.empty((0, 16200), dtype=torch.float32, device=device)
    eb = torch.empty((0, 16200), dtype=torch.uint8, device=device)
    ec = torch.empty(0, dtype=torch.uint8, device=device)
    ei = torch.empty(0, dtype=torch.int32, device=device)
    rows, count, _, _ = rescue_dev(e, eb, ec, ei, ei.clone(), code, budget_s=None)
    torch.cuda.synchronize()
    assert int(count.item()) == 0
