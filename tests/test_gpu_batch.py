"""Batched mixed-code engine (one persistent launch per batch): per codeword
bit-exact with the reference min-sum in FP32 -- the golden fixtures of every
code decoded together in mixed batches, live mixed frames against the
oracle, tiny slot/grid counts that force queueing and slot recycling -- and
identical to the per-code engines in FP16 / INT8 storage.  HIP-graph replays
are checked against eager runs."""
import numpy as np
import pytest

from tests import _fixtures as F

F.ensure_oracle_built()

from oracle import ldpc as OL          # noqa: E402
from atsc3_amd import synth            # noqa: E402
from atsc3_amd.codes import get_code   # noqa: E402

pytestmark = pytest.mark.gpu

MIX = [(64800, "6/15"), (64800, "8/15"), (64800, "10/15"), (64800, "12/15"),
       (16200, "4/15"), (16200, "9/15"), (64800, "4/15")]
SNR = {"4/15": -1.6, "6/15": 0.6, "8/15": 1.6, "9/15": 2.2, "10/15": 2.4, "12/15": 4.4}


def _mixed_llrs(nblk, seed, spread=0.6):
    """nblk codewords of random codes from MIX, SNRs around each threshold
    (iteration counts from 0 to the limit, some failures)."""
    rng = np.random.default_rng(seed)
    codes, llrs = [], []
    for b in range(nblk):
        n, rate = MIX[rng.integers(len(MIX))]
        cw = synth.codewords(rate, n, 1, seed * 1000 + b)
        snr = SNR[rate] + rng.uniform(-spread, 3 * spread)
        codes.append(get_code(rate, n))
        llrs.append(synth.llr_bpsk_like(cw, snr, seed * 1000 + b + 7)[0])
    return codes, llrs


def _batch(codes, llrs, device, iters, alpha, msg="float32", slots=None, grid=None,
           ladder=None, xcd=False):
    import torch
    from atsc3_amd.batch import BatchPlan
    plan = BatchPlan(codes, device, iters, alpha, msg, slots, grid, ladder=ladder, xcd=xcd)
    for b, l in enumerate(llrs):
        plan.llr_view(b).copy_(torch.from_numpy(np.ascontiguousarray(l, np.float32)))
    plan.run()
    torch.cuda.synchronize()
    plan.check()
    bits = plan.bits.cpu().numpy()
    out = [(bits[plan.off[b]:plan.off[b + 1]], bool(c), int(i), int(d))
           for b, (c, i, d) in enumerate(zip(plan.conv.cpu().numpy(), plan.its.cpu().numpy(),
                                             plan.bad.cpu().numpy()))]
    return out, plan


def _assert_block(got, want, tag):
    bits, conv, its, bad = got
    assert conv == bool(want[1]), f"{tag} conv"
    assert its == int(want[2]), f"{tag} its {its} vs {want[2]}"
    assert bad == int(want[3]), f"{tag} bad {bad} vs {want[3]}"
    assert np.array_equal(bits, want[0]), f"{tag} bits"


def test_fixtures_in_mixed_batches(device):
    """Every golden fixture block, grouped by (iters, alpha), decoded as one
    mixed-code batch, equals the stored reference results."""
    groups = {}
    for name in F.fixture_names():
        code, llr, iters, alpha, want = F.load(name)
        g = groups.setdefault((iters, float(alpha)), [])
        for b in range(len(llr)):
            g.append((name, b, code, llr[b], (want["bits"][b], want["conv"][b], want["its"][b],
                                               want["bad"][b])))
    for (iters, alpha), items in groups.items():
        print(f"[gpu] batch fixtures iters {iters} alpha {alpha}: {len(items)} blocks", flush=True)
        for slots, grid in ((None, None), (3, 16)):
            got, _ = _batch([x[2] for x in items], [x[3] for x in items], device, iters, alpha,
                            slots=slots, grid=grid)
            for x, g in zip(items, got):
                _assert_block(g, x[4], f"{x[0]}[{x[1]}] slots={slots}")


@pytest.mark.parametrize("nblk,seed,slots,grid", [
    (30, 1, None, None), (30, 2, 4, 32), (45, 3, 1, 8), (12, 4, 64, 1)])
def test_live_mixed_vs_oracle(device, nblk, seed, slots, grid):
    codes, llrs = _mixed_llrs(nblk, seed)
    print(f"[gpu] batch live nblk {nblk} seed {seed} slots {slots} grid {grid}", flush=True)
    got, plan = _batch(codes, llrs, device, 50, 0.75, slots=slots, grid=grid)
    for b, (c, l) in enumerate(zip(codes, llrs)):
        w = OL.minsum_f32(l[None], c, 50, 0.75)
        _assert_block(got[b], (w[0][0], w[1][0], w[2][0], w[3][0]), f"block {b} ({c.n} {c.rate})")
    assert plan.rounds() > 0


@pytest.mark.parametrize("msg", ["float16", "int8"])
def test_low_precision_matches_phase_engine(device, msg):
    """FP16 / INT8 storage: same arithmetic as the per-code engines, so the
    batched results equal decode_dev per code group (fused engine)."""
    import torch
    from atsc3_amd import ldpc
    codes, llrs = _mixed_llrs(36, 11)
    print(f"[gpu] batch {msg} vs per-code engine", flush=True)
    got, _ = _batch(codes, llrs, device, 50, 0.75, msg=msg, slots=7)
    for c in {id(c): c for c in codes}.values():
        idx = [b for b in range(len(codes)) if codes[b] is c]
        t = torch.from_numpy(np.stack([llrs[b] for b in idx]).astype(np.float32)).to(device)
        ref = [x.cpu().numpy() for x in ldpc.decode_dev(t, c, 50, 0.75, msg, "qc")]
        for r, b in enumerate(idx):
            _assert_block(got[b], (ref[0][r], ref[1][r], ref[2][r], ref[3][r]), f"{msg} block {b}")


def test_graph_replay_matches_eager(device):
    import torch
    from atsc3_amd.batch import BatchPlan
    codes, _ = _mixed_llrs(20, 21)
    plan = BatchPlan(codes, device, 50, 0.75, "float32", slots=6)
    src = torch.empty_like(plan.llr)
    g = plan.capture(pre=lambda: plan.llr.copy_(src))
    for rep in range(3):
        _, llrs = _mixed_llrs(20, 21)  # same codes, fresh noise per replay below
        rng = np.random.default_rng(100 + rep)
        flat = np.concatenate(llrs).astype(np.float32) * np.float32(rng.uniform(0.7, 1.3))
        src.copy_(torch.from_numpy(flat))
        print(f"[gpu] graph replay {rep}", flush=True)
        g.replay()
        torch.cuda.synchronize()
        plan.check()
        gb, gc, gi, gd = (x.clone() for x in (plan.bits, plan.conv, plan.its, plan.bad))
        junk = torch.randn(1 << 20, device=device).sum()     # eager work between replays
        torch.cuda.synchronize()
        del junk
        plan.llr.copy_(src)
        plan.run()
        torch.cuda.synchronize()
        plan.check()
        assert torch.equal(gb, plan.bits) and torch.equal(gc, plan.conv)
        assert torch.equal(gi, plan.its) and torch.equal(gd, plan.bad)
        # and the replay equals the oracle on a sample
        for b in (0, 7, 19):
            w = OL.minsum_f32(flat[plan.off[b]:plan.off[b + 1]][None], codes[b], 50, 0.75)
            assert int(gi[b]) == int(w[2][0]) and bool(gc[b]) == bool(w[1][0])
            assert np.array_equal(gb[plan.off[b]:plan.off[b + 1]].cpu().numpy(), w[0][0])


def test_grid_beyond_capacity_rejected(device):
    from atsc3_amd.batch import BatchPlan, max_grid
    cap = max_grid("float32")
    assert cap >= 256
    with pytest.raises(ValueError):
        BatchPlan([get_code("10/15", 64800)], device, grid=cap + 1)


# ---------------------------------------------------------------------------
# alpha ladder inside the kernel (lab/m9_fast.py:843-891)
# ---------------------------------------------------------------------------

LADDER = (1.0, 0.85, 0.75)


def _ladder_case():
    """24 blocks of the 16200-bit 8/15 code at 1.6 dB: with this ladder some
    converge at each rung, some never (and some of those take a later rung's
    fewer unsatisfied checks)."""
    code = get_code("8/15", 16200)
    cw = synth.codewords("8/15", 16200, 24, 204)
    return code, synth.llr_bpsk_like(cw, 1.6, 205)


@pytest.mark.parametrize("slots,grid", [(None, None), (3, 16)])
def test_ladder_in_kernel_vs_oracle(device, slots, grid):
    code, llr = _ladder_case()
    want = OL.ladder(llr, code, LADDER, 50)
    r1 = OL.minsum_f32(llr, code, 50, LADDER[0])
    r2 = OL.minsum_f32(llr, code, 50, LADDER[1])
    # the case exercises every branch of the ladder: codewords converging
    # only at rung 2, only at rung 3, never (some of those keeping a later
    # rung's result)
    assert r1[1].sum() < want[1].sum() < len(llr)
    assert (want[1] & ~r1[1] & r2[1]).any() and (want[1] & ~r1[1] & ~r2[1]).any()
    assert ((~want[1]) & (want[3] < r1[3])).any()
    print(f"[gpu] ladder in kernel slots={slots} grid={grid}", flush=True)
    got, _ = _batch([code] * len(llr), list(llr), device, 50, LADDER[0], slots=slots,
                    grid=grid, ladder=LADDER)
    for b, g in enumerate(got):
        _assert_block(g, (want[0][b], want[1][b], want[2][b], want[3][b]), f"blk {b}")


def test_ladder_mixed_codes_vs_oracle(device):
    """A mixed batch under the ladder: per codeword the oracle ladder of its
    own code."""
    code_a, llr_a = _ladder_case()
    code_b = get_code("6/15", 64800)
    llr_b = synth.llr_bpsk_like(synth.codewords("6/15", 64800, 3, 31), 0.5, 32)
    codes = [code_a] * 6 + [code_b] * 3
    llrs = list(llr_a[:6]) + list(llr_b)
    wa = OL.ladder(llr_a[:6], code_a, LADDER, 50)
    wb = OL.ladder(llr_b, code_b, LADDER, 50)
    print("[gpu] ladder mixed codes", flush=True)
    got, _ = _batch(codes, llrs, device, 50, LADDER[0], slots=4, ladder=LADDER)
    for b, g in enumerate(got):
        w, k = (wa, b) if b < 6 else (wb, b - 6)
        _assert_block(g, (w[0][k], w[1][k], w[2][k], w[3][k]), f"blk {b}")


@pytest.mark.parametrize("msg", ["float16", "int8"])
def test_ladder_low_precision_matches_host_ladder(device, msg):
    """FP16 / INT8 storage: the in-kernel ladder equals one decode per rung
    on the per-code engine with the selection done on the host."""
    import torch
    from atsc3_amd import ldpc
    code, llr = _ladder_case()
    t = torch.from_numpy(np.ascontiguousarray(llr, np.float32)).to(device)
    print(f"[gpu] ladder {msg}", flush=True)
    a = ldpc.ladder_dev(t, code, LADDER, 50, msg, engine="batch")
    b = ldpc.ladder_dev(t, code, LADDER, 50, msg, engine="qc")
    torch.cuda.synchronize()
    for x, y in zip(a, b):
        assert torch.equal(x, y)


def test_single_rung_ladder_is_plain_decode(device):
    import torch
    from atsc3_amd import ldpc
    code, llr = _ladder_case()
    t = torch.from_numpy(np.ascontiguousarray(llr, np.float32)).to(device)
    a = ldpc.ladder_dev(t, code, (0.85,), 50, engine="batch")
    b = ldpc.decode_dev(t, code, 50, 0.85, engine="qc")
    torch.cuda.synchronize()
    for x, y in zip(a, b):
        assert torch.equal(x, y)


def test_abort_marks_undecoded_codewords(device, monkeypatch):
    """A round limit lowered from the host (ATSC3_BATCH_MAX_ROUNDS) makes the
    compaction set the abort word; the kernel leaves normally, the status
    word is set, unfinished codewords read its = -1 and the host conversion
    raises instead of returning them."""
    import torch
    from atsc3_amd import ldpc
    from atsc3_amd.batch import BatchPlan
    code = get_code("8/15", 16200)
    llr = synth.llr_bpsk_like(synth.codewords("8/15", 16200, 6, 3), -1.0, 4)
    monkeypatch.setenv("ATSC3_BATCH_MAX_ROUNDS", "4")
    print("[gpu] forced abort (round limit 4)", flush=True)
    plan = BatchPlan([code] * 6, device, 50, 0.75, "float16")
    plan.llr.copy_(torch.from_numpy(llr.reshape(-1)))
    plan.run()
    torch.cuda.synchronize()
    with pytest.raises(RuntimeError):
        plan.check()
    assert (plan.its.cpu().numpy() == -1).all() and (plan.bad.cpu().numpy() == -1).all()
    with pytest.raises(RuntimeError):
        ldpc.min_sum_decode_batch(llr, code.checks, 16200, 50, 0.75, np.float32,
                                  msg_dtype="float16", engine="batch", device=device)
    monkeypatch.delenv("ATSC3_BATCH_MAX_ROUNDS")
    out = ldpc.min_sum_decode_batch(llr, code.checks, 16200, 50, 0.75, np.float32,
                                    msg_dtype="float16", engine="batch", device=device)
    assert (out[2] >= 0).all()
