"""XCD-partitioned mode of the batched engine (atsc3_minsum_batch_xcd_dev):
each XCD decodes its own part of the queue with barriers among its own
workgroups.  Per codeword it must give exactly the grid-wide mode's results:
the golden fixtures of every code in mixed batches, live mixed batches
against the oracle (slot and grid counts that leave XCDs without codewords,
give one XCD several parts, or force queueing), FP16 / INT8 storage equal to
the grid-wide mode, the in-kernel ladder, graph replays and the abort path.
"""
import numpy as np
import pytest

from tests import _fixtures as F
from tests.test_gpu_batch import LADDER, _assert_block, _batch, _ladder_case, _mixed_llrs

F.ensure_oracle_built()

from oracle import ldpc as OL          # noqa: E402
from atsc3_amd import synth            # noqa: E402
from atsc3_amd.codes import get_code   # noqa: E402

pytestmark = pytest.mark.gpu


def test_fixtures_in_mixed_batches_xcd(device):
    groups = {}
    for name in F.fixture_names():
        code, llr, iters, alpha, want = F.load(name)
        g = groups.setdefault((iters, float(alpha)), [])
        for b in range(len(llr)):
            g.append((name, b, code, llr[b], (want["bits"][b], want["conv"][b], want["its"][b],
                                               want["bad"][b])))
    for (iters, alpha), items in groups.items():
        for slots, grid in ((None, None), (16, 3), (16, 24)):
            print(f"[gpu] xcd fixtures iters {iters} alpha {alpha}: {len(items)} blocks "
                  f"slots {slots} grid {grid}", flush=True)
            got, _ = _batch([x[2] for x in items], [x[3] for x in items], device, iters, alpha,
                            slots=slots, grid=grid, xcd=True)
            for x, g in zip(items, got):
                _assert_block(g, x[4], f"{x[0]}[{x[1]}] xcd slots={slots} grid={grid}")


@pytest.mark.parametrize("nblk,seed,slots,grid", [
    (30, 21, None, None), (5, 22, None, None), (40, 23, 16, 64), (24, 24, 16, 5),
    (36, 25, 64, 512)])
def test_live_mixed_xcd_vs_oracle(device, nblk, seed, slots, grid):
    """nblk 5: three of the eight parts are empty; slots 16 over 8 XCDs:
    2 slots each, so codewords queue; grid 5: five XCDs hold workgroups, so
    three of them decode two parts each."""
    codes, llrs = _mixed_llrs(nblk, seed)
    print(f"[gpu] xcd live nblk {nblk} seed {seed} slots {slots} grid {grid}", flush=True)
    got, plan = _batch(codes, llrs, device, 50, 0.75, slots=slots, grid=grid, xcd=True)
    for b, (c, l) in enumerate(zip(codes, llrs)):
        w = OL.minsum_f32(l[None], c, 50, 0.75)
        _assert_block(got[b], (w[0][0], w[1][0], w[2][0], w[3][0]), f"block {b} ({c.n} {c.rate})")
    assert plan.rounds() > 0


@pytest.mark.parametrize("msg", ["float16", "int8"])
def test_low_precision_xcd_equals_grid_wide(device, msg):
    codes, llrs = _mixed_llrs(36, 31)
    print(f"[gpu] xcd {msg} vs grid-wide", flush=True)
    a, _ = _batch(codes, llrs, device, 50, 0.75, msg=msg, slots=16, grid=128, xcd=True)
    b, _ = _batch(codes, llrs, device, 50, 0.75, msg=msg, slots=7)
    for k, (x, y) in enumerate(zip(a, b)):
        _assert_block(x, y, f"{msg} block {k}")


def test_ladder_xcd_vs_oracle(device):
    code, llr = _ladder_case()
    want = OL.ladder(llr, code, LADDER, 50)
    print("[gpu] xcd ladder", flush=True)
    got, _ = _batch([code] * len(llr), list(llr), device, 50, LADDER[0], slots=16, grid=40,
                    ladder=LADDER, xcd=True)
    for b, g in enumerate(got):
        _assert_block(g, (want[0][b], want[1][b], want[2][b], want[3][b]), f"blk {b}")


def test_graph_replay_xcd_matches_eager(device):
    import torch
    from atsc3_amd.batch import BatchPlan
    codes, llrs = _mixed_llrs(20, 41)
    plan = BatchPlan(codes, device, 50, 0.75, "float16", xcd=True)
    for b, l in enumerate(llrs):
        plan.llr_view(b).copy_(torch.from_numpy(np.ascontiguousarray(l, np.float32)))
    print("[gpu] xcd graph replay", flush=True)
    plan.run()
    torch.cuda.synchronize()
    plan.check()
    want = [x.clone() for x in (plan.bits, plan.conv, plan.its, plan.bad)]
    g = plan.capture()
    for _ in range(3):
        for x in (plan.bits, plan.its):
            x.fill_(7)
        g.replay()
        torch.cuda.synchronize()
        plan.check()
        for x, y in zip((plan.bits, plan.conv, plan.its, plan.bad), want):
            assert torch.equal(x, y)


def test_abort_xcd_marks_undecoded_codewords(device, monkeypatch):
    import torch
    from atsc3_amd.batch import BatchPlan
    code = get_code("8/15", 16200)
    llr = synth.llr_bpsk_like(synth.codewords("8/15", 16200, 12, 3), -1.0, 4)
    monkeypatch.setenv("ATSC3_BATCH_MAX_ROUNDS", "4")
    print("[gpu] xcd forced abort (round limit 4)", flush=True)
    plan = BatchPlan([code] * 12, device, 50, 0.75, "float32", xcd=True)
    plan.llr.copy_(torch.from_numpy(llr.reshape(-1)))
    plan.run()
    torch.cuda.synchronize()
    with pytest.raises(RuntimeError):
        plan.check()
    assert (plan.its.cpu().numpy() == -1).all() and (plan.bad.cpu().numpy() == -1).all()
