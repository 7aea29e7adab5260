"""Resident min-sum, team mode (atsc3_minsum_resident_team_dev).

Each codeword is decoded by a team of workgroups on one XCD that split its
check and variable steps; the results must be those of the one-workgroup
resident kernel bit for bit, so every test compares with the oracle of the
resident decode (oracle/resident_ref.py, FP16 / INT8 storage), the
reference kernel's golden outputs (FP32 storage) or the C restatement of the
reference kernel (oracle/minsum_oracle.c):

* mixed codes in one queue (fixture blocks and synthetic blocks of the codes
  that take the generic paths) for team sizes 1, 3 (an uneven split of every
  XCD's workgroups: a smaller last team), 8 and 32 (a whole XCD per
  codeword), FP16, with the members' LDS copies of the posteriors and
  without them (direct: the team's rows read in L2);
* every golden fixture in FP32 with team size 8, one launch per (iters,
  alpha) setting, and the C oracle on the synthetic codes;
* INT8 with the oracle's steps given, and the device-derived steps equal to
  the explicit ones;
* fewer codewords than teams, one codeword, iters = 0, an empty batch, a
  queue order given as a permutation;
* the uneven-frame pipeline captured in a HIP graph and replayed: equal to
  the one-workgroup resident engine and, in FP32, to the batched engine.
"""
import numpy as np
import pytest

from oracle.resident_ref import int8_step, resident_minsum
from tests import _fixtures as F
from tests.test_resident import SYNTH, _case, _oracle

MIXED_KEYS = [("fixture", "g16k_r2"), ("synth", "13/15 64800 5.5"), ("fixture", "s16k_r9_near"),
              ("synth", "7/15 64800 0.8"), ("fixture", "g64k_r10_nuc20"),
              ("synth", "8/15 16200 2.0"), ("synth", "2/15 64800 -4.0"),
              ("fixture", "s16k_r4_a75")]


@pytest.fixture(scope="module", autouse=True)
def _oracle_built():
    F.ensure_oracle_built()


def _mixed_blocks(seed=5):
    blocks = []
    for kind, key in MIXED_KEYS:
        code, llr, iters, alpha = _case(kind, key)
        assert (iters, alpha) == (50, 0.75), key
        want = _oracle(kind, key)
        for b in range(len(llr)):
            blocks.append((code, llr[b], [w[b] for w in want], f"{key}#{b}"))
    order = np.random.default_rng(seed).permutation(len(blocks))
    return [blocks[i] for i in order]


def _run_plan(plan, blocks, device):
    import torch
    llr = np.concatenate([b[1] for b in blocks]).astype(np.float32)
    out = plan.run(torch.from_numpy(llr).to(device))
    torch.cuda.synchronize()
    plan.check()
    return [x.cpu().numpy() for x in out]


def _check_blocks(plan, blocks, got, tag):
    bits, conv, its, bad = got
    for i, (code, _, want, name) in enumerate(blocks):
        g = (bits[plan.off[i]:plan.off[i + 1]], bool(conv[i]), its[i], bad[i])
        assert g[1] == bool(want[1]) and g[2] == want[2] and g[3] == want[3], f"{tag} {name}"
        assert np.array_equal(g[0], want[0]), f"{tag} {name} bits"


@pytest.mark.gpu
@pytest.mark.parametrize("team,direct", [(1, False), (3, False), (8, False), (32, False),
                                         (1, True), (5, True), (16, True), (32, True)])
def test_team_fp16_mixed_codes_equal_oracle(device, team, direct):
    """direct: no LDS copies, every member reading the team's rows in L2."""
    from atsc3_amd.resident import ResidentPlan
    blocks = _mixed_blocks()
    plan = ResidentPlan([b[0] for b in blocks], device, 50, 0.75, msg_dtype="float16",
                        team=team, direct=direct)
    print(f"[gpu] team fp16 mixed {len(blocks)} blocks, team {team}, direct {direct}, "
          f"grid {plan.grid}", flush=True)
    _check_blocks(plan, blocks, _run_plan(plan, blocks, device), f"team {team} {direct}")


@pytest.mark.gpu
@pytest.mark.parametrize("direct", [False, True])
def test_team_fp32_equals_reference_fixtures(device, direct):
    """Every golden fixture, decoded in FP32 team mode (team 8), equals the
    reference kernel's outputs; one mixed launch per (iters, alpha)."""
    from atsc3_amd.resident import ResidentPlan
    groups = {}
    for name in F.fixture_names():
        code, llr, iters, alpha, want = F.load(name)
        for b in range(len(llr)):
            groups.setdefault((iters, alpha), []).append(
                (code, llr[b], [want[k][b] for k in ("bits", "conv", "its", "bad")],
                 f"{name}#{b}"))
    for (iters, alpha), blocks in sorted(groups.items()):
        plan = ResidentPlan([b[0] for b in blocks], device, iters, alpha, msg_dtype="float32",
                            team=8, direct=direct)
        print(f"[gpu] team fp32 fixtures iters {iters} alpha {alpha}: {len(blocks)} blocks, "
              f"direct {direct}", flush=True)
        _check_blocks(plan, blocks, _run_plan(plan, blocks, device), f"fp32 {iters} {alpha}")


@pytest.mark.gpu
def test_team_fp32_equals_c_oracle(device):
    """The synthetic codes of the generic paths (check degree > 16 and > 24,
    variable degree > 16, Type A) in one FP32 team launch (team 5)."""
    from oracle.ldpc import minsum_f32
    from atsc3_amd.resident import ResidentPlan
    blocks = []
    for r, n, s in SYNTH:
        code, llr, iters, alpha = _case("synth", f"{r} {n} {s}")
        want = minsum_f32(llr, code, iters, alpha)
        for b in range(len(llr)):
            blocks.append((code, llr[b], [np.asarray(w)[b] for w in want], f"{r} {n}#{b}"))
    for direct in (False, True):
        plan = ResidentPlan([b[0] for b in blocks], device, 50, 0.75, msg_dtype="float32",
                            team=5, direct=direct)
        print(f"[gpu] team fp32 synthetic codes {len(blocks)} blocks, direct {direct}",
              flush=True)
        _check_blocks(plan, blocks, _run_plan(plan, blocks, device), f"fp32 synth {direct}")


@pytest.mark.gpu
def test_team_int8_equals_oracle(device):
    """INT8 records with the oracle's per-codeword steps given: bit-exact;
    the steps derived on the device (each member derives its codeword's)
    give the same outputs."""
    from atsc3_amd.resident import ResidentPlan
    blocks = []
    for kind, key in [("fixture", "g16k_r4"), ("fixture", "s16k_r4_a75"),
                      ("fixture", "g64k_r10_nuc20"), ("synth", "13/15 64800 5.5"),
                      ("synth", "7/15 64800 0.8")]:
        code, llr, iters, alpha = _case(kind, key)
        assert (iters, alpha) == (50, 0.75), key
        steps = np.array([int8_step(r) for r in llr], np.float32)
        want = resident_minsum(llr, code, iters, alpha, "int8", steps=steps)
        for b in range(len(llr)):
            blocks.append((code, llr[b], [w[b] for w in want], f"{key}#{b}", steps[b]))
    steps = np.array([b[4] for b in blocks], np.float32)
    plan = ResidentPlan([b[0] for b in blocks], device, 50, 0.75, msg_dtype="int8",
                        msg_scale=steps, team=8)
    print(f"[gpu] team int8 {len(blocks)} blocks", flush=True)
    got = _run_plan(plan, blocks, device)
    _check_blocks(plan, [b[:4] for b in blocks], got, "int8")
    auto = ResidentPlan([b[0] for b in blocks], device, 50, 0.75, msg_dtype="int8", team=8)
    got2 = _run_plan(auto, blocks, device)
    for a, b in zip(got, got2):
        assert np.array_equal(a, b)
    l2 = ResidentPlan([b[0] for b in blocks], device, 50, 0.75, msg_dtype="int8", team=16,
                      direct=True)
    got3 = _run_plan(l2, blocks, device)
    for a, b in zip(got, got3):
        assert np.array_equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("msg", ["float16", "float32"])
def test_team_edge_cases(device, msg):
    """One codeword on a team of 32 (every other team finds the queue
    empty); iters = 0 (the syndrome of the channel decisions only); an
    empty batch; a queue order given as a permutation."""
    import torch
    from atsc3_amd.resident import ResidentPlan, decode_resident_dev
    code, llr, iters, alpha = _case("fixture", "g64k_r10_nuc20")
    t = torch.from_numpy(llr).to(device)
    print(f"[gpu] team edge cases {msg}", flush=True)
    ref = [x.cpu().numpy() for x in decode_resident_dev(t, code, iters, alpha, msg_dtype=msg)]
    one = decode_resident_dev(t[:1].contiguous(), code, iters, alpha, msg_dtype=msg, team=32)
    torch.cuda.synchronize()
    for a, b in zip(one, ref):
        assert np.array_equal(a.cpu().numpy(), b[:1])
    one = ResidentPlan([code], device, iters, alpha, msg_dtype=msg, team=32, direct=True)
    got = [x.cpu().numpy() for x in one.run(t[:1].reshape(-1))]
    one.check()
    for a, b in zip(got, ref):
        assert np.array_equal(a.reshape(b[:1].shape), b[:1])
    z_ref = [x.cpu().numpy() for x in decode_resident_dev(t, code, 0, alpha, msg_dtype=msg)]
    z = decode_resident_dev(t, code, 0, alpha, msg_dtype=msg, team=4)
    torch.cuda.synchronize()
    for a, b in zip(z, z_ref):
        assert np.array_equal(a.cpu().numpy(), b)
    empty = decode_resident_dev(t[:0].contiguous(), code, iters, alpha, msg_dtype=msg, team=8)
    assert all(x.shape[0] == 0 for x in empty)
    nblk = len(llr)
    order = np.random.default_rng(1).permutation(nblk)
    plan = ResidentPlan([code] * nblk, device, iters, alpha, msg_dtype=msg, team=2, order=order)
    got = [x.cpu().numpy() for x in plan.run(t.reshape(-1))]
    torch.cuda.synchronize()
    plan.check()
    assert np.array_equal(got[0].reshape(nblk, -1), ref[0])
    for a, b in zip(got[1:], ref[1:]):
        assert np.array_equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("msg", ["float32", "float16"])
def test_team_mixed_pipeline_graph_replay(device, msg):
    """Two uneven frames (every code and constellation of the mixed
    workload), demapped and decoded in one captured graph with the team
    engine, replayed three times with the outputs cleared in between: every
    replay equals the one-workgroup resident engine (and, in FP32, the
    batched engine)."""
    import torch
    from atsc3_amd.mixed import MixedPipeline, MixedSource, block_esn0, frame_mix
    specs = frame_mix(74, 3)
    outs = {}
    engines = ["resident", "team", "team-l2"] + (["batch"] if msg == "float32" else [])
    for eng in engines:
        print(f"[gpu] mixed pipeline {msg} {eng}", flush=True)
        p = MixedPipeline(specs, 2, device, 50, 0.75, msg,
                          engine="team" if eng.startswith("team") else eng,
                          team=8 if eng == "team" else 16, direct=eng == "team-l2")
        src = MixedSource(p, 8, 21)
        cells, s2, _ = src.frame(block_esn0(p.specs, 23))
        p.load(cells, s2)
        g = p.capture()
        for r in range(3 if eng.startswith("team") else 1):
            for x in (p.plan.bits, p.plan.conv, p.plan.its, p.plan.bad):
                x.fill_(0x5A if x.dtype == torch.uint8 else -7)
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            p.plan.check()
            cur = [x.clone() for x in (p.plan.bits, p.plan.conv, p.plan.its, p.plan.bad)]
            if eng.startswith("team") and r > 0:
                for a, b in zip(cur, outs[eng]):
                    assert torch.equal(a, b), f"{eng} replay {r}"
            outs[eng] = cur
        del g, p
    for eng in engines[1:]:
        for a, b in zip(outs[eng], outs["resident"]):
            assert torch.equal(a, b), eng


def _sampled_unsat_fraction(code, lam):
    """The team queue's difficulty key restated: the fraction of unsatisfied
    checks of the hard decisions over positions 0..63 of every check group."""
    q = code.qc()
    cg, be = q["cg_ptr"].astype(np.int64), q["be_word"].astype(np.int64)
    j = np.arange(64)
    par = np.zeros((len(cg) - 1, 64), np.int64)
    for g in range(len(cg) - 1):
        for b in range(cg[g], cg[g + 1]):
            s, t, miss = be[b] & 511, (be[b] >> 9) & 511, (be[b] >> 18) & 1023
            hard = (lam[t * 360 + (j + s) % 360] < 0).astype(np.int64)
            par[g] ^= np.where(j + 1 == miss, 0, hard)
    return np.float32(par.sum()) / np.float32(par.size)


@pytest.mark.gpu
@pytest.mark.parametrize("msg", ["float16", "float32"])
def test_team_auto_order(device, msg):
    """order="auto": the device ranks the queue by the sampled unsatisfied-
    check fraction of the channel decisions (equal to its numpy restatement,
    in decreasing order, ties by index); results equal the unranked decode
    (and, FP16, the oracle)."""
    from atsc3_amd.resident import ResidentPlan
    blocks = _mixed_blocks(seed=8)
    codes = [b[0] for b in blocks]
    plan = ResidentPlan(codes, device, 50, 0.75, msg_dtype=msg, team=8, order="auto",
                        direct=msg == "float16")
    print(f"[gpu] team auto order {msg} {len(blocks)} blocks", flush=True)
    got = _run_plan(plan, blocks, device)
    key = plan.order_key.cpu().numpy()
    order = plan.order.cpu().numpy()
    want_key = np.array([_sampled_unsat_fraction(b[0], b[1]) for b in blocks], np.float32)
    assert np.array_equal(key, want_key)
    assert sorted(order.tolist()) == list(range(len(blocks)))
    ranked = sorted(range(len(blocks)), key=lambda i: (-want_key[i], i))
    assert order.tolist() == ranked
    plain = ResidentPlan(codes, device, 50, 0.75, msg_dtype=msg, team=8)
    ref = _run_plan(plain, blocks, device)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
    if msg == "float16":
        _check_blocks(plan, blocks, got, "auto order")


@pytest.mark.gpu
def test_team_launches_on_two_streams(device):
    """Team decodes enqueued alternately on two streams (each team launch is
    ordered after the previous one on the device by an event, so two never
    hold compute units the other waits for): every result equals the
    one-stream decode; and decode_dev's automatic choice for a few FP32
    codewords (the team mode) equals the fused engine bit for bit."""
    import torch
    from atsc3_amd import ldpc
    code, llr, iters, alpha = _case("fixture", "g64k_r10_nuc20")
    t = torch.from_numpy(llr).to(device)
    print("[gpu] team launches on two streams", flush=True)
    ref = [x.cpu().numpy() for x in ldpc.decode_dev(t, code, iters, alpha, "float32", "team")]
    fused = [x.cpu().numpy() for x in ldpc.decode_dev(t, code, iters, alpha, "float32", "qc")]
    auto = [x.cpu().numpy() for x in ldpc.decode_dev(t, code, iters, alpha, "float32")]
    for a, b, c in zip(ref, fused, auto):
        assert np.array_equal(a, b) and np.array_equal(a, c)
    torch.cuda.synchronize()
    streams = (torch.cuda.Stream(device), torch.cuda.Stream(device))
    outs = []
    for k in range(8):
        with torch.cuda.stream(streams[k % 2]):
            outs.append(ldpc.decode_dev(t, code, iters, alpha, "float32", "team"))
    torch.cuda.synchronize()
    for o in outs:
        for a, b in zip(o, ref):
            assert np.array_equal(a.cpu().numpy(), b)


@pytest.mark.gpu
@pytest.mark.parametrize("msg", ["float32", "float16"])
def test_team_ladder_equals_rung_by_rung(device, msg):
    """The alpha ladder inside one team launch (1.0, 0.85, 0.75; 20
    iterations) equals the host-driven ladder -- one resident decode per
    rung on the rows still failing, a rung's result taken when it converged
    or left fewer unsatisfied checks -- on blocks spread over the waterfall
    (some converge at the first rung, some at a later one, some never); in
    FP32 also the batched engine's in-kernel ladder."""
    import torch
    from atsc3_amd import ldpc, synth
    from atsc3_amd.codes import get_code
    code = get_code("10/15", 64800)
    cw = synth.codewords("10/15", 64800, 24, 77)
    llr = np.concatenate([synth.llr_bpsk_like(cw[i:i + 1], 3.0 + 0.08 * i, 78 + i)
                          for i in range(24)]).astype(np.float32)
    t = torch.from_numpy(llr).to(device)
    lad = (1.0, 0.85, 0.75)
    print(f"[gpu] team ladder {msg}", flush=True)
    got = [x.cpu().numpy() for x in ldpc.ladder_dev(t, code, lad, 20, msg, engine="team")]
    ref = [x.cpu().numpy() for x in ldpc.ladder_dev(t, code, lad, 20, msg, engine="resident")]
    for a, b, nm in zip(got, ref, ("bits", "conv", "its", "bad")):
        assert np.array_equal(a, b), nm
    # the blocks exercise every branch of the ladder
    single = [x.cpu().numpy() for x in ldpc.decode_dev(t, code, 20, 1.0, msg, "resident")]
    assert single[1].any() and not single[1].all()
    assert (got[1].astype(bool) & ~single[1].astype(bool)).any()
    assert not got[1].all()
    if msg == "float32":
        bat = [x.cpu().numpy() for x in ldpc.ladder_dev(t, code, lad, 20, msg, engine="batch")]
        auto = [x.cpu().numpy() for x in ldpc.ladder_dev(t, code, lad, 20, msg)]
        for a, b, c in zip(got, bat, auto):
            assert np.array_equal(a, b) and np.array_equal(a, c)
