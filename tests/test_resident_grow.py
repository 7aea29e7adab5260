"""Resident min-sum, grow mode (atsc3_minsum_resident_team_dev with
ATSC3_TEAM_GROW).

Every codeword starts on one workgroup as the resident kernel decodes it;
once the queue is empty, a codeword still running takes team - 1 waiting
workgroups of its XCD and finishes as a team.  The results must be the
resident kernel's bit for bit, so the tests compare with the oracle of the
resident decode (oracle/resident_ref.py; FP16 / INT8 storage), the
reference kernel's golden outputs (FP32 storage) and the plain resident
engine:

* mixed codes in one queue, fewer codewords than workgroups (every
  codeword grows after its first iteration) and five times as many (the
  queue runs several rounds and only the last codewords grow), teams of 2,
  3 (requests of two seats: 32 workgroups of an XCD never fill them
  evenly), 8 and 32 (a whole XCD's workgroups for one codeword);
* every golden fixture in FP32 (teams of 4), three copies of each block;
* INT8 with the oracle's steps given and with the steps derived on the
  device;
* one codeword, iters 0 and 1 (no iteration left to grow into), an empty
  batch, a queue order given as a permutation;
* decode_dev's resident engine (grow mode, every format) on 300 blocks of
  64800 r10/15 over a spread of SNRs equals the plain resident kernel,
  with launches enqueued on two streams alternately.
"""
import numpy as np
import pytest

from oracle.resident_ref import int8_step, resident_minsum
from tests import _fixtures as F
from tests.test_resident import _case
from tests.test_resident_team import _check_blocks, _mixed_blocks, _run_plan


@pytest.fixture(scope="module", autouse=True)
def _oracle_built():
    F.ensure_oracle_built()


@pytest.mark.gpu
@pytest.mark.parametrize("team", [2, 3, 8, 32])
@pytest.mark.parametrize("reps", [1, 5])
def test_grow_fp16_mixed_codes_equal_oracle(device, team, reps):
    from atsc3_amd.resident import ResidentPlan
    blocks = _mixed_blocks() * reps
    plan = ResidentPlan([b[0] for b in blocks], device, 50, 0.75, msg_dtype="float16",
                        team=team, grow=True)
    print(f"[gpu] grow fp16 mixed {len(blocks)} blocks, team {team}, grid {plan.grid}",
          flush=True)
    _check_blocks(plan, blocks, _run_plan(plan, blocks, device), f"grow {team} x{reps}")


@pytest.mark.gpu
def test_grow_fp32_equals_reference_fixtures(device):
    """Every golden fixture, three copies of each block, decoded in FP32
    grow mode (teams of 4), equals the reference kernel's outputs; one
    mixed launch per (iters, alpha)."""
    from atsc3_amd.resident import ResidentPlan
    groups = {}
    for name in F.fixture_names():
        code, llr, iters, alpha, want = F.load(name)
        for b in range(len(llr)):
            groups.setdefault((iters, alpha), []).append(
                (code, llr[b], [want[k][b] for k in ("bits", "conv", "its", "bad")],
                 f"{name}#{b}"))
    for (iters, alpha), blocks in sorted(groups.items()):
        blocks = blocks * 3
        plan = ResidentPlan([b[0] for b in blocks], device, iters, alpha, msg_dtype="float32",
                            team=4, grow=True)
        print(f"[gpu] grow fp32 fixtures iters {iters} alpha {alpha}: {len(blocks)} blocks",
              flush=True)
        _check_blocks(plan, blocks, _run_plan(plan, blocks, device), f"fp32 {iters} {alpha}")


@pytest.mark.gpu
def test_grow_int8_equals_oracle(device):
    """INT8 records with the oracle's per-codeword steps: bit-exact (the
    owner hands its codeword's step to the team); the steps derived on the
    device give the same outputs."""
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
                        msg_scale=steps, team=8, grow=True)
    print(f"[gpu] grow int8 {len(blocks)} blocks", flush=True)
    got = _run_plan(plan, blocks, device)
    _check_blocks(plan, [b[:4] for b in blocks], got, "int8")
    auto = ResidentPlan([b[0] for b in blocks], device, 50, 0.75, msg_dtype="int8", team=8,
                        grow=True)
    got2 = _run_plan(auto, blocks, device)
    for a, b in zip(got, got2):
        assert np.array_equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("msg", ["float16", "float32"])
def test_grow_edge_cases(device, msg):
    """One codeword (31 waiting workgroups of its XCD join it); iters 0
    and 1; an empty batch; a queue order given as a permutation."""
    import torch
    from atsc3_amd.resident import ResidentPlan, decode_resident_dev
    code, llr, iters, alpha = _case("fixture", "g64k_r10_nuc20")
    t = torch.from_numpy(llr).to(device)
    print(f"[gpu] grow edge cases {msg}", flush=True)
    ref = [x.cpu().numpy() for x in decode_resident_dev(t, code, iters, alpha, msg_dtype=msg)]
    for team in (8, 32):
        one = decode_resident_dev(t[:1].contiguous(), code, iters, alpha, msg_dtype=msg,
                                  grow=team)
        torch.cuda.synchronize()
        for a, b in zip(one, ref):
            assert np.array_equal(a.cpu().numpy(), b[:1]), team
    for it in (0, 1):
        z_ref = [x.cpu().numpy() for x in decode_resident_dev(t, code, it, alpha, msg_dtype=msg)]
        z = decode_resident_dev(t, code, it, alpha, msg_dtype=msg, grow=4)
        torch.cuda.synchronize()
        for a, b in zip(z, z_ref):
            assert np.array_equal(a.cpu().numpy(), b), it
    empty = decode_resident_dev(t[:0].contiguous(), code, iters, alpha, msg_dtype=msg, grow=8)
    assert all(x.shape[0] == 0 for x in empty)
    nblk = len(llr)
    order = np.random.default_rng(3).permutation(nblk)
    plan = ResidentPlan([code] * nblk, device, iters, alpha, msg_dtype=msg, team=2, order=order,
                        grow=True)
    got = [x.cpu().numpy() for x in plan.run(t.reshape(-1))]
    torch.cuda.synchronize()
    plan.check()
    assert np.array_equal(got[0].reshape(nblk, -1), ref[0])
    for a, b in zip(got[1:], ref[1:]):
        assert np.array_equal(a, b)
    with pytest.raises(ValueError):
        ResidentPlan([code], device, iters, alpha, msg_dtype=msg, grow=True)
    with pytest.raises(ValueError):
        ResidentPlan([code], device, iters, alpha, msg_dtype=msg, team=4, grow=True,
                     ladder=(1.0, 0.75))


@pytest.mark.gpu
def test_grow_decode_dev_window_two_streams(device):
    """decode_dev's resident engine (grow mode, every format) on 300 blocks
    of 64800 r10/15 (iteration counts spread by SNR, so the queue's last
    codewords run long) equals the plain resident kernel; launches enqueued
    on two streams alternately give the same results."""
    import torch
    from atsc3_amd import ldpc, synth
    from atsc3_amd.codes import get_code
    from atsc3_amd.resident import decode_resident_dev
    code = get_code("10/15", 64800)
    cw = synth.codewords("10/15", 64800, 20, 91)
    rng = np.random.default_rng(92)
    snr = rng.uniform(3.3, 4.2, 300)
    llr = np.concatenate([synth.llr_bpsk_like(cw[i % 20:i % 20 + 1], snr[i], 1000 + i)
                          for i in range(300)]).astype(np.float32)
    t = torch.from_numpy(llr).to(device)
    for msg in ("float16", "int8", "float32"):
        print(f"[gpu] grow decode_dev window {msg}", flush=True)
        ref = [x.cpu().numpy() for x in decode_resident_dev(t, code, 50, 0.75, msg_dtype=msg)]
        print(f"[gpu] its median {np.median(ref[2])}, max {ref[2].max()}", flush=True)
        got = [x.cpu().numpy() for x in ldpc.decode_dev(t, code, 50, 0.75, msg, "resident")]
        for a, b in zip(got, ref):
            assert np.array_equal(a, b), msg
        streams = (torch.cuda.Stream(device), torch.cuda.Stream(device))
        outs = []
        for k in range(4):
            with torch.cuda.stream(streams[k % 2]):
                outs.append(ldpc.decode_dev(t, code, 50, 0.75, msg, "resident"))
        torch.cuda.synchronize()
        for o in outs:
            for a, b in zip(o, ref):
                assert np.array_equal(a.cpu().numpy(), b), msg
