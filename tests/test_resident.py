"""Resident FP16 min-sum (atsc3_minsum_resident_dev).

CPU: the numpy oracle of the resident decode (oracle/resident_ref.py), run
with FP32 storage, reproduces the upstream kernel's own golden outputs bit
for bit (bits, conv, its, bad) -- every code length, Type A and B, alpha
0.75/0.85/1.0, ties, +-inf/NaN LLRs; the packed tables the kernel reads
decode back to the code's QC base graph.

GPU: the kernel equals the oracle with FP16 storage bit for bit on the same
LLRs (golden-fixture LLRs, plus synthetic blocks of the codes that take the
generic paths: check degree > 16 and > 25, variable degree > 16, Type A),
one code per launch and mixed codes in one launch with fewer workgroups
than codewords; at the live-frame operating point (256-NUC r10/15, Rayleigh,
21 dB) every block decodes to the transmitted codeword, with mean iterations
within 3% of the plain FP32 decode; a HIP-graph replay equals the eager run.
"""
import numpy as np
import pytest

from oracle.resident_ref import resident_minsum
from tests import _fixtures as F

# fixtures fast enough for the CPU suite (the 64800-bit ones at 2-5 s each)
CPU_FIXTURES = [n for n in F.fixture_names() if n.startswith(("g16k", "s16k"))] + [
    "g64k_r10_nuc20", "g64k_r4", "g64k_r12"]
GPU_FIXTURES = ["g16k_r2", "g16k_r6", "g16k_r11", "s16k_r3_infnan", "s16k_r6_ties",
                "s16k_r4_a100", "s16k_r9_near", "g64k_r10_nuc20", "g64k_r10_nuc16", "g64k_r4"]
# (rate, n, snr_db): check degree 33 (third record word), 86 (Type A, > 64
# signs), variable degree 27 and 32 (generic variable path), Type A 2/15
SYNTH = [("13/15", 64800, 5.5), ("7/15", 64800, 0.8), ("6/15", 64800, 1.5),
         ("8/15", 16200, 2.0), ("2/15", 64800, -4.0), ("5/15", 64800, 0.0)]


@pytest.fixture(scope="module", autouse=True)
def _oracle_built():
    F.ensure_oracle_built()


@pytest.mark.parametrize("name", CPU_FIXTURES)
def test_oracle_fp32_storage_reproduces_reference_fixtures(name):
    code, llr, iters, alpha, want = F.load(name)
    got = resident_minsum(llr, code, iters, alpha, "float32")
    F.assert_same(got, want, name)


def test_oracle_fp16_storage_converges_to_the_same_codewords():
    """FP16 storage changes iteration counts at most slightly and decodes the
    converged blocks of a fixture to the same codewords as FP32."""
    code, llr, iters, alpha, want = F.load("g16k_r4")
    bits, conv, its, bad = resident_minsum(llr, code, iters, alpha, "float16")
    assert np.array_equal(conv, want["conv"])
    assert np.array_equal(bits[conv], want["bits"][conv])
    assert np.abs(its.astype(int) - want["its"]).max() <= 2


def test_oracle_detects_a_planted_order_change():
    """Negative control: summing a variable's messages in table order instead
    of ascending check index is a different FP32 computation; the golden
    outputs must reject it (the order the oracle follows is checked, not
    assumed).  The fixtures that see it are the ones that run all 50
    iterations without converging, where rounding differences accumulate."""
    missed = 0
    for name in ("g16k_r6", "s16k_r9_near"):
        code, llr, iters, alpha, want = F.load(name)
        got = resident_minsum(llr, code, iters, alpha, "float32", _plant="list_order")
        missed += not all(np.array_equal(np.asarray(g), want[k]) for g, k in
                          zip(got, ("bits", "conv", "its", "bad")))
    assert missed == 2


def test_int8_step_closed_forms():
    """6 * median|llr| / 127 with the lower median kept to its exponent and
    top 8 mantissa bits, placed in the middle of that bucket."""
    from oracle.resident_ref import int8_step
    f = np.float32
    mid = lambda v: np.array([np.array([v], f).view(np.uint32)[0] & 0xFFFF8000 | (1 << 14)],
                             np.uint32).view(f)[0]
    assert int8_step([-3.0, 1.0, 2.0]) == f(f(6) * mid(2.0)) / f(127)
    assert int8_step([4.0, -1.0, 3.0, 2.0]) == f(f(6) * mid(2.0)) / f(127)     # lower median
    assert mid(2.0) == f(2.0 + 2.0 ** -8)          # bucket [2, 2 + 2^-7): its middle
    assert int8_step([1.234567, -1.234567]) == f(f(6) * mid(1.234567)) / f(127)
    assert int8_step(np.zeros(7)) == f(1e-30)


@pytest.mark.parametrize("n,rate", [(64800, "10/15"), (64800, "2/15"), (64800, "7/15"),
                                    (64800, "13/15"), (16200, "8/15"), (16200, "4/15")])
def test_packed_tables_decode_to_the_base_graph(n, rate):
    from atsc3_amd.codes import get_code
    from atsc3_amd.resident import _host_tables
    code = get_code(rate, n)
    q = code.qc()
    h = _host_tables(code)
    cg, be = q["cg_ptr"], q["be_word"].astype(np.int64)
    gc = h["grp_c"].astype(np.int64)
    assert np.array_equal(gc & 127, np.diff(cg)) and np.array_equal((gc >> 7) & 2047, cg[:-1])
    ev = h["edge_v"].astype(np.int64).reshape(-1, 2)
    gv = h["grp_v"].astype(np.int64)
    seen = np.zeros(len(be), int)
    for t in range(q["nvg"]):
        dv, e0 = gv[t] & 63, (gv[t] >> 6) & 2047
        for e in range(e0, e0 + dv):
            u, g, k = ev[e, 0] & 511, (ev[e, 0] >> 9) & 255, (ev[e, 0] >> 17) & 127
            b = cg[g] + k
            assert b < cg[g + 1]
            s, tb, miss = be[b] & 511, (be[b] >> 9) & 511, (be[b] >> 18) & 1023
            assert tb == t and (s + u) % 360 == 0
            assert ev[e, 1] == (0 if miss == 0 else (miss - 1 + s) % 360 + 1)
            # sign bit of the check's record: shifted-in order (bit dc-1-k)
            # for checks without missing entries of degree <= 16, else k
            dcg = cg[g + 1] - cg[g]
            fast = dcg <= 16 and not ((be[cg[g]:cg[g + 1]] >> 18) & 1023).any()
            assert ev[e, 0] >> 24 == (dcg - 1 - k if fast else k)
            seen[b] += 1
    assert (seen == 1).all()
    # every 64-row task of both phases is covered once, in steps of G tasks
    # of one degree; fast steps only over full circulants of degree <= 16
    for steps, tasks, ngrp, deg_of in (("c_steps", "c_tasks", q["ncg"], gc & 127),
                                       ("v_steps", "v_tasks", q["nvg"], gv & 63)):
        st = h[steps].astype(np.int64)
        tk = h[tasks].astype(np.int64)
        cover = np.zeros(len(tk), int)
        for w in st:
            d, g, start, generic = w & 127, 1 << ((w >> 7) & 7), (w >> 10) & 4095, (w >> 22) & 1
            assert g in (1, 2, 4, 8) and (not generic or g == 1)
            cover[start:start + g] += 1
            grp = tk[start:start + g] & 255
            if not generic:
                assert (deg_of[grp] == d).all() and d <= 16
        assert (cover == 1).all()
        assert sorted((tk & 255) * 8 + (tk >> 8)) == sorted(
            t * 8 + c for t in range(ngrp) for c in range(6))
    # flattened variable steps: each task's head word names its group, row
    # block and first-region size, and its entries are the group's edge_v
    # words in list order
    sd, so = h["v_sdata"].astype(np.int64), h["v_soff"].astype(np.int64)
    vt = h["v_tasks"].astype(np.int64)
    for si, w in enumerate(h["v_steps"].astype(np.int64)):
        if (w >> 22) & 1:
            continue
        d, g, start = w & 127, 1 << ((w >> 7) & 7), (w >> 10) & 4095
        for q in range(g):
            hw = sd[so[si] + q * (d + 1)]
            t, c6 = vt[start + q] & 255, vt[start + q] >> 8
            assert (hw & 255, (hw >> 8) & 7) == (t, c6)
            assert (hw >> 11) & 63 == (gv[t] >> 17) & 63 and (hw >> 17) & 255 == (gv[t] >> 23) & 255
            e0 = (gv[t] >> 6) & 2047
            assert np.array_equal(sd[so[si] + q * (d + 1) + 1:so[si] + (q + 1) * (d + 1)],
                                  ev[e0:e0 + d, 0])
    # flattened check steps: task word, then the group's c_off offsets for
    # the task's row block (FP16 / INT8 tables)
    sd, so = h["c_sdata"].astype(np.int64), h["c_soff"].astype(np.int64)
    ct_, off = h["c_tasks"].astype(np.int64), h["c_off"].astype(np.int64).reshape(-1, 6)
    for si, w in enumerate(h["c_steps"].astype(np.int64)):
        if (w >> 22) & 1:
            continue
        d, g, start = w & 127, 1 << ((w >> 7) & 7), (w >> 10) & 4095
        for q in range(g):
            task = ct_[start + q]
            grp, c6 = task & 255, task >> 8
            base = so[si] + q * (d + 1)
            assert sd[base] == task
            assert np.array_equal(sd[base + 1:base + 1 + d], off[cg[grp]:cg[grp] + d, c6])


@pytest.mark.parametrize("n,rate", [(64800, "10/15"), (64800, "2/15"), (64800, "7/15"),
                                    (16200, "8/15")])
@pytest.mark.parametrize("f32", [False, True])
def test_posterior_rows(n, rate, f32):
    """Each variable group's posterior row (LDS or the global slot) is the
    same in the check-side and variable-side tables; rows are distinct within
    their space; FP16 keeps every row in LDS at its group index; FP32 puts
    the highest-degree groups in the 113 LDS rows."""
    from atsc3_amd.codes import get_code
    from atsc3_amd.resident import F32_LDS_ROWS, _host_tables
    code = get_code(rate, n)
    q = code.qc()
    h = _host_tables(code, f32)
    gv = h["grp_v"].astype(np.int64) & 0xFFFFFFFF
    row, glob, dv = (gv >> 23) & 255, gv >> 31, gv & 63
    ec = h["edge_c"].astype(np.int64)
    t_of_b = (q["be_word"].astype(np.int64) >> 9) & 511
    assert np.array_equal((ec >> 9) & 511, row[t_of_b])
    assert np.array_equal((ec >> 28) & 1, glob[t_of_b])
    for g in (0, 1):
        r = row[glob == g]
        assert len(np.unique(r)) == len(r) and (len(r) == 0 or r.max() == len(r) - 1)
    if not f32:
        assert (glob == 0).all() and np.array_equal(row, np.arange(q["nvg"]))
        # lane l of the wave on checks (g, c6*64 .. c6*64+63) reads the half
        # at c_off / 2 + l of rows of 424 halves, positions 360..423 being
        # copies of 0..63: it must be posterior (t, (j + s) mod 360), j = c6*64 + l
        from atsc3_amd.resident import PROW
        off = h["c_off"].astype(np.int64).reshape(-1, 6) // 2
        s = q["be_word"].astype(np.int64) & 511
        lane = np.arange(64)
        for c6 in range(6):
            a = off[:, c6, None] + lane[None, :]
            r, pos = a // PROW, a % PROW
            assert np.array_equal(r, np.broadcast_to(row[t_of_b][:, None], r.shape))
            unpadded = np.where(pos >= 360, pos - 360, pos)
            assert np.array_equal(unpadded, (c6 * 64 + lane[None, :] + s[:, None]) % 360)
    else:
        nl = int((glob == 0).sum())
        assert nl == min(q["nvg"], F32_LDS_ROWS) and h["global_rows"] == q["nvg"] - nl
        if (glob == 1).any():
            assert dv[glob == 0].min() >= dv[glob == 1].max()


# ---------------------------------------------------------------------------
# device
# ---------------------------------------------------------------------------

def _synth_llr(rate, n, snr, nblk=3, seed=11):
    from atsc3_amd import synth
    from atsc3_amd.codes import get_code
    cw = synth.codewords(rate, n, nblk, seed)
    return get_code(rate, n), synth.llr_bpsk_like(cw, snr, seed + 1)


def _cases():
    out = [("fixture", name) for name in GPU_FIXTURES]
    out += [("synth", f"{r} {n} {s}") for r, n, s in SYNTH]
    return out


def _case(kind, key):
    if kind == "fixture":
        code, llr, iters, alpha, _ = F.load(key)
        return code, llr, iters, alpha
    r, n, s = key.split()
    code, llr = _synth_llr(r, int(n), float(s))
    return code, llr, 50, 0.75


_ORACLE = {}


def _oracle(kind, key):
    if (kind, key) not in _ORACLE:
        code, llr, iters, alpha = _case(kind, key)
        _ORACLE[(kind, key)] = resident_minsum(llr, code, iters, alpha, "float16")
    return _ORACLE[(kind, key)]


def _check(got, want, tag):
    bits, conv, its, bad = got
    assert np.array_equal(np.asarray(conv).astype(bool), want[1]), f"{tag} conv"
    assert np.array_equal(np.asarray(its), want[2]), f"{tag} its"
    assert np.array_equal(np.asarray(bad), want[3]), f"{tag} bad"
    assert np.array_equal(np.asarray(bits), want[0]), f"{tag} bits"


@pytest.mark.gpu
@pytest.mark.parametrize("kind,key", _cases())
def test_kernel_equals_oracle(device, kind, key):
    import torch
    from atsc3_amd.resident import decode_resident_dev
    code, llr, iters, alpha = _case(kind, key)
    want = _oracle(kind, key)
    print(f"[gpu] resident {kind} {key} ({code.rate} {code.n}, {len(llr)} blocks)", flush=True)
    out = decode_resident_dev(torch.from_numpy(llr).to(device), code, iters, alpha)
    torch.cuda.synchronize()
    _check([o.cpu().numpy() for o in out], want, f"{kind} {key}")


@pytest.mark.gpu
@pytest.mark.parametrize("name", F.fixture_names())
def test_fp32_kernel_equals_reference_fixtures(device, name):
    """FP32 records and posteriors: the plain decoder, bit-exact with the
    reference kernel's own golden outputs."""
    import torch
    from atsc3_amd.resident import decode_resident_dev
    code, llr, iters, alpha, want = F.load(name)
    print(f"[gpu] resident fp32 fixture {name} ({code.rate} {code.n})", flush=True)
    out = decode_resident_dev(torch.from_numpy(llr).to(device), code, iters, alpha,
                              msg_dtype="float32")
    torch.cuda.synchronize()
    F.assert_same([o.cpu().numpy() for o in out], want, name)


@pytest.mark.gpu
@pytest.mark.parametrize("key", [f"{r} {n} {s}" for r, n, s in SYNTH])
def test_fp32_kernel_equals_c_oracle(device, key):
    import torch
    from oracle.ldpc import minsum_f32
    from atsc3_amd.resident import decode_resident_dev
    code, llr, iters, alpha = _case("synth", key)
    want = minsum_f32(llr, code, iters, alpha)
    print(f"[gpu] resident fp32 synth {key}", flush=True)
    out = decode_resident_dev(torch.from_numpy(llr).to(device), code, iters, alpha,
                              msg_dtype="float32")
    torch.cuda.synchronize()
    _check([o.cpu().numpy() for o in out], [np.asarray(w) for w in want], f"fp32 {key}")


@pytest.mark.gpu
def test_fp32_mixed_codes_one_launch(device):
    """Fixture blocks of every code in one FP32 launch on 5 workgroups."""
    import torch
    from atsc3_amd.resident import ResidentPlan
    blocks = []
    for name in F.fixture_names():
        code, llr, iters, alpha, want = F.load(name)
        if (iters, alpha) != (50, 0.75):
            continue
        for b in range(len(llr)):
            blocks.append((code, llr[b], {k: v[b] for k, v in want.items()}, f"{name}#{b}"))
    order = np.random.default_rng(9).permutation(len(blocks))
    blocks = [blocks[i] for i in order]
    plan = ResidentPlan([b[0] for b in blocks], device, 50, 0.75, grid=5, msg_dtype="float32")
    llr = np.concatenate([b[1] for b in blocks]).astype(np.float32)
    print(f"[gpu] resident fp32 mixed launch {len(blocks)} blocks", flush=True)
    bits, conv, its, bad = (x.cpu().numpy() for x in plan.run(torch.from_numpy(llr).to(device)))
    for i, (code, _, want, tag) in enumerate(blocks):
        assert bool(conv[i]) == bool(want["conv"]) and its[i] == want["its"], tag
        assert bad[i] == want["bad"], tag
        assert np.array_equal(bits[plan.off[i]:plan.off[i + 1]], want["bits"]), tag


_ORACLE8 = {}


@pytest.mark.gpu
@pytest.mark.parametrize("kind,key", [("fixture", "g16k_r4"), ("fixture", "s16k_r6_ties"),
                                      ("fixture", "s16k_r3_infnan"), ("fixture", "g64k_r10_nuc20"),
                                      ("synth", "13/15 64800 5.5"), ("synth", "7/15 64800 0.8"),
                                      ("synth", "6/15 64800 1.5")])
def test_int8_kernel_equals_oracle(device, kind, key):
    """INT8 record magnitudes with the oracle's per-codeword steps given
    explicitly: bit-exact."""
    import torch
    from oracle.resident_ref import int8_step
    from atsc3_amd.resident import decode_resident_dev
    code, llr, iters, alpha = _case(kind, key)
    steps = np.array([int8_step(r) for r in llr], np.float32)
    want = resident_minsum(llr, code, iters, alpha, "int8", steps=steps)
    print(f"[gpu] resident int8 {kind} {key}", flush=True)
    out = decode_resident_dev(torch.from_numpy(llr).to(device), code, iters, alpha,
                              msg_dtype="int8", msg_scale=steps)
    torch.cuda.synchronize()
    _check([o.cpu().numpy() for o in out], want, f"int8 {kind} {key}")


@pytest.mark.gpu
@pytest.mark.parametrize("engine", ["resident", "qc"])
def test_device_int8_step_is_the_oracle_rule(device, engine):
    """With no step given the engines derive 6 * median|llr| / 127 on the
    device; decoding with the oracle's int8_step given explicitly yields the
    same outputs (the resident engine and the fused engine, whose INT8 step
    code every QC engine shares)."""
    import torch
    from oracle.resident_ref import int8_step
    from atsc3_amd import ldpc
    code, llr = _synth_llr("10/15", 64800, 3.3, nblk=6, seed=31)
    llr = llr * np.float32(np.linspace(0.3, 3.0, 6))[:, None]      # different medians
    steps = np.array([int8_step(r) for r in llr], np.float32)
    t = torch.from_numpy(llr).to(device)
    print(f"[gpu] device int8 step ({engine})", flush=True)
    a = ldpc.decode_dev(t, code, 50, 0.75, "int8", engine)
    b = ldpc.decode_dev(t, code, 50, 0.75, "int8", engine,
                        msg_scale=torch.from_numpy(steps).to(device))
    torch.cuda.synchronize()
    for x, y in zip(a, b):
        assert torch.equal(x, y)


@pytest.mark.gpu
def test_mixed_codes_one_launch_few_workgroups(device):
    """Blocks of six codes interleaved in one queue, 3 workgroups: each
    workgroup decodes many codewords of different codes in turn."""
    import torch
    from atsc3_amd.resident import ResidentPlan
    keys = [("fixture", "g16k_r2"), ("synth", "13/15 64800 5.5"), ("fixture", "s16k_r9_near"),
            ("synth", "7/15 64800 0.8"), ("fixture", "g64k_r10_nuc20"),
            ("synth", "8/15 16200 2.0")]
    blocks = []
    for kind, key in keys:
        code, llr, iters, alpha = _case(kind, key)
        want = _oracle(kind, key)
        for b in range(len(llr)):
            blocks.append((code, llr[b], [w[b] for w in want], f"{key}#{b}"))
    order = np.random.default_rng(5).permutation(len(blocks))
    blocks = [blocks[i] for i in order]
    # every case above decodes with iters 50, alpha 0.75
    for kind, key in keys:
        _, _, iters, alpha = _case(kind, key)
        assert (iters, alpha) == (50, 0.75), key
    plan = ResidentPlan([b[0] for b in blocks], device, 50, 0.75, grid=3)
    llr = np.concatenate([b[1] for b in blocks]).astype(np.float32)
    print(f"[gpu] resident mixed launch {len(blocks)} blocks grid 3", flush=True)
    bits, conv, its, bad = plan.run(torch.from_numpy(llr).to(device))
    torch.cuda.synchronize()
    bits, conv, its, bad = (x.cpu().numpy() for x in (bits, conv, its, bad))
    for i, (code, _, want, tag) in enumerate(blocks):
        got = (bits[plan.off[i]:plan.off[i + 1]], bool(conv[i]), its[i], bad[i])
        assert got[1] == bool(want[1]) and got[2] == want[2] and got[3] == want[3], tag
        assert np.array_equal(got[0], want[0]), tag


@pytest.mark.gpu
def test_live_frame_operating_point(device):
    """256-NUC r10/15 under Rayleigh fading at 21 dB, 512 blocks: every block
    returns the transmitted codeword; mean iterations within 3% of plain FP32."""
    import torch
    from atsc3_amd import ldpc, synth
    from atsc3_amd.codes import get_code
    from atsc3_amd.demap import Demapper
    from atsc3_amd.resident import decode_resident_dev
    code = get_code("10/15", 64800)
    st = synth.FadingStream("10/15", 64800, "256QAM", 64, 7, device)
    dm = Demapper(64800, "256QAM", "10/15", device)
    print("[gpu] resident live-frame point 512 blocks", flush=True)
    z, w, idx, n0 = st.batch(512, 21.0)
    llr, _ = dm.demap_dev(z, w, n0)
    bits, conv, its, bad = decode_resident_dev(llr, code, 50, 0.75)
    b32, c32, i32, _ = ldpc.decode_dev(llr, code, 50, 0.75, "float32", "phase")
    torch.cuda.synchronize()
    cw = st.cw_dev.index_select(0, idx)
    assert bool(conv.all()) and bool(c32.all())
    assert torch.equal(bits, cw)
    m16, m32 = its.float().mean().item(), i32.float().mean().item()
    assert abs(m16 - m32) <= 0.03 * m32, (m16, m32)


@pytest.mark.gpu
def test_graph_replay_equals_eager(device):
    import torch
    from atsc3_amd.resident import ResidentPlan
    code, llr, iters, alpha = _case("fixture", "g64k_r10_nuc20")
    plan = ResidentPlan([code] * len(llr), device, iters, alpha)
    src = torch.from_numpy(llr.reshape(-1)).to(device)
    print("[gpu] resident graph replay", flush=True)
    eager = [x.clone() for x in plan.run(src)]
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        plan.run(src)                        # warm-up on the capture stream
        torch.cuda.synchronize()
        with torch.cuda.graph(g, stream=s):
            plan.run(src)
    torch.cuda.synchronize()
    for r in range(3):
        for x in (plan.bits, plan.conv, plan.its, plan.bad):
            x.fill_(0xFF if x.dtype == torch.uint8 else -7)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for a, b in zip((plan.bits, plan.conv, plan.its, plan.bad), eager):
            assert torch.equal(a, b), f"replay {r}"


@pytest.mark.gpu
def test_mixed_pipeline_engines_agree_fp32(device):
    """Two uneven frames (every code and constellation of the mixed
    workload) demapped and decoded in one captured graph: the resident
    engine's FP32 mode equals the batched engine (both bit-exact)."""
    import torch
    from atsc3_amd.mixed import MixedPipeline, MixedSource, block_esn0, frame_mix
    specs = frame_mix(74, 3)
    outs = []
    for eng in ("batch", "resident"):
        print(f"[gpu] mixed pipeline fp32 {eng}", flush=True)
        p = MixedPipeline(specs, 2, device, 50, 0.75, "float32", engine=eng)
        src = MixedSource(p, 8, 21)
        cells, s2, _ = src.frame(block_esn0(p.specs, 23))
        p.load(cells, s2)
        g = p.capture()
        g.replay()
        torch.cuda.synchronize()
        p.plan.check()
        outs.append([x.clone() for x in (p.plan.bits, p.plan.conv, p.plan.its, p.plan.bad)])
        del g, p
    for a, b in zip(*outs):
        assert torch.equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("msg", ["float16", "float32", "int8"])
def test_edge_cases_small_batches(device, msg):
    """An empty batch returns empty outputs; one codeword on a full grid
    (255 workgroups find the queue empty at once); iters = 0 (syndrome of
    the LLRs only: bits are their hard decisions) -- each equal to the
    oracle in the same storage mode."""
    import torch
    from atsc3_amd.codes import get_code
    from atsc3_amd.resident import decode_resident_dev, default_grid
    code = get_code("8/15", 16200)
    out = decode_resident_dev(torch.empty((0, code.n), device=device), code, 50, 0.75,
                              msg_dtype=msg)
    assert all(o.shape[0] == 0 for o in out)
    code, llr = _synth_llr("8/15", 16200, 2.0, nblk=1, seed=5)
    for iters in (50, 0):
        print(f"[gpu] resident edge {msg} 1 block iters {iters} grid {default_grid()}",
              flush=True)
        want = resident_minsum(llr, code, iters, 0.75, msg)
        got = decode_resident_dev(torch.from_numpy(llr).to(device), code, iters, 0.75,
                                  grid=default_grid(), msg_dtype=msg)
        torch.cuda.synchronize()
        _check([o.cpu().numpy() for o in got], want, f"{msg} iters {iters}")
        if iters == 0:
            assert np.array_equal(want[0], (llr < 0).astype(np.uint8))


@pytest.mark.gpu
def test_two_stream_pipeline_equals_one_stream(device):
    """The receiver pipeline of the bench headline: LLRs produced on one
    stream, decoded on another behind an event (one resident plan per
    stream), three windows in flight -- every window's outputs equal the
    one-stream decode of the same LLRs."""
    import torch
    from atsc3_amd.resident import decode_resident_dev
    code, llr = _synth_llr("8/15", 16200, 2.0, nblk=24, seed=9)
    src = torch.from_numpy(llr).to(device)
    want = [o.cpu().numpy() for o in decode_resident_dev(src, code, 50, 0.75)]
    torch.cuda.synchronize()
    s_prod, s_dec = torch.cuda.Stream(device), torch.cuda.Stream(device)
    outs = []
    print("[gpu] resident two-stream pipeline, 3 windows of 24 blocks", flush=True)
    for _ in range(3):
        with torch.cuda.stream(s_prod):
            win = src * 1.0                     # the window's LLRs, made on s_prod
            done = torch.cuda.Event()
            done.record(s_prod)
        with torch.cuda.stream(s_dec):
            s_dec.wait_event(done)
            win.record_stream(s_dec)
            outs.append(decode_resident_dev(win, code, 50, 0.75))
    torch.cuda.synchronize()
    for k, out in enumerate(outs):
        _check([o.cpu().numpy() for o in out], want, f"window {k}")


@pytest.mark.gpu
@pytest.mark.parametrize("msg", ["float16", "float32"])
def test_soft_output_equals_oracle(device, msg):
    """atsc3_minsum_resident_llr_dev: the same hard results as the plain
    entry, and every bit's final posterior equal to the oracle's (FP16 mode:
    the stored FP16 value; FP32 mode: the FP32 posterior of the reference's
    arithmetic), bit for bit; bits are the posteriors' signs."""
    import torch
    from atsc3_amd import synth
    from atsc3_amd.codes import get_code
    from atsc3_amd.resident import ResidentPlan
    code = get_code("6/15", 16200)
    cw = synth.codewords("6/15", 16200, 6, 811)
    llr = np.concatenate([synth.llr_bpsk_like(cw[:3], 1.5, 812),
                          synth.llr_bpsk_like(cw[3:], 1.0, 813)]).astype(np.float32)
    print(f"[gpu] soft output {msg}", flush=True)
    plan = ResidentPlan([code] * 6, device, 30, 0.75, 4, msg)
    t = torch.from_numpy(llr.ravel()).to(device)
    post = torch.empty(plan.total, dtype=torch.float32, device=device)
    outs = [x.clone() for x in plan.run(t, post=post)]
    ref = [x.clone() for x in plan.run(t)]
    torch.cuda.synchronize()
    for a, b in zip(outs, ref):
        assert torch.equal(a, b)
    want = resident_minsum(llr, code, 30, 0.75, msg, return_post=True)
    p = post.cpu().numpy().reshape(6, -1)
    assert np.array_equal(p.view(np.uint32), want[4].view(np.uint32))
    assert np.array_equal(outs[0].cpu().numpy().reshape(6, -1), (p < 0).astype(np.uint8))
    assert np.array_equal(outs[2].cpu().numpy(), want[2])


@pytest.mark.gpu
@pytest.mark.parametrize("msg", ["float16", "float32"])
def test_stop_test_at_every_iteration_limit(device, msg):
    """The speculative stop test (once a check phase leaves at most a few
    unsatisfied checks, the next iteration counts them alone first and stops
    there when none is left) against the oracle at every iteration limit
    from 0 to past convergence: the stop can land on the limit itself
    (converged at its = iters), just after it (bad counted by the check
    phase with no update), or anywhere before; conv, its, bad and bits must
    be the plain schedule's each time, for the one-workgroup kernel and in
    grow mode."""
    import torch
    from atsc3_amd.resident import decode_resident_dev
    code, llr = _synth_llr("9/15", 16200, 2.8, nblk=6, seed=41)
    t = torch.from_numpy(llr).to(device)
    full = resident_minsum(llr, code, 50, 0.75, msg)
    top = int(np.max(full[2])) + 2
    assert 3 <= top <= 30, f"test blocks converge at {full[2]}"
    for iters in range(top):
        want = resident_minsum(llr, code, iters, 0.75, msg)
        print(f"[gpu] stop test {msg} iters {iters}", flush=True)
        for grow in (None, 8):
            got = [x.cpu().numpy() for x in decode_resident_dev(t, code, iters, 0.75,
                                                                msg_dtype=msg, grow=grow)]
            _check(got, want, f"{msg} iters {iters} grow {grow}")
