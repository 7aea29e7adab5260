"""The bench's workloads run as tests (smaller where noted), so what the
benchmark reports is also checked on every GPU run:

* the autotuner on an uneven 74-block frame (every candidate, the
  XCD-partitioned mode and the resident engine included, captured and
  timed; the choice it returns decodes the frame exactly as the default
  configuration and as the XCD-partitioned mode);
* an Es/N0 sweep of the live-frame workload (reduced: 256 blocks per point)
  through demap + resident decode, FP32 and FP16: FER falls with Es/N0, the
  operating point's blocks all decode, and FP16 stays within FP32's errors;
* the LDM two-layer chain over the bench's 50 frames (1,877 core and 5,631
  enhanced blocks) at 23 dB: every block of both layers decodes.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_autotune_on_uneven_frame(device, tmp_path, monkeypatch):
    import torch
    from atsc3_amd import autotune
    from atsc3_amd.mixed import MixedPipeline, MixedSource, block_esn0, frame_mix
    monkeypatch.setenv("ATSC3_TUNE_CACHE", str(tmp_path / "tune.json"))
    specs = frame_mix(74, 2024)
    print("[gpu] autotune on a 74-block uneven frame", flush=True)

    def build(slots, grid):
        return MixedPipeline(specs, 1, device, 50, 0.75, "float32",
                             None if slots in ("resident", "xcd") else slots, grid,
                             engine=slots if slots in ("resident", "xcd") else "batch")

    def prepare(p):
        src = MixedSource(p, 4, 31)
        cells, s2, _ = src.frame(block_esn0(p.specs, 33))
        p.load(cells, s2)
    slots, grid, table = autotune.tune("test74", 74, "float32", build, prepare, reps=2,
                                       use_cache=True, resident=True, xcd=True)
    cands = autotune.candidates(74, "float32", xcd=True)
    assert sum(c[0] == "xcd" for c in cands) == 2
    assert len(table) == len(cands) + 1 and all(v > 0 for v in table.values())
    best = min(table, key=table.get)
    assert best == ("resident" if slots == "resident" else f"{slots}x{grid}")
    # cached: the second call returns the same choice without timing
    assert autotune.tune("test74", 74, "float32", build, prepare, resident=True,
                         xcd=True)[:2] == (slots, grid)
    outs = []
    for sl, gr in ((slots, grid), (None, None), ("xcd", None)):
        p = build(sl if sl is not None else 74, gr)
        prepare(p)
        p.run()
        p.plan.check()
        torch.cuda.synchronize()
        outs.append([x.clone() for x in (p.plan.bits, p.plan.conv, p.plan.its, p.plan.bad)])
    for other in outs[1:]:
        for a, b in zip(outs[0], other):
            assert torch.equal(a, b)


def test_live_frame_sweep_reduced(device):
    import torch
    from atsc3_amd import ldpc, synth
    from atsc3_amd.codes import get_code
    from atsc3_amd.demap import Demapper
    code = get_code("10/15", 64800)
    stream = synth.FadingStream("10/15", 64800, "256QAM", 32, 77, device)
    dm = Demapper(64800, "256QAM", "10/15", device)
    errs = {}
    for es in (18.0, 19.0, 20.0, 22.0):
        print(f"[gpu] live-frame sweep {es} dB", flush=True)
        z, w, idx, n0 = stream.batch(256, es)
        llr, _ = dm.demap_dev(z, csi=w, sigma2=n0)
        ref = stream.cw_dev.index_select(0, idx)
        for fmt in ("float32", "float16"):
            out = ldpc.decode_dev(llr, code, 50, 0.75, fmt, "resident")
            errs[(es, fmt)] = int((out[0] != ref).any(1).sum().item())
            assert out[2].max().item() <= 50
    for fmt in ("float32", "float16"):
        seq = [errs[(es, fmt)] for es in (18.0, 19.0, 20.0, 22.0)]
        assert seq == sorted(seq, reverse=True), (fmt, seq)
        assert seq[0] > 0 and seq[-1] == 0, (fmt, seq)
    for es in (18.0, 19.0, 20.0, 22.0):
        assert errs[(es, "float16")] <= errs[(es, "float32")] + 3


def test_ldm_fifty_frames(device):
    import torch
    from atsc3_amd import synth
    from atsc3_amd.ldm import LdmPipeline, LdmStream
    pipe = LdmPipeline(device=device, msg_dtype="float16", engine="resident")
    stream = LdmStream(pipe, 16, 5)
    ncore = 50 * 1_216_939 // pipe.ncell_c
    assert ncore == 1877
    n0 = synth.n0_of(23.0)
    print(f"[gpu] LDM 50 frames: {ncore} core + {ncore * pipe.ratio} enhanced blocks", flush=True)
    y, ic, ie = stream.batch(ncore, n0, start=99)
    out = pipe.run(y, n0)
    torch.cuda.synchronize()
    ec = int((out["core"][0] != stream.cw_c_dev.index_select(0, ic)).any(1).sum().item())
    ee = int((out["enh"][0] != stream.cw_e_dev.index_select(0, ie)).any(1).sum().item())
    assert ec == 0 and ee <= 1, (ec, ee)
    assert out["enh"][0].shape[0] == ncore * pipe.ratio
