#!/usr/bin/env python3
"""Headline benchmark: FEC blocks/s of demap + min-sum decode on one MI355X.

The headline ``value`` is the FP16 message form: each check's messages kept
as one FP16 record, posteriors as FP16 in LDS, sums in FP32, decoded by the
resident engine (one codeword per compute unit).  The plain bit-exact FP32
decoder (``config.plain_fp32``, also through the host-pointer drop-in), the
other message forms (``config.forms``) and the CPU path (``cpu_baseline``)
are measured beside it on the same blocks.

Workload (the live-frame configuration): 256-NUC rate 10/15, N = 64800
(8100 cells per block), per-cell Rayleigh fading (|h| Rayleigh sigma = 1,
uniform phase) with perfect CSI, complex AWGN at Es/N0 = E|x|^2 / N0.
Frames carry 207,900 data cells (242 ms); blocks straddle frame boundaries
(stream model), so a window of 100 frames holds 2566 complete blocks.  One
step = demap (fused de-interleave) + decode of that whole window: two
kernel launches on one stream, inputs resident in HBM.

Operating point: the metric is quoted at FER <= 1e-4.  The bench sweeps
Es/N0 17..23 dB (one window each, plain FP32), takes the lowest Es/N0 with
no block error, and confirms it on >= --fer-blocks blocks with zero errors
for both plain FP32 and the headline format (rule of three: FER < 1e-4 at
95% for >= 30,000 blocks); any error moves the point up 1 dB.  The timed
steps then run at that Es/N0.

Printed: ONE JSON line on stdout (progress on stderr), including
  roofline     the decode kernel: algorithmic bytes per launch
               sum_b [4N + N + its_b (2 E s + 4N)] (s = 2 bytes per FP16
               message) over its measured average duration (HIP events on
               the launch stream), vs 8 TB/s; the resident engine's own
               record-traffic model beside it
  cpu_baseline demap + decode on the host cores in the upstream's
               multi-process topology on a bounded sample of the same blocks
  config.ldm   the two-layer LDM frame (1,216,939 cells per 242 ms frame,
               core QPSK 4/15 + enhanced 64-NUC 10/15 at -5 dB injection,
               AWGN): core demap + decode, cancellation, enhanced demap +
               decode of 50 frames per step, as a real-time factor;
               ``--workload ldm`` makes it the headline line instead
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("FEC blocks/s at FER≤1e-4: batched FP16 vs plain GPU vs CPU; "
          "% of HBM peak")
HBM_PEAK_GBS = 8000.0
# VALU issue peak of one MI355X: 256 compute units x one wave64 VALU
# instruction per cycle each (4 SIMDs, 16 lanes) x 2.4 GHz, counted in lane
# operations
VALU_PEAK_LANE_OPS = 256 * 64 * 2.4e9
# VALU instructions per cell of demap_kernel<8, contiguous> (256-NUC: 64
# base points), SQ_INSTS_VALU / SQ_WAVES of a rocprofv3 counter run (one
# cell per lane); the static count of the compiled kernel is the same
DEMAP_VALU_PER_CELL_256 = 699
FRAME_CELLS = 207_900
FRAME_SEC = 0.242
RATE, N, MOD = "10/15", 64800, "256QAM"
SWEEP = (17.0, 18.0, 19.0, 20.0, 21.0, 22.0, 23.0)
FORM_SWEEP = (18.5, 19.0, 19.5, 19.75, 20.0, 20.25, 20.5, 20.75, 21.0)
LDM_FRAME_CELLS = 1_216_939
LDM_ESN0 = 18.0
LDM_SWEEP = (18.0, 19.0, 20.0, 21.0, 22.0, 23.0)


# the plain FP32 decoder of the bench (bit-exact with the reference kernel):
# the resident engine's FP32 mode
PLAIN_ENGINE = "resident"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _cpu_model():
    """Host CPU as /proc/cpuinfo describes it: model name, family, model,
    stepping, and the sockets x cores it reports."""
    try:
        txt = open("/proc/cpuinfo").read()
    except OSError:
        return "unknown"
    first = txt.split("\n\n")[0]
    kv = {}
    for line in first.splitlines():
        if ":" in line:
            k, v = line.split(":", 1)
            kv[k.strip()] = v.strip()
    sockets = len({ln.split(":")[1].strip() for ln in txt.splitlines()
                   if ln.startswith("physical id")}) or 1
    return (f"{kv.get('model name', 'unknown')} (family {kv.get('cpu family', '?')} model "
            f"{kv.get('model', '?')} stepping {kv.get('stepping', '?')}, {sockets} socket(s), "
            f"{kv.get('cpu cores', '?')} cores per socket, {txt.count('processor')} logical "
            f"processors listed)")


def model_bytes(its, n, e, s_msg):
    """Compulsory HBM bytes of one decode launch of the engines that keep
    per-edge messages and FP32 posteriors in device memory (per-block sum):
    LLRs in, bits out, and per iteration the messages read and written once
    and the posteriors written once."""
    its = np.asarray(its, np.float64)
    return float(np.sum(4.0 * n + n + its * (2.0 * e * s_msg + 4.0 * n)))


def resident_model_bytes(its, n, m):
    """Compulsory bytes of one resident-engine launch (posteriors in LDS):
    LLRs in, bits out, and per iteration each check's 8-byte record read
    and written by the check phase and read once by the variable phase,
    plus the LLRs the variable phase re-reads (tot = lam + sum)."""
    its = np.asarray(its, np.float64)
    return float(np.sum(4.0 * n + n + its * (3.0 * 8.0 * m + 4.0 * n)))


def load_traffic(workload_key):
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            d = json.load(f)
        return d.get(workload_key)
    except (OSError, ValueError):
        return None


def fer_gap_db(ref, other, target):
    """Es/N0 difference (dB) at which two FER curves cross ``target``
    (log-linear interpolation between sweep points; None if a curve does not
    cross it inside the sweep)."""
    def cross(curve):
        pts = [(c[0], c[1] / c[2]) for c in curve]
        for (e0, f0), (e1, f1) in zip(pts, pts[1:]):
            if f0 >= target > f1:
                if f1 <= 0:
                    return e1
                import math
                a = (math.log(f0) - math.log(target)) / (math.log(f0) - math.log(f1))
                return e0 + a * (e1 - e0)
        return None
    a, b = cross(ref), cross(other)
    return None if a is None or b is None else round(b - a, 3)


def ldm_bench(args, dev, msg_formats=("float32", "float16", "float16_resident")):
    """Two-layer LDM chain over ``args.ldm_frames`` frames per step.

    Core: QPSK 4/15 (32400 cells per block), enhanced: 64-NUC 10/15 (10800
    cells per block), three enhanced blocks per core block; total power 1,
    injection -5 dB, complex AWGN.  One step = the whole chain on one stream.
    """
    import torch
    from atsc3_amd import ldpc
    from atsc3_amd.ldm import LdmPipeline, LdmStream
    from atsc3_amd import synth

    res = {}
    for fmt in msg_formats:
        msg, engine = (("float16", "resident") if fmt == "float16_resident"
                       else (fmt, PLAIN_ENGINE if fmt == "float32" else "auto"))
        pipe = LdmPipeline(device=dev, msg_dtype=msg, engine=engine)
        if "stream" not in res:
            res["stream"] = LdmStream(pipe, 64, args.seed + 17)
        stream = res["stream"]
        ncore = args.ldm_frames * LDM_FRAME_CELLS // pipe.ncell_c
        nenh = ncore * pipe.ratio

        def errs(out, ic, ie):
            ec = int((out["core"][0] != stream.cw_c_dev.index_select(0, ic)).any(1).sum().item())
            ee = int((out["enh"][0] != stream.cw_e_dev.index_select(0, ie)).any(1).sum().item())
            return ec, ee

        sweep = {}
        for es in LDM_SWEEP:
            y, ic, ie = stream.batch(ncore, synth.n0_of(es), start=int(es) * 7919)
            out = pipe.run(y, synth.n0_of(es))
            ec, ee = errs(out, ic, ie)
            sweep[f"{es:g}"] = {"core_errors": ec, "core_blocks": ncore, "enh_errors": ee,
                                "enh_blocks": nenh,
                                "enh_mean_its": round(float(out["enh"][2].float().mean().item()), 2)}
            log(f"[ldm] {fmt} Es/N0 {es:g} dB: core {ec}/{ncore}, enhanced {ee}/{nenh} "
                f"block errors, enhanced mean its {sweep[f'{es:g}']['enh_mean_its']}")
        def timed(es, start):
            n0 = synth.n0_of(es)
            y, ic, ie = stream.batch(ncore, n0, start=start)
            for _ in range(args.warmup):
                out = pipe.run(y, n0)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                out = pipe.run(y, n0)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) * 1e3 / args.steps
            ec, ee = errs(out, ic, ie)
            rt = args.ldm_frames * FRAME_SEC / (ms / 1e3)
            return {"esn0_db": es, "ms_per_step": round(ms, 3), "realtime_factor": round(rt, 2),
                    "blocks_per_s": round((ncore + nenh) / (ms / 1e3), 1),
                    "core_errors": ec, "enh_errors": ee,
                    "core_mean_its": round(float(out["core"][2].float().mean().item()), 2),
                    "enh_mean_its": round(float(out["enh"][2].float().mean().item()), 2)}

        # the lowest swept Es/N0 where both layers reach FER <= 1e-4: over
        # >= --fer-blocks blocks of each layer (the sweep's batch included),
        # errors / blocks <= 1e-4 in the core and in the enhanced layer
        op_ldm, confirm = None, None
        if fmt in ("float32", "float16_resident"):
            for es in LDM_SWEEP:
                sw = sweep[f"{es:g}"]
                if sw["core_errors"] > 1e-4 * args.fer_blocks or \
                        sw["enh_errors"] > 1e-4 * args.fer_blocks * pipe.ratio:
                    continue
                n0 = synth.n0_of(es)
                cb, eb, ce_, ee_ = ncore, nenh, sw["core_errors"], sw["enh_errors"]
                k = 1
                while cb < args.fer_blocks and ce_ <= 1e-4 * args.fer_blocks and \
                        ee_ <= 1e-4 * args.fer_blocks * pipe.ratio:
                    y, ic, ie = stream.batch(ncore, n0, start=int(es) * 7919 + k * ncore)
                    c1, e1 = errs(pipe.run(y, n0), ic, ie)
                    ce_, ee_, cb, eb, k = ce_ + c1, ee_ + e1, cb + ncore, eb + nenh, k + 1
                log(f"[ldm] {fmt} confirm {es:g} dB: core {ce_}/{cb}, enhanced {ee_}/{eb}")
                if cb >= args.fer_blocks and ce_ / cb <= 1e-4 and ee_ / eb <= 1e-4:
                    op_ldm = es
                    confirm = {"core_blocks": cb, "core_errors": ce_, "enh_blocks": eb,
                               "enh_errors": ee_}
                    break
        worst = timed(args.ldm_esn0, 10 ** 5)
        res[fmt] = {"engine": engine, "core_blocks": ncore, "enh_blocks": nenh,
                    **worst, "fer_sweep": sweep}
        if op_ldm is not None:
            at = timed(op_ldm, 2 * 10 ** 5)
            res[fmt]["at_fer_1e-4"] = {**at, "confirmed": confirm}
            log(f"[ldm] {fmt}: both layers FER <= 1e-4 from {op_ldm:g} dB: "
                f"{at['ms_per_step']:.2f} ms per {args.ldm_frames} frames = "
                f"{at['realtime_factor']:.1f}x real time")
        log(f"[ldm] {fmt}: {worst['ms_per_step']:.2f} ms per {args.ldm_frames} frames = "
            f"{worst['realtime_factor']:.1f}x real time (core {worst['core_errors']}/{ncore}, "
            f"enhanced {worst['enh_errors']}/{nenh} block errors at {args.ldm_esn0:g} dB)")
        del pipe
        torch.cuda.empty_cache()
    res.pop("stream", None)
    return res


def dropin_bench(llr_dev, code, nblk=512, threads=8, per_call=8, reps=3):
    """The host-pointer drop-in ldpc_minsum_f32 driven the way the upstream
    drives its compiled kernel (lab/m9_fast.py:893-925, lab/m3_ldpc.py:485-506):
    a thread pool of ``threads`` workers, each call an 8-block sub-batch with
    its own host buffers; beside one call for all blocks and the
    device-resident decode_dev (FP32, no host transfer) of the same blocks."""
    import torch
    from concurrent.futures import ThreadPoolExecutor
    from atsc3_amd import ldpc
    llr = llr_dev[:nblk].contiguous()
    host = llr.cpu().numpy()
    subs = [host[i:i + per_call] for i in range(0, nblk, per_call)]

    def one(a):
        return ldpc.minsum_f32_host(a, code, code.n, 50, 0.75)
    res = {}
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(one, subs))                       # warm-up: pools and tables
        t0 = time.perf_counter()
        for _ in range(reps):
            outs = list(ex.map(one, subs))
        res["threaded_s"] = (time.perf_counter() - t0) / reps
    one(host)
    t0 = time.perf_counter()
    for _ in range(reps):
        whole = one(host)
    res["one_call_s"] = (time.perf_counter() - t0) / reps
    ldpc.decode_dev(llr, code, 50, 0.75, "float32", "resident")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        dv = ldpc.decode_dev(llr, code, 50, 0.75, "float32", "resident")
    torch.cuda.synchronize()
    res["decode_dev_s"] = (time.perf_counter() - t0) / reps
    bits = np.concatenate([o[0] for o in outs])
    same = bool(np.array_equal(bits, dv[0].cpu().numpy())
                and np.array_equal(np.concatenate([o[2] for o in outs]), dv[2].cpu().numpy())
                and np.array_equal(whole[0], bits))
    out = {"blocks": nblk, "threads": threads, "blocks_per_call": per_call,
           "threaded_blocks_per_s": round(nblk / res["threaded_s"], 1),
           "one_call_blocks_per_s": round(nblk / res["one_call_s"], 1),
           "decode_dev_blocks_per_s": round(nblk / res["decode_dev_s"], 1),
           "results_equal_decode_dev": same}
    log(f"[bench] drop-in ldpc_minsum_f32: {threads} threads x {per_call}-block calls "
        f"{out['threaded_blocks_per_s']:.0f} blocks/s, one {nblk}-block call "
        f"{out['one_call_blocks_per_s']:.0f}, decode_dev {out['decode_dev_blocks_per_s']:.0f}; "
        f"equal {same}")
    return out


def utc_bench(llr_dev, code, nblk=512, reps=3):
    """The upstream's other decode contract (GpuMinSum.decode,
    lab/m9_gpu.py:138-211; min_sum_decode, lab/m3_ldpc.py:245-286) on the
    same blocks, device-resident, in FP32 and FP64: blocks/s and mean
    iterations (its >= 1 under this contract)."""
    import torch
    from atsc3_amd import ldpc
    out = {"blocks": nblk}
    for name, dt in (("float32", torch.float32), ("float64", torch.float64)):
        t = llr_dev[:nblk].to(dt).contiguous()
        r = ldpc.decode_utc_dev(t, code, 50, 0.75)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            r = ldpc.decode_utc_dev(t, code, 50, 0.75)
        torch.cuda.synchronize()
        secs = (time.perf_counter() - t0) / reps
        out[name] = {"blocks_per_s": round(nblk / secs, 1), "ms": round(secs * 1e3, 3),
                     "mean_its": round(float(r[2].float().mean().item()), 3)}
    log(f"[bench] update-then-check (GpuMinSum contract) {nblk} blocks: FP32 "
        f"{out['float32']['blocks_per_s']:.0f} blocks/s, FP64 {out['float64']['blocks_per_s']:.0f}")
    return out


def mixed_bench(args, dev, formats=("float32", "float16", "int8")):
    """Uneven frames: 74 blocks per frame mixing 64800-bit rates 6/8/10/12/15
    and 16200-bit rates 4/9/15 on QPSK..256-NUC, per-block AWGN Es/N0 drawn
    around each (code, constellation)'s threshold (iterations 3..50).  Per
    message format: autotuned (slots, grid), FER over fresh frames, and the
    time of one HIP-graph replay (all demaps + the persistent decode)."""
    import torch
    from atsc3_amd import autotune
    from atsc3_amd.mixed import MixedPipeline, MixedSource, block_esn0, frame_mix

    specs = frame_mix(74, args.seed)
    counts = {}
    for n, r, m in specs:
        k = f"{n} {r} {m}"
        counts[k] = counts.get(k, 0) + 1
    res = {"blocks_per_frame": len(specs), "mix": counts, "frames_per_launch": args.mixed_frames}
    F = args.mixed_frames
    nblk = len(specs) * F

    def build(fmt):
        def make(slots, grid):
            if slots in ("team", "team-l2"):
                return MixedPipeline(specs, F, dev, 50, 0.75, fmt, engine="team", team=grid,
                                     direct=slots == "team-l2")
            return MixedPipeline(
                specs, F, dev, 50, 0.75, fmt, None if slots in ("resident", "xcd") else slots,
                grid, engine=slots if slots in ("resident", "xcd") else "batch")
        return make

    def prepare(p):
        src = MixedSource(p, 8, args.seed + 5)
        cells, s2, _ = src.frame(block_esn0(p.specs, args.seed + 9))
        p.load(cells, s2)

    ref_bits = None
    for fmt in formats:
        slots, grid, table = autotune.tune(f"mixed{len(specs)}x{F}", nblk, fmt, build(fmt),
                                           prepare, use_cache=False, log=log, resident=True,
                                           xcd=True, team=True)
        pipe = build(fmt)(slots, grid)
        src = MixedSource(pipe, 8, args.seed + 5)
        errs, blocks, its, disagree = 0, 0, [], 0
        nfr = max(1, args.mixed_fer_frames // F)
        for w in range(nfr):
            cells, s2, idx = src.frame(block_esn0(pipe.specs, args.seed + 1000 + w), start=3 * w)
            pipe.load(cells, s2)
            pipe.run()
            errs += int(src.errors(idx).sum().item())
            blocks += nblk
            its.append(pipe.plan.its.clone())
            if w == 0:
                if fmt == "float32":
                    ref_bits = pipe.plan.bits.clone()
                elif ref_bits is not None:
                    disagree = float((pipe.plan.bits != ref_bits).float().mean().item())
        pipe.plan.check()
        its = torch.cat(its).float()
        prepare(pipe)
        g = pipe.capture()
        ms_graph = autotune.time_graph(g, max(args.steps, 5))
        pipe.plan.check()
        # the same work without the graph (launch overheads included)
        pipe.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(max(args.steps, 5)):
            pipe.run()
        torch.cuda.synchronize()
        ms_eager = (time.perf_counter() - t0) * 1e3 / max(args.steps, 5)
        from atsc3_amd.codes import get_code
        kbits = sum(get_code(r, n).K for n, r, m in pipe.specs)
        # the compulsory-traffic model over the frame's blocks at the
        # iteration counts of the timed (prepared) frame, each at its own N, E
        its_t = pipe.plan.its.cpu().numpy()
        s_fmt = {"float32": 4, "float16": 2, "int8": 1}[fmt]
        frame_bytes = sum(model_bytes([its_t[b]], n, get_code(r, n).nedges, s_fmt)
                          for b, (n, r, m) in enumerate(pipe.specs))
        res[fmt] = {"engine": pipe.engine, "slots": slots, "grid": grid, "tune_ms": table,
                    "ms_per_launch_graph": round(ms_graph, 4),
                    "ms_per_launch_eager": round(ms_eager, 4),
                    "blocks_per_s": round(nblk / (ms_graph / 1e3), 1),
                    "info_gbit_s": round(kbits / (ms_graph / 1e3) / 1e9, 3),
                    "roofline": {"bound": "hbm", "model_bytes": round(frame_bytes),
                                 "achieved": round(frame_bytes / (ms_graph / 1e3) / 1e9, 1),
                                 "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": round(frame_bytes / (ms_graph / 1e3) / 1e9
                                               / HBM_PEAK_GBS, 4),
                                 "note": "per block 4N + N + its * (2 E s + 4N) at the timed "
                                         "frame's iteration counts, over one graph replay "
                                         "(demaps included)"},
                    "fer_blocks": blocks, "fer_errors": errs,
                    "its_min": int(its.min().item()), "its_max": int(its.max().item()),
                    "its_mean": round(float(its.mean().item()), 2),
                    "bit_disagreement_vs_fp32": disagree, "rounds": pipe.plan.rounds()}
        log(f"[mixed] {fmt}: {pipe.engine} slots {slots} grid {grid}: graph {ms_graph:.3f} ms "
            f"(eager {ms_eager:.3f}) per {nblk} blocks = {nblk / (ms_graph / 1e3):.0f} blocks/s; "
            f"{errs}/{blocks} block errors, its {its.min().item():.0f}..{its.max().item():.0f}")
        del pipe, g
        torch.cuda.empty_cache()
    return res


def learn_bench(args, dev, steps, eval_blocks, ebn0_points=(1.0, 1.5, 2.0, 2.5, 3.0)):
    """Differentiation: fit per-iteration, per-check-degree normalization
    factors of the 20-iteration unrolled min-sum (16200 r6/15, QPSK, AWGN,
    Eb/N0 ~ U[1, 3] dB, batch 512, BCE on the output LLRs), then FER of the
    learned table vs the plain one on fresh blocks."""
    import torch
    from atsc3_amd.learn import Trainer

    tr = Trainer(n=16200, rate="6/15", mod="QPSK", iters=20, batch=512, ebn0=(1.0, 3.0),
                 lr=args.learn_lr, seed=args.seed, device=dev, ncw=1024)
    tr.step()                       # warm-up (first launches)
    torch.cuda.synchronize()
    secs = tr.train(steps, log=log)
    per_pt = max(1, eval_blocks // len(ebn0_points))
    fer = {}
    for eb in ebn0_points:
        f, e = tr.evaluate(eb, per_pt)
        fer[f"{eb:g}"] = {"blocks": per_pt, **{k: int(v) for k, v in e.items()}}
        log(f"[learn] Eb/N0 {eb:g} dB: " + ", ".join(f"{k} {v}/{per_pt}" for k, v in e.items()))
    a = tr.alphas.detach().cpu().numpy()
    return {"steps": steps + 1, "batch": 512, "iters": 20, "classes": tr.graph.classes,
            "ms_per_step": round(secs * 1e3 / steps, 3),
            "loss_first": round(tr.history[0], 5), "loss_last": round(float(np.mean(tr.history[-10:])), 5),
            "alphas": [[round(float(x), 4) for x in row] for row in a],
            "fer_block_errors": fer}


def ldm_cti_bench(args, dev, fmt="float16", nrows=1024, every=4, engine="auto"):
    """The LDM chain fed as the upstream receives it (lab/m44_ldm.py:1187-1304):
    the combined two-layer cells go through a convolutional time interleaver
    (Nrows 1024), are cut into frames of 1,216,939 cells and pushed one frame
    at a time into the device ring; every ``every`` frames the complete FEC
    Blocks are de-interleaved on the device and decoded by the two-layer
    chain.  Real time = decoded cells / the frame cell rate.  The chunks are
    small (~150 core blocks), so the default engine is the batched one,
    which spreads each codeword over the chip; the resident engine (one
    codeword per compute unit) measured 74x against 123x real time here."""
    import torch
    from atsc3_amd.ldm import LdmPipeline, LdmStream
    from atsc3_amd.ti import CtiStreamDev
    from atsc3_amd import synth
    pipe = LdmPipeline(device=dev, msg_dtype=fmt,
                       engine=engine if fmt == "float16" else "auto")
    stream = LdmStream(pipe, 64, args.seed + 29)
    plp, ncell, nframes = LDM_FRAME_CELLS, pipe.ncell_c, args.ldm_frames
    c0, s0 = 1234, 345
    total = nframes * plp
    nblk = (total - c0) // ncell
    n0 = synth.n0_of(args.ldm_esn0_cti)
    y, ic, ie = stream.batch(nblk, n0, start=3 * 10 ** 5)
    pre = torch.zeros(total, dtype=torch.complex64, device=dev)
    pre[c0:c0 + nblk * ncell] = y.reshape(-1)
    i = torch.arange(total, device=dev)
    q = i + nrows * ((s0 + i) % nrows)
    ok = q < total
    tx = torch.zeros(total, dtype=torch.complex64, device=dev)
    tx[q[ok]] = pre[i[ok]]
    frames = tx.view(nframes, plp)
    del pre, i, q, ok, y
    cti = CtiStreamDev(nrows, ncell, plp, cap_frames=every + 3, device=dev)
    cti.reset(0, s0, c0 + nrows * ((s0 + c0) % nrows))
    ic_h, ie_h = ic, ie.view(nblk, pipe.ratio)
    done, ec, ee = 0, 0, 0
    pipe.run(stream.batch(2, n0)[0], n0)             # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for f in range(nframes):
        cti.push(f, frames[f])
        if (f + 1) % every and f + 1 < nframes:
            continue
        first, cells = cti.take_blocks()
        nb = cells.shape[0]
        if nb == 0:
            continue
        out = pipe.run(cells, n0)
        sel = ic_h[first:first + nb]
        ec += (out["core"][0] != stream.cw_c_dev.index_select(0, sel)).any(1).sum()
        esel = ie_h[first:first + nb].reshape(-1)
        ee += (out["enh"][0] != stream.cw_e_dev.index_select(0, esel)).any(1).sum()
        done += nb
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3
    rt = done * ncell / plp * FRAME_SEC / (ms / 1e3)
    res = {"messages": fmt, "engine": pipe.engine, "nrows": nrows, "frames_pushed": nframes,
           "take_every": every,
           "core_blocks_decoded": done, "enh_blocks_decoded": done * pipe.ratio,
           "esn0_db": args.ldm_esn0_cti, "core_errors": int(ec), "enh_errors": int(ee),
           "ms_total": round(ms, 2), "realtime_factor": round(rt, 2),
           "zero_filled_frames": int(cti.stats["zero_filled_frames"]),
           "overruns": int(cti.stats["buffer_overrun"])}
    log(f"[ldm] CTI stream {fmt} ({pipe.engine}): {done} core blocks over {nframes} frames in {ms:.1f} ms = "
        f"{rt:.1f}x real time; errors core {int(ec)}, enhanced {int(ee)}")
    return res


def frame_chain_bench(args, dev, esn0=19.0, frames=(1, 8)):
    """One RF33-like frame through the whole device chain, per frame: CPE of
    the 211,472-cell pool, the 74-block 64QAM 11/15 PLP demapped through the
    hybrid time de-interleaver's gather, the alpha ladder inside one decode
    launch (the team mode for one frame, the batched engine for eight), the
    sum-product rescue, BCH + descrambling; the QPSK 2/15 PLP
    demapped and decoded beside it.  Synthetic frames (random information
    bits, so the BCH verdict is not the check here; block errors are counted
    against the sent codewords), per-symbol gain and AWGN.  Timed for one
    frame per step (latency) and ``frames[-1]`` frames per step."""
    import torch
    from atsc3_amd import bicm, ldpc, synth
    from atsc3_amd.baseband import Baseband
    from atsc3_amd.codes import get_code
    from atsc3_amd.cpe import Cpe, dummy_values
    from atsc3_amd.demap import Demapper
    from atsc3_amd.rescue import rescue_dev
    from atsc3_amd.ti import HtiGather, hti_index
    rng = np.random.default_rng(args.seed + 41)
    code0, code16 = get_code("11/15", 16200), get_code("2/15", 16200)
    n0c, dstart = 74 * 2700, 74 * 2700 + 8100
    npool = dstart + 3572
    dv = dummy_values(npool, dstart)
    region = np.full(npool, 2, np.int8)
    region[:n0c] = 0
    region[n0c:dstart] = 1
    symbol_of = np.concatenate([np.full(2700, -1), np.arange(npool - 2700) // 5800])
    idx = hti_index(n0c, 1, 2700, 74)
    out = {}
    for nf in frames:
        cw0 = synth.codewords("11/15", 16200, 74 * nf, args.seed + 43 + nf)
        cw16 = synth.codewords("2/15", 16200, nf, args.seed + 44 + nf)
        x0 = synth.modulate(cw0, 16200, "64QAM", "11/15").reshape(nf, 74, 2700)
        x16 = synth.modulate(cw16, 16200, "QPSK", "2/15")
        pools = np.empty((nf, npool), np.complex128)
        for f in range(nf):
            tx = np.empty(n0c, np.complex128)
            tx[idx.reshape(-1)] = x0[f].reshape(-1)       # interleave = scatter through the map
            pools[f] = np.concatenate([tx, x16[f], dv[dstart:]])
        syms = np.unique(symbol_of)
        g = ((1 + rng.uniform(-0.1, 0.1, (nf, len(syms))))
             * np.exp(1j * rng.uniform(-0.1, 0.1, (nf, len(syms)))))
        pos = np.searchsorted(syms, symbol_of)
        nz = (rng.standard_normal(pools.shape) + 1j * rng.standard_normal(pools.shape))
        y = (g[:, pos] * pools + nz * np.sqrt(synth.n0_of(esn0) / 2)).astype(np.complex64)
        # nf frames as one pool of nf * npool cells (symbols numbered per frame)
        sym_all = (np.arange(nf)[:, None] * (syms.max() + 2) + symbol_of[None, :] + 1).reshape(-1)
        cpe = Cpe(sym_all, np.tile(region, nf),
                  [bicm.points("64QAM", "11/15"), bicm.points("QPSK", "2/15")],
                  np.tile(dv, nf), dev)
        yd = torch.from_numpy(y.reshape(-1)).to(dev)
        hti = HtiGather(n0c, 1, 2700, 74, 0, dev, frames=nf, stride=npool)
        dm0 = Demapper(16200, "64QAM", "11/15", dev)
        dm16 = Demapper(16200, "QPSK", "2/15", dev)
        bb = Baseband("11/15", 16200, dev)
        sel16 = torch.from_numpy((np.arange(nf)[:, None] * npool + n0c
                                  + np.arange(8100)[None, :]).reshape(-1)).to(dev)
        cur = torch.cuda.current_stream(dev)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(6)]

        def step():
            ev[0].record(cur)
            z, _ = cpe.run(yd)
            ev[1].record(cur)
            llr0, _ = hti.demap(dm0, z)
            ev[2].record(cur)
            bits, conv, its, bad = ldpc.ladder_dev(llr0, code0, (1.0, 0.85, 0.75), 50)
            ev[3].record(cur)
            rescue_dev(llr0, bits, conv, its, bad, code0, budget_s=0.12, skip_if_idle=True)
            ev[4].record(cur)
            bb.run(bits)
            llr16, _ = dm16.demap_dev(z.index_select(0, sel16).view(nf, 8100))
            r16 = ldpc.decode_dev(llr16, code16, 50, 0.75)
            ev[5].record(cur)
            return bits, r16
        step()
        torch.cuda.synchronize()
        acc = np.zeros(5)
        for _ in range(args.steps):
            bits, r16 = step()
            torch.cuda.synchronize()
            acc += [ev[k].elapsed_time(ev[k + 1]) for k in range(5)]
        acc /= args.steps
        e0 = int((bits.cpu() != torch.from_numpy(cw0)).any(1).sum())
        e16 = int((r16[0].cpu() != torch.from_numpy(cw16)).any(1).sum())
        tot = float(acc.sum())
        out[f"{nf}_frames"] = {
            "ms_per_step": round(tot, 3), "ms_per_frame": round(tot / nf, 3),
            "stages_ms": {k: round(float(v), 3) for k, v in
                          zip(("cpe", "hti_demap", "ladder_decode", "rescue",
                               "baseband_plp16"), acc)},
            "realtime_factor": round(nf * FRAME_SEC / (tot / 1e3), 1),
            "block_errors_plp0": e0, "blocks_plp0": 74 * nf, "block_errors_plp16": e16}
        log(f"[bench] frame chain {nf} frame(s)/step at {esn0:g} dB: {tot:.3f} ms "
            f"({', '.join(f'{k} {v:.3f}' for k, v in out[f'{nf}_frames']['stages_ms'].items())}); "
            f"errors {e0}/{74 * nf} + {e16}/{nf}")
        del cpe, yd, hti
        torch.cuda.empty_cache()
    out["esn0_db"] = esn0
    return out


def rescue_bench(args, dev, rate="10/15", snr=2.8, nblk=64, max_blocks=8, iters=100):
    """Near-threshold batch of one 64800-bit code: the alpha ladder inside the
    batched kernel (against one host-synchronised decode per rung), then the
    float64 sum-product rescue of the closest near misses (select, decode,
    write-back on the device), beside the numpy decoder of the upstream's
    statement on the host."""
    import torch
    from atsc3_amd import ldpc, synth
    from atsc3_amd.codes import get_code
    from atsc3_amd.rescue import rescue_dev
    from oracle import sp_ref
    code = get_code(rate, 64800)
    cw = synth.codewords(rate, 64800, nblk, args.seed + 5)
    llr_h = synth.llr_bpsk_like(cw, snr, args.seed + 6)
    llr = torch.from_numpy(llr_h).to(dev)
    cur = torch.cuda.current_stream(dev)
    ladder = (1.0, 0.85, 0.75)
    ms = {}
    for eng in ("batch", "qc"):
        ldpc.ladder_dev(llr, code, ladder, 50, engine=eng)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            lad = ldpc.ladder_dev(llr, code, ladder, 50, engine=eng)
        torch.cuda.synchronize()
        ms[eng] = (time.perf_counter() - t0) * 1e3 / args.steps
    bits0, conv0, its0, bad0 = ldpc.ladder_dev(llr, code, ladder, 50, engine="batch")
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    times = []
    for step in range(args.warmup + args.steps):
        bits, conv, its, bad = (x.clone() for x in (bits0, conv0, its0, bad0))
        e0.record(cur)
        rows, count, c2, s2 = rescue_dev(llr, bits, conv, its, bad, code, 1500, max_blocks,
                                         iters, budget_s=None)
        e1.record(cur)
        torch.cuda.synchronize()
        if step >= args.warmup:
            times.append(e0.elapsed_time(e1))
    r_ms = float(np.median(times))
    offered = int(count.item())
    rescued = int(c2[:offered].sum().item())
    conv_h = conv.cpu().numpy().astype(bool)
    wrong = int(sum(not np.array_equal(bits[b].cpu().numpy(), cw[b])
                    for b in np.flatnonzero(conv_h)))
    its_h = its.cpu().numpy()
    sp_its = int(its_h[rows.cpu().numpy()[:offered]].max()) if offered else 0
    # the upstream's numpy decoder (restated by the oracle) on one block, a
    # bounded number of iterations
    k = 3
    t0 = time.perf_counter()
    sp_ref.sum_product(llr_h[:1], code.checks, 64800, k)
    cpu_ms_blk_it = (time.perf_counter() - t0) * 1e3 / (k + 1)
    out = {"code": f"{rate} N=64800", "snr_db": snr, "blocks": nblk,
           "ladder_ms": {"in_kernel": round(ms["batch"], 3), "host_per_rung": round(ms["qc"], 3)},
           "ladder_converged": int(conv0.sum().item()),
           "rescue_offered": offered, "rescue_converged": rescued, "rescue_iters_max": sp_its,
           "converged_after": int(conv_h.sum()), "wrong_codewords": wrong,
           "rescue_ms": round(r_ms, 3), "rescue_iters_limit": iters,
           "cpu_numpy_ms_per_block_iteration": round(cpu_ms_blk_it, 1),
           "cpu_sample": f"float64 numpy sum-product, 1 block x {k} iterations, 1 core"}
    log(f"[bench] rescue: ladder {ms['batch']:.2f} ms in kernel vs {ms['qc']:.2f} ms per-rung; "
        f"{int(conv0.sum().item())}/{nblk} converged, {offered} offered, {rescued} rescued in "
        f"{r_ms:.2f} ms (<= {sp_its} its); numpy {cpu_ms_blk_it:.1f} ms per block-iteration")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--frames", type=int, default=100)
    ap.add_argument("--msg", default="float16", choices=["float32", "float16", "int8"],
                    help="message storage of the headline decode (float16 = the batched "
                         "FP16 form; plain FP32 is always measured beside it)")
    ap.add_argument("--engine", default="resident", choices=["resident", "auto"],
                    help="decoder of the FP16 headline: resident (one workgroup per "
                         "codeword, FP16 posteriors in LDS) or auto (the batched engine)")
    ap.add_argument("--esn0", type=float, default=None, help="skip the sweep")
    ap.add_argument("--fer-blocks", type=int, default=30000)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the FP16 / INT8 message-storage comparison")
    ap.add_argument("--form-fer-blocks", type=int, default=30000,
                    help="blocks per FER point of the message-storage forms (rounded up "
                         "to whole windows)")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--workload", default="live", choices=["live", "ldm", "mixed", "learn"])
    ap.add_argument("--learn-steps", type=int, default=300)
    ap.add_argument("--learn-eval", type=int, default=100000)
    ap.add_argument("--learn-lr", type=float, default=0.01)
    ap.add_argument("--no-learn", action="store_true",
                    help="skip the short alpha-fitting summary in the live-frame line")
    ap.add_argument("--mixed-frames", type=int, default=1,
                    help="frames of the uneven workload per batched launch")
    ap.add_argument("--mixed-fer-frames", type=int, default=200)
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the uneven-frame summary in the live-frame line")
    ap.add_argument("--ldm-frames", type=int, default=50)
    ap.add_argument("--no-frame-chain", action="store_true",
                    help="skip the whole-chain per-frame timing")
    ap.add_argument("--no-rescue", action="store_true",
                    help="skip the ladder + sum-product rescue summary")
    ap.add_argument("--ldm-esn0", type=float, default=LDM_ESN0)
    ap.add_argument("--ldm-esn0-cti", type=float, default=21.0,
                    help="Es/N0 of the CTI-streamed LDM leg (both layers decode)")
    ap.add_argument("--no-ldm", action="store_true",
                    help="skip the LDM summary in the live-frame line")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 1 GiB device-to-device copy first (reference "
                         "kernel for scaling rocprof FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; "
                          "--gpus must be 1 (blocks shard with no exchange: see "
                          "the sharding design notes)"}))
        sys.exit(2)

    import torch
    from atsc3_amd import ldpc, synth
    from atsc3_amd.codes import get_code
    from atsc3_amd.demap import Demapper
    from atsc3_amd.resident import GROW_TEAM

    dev = torch.device("cuda:0")
    if args.copy_calibration:
        a = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    if args.workload == "ldm":
        fname = ("float16_resident" if args.msg == "float16" and args.engine == "resident"
                 else args.msg)
        res = ldm_bench(args, dev, (fname,))[fname]
        res["cti_stream"] = ldm_cti_bench(args, dev, args.msg)
        print(json.dumps({
            "metric": "LDM frame real-time factor (core + enhanced demap, decode, SIC)",
            "value": res["realtime_factor"], "unit": "x real time", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": res["ms_per_step"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic",
            "config": {"workload": (f"LDM: {args.ldm_frames} frames x {LDM_FRAME_CELLS} cells, "
                                    f"core QPSK 4/15 + enhanced 64-NUC 10/15 N=64800 at -5 dB "
                                    f"injection, AWGN"), "messages": args.msg, **res},
            "roofline": None, "cpu_baseline": None}), flush=True)
        return
    if args.workload == "learn":
        res = learn_bench(args, dev, args.learn_steps, args.learn_eval)
        tot = {k: sum(v[k] for v in res["fer_block_errors"].values())
               for k in ("learned", "plain_unrolled", "plain_minsum", "plain_minsum_50",
                         "plain_ladder_50", "learned_fallback")}
        nb = sum(v["blocks"] for v in res["fer_block_errors"].values())
        print(json.dumps({
            "metric": "training step time, unrolled min-sum + soft demapper (forward + backward)",
            "value": res["ms_per_step"], "unit": "ms/step", "n_gpus": 1, "steps": res["steps"],
            "warmup": 1, "ms_per_step": res["ms_per_step"], "higher_is_better": False,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": ("differentiation: 20-iteration unrolled normalized min-sum, "
                                    "16200 r6/15 QPSK AWGN Eb/N0 U[1,3] dB, batch 512, "
                                    "per-iteration per-degree factors"),
                       "fer_total": {k: v / nb for k, v in tot.items()}, "eval_blocks": nb, **res},
            "roofline": None, "cpu_baseline": None}), flush=True)
        return
    if args.workload == "mixed":
        res = mixed_bench(args, dev, (args.msg,))
        r = res[args.msg]
        print(json.dumps({
            "metric": "FEC blocks/s, uneven frames (mixed codes and constellations)",
            "value": r["blocks_per_s"], "unit": "FEC blocks/s", "n_gpus": 1,
            "steps": max(args.steps, 5), "warmup": 1, "ms_per_step": r["ms_per_launch_graph"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic",
            "config": {"workload": (f"uneven frames: {res['blocks_per_frame']} blocks per frame x "
                                    f"{args.mixed_frames} frames per launch, 64800 r6/8/10/12 + "
                                    f"16200 r4/9 on QPSK..256-NUC, per-block AWGN"),
                       "messages": args.msg, **res},
            "roofline": None, "cpu_baseline": None}), flush=True)
        return
    code = get_code(RATE, N)
    E = code.nedges
    ncell = N // 8
    nblk = args.frames * FRAME_CELLS // ncell
    s_msg = {"float32": 4, "float16": 2, "int8": 1}[args.msg]
    engine = args.engine if args.msg == "float16" else "auto"
    log(f"[bench] {nblk} blocks per step ({args.frames} frames x {FRAME_CELLS} cells)")
    t0 = time.time()
    stream = synth.FadingStream(RATE, N, MOD, 128, args.seed, dev)
    dm = Demapper(N, MOD, RATE, dev)
    log(f"[bench] codewords ready in {time.time() - t0:.1f} s")

    def run_window(z, w, n0, msg, engine="auto"):
        llr, _ = dm.demap_dev(z, csi=w, sigma2=n0)
        return ldpc.decode_dev(llr, code, 50, 0.75, msg, engine)

    def errors(idx, out):
        bits = out[0]
        ref = stream.cw_dev.index_select(0, idx)
        return int((bits != ref).any(dim=1).sum().item())

    def _msg_pass(es, k):
        z, w, idx, n0 = stream.batch(nblk, es, start=k * nblk)
        return idx, run_window(z, w, n0, args.msg, engine)

    # -- operating point ------------------------------------------------------
    fer = {}
    if args.esn0 is None:
        op = None
        for es in SWEEP:
            z, w, idx, n0 = stream.batch(nblk, es)
            out = run_window(z, w, n0, "float32", PLAIN_ENGINE)
            e = errors(idx, out)
            fer[es] = dict(blocks=nblk, errors=e,
                           its_mean=float(out[2].float().mean().item()))
            log(f"[bench] sweep Es/N0 {es:.0f} dB: {e}/{nblk} block errors, "
                f"mean its {fer[es]['its_mean']:.2f}")
            if e == 0 and op is None:
                op = es
        if op is None:
            op = SWEEP[-1]
        while True:
            # FER <= 1e-4 for the plain decoder AND the headline format: zero
            # block errors over >= --fer-blocks blocks (same blocks for both)
            tot_b, tot_e = fer[op]["blocks"], fer[op]["errors"]
            e_msg = errors(*_msg_pass(op, 0)) if args.msg != "float32" else tot_e
            k = 1
            while tot_b < args.fer_blocks and tot_e == 0 and e_msg == 0:
                z, w, idx, n0 = stream.batch(nblk, op, start=k * nblk)
                llr_k, _ = dm.demap_dev(z, csi=w, sigma2=n0)
                tot_e += errors(idx, ldpc.decode_dev(llr_k, code, 50, 0.75, "float32",
                                                     PLAIN_ENGINE))
                if args.msg != "float32":
                    e_msg += errors(idx, ldpc.decode_dev(llr_k, code, 50, 0.75, args.msg,
                                                         engine))
                tot_b += nblk
                k += 1
            fer[op].update(blocks=tot_b, errors=tot_e, **({f"errors_{args.msg}": e_msg}
                                                          if args.msg != "float32" else {}))
            log(f"[bench] confirm Es/N0 {op:.0f} dB: {tot_e}/{tot_b} block errors (float32), "
                f"{e_msg}/{tot_b} ({args.msg})")
            if (tot_e == 0 and e_msg == 0) or op >= SWEEP[-1]:
                break
            op += 1.0
            if op not in fer:
                z, w, idx, n0 = stream.batch(nblk, op)
                out = run_window(z, w, n0, "float32", PLAIN_ENGINE)
                fer[op] = dict(blocks=nblk, errors=errors(idx, out),
                               its_mean=float(out[2].float().mean().item()))
    else:
        op = args.esn0
    # -- timed steps ------------------------------------------------------------
    z, w, idx, n0 = stream.batch(nblk, op, start=10 ** 6)
    cur = torch.cuda.current_stream(dev)
    # The receiver pipeline: each step demaps its window on one stream and
    # decodes it on another, so the demap of step k + 1 runs while step k
    # decodes (the resident decoder leaves VALU issue slots free). Both
    # streams keep the default priority. Every step still does the full
    # demap and decode of its window; the sequential figure is reported too.
    s_dm = torch.cuda.Stream(dev)
    s_dec = torch.cuda.Stream(dev)

    def pipelined(nsteps, ev=None, msg=args.msg, eng=engine):
        out = None
        for k in range(nsteps):
            with torch.cuda.stream(s_dm):
                llr, _ = dm.demap_dev(z, csi=w, sigma2=n0)
                done = torch.cuda.Event()
                done.record(s_dm)
            with torch.cuda.stream(s_dec):
                s_dec.wait_event(done)
                llr.record_stream(s_dec)
                if ev is not None:
                    ev[k][0].record(s_dec)
                out = ldpc.decode_dev(llr, code, 50, 0.75, msg, eng)
                if ev is not None:
                    ev[k][1].record(s_dec)
        return out

    if args.warmup:
        out = pipelined(args.warmup)
    torch.cuda.synchronize()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(args.steps)]
    t_start = time.perf_counter()
    out = pipelined(args.steps, ev)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    # the decode call on its stream, overlapped with the next demap
    decode_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in ev]))
    ms_per_step = wall * 1e3 / args.steps
    # the same steps one after the other on one stream; the demap's own time
    # comes from here (in the pipeline its stream runs ahead and its events
    # also span the wait for compute units the decode holds)
    ev_s = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(args.steps)]
    t_seq = time.perf_counter()
    for k in range(args.steps):
        ev_s[k][0].record(cur)
        llr_s, _ = dm.demap_dev(z, csi=w, sigma2=n0)
        ev_s[k][1].record(cur)
        out_s = ldpc.decode_dev(llr_s, code, 50, 0.75, args.msg, engine)
    torch.cuda.synchronize()
    seq_ms_per_step = (time.perf_counter() - t_seq) * 1e3 / args.steps
    demap_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in ev_s]))
    seq_errors = errors(idx, out_s)
    its = out[2].cpu().numpy()
    conv = out[1].cpu().numpy().astype(bool)
    step_errors = errors(idx, out)
    value = nblk / (ms_per_step / 1e3)
    # roofline: the compulsory-traffic model of the decode path, per block
    # 4N + N + its * (2 E s + 4N) with s the stored message width, over the
    # decode's measured duration; the resident engine's own model (records
    # instead of per-edge messages) is reported beside it
    achieved = model_bytes(its, N, E, s_msg) / (decode_ms / 1e3) / 1e9
    engine_model = None
    if engine == "resident":
        eb = resident_model_bytes(its, N, code.nc)
        engine_model = {"bytes_per_launch": round(eb), "achieved": round(eb / (decode_ms / 1e3) / 1e9, 1),
                        "frac": round(eb / (decode_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                        "model": "per block 4N + N + its * (3 * 8 * M + 4N): each check's 8-byte "
                                 "record read and written by the check phase and read by the "
                                 "variable phase, LLRs re-read"}
        key = f"live_frame_{args.msg}_resident_{args.frames}f_es{op:g}"
    else:
        key = f"live_frame_{args.msg}_{args.frames}f_es{op:g}"
    traffic = load_traffic(key)          # measured HBM bytes per decode call
    log(f"[bench] Es/N0 {op:g} dB, {args.msg} ({engine}): {ms_per_step:.3f} ms/step (demap {demap_ms:.3f} ms, "
        f"decode {decode_ms:.3f} ms), mean its {its.mean():.2f}, "
        f"conv {conv.mean():.4f}, errors {step_errors}; one stream: {seq_ms_per_step:.3f} ms/step, "
        f"errors {seq_errors}")
    # the Baseband Packet step on the decoded window (BCH syndrome + descramble);
    # the synthetic codewords carry random (non-BCH) information bits, so this
    # is a throughput figure
    from atsc3_amd.baseband import Baseband
    bbp = Baseband(RATE, N, dev)
    bbp.run(out[0])
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(cur)
    for _ in range(args.steps):
        bbp.run(out[0])
    e1.record(cur)
    torch.cuda.synchronize()
    baseband_ms = e0.elapsed_time(e1) / args.steps
    # the same window read through the hybrid time de-interleaver's gather
    # (the window's cells taken as the transmitted order of `frames` TI
    # blocks): the gather fused into the demap read against the plain demap
    from atsc3_amd.ti import HtiGather
    ncols = nblk // args.frames
    hti = HtiGather(args.frames * ncols * ncell, args.frames, ncell, ncols, 0, dev)
    pool, pool_w = z.reshape(-1), w.reshape(-1)
    demap_times = {}
    # sigma2 and the outputs resident on the device: the events time the
    # demap kernels, not per-call host-to-device copies of sigma2
    s2_dev = torch.full((nblk,), float(n0), dtype=torch.float64, device=dev)
    llr_dm = torch.empty((nblk, N), dtype=torch.float32, device=dev)
    llr_hti = torch.empty((hti.nblk, N), dtype=torch.float32, device=dev)
    for name, fn in (("plain", lambda: dm.demap_dev(z, csi=w, sigma2=s2_dev, out=llr_dm,
                                                    want_sigma2=False)),
                     ("hti_gather", lambda: hti.demap(dm, pool, pool_w, s2_dev[:hti.nblk],
                                                      llr_hti))):
        fn()
        torch.cuda.synchronize()
        e0.record(cur)
        for _ in range(args.steps):
            fn()
        e1.record(cur)
        torch.cuda.synchronize()
        demap_times[name] = round(e0.elapsed_time(e1) / args.steps, 4)
    log(f"[bench] demap {demap_times['plain']:.3f} ms, through the HTI gather "
        f"{demap_times['hti_gather']:.3f} ms ({hti.nblk} blocks)")
    # the demap against both of its bounds: bytes per cell (the complex64
    # cell, its f32 CSI weight, nb f32 LLRs written; sigma2 per block and the
    # de-interleave table, shared by every block and L2-resident, left out)
    # over HBM, and VALU instructions per cell over the VALU issue rate
    cells = nblk * ncell
    nbits = N // ncell
    dm_bytes = 8 + 4 + 4 * nbits
    t_dm = demap_times["plain"] / 1e3
    dm_hbm = cells * dm_bytes / t_dm / 1e9
    dm_valu = cells * DEMAP_VALU_PER_CELL_256 / t_dm
    demap_roofline = {
        "cells": cells, "ms": demap_times["plain"],
        "hbm": {"bytes_per_cell": dm_bytes, "achieved": round(dm_hbm, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(dm_hbm / HBM_PEAK_GBS, 4)},
        "valu": {"insts_per_cell": DEMAP_VALU_PER_CELL_256,
                 "achieved": round(dm_valu / 1e12, 3), "peak": round(VALU_PEAK_LANE_OPS / 1e12, 3),
                 "unit": "T lane-instructions/s", "frac": round(dm_valu / VALU_PEAK_LANE_OPS, 4)},
    }
    demap_roofline["bound"] = ("valu" if demap_roofline["valu"]["frac"]
                               >= demap_roofline["hbm"]["frac"] else "hbm")
    log(f"[bench] demap roofline: {dm_hbm:.0f} GB/s = {dm_hbm / HBM_PEAK_GBS:.3f} of HBM, "
        f"VALU {dm_valu / VALU_PEAK_LANE_OPS:.3f} of issue: {demap_roofline['bound']}-bound")
    # common-phase-error correction of the same cells taken as a pool of
    # OFDM symbols of 5,700 cells (3 rounds, 256-NUC alphabet)
    from atsc3_amd.cpe import Cpe
    from atsc3_amd import bicm as _bicm
    npool = pool.numel()
    cpe = Cpe(np.arange(npool) // 5700, np.zeros(npool, np.int8),
              [_bicm.points(MOD, RATE)], None, dev)
    cpe_ms = {}
    for decim in (1, 4):
        cpe.run(pool, decim=decim)
        torch.cuda.synchronize()
        e0.record(cur)
        for _ in range(args.steps):
            cpe.run(pool, decim=decim)
        e1.record(cur)
        torch.cuda.synchronize()
        cpe_ms[f"decim{decim}"] = round(e0.elapsed_time(e1) / args.steps, 4)
    log(f"[bench] CPE over {npool} cells ({len(cpe.symbols)} symbols): {cpe_ms} ms")
    del cpe
    # plain FP32 (the bit-exact decoder) on the same resident window, through
    # the same two-stream pipeline as the headline, and on one stream
    pipelined(1, msg="float32", eng=PLAIN_ENGINE)
    run_window(z, w, n0, "float32", PLAIN_ENGINE)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out32 = pipelined(args.steps, msg="float32", eng=PLAIN_ENGINE)
    torch.cuda.synchronize()
    ms32 = (time.perf_counter() - t0) * 1e3 / args.steps
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out32s = run_window(z, w, n0, "float32", PLAIN_ENGINE)
    torch.cuda.synchronize()
    ms32s = (time.perf_counter() - t0) * 1e3 / args.steps
    plain = {"engine": PLAIN_ENGINE,
             "blocks_per_s": round(nblk / (ms32 / 1e3), 1), "ms_per_step": round(ms32, 4),
             "sequential_ms_per_step": round(ms32s, 4),
             "errors": errors(idx, out32), "errors_sequential": errors(idx, out32s),
             "mean_its": round(float(out32[2].float().mean().item()), 3)}
    log(f"[bench] plain float32: {ms32:.3f} ms/step = {nblk / (ms32 / 1e3):.0f} blocks/s "
        f"(one stream {ms32s:.3f} ms/step)")
    llr_w, _ = dm.demap_dev(z, csi=w, sigma2=n0)
    plain["dropin"] = dropin_bench(llr_w, code)
    plain["update_then_check"] = utc_bench(llr_w, code)
    del llr_w

    # -- message-storage forms at the same operating point ----------------------
    # Throughput of every (form, engine) at the operating point; then FER
    # curves over FORM_SWEEP (and the operating point) with >= --form-fer-
    # blocks blocks per point, every window demapped once and decoded by
    # every form (the same blocks for all); the Es/N0 gap to plain FP32 where
    # each curve crosses FER 1e-2 / 1e-3 / 1e-4; and the output-LLR error of
    # the FP16 / INT8 resident forms against the FP32 decoder per Es/N0:
    # each form's final posteriors against FP32's on the blocks where both
    # stopped at the same iteration, max|p - p32| / max|p32| per block.
    forms = {}
    if not args.no_forms:
        from atsc3_amd.resident import ResidentPlan
        form_defs = (("float32", "float32", ("phase", "batch", "resident")),
                     ("float16", "float16", ("phase", "batch")),
                     ("float16_resident", "float16", ("resident",)),
                     ("int8", "int8", ("phase", "batch")),
                     ("int8_resident", "int8", ("resident",)))
        for name, fmt, engines in form_defs:
            eng_ms = {}
            for eng in engines:
                run_window(z, w, n0, fmt, eng)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(args.steps):
                    fo = run_window(z, w, n0, fmt, eng)
                torch.cuda.synchronize()
                eng_ms[eng] = (time.perf_counter() - t0) * 1e3 / args.steps
            best = min(eng_ms, key=eng_ms.get)
            forms[name] = {"blocks_per_s": round(nblk / (eng_ms[best] / 1e3), 1),
                           "ms_per_step": round(eng_ms[best], 3), "engine": best,
                           "messages": fmt,
                           "ms_by_engine": {k: round(v, 3) for k, v in eng_ms.items()},
                           "mean_its": round(float(fo[2].float().mean().item()), 3)}
            log(f"[bench] form {name}: {nblk / (eng_ms[best] / 1e3):.0f} blocks/s ({best}; "
                f"{', '.join(f'{k} {v:.2f} ms' for k, v in eng_ms.items())})")
        plans = {f: ResidentPlan([code] * nblk, dev, 50, 0.75, None, f)
                 for f in ("float32", "float16", "int8")}
        posts = {f: torch.empty(plans[f].total, dtype=torch.float32, device=dev)
                 for f in plans}
        nwin = max(1, -(-args.form_fer_blocks // nblk))
        points = sorted(set(FORM_SWEEP) | {float(op)})
        curves = {name: [] for name, _, _ in form_defs}
        llr_err = {"float16_resident": {}, "int8_resident": {}}
        for es in points:
            ce = {name: 0 for name in curves}
            cits = {name: 0.0 for name in curves}
            for k in range(nwin):
                z3, w3, idx3, n03 = stream.batch(nblk, es, start=(int(es * 4) * 7919 + k) * nblk)
                llr3, _ = dm.demap_dev(z3, csi=w3, sigma2=n03)
                for name, fmt, _ in form_defs:
                    eng = forms[name]["engine"]
                    if eng == "resident" and k == 0:
                        o3 = plans[fmt].run(llr3.reshape(-1), post=posts[fmt])
                        o3 = (o3[0].view(nblk, N),) + tuple(o3[1:])
                    else:
                        o3 = ldpc.decode_dev(llr3, code, 50, 0.75, fmt, eng)
                    ce[name] += errors(idx3, o3)
                    cits[name] += float(o3[2].float().mean().item())
                    if k == 0 and name in ("float32", "float16_resident", "int8_resident"):
                        forms[name]["_its0"] = o3[2].clone()
                if k == 0:
                    p32 = posts["float32"].view(nblk, N)
                    ref_max = p32.abs().amax(1).clamp_min(1e-30)
                    for name, fmt in (("float16_resident", "float16"), ("int8_resident", "int8")):
                        same = forms[name]["_its0"] == forms["float32"]["_its0"]
                        e = ((posts[fmt].view(nblk, N) - p32).abs().amax(1) / ref_max)[same]
                        e = e.cpu().numpy()
                        llr_err[name][f"{es:g}"] = {
                            "blocks_same_its": int(same.sum().item()), "blocks": nblk,
                            "median": float(np.median(e)) if e.size else None,
                            "p99": float(np.quantile(e, 0.99)) if e.size else None,
                            "max": float(e.max()) if e.size else None}
            for name in curves:
                curves[name].append([es, ce[name], nblk * nwin, round(cits[name] / nwin, 3)])
            log(f"[bench] Es/N0 {es:g} dB ({nblk * nwin} blocks): "
                + ", ".join(f"{nm} {ce[nm]}" for nm in curves) + " block errors")
        del plans, posts
        for name in curves:
            f = forms[name]
            f.pop("_its0", None)
            f["fer_curve"] = curves[name]
            at_op = next(c for c in curves[name] if c[0] == float(op))
            f["fer_blocks"], f["fer_errors"] = at_op[2], at_op[1]
            f["meets_fer_1e-4_at_op"] = bool(at_op[2] >= 30000 and at_op[1] / at_op[2] <= 1e-4)
            if name != "float32":
                for t, tag in ((1e-2, "1e-2"), (1e-3, "1e-3"), (1e-4, "1e-4")):
                    f[f"gap_db_at_fer_{tag}"] = fer_gap_db(curves["float32"], curves[name], t)
            if name in llr_err:
                f["llr_error_vs_fp32"] = llr_err[name]
        for name in curves:
            log(f"[bench] form {name}: gaps to FP32 (dB) at FER 1e-2/1e-3/1e-4: "
                + "/".join(str(forms[name].get(f"gap_db_at_fer_{t}")) for t in ("1e-2", "1e-3", "1e-4"))
                + f"; at {op:g} dB {forms[name]['fer_errors']}/{forms[name]['fer_blocks']}")

    # -- CPU baseline (same blocks, bounded sample) ----------------------------
    cpu = None
    if not args.no_cpu:
        from oracle import cpu_pool
        # the host cores this process may use, capped at 16: the CPU share of
        # one GPU on the MI355X boxes this runs on (worker pools are sized to it)
        n_aff = len(os.sched_getaffinity(0))
        procs = min(n_aff, 16)
        zh = z.cpu().numpy()
        wh = w.cpu().numpy()
        t1, *_ = cpu_pool.run_pool(zh[:procs], wh[:procs], n0, RATE, N, MOD, 50,
                                   0.75, procs, chunk=1)
        # t1: wall time of one block per worker -> size the sample for about
        # --cpu-seconds of wall time on `procs` workers
        ns = int(min(nblk, procs * args.cpu_seconds / max(t1, 1e-3)))
        ns = max(procs, (ns // procs) * procs)
        log(f"[bench] CPU baseline: {procs} processes, sample {ns} blocks")
        secs, bits, cconv, cits, cbad, which = cpu_pool.run_pool(
            zh[:ns], wh[:ns], n0, RATE, N, MOD, 50, 0.75, procs, chunk=2)
        # against the plain FP32 decode of the same blocks (the CPU demaps in
        # numpy float32, so LLRs differ from the GPU's in the last bits)
        g_bits = out32[0][:ns].cpu().numpy()
        agree = float((bits == g_bits).all(axis=1).mean())
        its_agree = float((np.asarray(cits) == out32[2][:ns].cpu().numpy()).mean())
        t_one, *_ = cpu_pool.run_pool(zh[:2], wh[:2], n0, RATE, N, MOD, 50, 0.75, 1,
                                      chunk=1)
        cpu_model = _cpu_model()
        cpu = {"value": round(ns / secs, 2), "unit": "FEC blocks/s", "cores": procs,
               "kind": "reference" if which == "ref" else "port",
               "sample": (f"{ns} of the timed window's blocks (same cells, CSI, sigma2): "
                          f"numpy f32 demap + de-interleave + FP32 min-sum C kernel, "
                          f"{procs} spawn processes over shared memory"),
               "one_core_blocks_per_s": round(2 / t_one, 2),
               "cpu_model": cpu_model, "host_cores_visible": n_aff,
               "bits_agree_with_gpu_fp32": agree, "its_agree_with_gpu_fp32": its_agree}
        log(f"[bench] CPU: {ns / secs:.1f} blocks/s on {procs} cores ({which}); "
            f"1 core {2 / t_one:.2f} blocks/s; GPU/CPU bit agreement {agree:.4f}")

    mixed = None
    if not args.no_mixed:
        del z, w, llr_s, out_s
        torch.cuda.empty_cache()
        mixed = mixed_bench(args, dev)
    learn = None
    if not args.no_learn:
        learn = learn_bench(args, dev, 100, args.learn_eval, (1.5, 2.5))
    ldm = None
    if not args.no_ldm:
        torch.cuda.empty_cache()
        ldm = ldm_bench(args, dev)
        torch.cuda.empty_cache()
        ldm["cti_stream"] = ldm_cti_bench(args, dev)
    rescue = None
    if not args.no_rescue:
        torch.cuda.empty_cache()
        rescue = rescue_bench(args, dev)
    frame_chain = None
    if not args.no_frame_chain:
        torch.cuda.empty_cache()
        frame_chain = frame_chain_bench(args, dev)

    info_gbps = value * code.K / 1e9
    rt_factor = value / (FRAME_CELLS / ncell / FRAME_SEC)
    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "FEC blocks/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": ({"float16": "fp16 messages+posteriors, fp32 accumulate",
                   "int8": "int8 messages, fp16 posteriors, fp32 accumulate",
                   "float32": "f32"}[args.msg] if engine == "resident"
                  else {"float16": "fp16 messages, fp32 posteriors+accumulate",
                        "int8": "int8 messages, fp32 posteriors+accumulate",
                        "float32": "f32"}[args.msg]),
        "data": "synthetic",
        "config": {"workload": (f"live-frame: 256-NUC r10/15 N=64800, Rayleigh fading with "
                                f"perfect CSI, {args.frames} frames x {FRAME_CELLS} cells = "
                                f"{nblk} blocks per step, demap + decode"),
                   "esn0_db": op, "messages": args.msg, "accumulation": "float32",
                   "engine": engine,
                   "headline": (f"{args.msg} check-record messages on the resident engine "
                                f"(one codeword per compute unit, posteriors in LDS"
                                + f"; grow mode: the codewords still running once the queue "
                                  f"is empty finish on teams of {GROW_TEAM} compute units of "
                                  f"their XCD)"
                                if engine == "resident" else f"{args.msg} messages, {engine} engine"),
                   "posteriors": ("float16 in LDS" if engine == "resident" else "float32"),
                   "plain_fp32": plain, "iters_max": 50, "alpha": 0.75,
                   "baseband_ms": round(baseband_ms, 4),
                   "demap_ms_plain_vs_hti_gather": demap_times,
                   "demap_roofline": demap_roofline,
                   "cpe_ms": cpe_ms,
                   "blocks_per_step": nblk, "mean_its": round(float(its.mean()), 3),
                   "info_gbit_s": round(info_gbps, 3), "realtime_factor": round(rt_factor, 1),
                   "demap_ms": round(demap_ms, 4), "decode_ms": round(decode_ms, 4),
                   "pipeline": "demap of step k+1 on its own stream during the decode of step k",
                   "sequential_ms_per_step": round(seq_ms_per_step, 4),
                   "sequential_blocks_per_s": round(nblk / (seq_ms_per_step / 1e3), 2),
                   "fer_sweep": {f"{k:g}": v for k, v in fer.items()},
                   "forms": forms, "mixed": mixed, "ldm": ldm, "learn": learn,
                   "rescue": rescue, "frame_chain": frame_chain},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "kernel": (f"resident_grow_kernel<{ {'float32': 0, 'float16': 1, 'int8': 2}[args.msg]}>"
                                if engine == "resident" else "decode launch"),
                     "model": (f"per block 4N + N + its * (2 E s + 4N), s = {s_msg} bytes per "
                               f"stored message, N = {N}, E = {E}"),
                     "bytes_per_launch": round(model_bytes(its, N, E, s_msg)),
                     "duration_ms": round(decode_ms, 4),
                     "engine_model": engine_model},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
