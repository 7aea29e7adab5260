"""Autotuner for the batched decoder's launch shape.

Two knobs set how the persistent kernel maps onto the device:
  slots  codewords decoded concurrently ("blocks per wave"): more slots give
         every phase more parallel work and fewer rounds; fewer keep the
         working set (tot rows + messages of the active codewords) small
         enough to stay in the last-level cache;
  grid   workgroups of the persistent kernel (at most the co-resident
         capacity): more hide latency, fewer make the grid barriers cheaper.

``tune`` times every candidate on the caller's real workload (a factory that
builds the pipeline for a (slots, grid) pair; each candidate is captured in a
HIP graph and replayed), keeps the fastest, and caches the choice in a JSON
file keyed by device name and workload signature (``ATSC3_TUNE_CACHE``, else
~/.cache/atsc3_amd/batch_tune.json).  With ``resident=True`` the resident
engine (one workgroup per codeword, one per compute unit) is a candidate
too, built as slots = "resident": it wins when a launch carries many
codewords, the batched engine when a few slow codewords would leave most
compute units idle.  With ``xcd=True`` the batched engine's XCD-partitioned
mode is a candidate too (slots = "xcd": every codeword of the largest part
on each XCD): it wins when the rounds of a few slow codewords dominate, since
its barriers span one XCD and need no L2 write-back.  With ``team=True`` the
resident engine's team mode is a candidate for each team size in
``TEAM_SIZES`` (slots = "team", grid = the team size: workgroups of one XCD
per codeword; slots = "team-l2": the same without the members' LDS copies
of the posteriors): it wins when a few slow codewords would otherwise run
their iterations on one compute unit each.
"""
from __future__ import annotations

import json
import os
import time

import torch

from .batch import max_grid


def cache_path():
    return os.environ.get("ATSC3_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "atsc3_amd",
                                       "batch_tune.json"))


def _load():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(d):
    p = cache_path()
    try:
        os.makedirs(os.path.dirname(p), exist_ok=True)
        tmp = p + ".tmp"
        with open(tmp, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        os.replace(tmp, p)
    except OSError:
        pass


# team sizes tried: a partial team takes an XCD's remaining workgroups (32
# per XCD: 12 gives teams of 12, 12 and 8); on the 74-block uneven frame
# FP32 replays measured 0.94-1.05 ms with teams of 12 against 1.03-1.12
# with 8 or 16 (FP16: 0.92 with 12, 0.91 with 16)
TEAM_SIZES = (4, 8, 12, 16)


def candidates(nblk, msg_dtype, xcd=False, team=False):
    cap = max_grid(msg_dtype)
    grids = sorted({g for g in (cap, cap // 2, cap // 4, 128) if 1 <= g <= cap}, reverse=True)
    slots = sorted({max(1, s) for s in (nblk, (nblk + 1) // 2, (nblk + 3) // 4)}, reverse=True)
    out = [(s, g) for s in slots for g in grids]
    if xcd:
        capx = max_grid(msg_dtype, xcd=True)
        out += [("xcd", g) for g in sorted({capx, capx // 2} - {0}, reverse=True)]
    if team:
        out += [(kind, t) for kind in ("team", "team-l2") for t in TEAM_SIZES]
    return out


def time_graph(g, reps):
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    g.replay()
    torch.cuda.synchronize()
    e0.record(st)
    for _ in range(reps):
        g.replay()
    e1.record(st)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def tune(signature, nblk, msg_dtype, build, prepare=None, reps=5, use_cache=True, log=None,
         resident=False, xcd=False, team=False):
    """build(slots, grid) -> object with .capture() and .plan; prepare(obj)
    loads the tuning input.  Returns (slots, grid, table of ms per candidate);
    slots == "resident" (grid None: one workgroup per compute unit) when the
    resident engine won, slots == "xcd" when the XCD-partitioned mode did,
    slots == "team" / "team-l2" (grid = team size) when the resident team
    mode did."""
    dev = torch.cuda.get_device_name()
    key = (f"{dev}|{signature}|{msg_dtype}|{'r' if resident else 'b'}{'x' if xcd else ''}"
           f"{'t' if team else ''}")
    cache = _load() if use_cache else {}
    if key in cache:
        c = cache[key]
        return c["slots"], c["grid"], c.get("table", {})
    table = {}
    best = None
    cands = candidates(nblk, msg_dtype, xcd, team) + ([("resident", None)] if resident else [])
    for slots, grid in cands:
        obj = build(slots, grid)
        if prepare is not None:
            prepare(obj)
        g = obj.capture()
        ms = time_graph(g, reps)
        obj.plan.check()
        table["resident" if slots == "resident" else f"{slots}x{grid}"] = round(ms, 4)
        if log:
            log(f"[tune] {signature} {msg_dtype}: slots {slots} grid {grid}: {ms:.3f} ms")
        if best is None or ms < best[0]:
            best = (ms, slots, grid)
        del g, obj
        torch.cuda.synchronize()
    torch.cuda.empty_cache()
    _, slots, grid = best
    if use_cache:
        cache = _load()
        cache[key] = {"slots": slots, "grid": grid, "table": table,
                      "when": time.strftime("%Y-%m-%d %H:%M:%S")}
        _store(cache)
    return slots, grid, table
