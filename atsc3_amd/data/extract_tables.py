#!/usr/bin/env python3
"""Transcribe the A/322 numeric tables this path needs into ``atsc3_amd/data``.

The tables are ATSC A/322 standard data (Annex A LDPC address rows, Section
6.1.3 code parameters, Section 6.2 bit-interleaver tables, Annex C NUC
position vectors).  The upstream receiver ships them as Python literals and a
numeric ``.npz``; this script reads those files **as text** (``ast``) or with
a loader that executes nothing (``numpy.load(allow_pickle=False)``), never by
importing them, and writes plain JSON / NPZ data files.

Sources (paths relative to the upstream tree):
  lab/spec_ldpc.py:231      LDPC_16200   (Annex A.2 rows)
  lab/spec_ldpc.py:601      LDPC_PARAMS  (16200 K/M1/M2/Q1/Q2/Qldpc/type)
  lab/spec_ldpc.py:620      LDPC_PARAMS_64800
  lab/spec_ldpc64k.py:21    LDPC_64800   (Annex A.1 rows + parameters)
  lab/spec_bitint.py:57-240 RATES, MOD_BITS, TYPE_A, TYPE_B, DEMUX, BI_TYPE,
                            GROUPWISE, GROUPWISE_HDR
  lab/spec_bitint.py:380-391 GROUPWISE_LABELS (the printed rate label of
                            each Annex B block, for the table gate H8)
  lab/spec_bank/nuc_a322.npz  Annex C.1.2-C.1.7 position vectors ``w``

The fingerprints the upstream tree records for the LDPC payload
(lab/spec_ldpc.py:636-638, lab/spec_ldpc64k.py:18-19) are re-checked here and
again by the test suite against the written files.

Usage:  python atsc3_amd/data/extract_tables.py [/path/to/upstream]
"""
from __future__ import annotations

import ast
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _literals(path, names):
    """Evaluate the top-level literal assignments ``names`` of a source file."""
    tree = ast.parse(open(path, encoding="utf-8").read(), filename=path)
    out = {}
    for node in tree.body:
        if isinstance(node, ast.Assign) and len(node.targets) == 1:
            tgt = node.targets[0]
            if isinstance(tgt, ast.Name) and tgt.id in names:
                out[tgt.id] = ast.literal_eval(node.value)
    missing = set(names) - set(out)
    if missing:
        raise SystemExit(f"{path}: literals not found: {sorted(missing)}")
    return out


def main(root):
    lab = os.path.join(root, "lab")
    sl = _literals(os.path.join(lab, "spec_ldpc.py"),
                   ["LDPC_16200", "LDPC_PARAMS", "LDPC_PARAMS_64800",
                    "A2_INT_COUNT", "A2_INT_SUM", "A2_SHA256"])
    s64 = _literals(os.path.join(lab, "spec_ldpc64k.py"),
                    ["LDPC_64800", "A1_INT_COUNT", "A1_INT_SUM"])
    bi = _literals(os.path.join(lab, "spec_bitint.py"),
                   ["RATES", "MOD_BITS", "TYPE_A", "TYPE_B", "DEMUX",
                    "BI_TYPE", "GROUPWISE", "GROUPWISE_HDR", "GROUPWISE_LABELS"])

    # -- LDPC fingerprints, exactly as the upstream records define them ------
    flat16 = [x for r in sl["LDPC_16200"] for row in sl["LDPC_16200"][r]["rows"]
              for x in row]
    blob = json.dumps({r: sl["LDPC_16200"][r]["rows"] for r in sl["LDPC_16200"]},
                      sort_keys=True).encode()
    assert len(flat16) == sl["A2_INT_COUNT"], len(flat16)
    assert sum(flat16) == sl["A2_INT_SUM"], sum(flat16)
    assert hashlib.sha256(blob).hexdigest() == sl["A2_SHA256"]
    flat64 = [x for r in s64["LDPC_64800"] for row in s64["LDPC_64800"][r]["rows"]
              for x in row]
    assert len(flat64) == s64["A1_INT_COUNT"], len(flat64)
    assert sum(flat64) == s64["A1_INT_SUM"], sum(flat64)

    codes = {}
    for rate, p in sl["LDPC_PARAMS"].items():
        codes[f"16200:{rate}"] = dict(
            n=16200, rate=rate, K=p["Kldpc"], type=p["type"],
            M1=p["M1"], M2=p["M2"], Q1=p["Q1"], Q2=p["Q2"], Qldpc=p["Qldpc"],
            rows=sl["LDPC_16200"][rate]["rows"])
    for rate, p in s64["LDPC_64800"].items():
        q = sl["LDPC_PARAMS_64800"][rate]
        for k in ("Kldpc", "type", "M1", "M2", "Q1", "Q2", "Qldpc"):
            assert p[k] == q[k], (rate, k)
        codes[f"64800:{rate}"] = dict(
            n=64800, rate=rate, K=p["Kldpc"], type=p["type"],
            M1=p["M1"], M2=p["M2"], Q1=p["Q1"], Q2=p["Q2"], Qldpc=p["Qldpc"],
            rows=p["rows"])
    ldpc = dict(
        source="ATSC A/322 Annex A.1/A.2 and Section 6.1.3 parameters",
        fingerprints=dict(
            a2_count=sl["A2_INT_COUNT"], a2_sum=sl["A2_INT_SUM"],
            a2_sha256=sl["A2_SHA256"], a1_count=s64["A1_INT_COUNT"],
            a1_sum=s64["A1_INT_SUM"]),
        codes=codes)
    with open(os.path.join(HERE, "ldpc_tables.json"), "w") as f:
        json.dump(ldpc, f, separators=(",", ":"))

    bitint = dict(
        source="ATSC A/322 Section 6.2 Tables 6.8-6.11, 6.14 and Annex B",
        rates=bi["RATES"], mod_bits=bi["MOD_BITS"],
        type_a=[[n, m, *v] for (n, m), v in bi["TYPE_A"].items()],
        type_b=[[n, m, *v] for (n, m), v in bi["TYPE_B"].items()],
        demux=[[n, m, *v] for (n, m), v in bi["DEMUX"].items()],
        bi_type=[[n, nb, r, t] for (n, nb, r), t in bi["BI_TYPE"].items()],
        groupwise=[[n, m, r, list(p)] for (n, m, r), p in bi["GROUPWISE"].items()],
        groupwise_hdr=[[n, m, list(h)] for (n, m), h in bi["GROUPWISE_HDR"].items()],
        groupwise_labels=[[n, m, [list(x) for x in labs]]
                          for (n, m), labs in bi["GROUPWISE_LABELS"].items()])
    with open(os.path.join(HERE, "bitint_tables.json"), "w") as f:
        json.dump(bitint, f, separators=(",", ":"))

    nuc = np.load(os.path.join(lab, "spec_bank", "nuc_a322.npz"),
                  allow_pickle=False)
    cols = {"C_1_2": (16, ["2/15", "3/15", "4/15", "5/15", "6/15", "7/15"]),
            "C_1_3": (16, ["8/15", "9/15", "10/15", "11/15", "12/15", "13/15"]),
            "C_1_4": (64, ["2/15", "3/15", "4/15", "5/15", "6/15", "7/15"]),
            "C_1_5": (64, ["8/15", "9/15", "10/15", "11/15", "12/15", "13/15"]),
            "C_1_6": (256, ["2/15", "3/15", "4/15", "5/15", "6/15", "7/15"]),
            "C_1_7": (256, ["8/15", "9/15", "10/15", "11/15", "12/15", "13/15"])}
    out = {}
    for tab, (m, rates) in cols.items():
        for r in rates:
            key = f"{tab}__{r.replace('/', '_')}"
            w = np.asarray(nuc[key], np.complex128)
            assert w.shape == (m // 4,), (key, w.shape)
            out[f"w{m}_{r.replace('/', '_')}"] = w
    np.savez(os.path.join(HERE, "nuc_positions.npz"), **out)
    print(f"wrote {len(codes)} LDPC codes, {len(bitint['groupwise'])} "
          f"group-wise permutations, {len(out)} NUC position vectors")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
