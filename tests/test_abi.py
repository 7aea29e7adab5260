"""The C-ABI library: it loads without a GPU, exports every function that
include/atsc3_fec.h declares, the ctypes layer binds exactly those, and the
pure size queries answer (no compute call is made)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "atsc3_fec.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\bint(?:32|64)_t\s+(\w+)\s*\(", txt)))


def return_types():
    """name -> declared return type ("int32_t" or "int64_t")."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return {n: r for r, n in re.findall(r"\b(int(?:32|64)_t)\s+(\w+)\s*\(", txt)}


def prototypes():
    """name -> list of parameter type strings, from the header."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    out = {}
    for name, params in re.findall(r"\bint(?:32|64)_t\s+(\w+)\s*\(([^)]*)\)\s*;", txt):
        params = " ".join(params.split())
        if params in ("", "void"):
            out[name] = []
            continue
        types = []
        for p in params.split(","):
            p = p.strip()
            m = re.match(r"(.*?)(\w+)$", p)
            types.append(" ".join(m.group(1).replace("*", " * ").split()))
        out[name] = types
    return out


def _ctype_of(t):
    import ctypes
    if "*" in t:
        return ctypes.c_void_p
    return {"int32_t": ctypes.c_int32, "int64_t": ctypes.c_int64, "float": ctypes.c_float,
            "double": ctypes.c_double}[t.replace("const ", "")]


def _lib():
    from atsc3_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    return _lib


def test_header_declarations_parsed():
    names = declared()
    assert "ldpc_minsum_f32" in names and "ldpc_kernel_abi" in names
    assert len(names) >= 27


def test_library_exports_every_declared_symbol():
    _l = _lib()
    L = _l.lib()
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_ctypes_layer_binds_the_header():
    _l = _lib()
    assert sorted(_l.exported_symbols()) == declared()


def test_ctypes_signatures_match_prototypes():
    """Every bound signature has the header's parameter count and kinds
    (a short list would let ctypes pass a 64-bit handle as a C int)."""
    from atsc3_amd import _lib as L
    protos = prototypes()
    assert set(protos) == set(L._SIGS)
    for name, args in L._SIGS.items():
        want = [_ctype_of(t) for t in protos[name]]
        assert list(args) == want, name


def test_ctypes_return_types_match_header():
    import ctypes
    from atsc3_amd import _lib as L
    for name, r in return_types().items():
        want = ctypes.c_int64 if r == "int64_t" else ctypes.c_int32
        assert L.restype_of(name) is want, name


def test_size_queries():
    L = _lib().lib()
    assert L.ldpc_kernel_abi() == 1
    assert L.atsc3_minsum_qc_ctl_words(10) == 42
    assert L.atsc3_demap_ws_doubles(0, 0) >= 0
    assert L.atsc3_ldm_ws_doubles(2, 32400) == 3 * 2 * ((32400 + 255) // 256)
    assert L.atsc3_minsum_batch_ctl_words(100, 10) > 13 * 10 + 100
    assert L.atsc3_minsum_batch_ctl_words(-1, 10) < 0
    assert L.atsc3_soft_demap_bwd_grid(3, 256) == 3
    assert L.atsc3_minsum_batch_status_word() >= 0
    # the XCD-partitioned mode keeps one list region per possible XCD
    assert L.atsc3_minsum_batch_xcd_ctl_words(100, 16) > 16 * 10 * 16
    assert L.atsc3_minsum_batch_xcd_ctl_words(100, 0) < 0


def test_batch_xcd_entry_rejects_bad_arguments():
    """atsc3_minsum_batch_xcd_dev checks its partition arguments before any
    device call: no order or part table, no part, fewer than 16 slots."""
    L = _lib().lib()
    E_ARG, OK = -2, 0
    buf = ctypes.create_string_buffer(64)
    p = ctypes.cast(buf, ctypes.c_void_p)

    def call(nblk=4, order=p, part_off=p, nparts=8, slots=16):
        return L.atsc3_minsum_batch_xcd_dev(p, nblk, p, p, p, 1, 180, 500, 13, 50, 0.75, None, 0,
                                            0, None, order, part_off, nparts, slots, 64, p, p,
                                            p, p, p, p, p, None)
    assert call(nblk=-1) == E_ARG
    assert call(nblk=0, order=None) == OK
    assert call(order=None) == E_ARG
    assert call(part_off=None) == E_ARG
    assert call(nparts=0) == E_ARG
    assert call(slots=15) == E_ARG


def test_xcd_parts_balance_code_bits():
    """The host's split of a queue into per-XCD parts: a permutation, parts
    in input order, longest codewords dealt first to the lightest part."""
    from atsc3_amd.batch import xcd_parts
    rng = np.random.default_rng(5)
    for nblk in (0, 1, 5, 8, 74, 300):
        n_of = rng.choice([16200, 64800], nblk)
        order, off = xcd_parts(n_of)
        assert off[0] == 0 and off[-1] == nblk and len(off) == 9 and (np.diff(off) >= 0).all()
        assert sorted(order.tolist()) == list(range(nblk))
        load = [int(n_of[order[a:b]].sum()) for a, b in zip(off[:-1], off[1:])]
        for a, b in zip(off[:-1], off[1:]):
            assert (np.diff(order[a:b]) > 0).all()
        if nblk:
            # greedy longest-first: no part exceeds the lightest by more than
            # one longest codeword
            assert max(load) - min(load) <= int(n_of.max())
    order, off = xcd_parts([64800] * 3 + [16200] * 4)
    assert np.diff(off).tolist() == [1, 1, 1, 1, 1, 1, 1, 0]


def test_resident_entry_rejects_bad_arguments():
    """atsc3_minsum_resident_dev validates its arguments before any device
    call: negative sizes, a negative or NaN alpha (records hold magnitudes),
    an unknown message format, NULL buffers, sizes beyond the kernel's
    capacities and a short workspace are refused with the header's codes; an
    empty batch is a no-op that needs no buffers."""
    L = _lib().lib()
    E_DEGREE, E_ARG, E_SIZE, OK = -1, -2, -4, 0
    buf = ctypes.create_string_buffer(64)
    p = ctypes.cast(buf, ctypes.c_void_p)
    ws_need = L.atsc3_minsum_resident_ws_bytes(4, 21600, 0, 1)
    assert ws_need > 4 * 21600 * 8

    def call(nblk=1, iters=50, alpha=0.75, fmt=1, ptr=p, max_n=64800, max_dc=14, max_dv=13,
             ws_bytes=None):
        return L.atsc3_minsum_resident_dev(ptr, nblk, ptr, ptr, ptr, 1, 21600, 0, max_n,
                                           max_dc, max_dv, iters, alpha, fmt, None, 4, ptr,
                                           ws_need if ws_bytes is None else ws_bytes,
                                           ptr, ptr, ptr, ptr, None)
    assert call(nblk=-1) == E_ARG
    assert call(iters=-1) == E_ARG
    assert call(alpha=-0.5) == E_ARG
    assert call(alpha=float("nan")) == E_ARG
    assert call(fmt=3) == E_ARG
    assert call(nblk=0, ptr=None) == OK
    assert call(ptr=None) == E_ARG
    assert call(max_n=64801) == E_SIZE
    assert call(max_dc=90) == E_DEGREE
    assert call(max_dv=33) == E_DEGREE
    assert call(ws_bytes=ws_need - 1) == E_SIZE


def test_demap_groups_entry_rejects_bad_arguments():
    """atsc3_demap_groups_dev: negative counts and a missing table are refused,
    more blocks than one launch's grid dimension too; nothing to do is a
    no-op.  No device call on any of these."""
    L = _lib().lib()
    E_ARG, OK = -2, 0
    buf = ctypes.create_string_buffer(64)
    p = ctypes.cast(buf, ctypes.c_void_p)
    assert L.atsc3_demap_groups_dev(p, -1, 1, 10, None) == E_ARG
    assert L.atsc3_demap_groups_dev(p, 1, -1, 10, None) == E_ARG
    assert L.atsc3_demap_groups_dev(p, 1, 1, -1, None) == E_ARG
    assert L.atsc3_demap_groups_dev(None, 1, 1, 10, None) == E_ARG
    assert L.atsc3_demap_groups_dev(p, 1, 65536, 10, None) == E_ARG
    assert L.atsc3_demap_groups_dev(None, 0, 0, 0, None) == OK


def test_resident_team_entry_rejects_bad_arguments():
    """atsc3_minsum_resident_team_dev: the resident entry's checks, a team
    size outside 1..32, a ladder of more than 4 rungs, without alphas or with
    a negative one, and a short workspace, all before any device call; an
    empty batch is a no-op."""
    L = _lib().lib()
    E_ARG, E_SIZE, OK = -2, -4, 0
    buf = ctypes.create_string_buffer(64)
    p = ctypes.cast(buf, ctypes.c_void_p)
    ws_need = L.atsc3_minsum_resident_team_ws_bytes(16, 8, 21600, 0, 1)

    lad = (ctypes.c_float * 4)(1.0, 0.85, 0.75, -1.0)

    def call(nblk=1, team=8, alpha=0.75, fmt=1, ptr=p, ws_bytes=None, ladder=None, nrungs=0):
        return L.atsc3_minsum_resident_team_dev(ptr, nblk, ptr, ptr, ptr, 1, 21600, 0, 64800,
                                                14, 13, 50, alpha, fmt, None, None, ladder,
                                                nrungs, 16, team, ptr,
                                                ws_need if ws_bytes is None else ws_bytes,
                                                ptr, ptr, ptr, ptr, None)
    assert call(nblk=-1) == E_ARG
    assert call(alpha=-1.0) == E_ARG
    assert call(fmt=3) == E_ARG
    assert call(team=0) == E_ARG
    assert call(team=33) == E_ARG
    assert call(nblk=0, ptr=None) == OK
    assert call(ptr=None) == E_ARG
    assert call(ws_bytes=ws_need - 1) == E_SIZE
    assert call(ladder=lad, nrungs=5) == E_ARG          # more than 4 rungs
    assert call(ladder=None, nrungs=2) == E_ARG         # rungs without a ladder
    assert call(ladder=lad, nrungs=4) == E_ARG          # a negative rung alpha
    assert call(nblk=0, ptr=None, ladder=lad, nrungs=3) == OK
    assert L.atsc3_minsum_resident_team_ws_bytes(16, 0, 21600, 0, 1) == E_ARG
    assert L.atsc3_minsum_resident_team_ws_bytes(16, 33, 21600, 0, 1) == E_ARG


def test_resident_grow_entry_arguments_and_workspace():
    """Grow mode (ATSC3_TEAM_GROW OR-ed into team): no ladder (E_ARG before
    any device call); its workspace, one slot per workgroup (records and
    published rows; FP32 also the global rows beyond LDS) plus the request
    tables, is what the decode demands (one byte less: E_SIZE)."""
    L = _lib().lib()
    E_ARG, E_SIZE = -2, -4
    GROW = 0x200
    buf = ctypes.create_string_buffer(64)
    p = ctypes.cast(buf, ctypes.c_void_p)
    lad = (ctypes.c_float * 2)(1.0, 0.75)
    for fmt in (0, 1, 2):
        need = L.atsc3_minsum_resident_team_ws_bytes(16, 8 | GROW, 21600, 67, fmt)
        assert need > 0

        def call(team=8 | GROW, ws_bytes=need, ladder=None, nrungs=0):
            return L.atsc3_minsum_resident_team_dev(p, 1, p, p, p, 1, 21600, 67, 64800, 14, 13,
                                                    50, 0.75, fmt, None, None, ladder, nrungs,
                                                    16, team, p, ws_bytes, p, p, p, p, None)
        assert call(ladder=lad, nrungs=2) == E_ARG
        assert call(ws_bytes=need - 1) == E_SIZE
        assert call(team=0 | GROW) == E_ARG
        assert call(team=33 | GROW) == E_ARG
        for grid in (1, 64, 256):
            ws = L.atsc3_minsum_resident_team_ws_bytes(grid, 4 | GROW, 21600, 67, fmt)
            # per workgroup: records (2 x 8 bytes per check; FP32 4 more) and
            # the LDS rows' copies (FP16: 180 padded rows of 424 halves; FP32:
            # 113 rows of 360 floats, then 67 global rows)
            per = 2 * 8 * 21600 + (4 * 21600 + 180 * 360 * 4 if fmt == 0 else 180 * 424 * 2)
            assert ws >= grid * per
            # the grow header: 16 XCD lines and 16 x 64 requests of 8 words
            assert ws >= grid * per + 4 * (16 * 32 + 16 * 64 * 8)


def test_resident_team_workspace_holds_every_team():
    """Teams are formed per XCD (ranks r of XCD x: team r // team), so a
    launch can have up to sum_x ceil(n_x / team) teams; the workspace has a
    slot for each of them for any split of the grid over 16 XCD ids: each
    extra slot costs one team's records and posterior rows."""
    L = _lib().lib()
    rng = np.random.default_rng(3)
    for grid in (1, 7, 64, 256):
        for team in (1, 3, 8, 32):
            ws = {fmt: L.atsc3_minsum_resident_team_ws_bytes(grid, team, 21600, 67, fmt)
                  for fmt in (0, 1, 2)}
            one = {fmt: L.atsc3_minsum_resident_team_ws_bytes(grid, team, 21600, 67, fmt)
                   - L.atsc3_minsum_resident_team_ws_bytes(max(grid - team, 1), team, 21600,
                                                           67, fmt) for fmt in (0, 1)}
            for _ in range(50):
                counts = np.bincount(rng.integers(0, 16, grid), minlength=16)
                teams = int(sum(-(-c // team) for c in counts))
                assert teams <= grid // team + 16
            # FP16 slot: records (2 x 8 bytes per check) + 180 rows of 360 halves
            per_slot = 2 * 8 * 21600 + 180 * 360 * 2
            assert ws[1] >= (grid // team + 16) * per_slot
            assert ws[0] > ws[1] and ws[2] == ws[1]
            if grid > team:
                assert one[1] >= per_slot


def test_dropin_degree_limit_writes_nothing():
    """ldpc_minsum_f32 returns -1 when a check degree exceeds 128, as
    lab/ldpc_kernel.c:73-80 does, before any device work, and leaves every
    output untouched (lab/m3_ldpc.py:498-500 then falls back)."""
    import numpy as np
    from atsc3_amd.ldpc import _GenericCode
    L = _lib().lib()
    n = 200
    checks = [list(range(129)), [0, 150, 199], [5, 6]]
    code = _GenericCode(checks, n)
    cptr, cidx, vptr, vedge = code.csr()
    llr = np.ones((2, n), np.float32)
    bits = np.full((2, n), 7, np.uint8)
    conv = np.full(2, 7, np.uint8)
    its = np.full(2, -9, np.int32)
    bad = np.full(2, -9, np.int32)
    work = np.empty(n + len(cidx), np.float32)
    p = lambda a: a.ctypes.data
    rc = L.ldpc_minsum_f32(p(llr), 2, n, len(checks), len(cidx), p(cptr), p(cidx), p(vptr),
                           p(vedge), 50, 0.75, p(bits), p(conv), p(its), p(bad), p(work))
    assert rc == -1
    assert (bits == 7).all() and (conv == 7).all() and (its == -9).all() and (bad == -9).all()
    # degree exactly 128 passes the check (the empty batch needs no device)
    checks[0] = list(range(128))
    code = _GenericCode(checks, n)
    cptr, cidx, vptr, vedge = code.csr()
    assert L.ldpc_minsum_f32(p(llr), 0, n, len(checks), len(cidx), p(cptr), p(cidx), p(vptr),
                             p(vedge), 50, 0.75, p(bits), p(conv), p(its), p(bad), p(work)) == 0


def test_dropin_recognises_every_a322_code():
    """The drop-in finds the resident engine's tables of all 24 A/322 codes
    by the hash of the CSR arrays a caller passes (the C hash equals
    atsc3_amd.resident.csr_hash), and routes any other graph to the CSR
    kernels.  No device call."""
    import numpy as np
    from atsc3_amd.codes import RATES, get_code
    from atsc3_amd.ldpc import _GenericCode
    from atsc3_amd.resident import DROPIN_TABLES
    L = _lib().lib()
    if not os.path.exists(os.path.join(ROOT, "atsc3_amd", DROPIN_TABLES)):
        pytest.skip("drop-in tables not built")
    p = lambda a: a.ctypes.data
    for n in (16200, 64800):
        for rate in RATES:
            cptr, cidx, vptr, vedge = get_code(rate, n).csr()
            assert L.atsc3_dropin_engine(n, len(cptr) - 1, len(cidx), p(cptr), p(cidx), p(vptr),
                                         p(vedge)) == 1, (n, rate)
    # one edge moved: not an A/322 code any more
    cptr, cidx, vptr, vedge = get_code("10/15", 64800).csr()
    cidx2 = cidx.copy()
    cidx2[[5, 6]] = cidx2[[6, 5]]
    assert L.atsc3_dropin_engine(64800, len(cptr) - 1, len(cidx), p(cptr), p(cidx2), p(vptr),
                                 p(vedge)) == 0
    g = _GenericCode([[0, 1, 2], [2, 3], [1, 3, 4]], 5)
    a, b, c, d = g.csr()
    assert L.atsc3_dropin_engine(5, 3, len(b), p(a), p(b), p(c), p(d)) == 0
