"""Host-side code construction: tables, H, encoder, QC layout (CPU)."""
import hashlib
import json

import numpy as np
import pytest

from atsc3_amd import codes
from atsc3_amd.codes import get_code, Z

# (n, rate): check-degree and variable-degree histograms.  16200 r11/15 and
# r2/15 are the upstream's own counts (lab/m9_gpu.py:7-10: 4320 checks of
# degree 14..16 and 66,959 edges; 14,040 checks of degree 3..4).
DEGREES = {
    (64800, "10/15"): ({12: 2521, 13: 16559, 14: 2520},
                       {1: 1, 2: 21599, 3: 32400, 9: 2160, 14: 8640}, 280799),
    (64800, "4/15"): ({4: 4680, 5: 41040, 30: 361, 31: 1439},
                      {1: 45720, 11: 1, 12: 14399, 13: 4680}, 279359),
    (64800, "6/15"): ({6: 1, 7: 29519, 8: 9360},
                      {1: 1, 2: 38879, 3: 7560, 4: 13680, 27: 4680}, 281519),
    (64800, "8/15"): ({9: 17281, 10: 12959},
                      {1: 1, 2: 30239, 3: 21600, 4: 5760, 19: 7200}, 285119),
    (64800, "12/15"): ({21: 721, 22: 12239},
                       {1: 1, 2: 12959, 3: 42480, 14: 9360}, 284399),
    (16200, "4/15"): ({4: 9000, 5: 1800, 13: 1, 14: 1079},
                      {1: 10800, 8: 1, 9: 4679, 10: 720}, 60119),
    (16200, "6/15"): ({7: 6121, 8: 3599},
                      {1: 1, 2: 9719, 3: 1800, 4: 3600, 30: 1080}, 71639),
    (16200, "9/15"): (None, {1: 1, 2: 6479, 3: 6840, 9: 1080, 10: 720, 16: 1080}, 67679),
}


def test_table_fingerprints():
    """Annex A payload fingerprints recorded upstream
    (lab/spec_ldpc.py:636-638, lab/spec_ldpc64k.py:18-19)."""
    fp = codes.fingerprints()
    r16 = {r: codes.table_rows(r, 16200) for r in codes.RATES}
    flat16 = [x for r in r16 for row in r16[r] for x in row]
    assert len(flat16) == fp["a2_count"] == 1819
    assert sum(flat16) == fp["a2_sum"] == 6754166
    blob = json.dumps(r16, sort_keys=True).encode()
    assert hashlib.sha256(blob).hexdigest() == fp["a2_sha256"]
    flat64 = [x for r in codes.RATES for row in codes.table_rows(r, 64800) for x in row]
    assert len(flat64) == fp["a1_count"] == 7865
    assert sum(flat64) == fp["a1_sum"] == 113641173


@pytest.mark.parametrize("n", [16200, 64800])
@pytest.mark.parametrize("rate", codes.RATES)
def test_structure_and_encoder(n, rate):
    """lab/m3_ldpc.py:637-674: check count N-K, every column covered,
    H . encode(s) = 0, systematic."""
    c = get_code(rate, n)
    assert c.nc == n - c.K
    vd = c.var_degrees()
    assert (vd > 0).all()
    rng = np.random.default_rng(hash((n, rate)) & 0xffff)
    s = rng.integers(0, 2, (3, c.K), dtype=np.uint8)
    cw = c.encode(s)
    assert (cw[:, :c.K] == s).all()
    assert (c.syndrome_count(cw) == 0).all()
    # a single flipped bit must be detected (negative control)
    cw[0, rng.integers(0, n)] ^= 1
    assert c.syndrome_count(cw)[0] > 0


@pytest.mark.parametrize("key", list(DEGREES))
def test_degree_profiles(key):
    n, rate = key
    cdeg, vdeg, E = DEGREES[key]
    c = get_code(rate, n)
    assert c.nedges == E
    if cdeg is not None:
        u, k = np.unique(c.check_degrees(), return_counts=True)
        assert dict(zip(u.tolist(), k.tolist())) == cdeg
    u, k = np.unique(c.var_degrees(), return_counts=True)
    assert dict(zip(u.tolist(), k.tolist())) == vdeg


def test_csr_matches_checks():
    c = get_code("11/15", 16200)
    cptr, cidx, vptr, vedge = c.csr()
    assert len(c.checks) == 4320 and c.nedges == 66959
    deg = np.diff(cptr)
    assert deg.min() == 14 and deg.max() == 16
    for i in (0, 1, 2000, 4319):
        assert cidx[cptr[i]:cptr[i + 1]].tolist() == c.checks[i]
    # vedge is the stable argsort of cidx: per variable ascending edge id
    for v in (0, 5, 12000, 16199):
        e = vedge[vptr[v]:vptr[v + 1]]
        assert (cidx[e] == v).all() and (np.diff(e) > 0).all()


def _edge_check(cptr):
    return np.repeat(np.arange(len(cptr) - 1), np.diff(cptr))


@pytest.mark.parametrize("n,rate", [(64800, "10/15"), (64800, "4/15"), (64800, "6/15"),
                                    (16200, "6/15"), (16200, "2/15"), (16200, "11/15"),
                                    (64800, "7/15"), (16200, "9/15")])
def test_qc_layout_reproduces_csr_orders(n, rate):
    """Simulate the QC kernel's lane traversal on the host and require:
    * check side: check (g, j) meets exactly the columns of H's check
      c(g, j), each through a message slot b*360+j;
    * variable side: lane p of group t adds exactly the slots the check side
      wrote for variable t*360+p, in ascending check index (the vedge order
      of lab/m3_ldpc.py:459-482)."""
    c = get_code(rate, n)
    q = c.qc()
    cptr, cidx, vptr, vedge = c.csr()
    echk = _edge_check(cptr)
    slot_of = {}                                   # (check, var) -> slot
    for g in range(q["ncg"]):
        ws = q["be_word"][q["cg_ptr"][g]:q["cg_ptr"][g + 1]]
        s = ws & 511
        t = (ws >> 9) & 511
        miss = (ws >> 18) & 1023
        b = np.arange(q["cg_ptr"][g], q["cg_ptr"][g + 1])
        for j in range(Z):
            ok = miss != j + 1
            vs = t[ok] * Z + (j + s[ok]) % Z
            chk = int(c.check_index(g, j))
            assert sorted(vs.tolist()) == c.checks[chk]
            for v, bb in zip(vs.tolist(), b[ok].tolist()):
                slot_of[(chk, v)] = bb * Z + j
    assert len(slot_of) == c.nedges
    for t in range(q["nvg"]):
        ws = q["ve_word"][q["vg_ptr"][t]:q["vg_ptr"][t + 1]]
        b = ws & 2047
        u = (ws >> 11) & 511
        reg = (ws >> 20) & 1
        vmiss = (ws >> 21) & 1023
        for p in range(0, Z, 7 if n == 64800 else 3):
            v = t * Z + p
            jj = p + u
            wrapped = jj >= Z
            jj = jj % Z
            valid = vmiss != p + 1
            order = []
            for r in (0, 1):
                for w in (True, False):
                    sel = valid & (reg == r) & (wrapped == w)
                    order += (b[sel] * Z + jj[sel]).tolist()
            want = [slot_of[(int(echk[e]), v)] for e in vedge[vptr[v]:vptr[v + 1]]]
            assert order == want, (t, p)


@pytest.mark.parametrize("n", [16200, 64800])
@pytest.mark.parametrize("rate", codes.RATES)
def test_kernel_tables_reproduce_csr_orders(n, rate):
    """The packed kernel tables (kernel_tables) traversed the way the HIP
    kernel traverses them -- step by step, group by group: message slot
    b*384 + p is written by exactly one (check, variable) edge, every check
    and variable group is visited exactly once per phase, and lane p of each
    variable group reads its slots in ascending check index (per region:
    wrapped suffix, then prefix)."""
    c = get_code(rate, n)
    q = c.qc()
    k = c.kernel_tables()
    R = c.ROW
    cptr, cidx, vptr, vedge = c.csr()
    echk = _edge_check(cptr)

    def steps(words, info, cases):
        for sw in words:
            D, G, start, gen = sw & 63, (sw >> 6) & 7, (sw >> 9) & 4095, (sw >> 21) & 1
            assert not gen or G == 1
            # every fast step must have a (G, D) case in the kernel's switch
            assert gen or (G, D) in cases, (G, D)
            for gi in range(G):
                yield D, gen, int(info[start + gi])

    slot_of = {}
    seen = set()
    ccases = {(1, d) for d in range(1, 17)} | {(2, d) for d in range(1, 9)}
    vcases = ({(1, d) for d in range(1, 33)} | {(2, d) for d in range(1, 9)}
              | {(4, d) for d in range(1, 5)})
    for D, gen, info in steps(k["c_steps"], k["cg_info"], ccases):
        dc, b0, gm = info & 511, (info >> 9) & 2047, (info >> 20) & 1
        assert dc == D or (gen and D == min(dc, 63))
        assert (gen or not gm) and (gen or dc <= 16)
        g = int(np.searchsorted(q["cg_ptr"], b0, side="right") - 1)
        assert g not in seen
        seen.add(g)
        for j in range(Z):
            chk = int(c.check_index(g, j))
            vs = []
            for kk in range(dc):
                b = b0 + kk
                if k["be_miss"][b] == j + 1:
                    continue
                s, toff = k["be_off"][2 * b], k["be_off"][2 * b + 1]
                v = toff // R * Z + (j + s) % Z
                vs.append(v)
                assert (chk, v) not in slot_of
                slot_of[(chk, v)] = b * R + (j + s) % Z
            assert sorted(vs) == c.checks[chk]
    assert len(seen) == k["ncg"] and len(slot_of) == c.nedges
    seen = set()
    for D, gen, info in steps(k["v_steps"], k["vg_info"], vcases):
        dv, n0, e0, t = info & 63, (info >> 6) & 63, (info >> 12) & 2047, (info >> 24) & 255
        assert dv == D and t not in seen
        seen.add(t)
        rows = k["ve_off"][2 * e0:2 * (e0 + dv):2]
        us = k["ve_off"][2 * e0 + 1:2 * (e0 + dv):2]
        vm = k["ve_miss"][e0:e0 + dv]
        assert gen or not vm.any()
        for p in range(0, Z, 7):
            wr = (p + us) >= Z
            order = []
            for lo, hi in ((0, n0), (n0, dv)):
                idx = np.arange(lo, hi)
                w = int(wr[lo:hi].sum())
                assert (wr[lo:hi][:hi - lo - w] == 0).all() and wr[lo:hi][hi - lo - w:].all()
                for kk in list(idx[hi - lo - w:]) + list(idx[:hi - lo - w]):
                    if vm[kk] != p + 1:
                        order.append(int(rows[kk]) + p)
            v = t * Z + p
            want = [slot_of[(int(echk[e]), v)] for e in vedge[vptr[v]:vptr[v + 1]]]
            assert order == want, (t, p)
    assert len(seen) == k["nvg"]
