"""A/322 QC-LDPC codes: parameters, parity-check structure, encoder, and the
device-side table layouts the HIP decoder consumes.

Reference numbering.  Checks are numbered by parity-accumulator index and
variables by codeword index, exactly as the upstream receiver builds H
(lab/m3_ldpc.py:63-112 ``_addr``/``_pi``/``_params``, :166-212
``parity_check``).  The decoder's bit-exactness contract is stated in that
numbering: the first-minimum tie rule is by ascending variable index inside a
check, and each variable's message sum runs in ascending check index
(lab/ldpc_kernel.c:116-142).

Three table layouts are produced for one code:

* ``checks``      -- list of sorted column lists, one per check
                     (lab/m3_ldpc.py:166-212).
* ``csr()``       -- ``(cptr, cidx, vptr, vedge)`` int32, the C-kernel tables
                     of lab/m3_ldpc.py:459-482 (``vedge`` is the stable argsort
                     of ``cidx``: each variable's edges in ascending check).
* ``qc()``        -- the 360-lifted base graph.  Check ``c`` lives in check
                     group ``g`` at position ``j``; variable ``v`` in group
                     ``v // 360`` at position ``v % 360``.  Every (check group,
                     variable group) block of H is a 360x360 circulant
                     permutation, or one with a single missing entry (the
                     dual-diagonal corner).  This is what lets the HIP kernel
                     read 360 consecutive (rotated) words per base edge.
"""
from __future__ import annotations

import functools
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
Z = 360                                    # A/322 lifting size

with open(os.path.join(HERE, "data", "ldpc_tables.json")) as _f:
    _LDPC = json.load(_f)

RATES = ["2/15", "3/15", "4/15", "5/15", "6/15", "7/15", "8/15", "9/15",
         "10/15", "11/15", "12/15", "13/15"]


def params(rate, n):
    """(K, type, M1, M2, Q1, Q2, rows) for (rate, N) -- lab/m3_ldpc.py:85-112.

    Type B codes are expressed with M1 = Q1 = 0 and M2 = N-K, Q2 = Qldpc.
    """
    key = f"{int(n)}:{rate}"
    if key not in _LDPC["codes"]:
        raise ValueError(f"no A/322 LDPC table for N={n} rate {rate}")
    p = _LDPC["codes"][key]
    K = p["K"]
    if p["type"] == "A":
        M1, M2, Q1, Q2 = p["M1"], p["M2"], p["Q1"], p["Q2"]
    else:
        M1, M2, Q1, Q2 = 0, n - K, 0, p["Qldpc"] or (n - K) // Z
    return K, p["type"], M1, M2, Q1, Q2, p["rows"]


def fingerprints():
    return dict(_LDPC["fingerprints"])


def table_rows(rate, n):
    return _LDPC["codes"][f"{int(n)}:{rate}"]["rows"]


# ---------------------------------------------------------------------------
# accumulator addressing and parity interleave (A/322 6.1.3)
# ---------------------------------------------------------------------------

def _addr(x, mm, M1, M2, Q1, Q2):
    """Accumulator address of table entry x for bit mm of its 360-group.

    Same arithmetic as lab/m3_ldpc.py:63-66 (A/322 6.1.3.1 step ii); works on
    numpy arrays element-wise.
    """
    x = np.asarray(x)
    mm = np.asarray(mm)
    lo = (x + mm * Q1) % M1 if M1 else np.zeros_like(x + mm)
    hi = M1 + ((x - M1 + mm * Q2) % M2)
    return np.where(x < M1, lo, hi)


def _pi(a, M1, Q1, Q2, K):
    """Parity-accumulator index -> codeword index (lab/m3_ldpc.py:69-82,
    "standard" variant)."""
    a = np.asarray(a)
    r1 = K + Z * (a % max(Q1, 1)) + a // max(Q1, 1)
    b = a - M1
    r2 = K + M1 + Z * (b % Q2) + b // Q2
    return np.where(a < M1, r1, r2)


# ---------------------------------------------------------------------------
# the code object
# ---------------------------------------------------------------------------

class LdpcCode:
    """One A/322 LDPC code (N in {16200, 64800}, rate in ``RATES``)."""

    def __init__(self, rate, n):
        self.rate, self.n = rate, int(n)
        (self.K, self.type, self.M1, self.M2, self.Q1, self.Q2,
         self.rows) = params(rate, n)
        self.nc = self.M1 + self.M2
        assert self.nc == self.n - self.K

    # -- edge list in reference numbering ----------------------------------
    @functools.cached_property
    def _edges(self):
        """(check, variable) pairs of H before GF(2) pair cancellation.

        Follows lab/m3_ldpc.py:166-194: information edges through
        ``_addr``; Type A step-(vii) feedback of codeword bits K+i through the
        trailing Q1 rows; the dual-diagonal ``_pi(a)``/``_pi(a-1)`` edges
        (region 1 for Type A, everything for Type B) and the identity
        ``_pi(a)`` for Type A region 2.
        """
        K, M1, M2, Q1, Q2 = self.K, self.M1, self.M2, self.Q1, self.Q2
        cs, vs = [], []
        ninfo = K // Z
        for t in range(ninfo):
            row = np.asarray(self.rows[t], np.int64)
            mm = np.arange(Z)
            c = _addr(row[None, :], mm[:, None], M1, M2, Q1, Q2)
            cs.append(c.ravel())
            vs.append(np.repeat(Z * t + mm, len(row)))
        if self.type == "A":
            for t in range(M1 // Z):
                row = np.asarray(self.rows[ninfo + t], np.int64)
                mm = np.arange(Z)
                c = _addr(row[None, :], mm[:, None], M1, M2, Q1, Q2)
                cs.append(c.ravel())
                vs.append(np.repeat(K + Z * t + mm, len(row)))
            a = np.arange(M1)
            cs.append(a)
            vs.append(_pi(a, M1, Q1, Q2, K))
            a1 = np.arange(1, M1)
            cs.append(a1)
            vs.append(_pi(a1 - 1, M1, Q1, Q2, K))
            a2 = np.arange(M1, M1 + M2)
            cs.append(a2)
            vs.append(_pi(a2, M1, Q1, Q2, K))
        else:
            a = np.arange(M2)
            cs.append(a)
            vs.append(_pi(a, 0, 0, Q2, K))
            a1 = np.arange(1, M2)
            cs.append(a1)
            vs.append(_pi(a1 - 1, 0, 0, Q2, K))
        return np.concatenate(cs).astype(np.int64), np.concatenate(vs).astype(np.int64)

    @functools.cached_property
    def _hedges(self):
        """Edges of H after GF(2) cancellation, sorted check-major then by
        variable (lab/m3_ldpc.py:196-211: sort, drop equal pairs)."""
        c, v = self._edges
        key = c * self.n + v
        key.sort()
        uniq, cnt = np.unique(key, return_counts=True)
        keep = uniq[cnt % 2 == 1]          # a column present twice cancels
        return (keep // self.n).astype(np.int64), (keep % self.n).astype(np.int64)

    @functools.cached_property
    def checks(self):
        """list[nc] of sorted column lists -- lab/m3_ldpc.py:166-212."""
        c, v = self._hedges
        cptr = np.searchsorted(c, np.arange(self.nc + 1))
        return [v[cptr[i]:cptr[i + 1]].tolist() for i in range(self.nc)]

    @functools.lru_cache(maxsize=None)
    def csr(self):
        """(cptr, cidx, vptr, vedge) int32 -- lab/m3_ldpc.py:459-482."""
        c, v = self._hedges
        cptr = np.searchsorted(c, np.arange(self.nc + 1)).astype(np.int32)
        cidx = v.astype(np.int32)
        vedge = np.argsort(cidx, kind="stable").astype(np.int32)
        counts = np.bincount(cidx, minlength=self.n)
        vptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
        return cptr, cidx, vptr, vedge

    @property
    def nedges(self):
        return int(self._hedges[0].size)

    def check_degrees(self):
        return np.diff(self.csr()[0])

    def var_degrees(self):
        return np.diff(self.csr()[2])

    # -- QC base graph ------------------------------------------------------
    def check_group_pos(self, c):
        """check index -> (group, position)."""
        c = np.asarray(c)
        g1, j1 = c % max(self.Q1, 1), c // max(self.Q1, 1)
        d = c - self.M1
        g2, j2 = self.Q1 + d % self.Q2, d // self.Q2
        r1 = c < self.M1
        return np.where(r1, g1, g2), np.where(r1, j1, j2)

    def check_index(self, g, j):
        """(group, position) -> check index (inverse of check_group_pos)."""
        g = np.asarray(g)
        j = np.asarray(j)
        return np.where(g < self.Q1, g + self.Q1 * j,
                        self.M1 + (g - self.Q1) + self.Q2 * j)

    @functools.lru_cache(maxsize=None)
    def qc(self):
        """The 360-lifted base graph as flat int32 tables.

        Returns a dict:
          ncg, nvg          number of check / variable groups
          cg_ptr[ncg+1]     base-edge ranges per check group
          be_word[nbe]      per base edge b (check-group-major):
                            bits 0-8 shift s, 9-17 variable group t,
                            18-27 missing check position + 1 (0 = full).
                            Check (g, j) meets variable t*360 + (j+s) % 360;
                            its message lives at m[b*360 + j].
          vg_ptr[nvg+1]     entry ranges per variable group
          ve_word[nbe]      per entry, in summation order at position 0:
                            bits 0-10 base edge b, 11-19 u = (360-s) % 360,
                            20 region (0: checks < M1, 1: the rest),
                            21-30 missing variable position + 1 (0 = full).
                            Variable (t, p) meets check position (p+u) % 360.
          max_cdeg, max_vdeg, nbe
        A variable's messages are summed in ascending check index: region 0
        entries before region 1; inside a region the list is sorted by
        (u, g), and at position p the entries with u >= 360-p (whose check
        position wrapped past 359) come first -- a rotation of that list.
        """
        c, v = self._hedges
        g, j = self.check_group_pos(c)
        t, p = v // Z, v % Z
        s = (p - j) % Z
        key = (g * (self.n // Z) + t) * Z + s
        order = np.lexsort((j, key))
        key_s = key[order]
        uk, start, cnt = np.unique(key_s, return_index=True, return_counts=True)
        nbe = len(uk)
        if np.any((cnt != Z) & (cnt != Z - 1)):
            raise ValueError("H is not 360-lifted QC in the A/322 grouping")
        bg = uk // ((self.n // Z) * Z)
        bt = (uk // Z) % (self.n // Z)
        bs = uk % Z
        miss = np.zeros(nbe, np.int64)
        jj = j[order]
        for b in np.flatnonzero(cnt == Z - 1):
            present = np.zeros(Z, bool)
            present[jj[start[b]:start[b] + cnt[b]]] = True
            miss[b] = int(np.flatnonzero(~present)[0]) + 1
        ncg = self.nc // Z
        nvg = self.n // Z
        # base edges are already check-group-major (uk sorted by g first)
        cg_ptr = np.searchsorted(bg, np.arange(ncg + 1)).astype(np.int32)
        be_word = (bs | (bt << 9) | (miss << 18)).astype(np.int32)
        # variable side
        u = (Z - bs) % Z
        region = (bg >= self.Q1).astype(np.int64)   # Type B: all ones
        vmiss = np.where(miss > 0, ((miss - 1) + bs) % Z + 1, 0)
        vorder = np.lexsort((bg, u, region, bt))
        vt = bt[vorder]
        vg_ptr = np.searchsorted(vt, np.arange(nvg + 1)).astype(np.int32)
        ve_word = (vorder | (u[vorder] << 11) | (region[vorder] << 20)
                   | (vmiss[vorder] << 21)).astype(np.int32)
        assert nbe < (1 << 11)
        cdeg = np.diff(cg_ptr)
        vdeg = np.diff(vg_ptr)
        return dict(ncg=ncg, nvg=nvg, nbe=nbe, cg_ptr=cg_ptr, be_word=be_word,
                    vg_ptr=vg_ptr, ve_word=ve_word,
                    max_cdeg=int(cdeg.max()), max_vdeg=int(vdeg.max()))

    ROW = 384          # padded row length of the device message / tot arrays

    @functools.lru_cache(maxsize=None)
    def kernel_tables(self):
        """Packed int32 tables of the QC HIP kernel (derived from ``qc()``).

        Device layout: the message of base edge b at variable position p is
        msg[b*ROW + p]; the posterior of variable (t, p) is tot[t*ROW + p]
        (rows padded to ROW = 384 so every wave64 row access is 128-byte
        aligned).

          cg_info[ncg]   dc | b0 << 9 | has_missing << 20   (check groups)
          be_off[2*nbe]  (s, t*ROW) per base edge: check (g, j) meets variable
                         position p = (j + s) % 360 of group t
          be_miss[nbe]   missing check position + 1 (0 = full circulant)
          vg_info[nvg]   dv | n0 << 6 | e0 << 12 | has_missing << 23, n0 =
                         entries in region 0 (checks < M1; 0 for Type B)
          ve_off[2*nbe]  (b*ROW, u) per entry: variable (t, p) meets check
                         position (p + u) % 360; entries in summation order
                         at p = 0 (region, u, g)
          ve_miss[nbe]   missing variable position + 1 (0 = full)
        """
        q = self.qc()
        R = self.ROW
        w = q["be_word"].astype(np.int64)
        s, t, miss = w & 511, (w >> 9) & 511, (w >> 18) & 1023
        cdeg = np.diff(q["cg_ptr"]).astype(np.int64)
        b0 = q["cg_ptr"][:-1].astype(np.int64)
        gmiss = np.zeros(q["ncg"], np.int64)
        for b in np.flatnonzero(miss):
            gmiss[np.searchsorted(q["cg_ptr"], b, side="right") - 1] = 1
        cg_info = cdeg | (b0 << 9) | (gmiss << 20)
        be_off = np.stack([s, t * R], 1).ravel()
        vw = q["ve_word"].astype(np.int64)
        vb, vu, vreg, vmiss = vw & 2047, (vw >> 11) & 511, (vw >> 20) & 1, (vw >> 21) & 1023
        vdeg = np.diff(q["vg_ptr"]).astype(np.int64)
        e0 = q["vg_ptr"][:-1].astype(np.int64)
        n0 = np.add.reduceat(1 - vreg, e0) if len(vreg) else np.zeros(0, np.int64)
        n0 = np.where(vdeg > 0, n0, 0)
        vgm = np.add.reduceat((vmiss > 0).astype(np.int64), e0)
        assert vdeg.max() < 64 and n0.max() < 64 and e0.max() < 2048
        vg_info = vdeg | (n0 << 6) | (e0 << 12) | ((vgm > 0).astype(np.int64) << 23)
        ve_off = np.stack([vb * R, vu], 1).ravel()
        # processing order and steps: groups of one phase are independent, so
        # they are sorted by degree and packed G at a time into steps (more
        # independent loads in flight per wave for low-degree groups)
        tvar = np.arange(q["nvg"], dtype=np.int64)
        vg_info_t = vg_info | (tvar << 24)
        c_order, c_steps = self._pack_steps(cdeg, gmiss, self.CHECK_PACK, 16)
        v_order, v_steps = self._pack_steps(vdeg, (vgm > 0).astype(np.int64),
                                            self.VAR_PACK, 32)
        i32 = lambda a: np.ascontiguousarray(a, np.int32)
        return dict(ncg=q["ncg"], nvg=q["nvg"], nbe=q["nbe"], type_a=int(self.type == "A"),
                    cg_info=i32(cg_info[c_order]), be_off=i32(be_off), be_miss=i32(miss),
                    vg_info=i32(vg_info_t[v_order]), ve_off=i32(ve_off), ve_miss=i32(vmiss),
                    c_steps=i32(c_steps), v_steps=i32(v_steps),
                    max_cdeg=q["max_cdeg"], max_vdeg=q["max_vdeg"])

    # groups per step by degree: check steps (degree <= 8 -> 2 groups),
    # variable steps (<= 4 -> 4 groups, <= 8 -> 2 groups)
    CHECK_PACK = ((8, 2),)
    VAR_PACK = ((4, 4), (8, 2))

    @staticmethod
    def _pack_steps(deg, special, pack, fast_max):
        """Order groups by (special, degree) and cut them into steps.

        step word: D | G << 6 | start << 9 | generic << 21, where the step
        covers processing-order entries [start, start + G) all of degree D;
        generic steps (a missing entry, or D > fast_max) have G = 1 and a
        degree field saturated at 63 (the kernel reads the degree from the
        group word).
        """
        deg = np.asarray(deg, np.int64)
        generic = (np.asarray(special) > 0) | (deg > fast_max)
        order = np.lexsort((np.arange(len(deg)), -deg, generic))
        steps = []
        i = 0
        while i < len(order):
            g = order[i]
            d = int(deg[g])
            if generic[g]:
                steps.append(min(d, 63) | (1 << 6) | (i << 9) | (1 << 21))
                i += 1
                continue
            gmax = 1
            for lim, gg in pack:
                if d <= lim:
                    gmax = max(gmax, gg)
            run = 1
            while (run < gmax and i + run < len(order) and not generic[order[i + run]]
                   and deg[order[i + run]] == d):
                run += 1
            k = 1 << (run.bit_length() - 1)        # the kernel has G in {1, 2, 4}
            steps.append(d | (k << 6) | (i << 9))
            i += k
        return order, np.asarray(steps, np.int64)

    # -- encoder ------------------------------------------------------------
    def encode(self, s):
        """Systematic A/322 encoder, vectorised over a batch.

        s: uint8 (K,) or (B, K).  Follows lab/m3_ldpc.py:119-159: accumulate
        information bits into parity accumulators through ``_addr``, running
        XOR, parity interleave ``_pi``; Type A runs the two-stage form (step
        vii feeds the first M1 codeword parity bits back through the trailing
        Q1 rows).
        """
        s = np.asarray(s, np.uint8)
        one = s.ndim == 1
        s = np.atleast_2d(s)
        B = s.shape[0]
        assert s.shape[1] == self.K
        K, M1, M2, Q1, Q2 = self.K, self.M1, self.M2, self.Q1, self.Q2
        lam = np.zeros((B, self.n), np.uint8)
        lam[:, :K] = s
        A = self._acc_matrix_info()           # (K, nacc) sparse via indices
        p = self._accumulate(s, *A)
        if self.type == "A":
            p[:, :M1] = np.bitwise_xor.accumulate(p[:, :M1], axis=1)
            a = np.arange(M1)
            lam[:, _pi(a, M1, Q1, Q2, K)] = p[:, :M1]
            fb = lam[:, K:K + M1]
            p ^= self._accumulate(fb, *self._acc_matrix_feedback())
            a2 = np.arange(M1, M1 + M2)
            lam[:, _pi(a2, M1, Q1, Q2, K)] = p[:, M1:]
        else:
            p = np.bitwise_xor.accumulate(p, axis=1)
            a = np.arange(M2)
            lam[:, _pi(a, 0, 0, Q2, K)] = p
        return lam[0] if one else lam

    def _acc_matrix_info(self):
        K, M1, M2, Q1, Q2 = self.K, self.M1, self.M2, self.Q1, self.Q2
        src, dst = [], []
        for t in range(K // Z):
            row = np.asarray(self.rows[t], np.int64)
            mm = np.arange(Z)
            dst.append(_addr(row[None, :], mm[:, None], M1, M2, Q1, Q2).ravel())
            src.append(np.repeat(Z * t + mm, len(row)))
        return np.concatenate(src), np.concatenate(dst)

    def _acc_matrix_feedback(self):
        K, M1, M2, Q1, Q2 = self.K, self.M1, self.M2, self.Q1, self.Q2
        ninfo = K // Z
        src, dst = [], []
        for t in range(M1 // Z):
            row = np.asarray(self.rows[ninfo + t], np.int64)
            mm = np.arange(Z)
            dst.append(_addr(row[None, :], mm[:, None], M1, M2, Q1, Q2).ravel())
            src.append(np.repeat(Z * t + mm, len(row)))
        return np.concatenate(src), np.concatenate(dst)

    def _accumulate(self, bits, src, dst, chunk=64):
        """p[b, d] = XOR over edges (src -> d) of bits[b, src]."""
        B = bits.shape[0]
        out = np.zeros((B, self.nc), np.uint8)
        order = np.argsort(dst, kind="stable")
        ds, ss = dst[order], src[order]
        starts = np.searchsorted(ds, np.arange(self.nc))
        nz = np.bincount(ds, minlength=self.nc) > 0
        if not len(ss):
            return out
        for lo in range(0, B, chunk):
            vals = bits[lo:lo + chunk][:, ss]
            out[lo:lo + chunk][:, nz] = np.bitwise_xor.reduceat(
                vals, starts[nz], axis=1)
        return out

    def syndrome_count(self, bits):
        """Unsatisfied checks of hard bits (B, n) or (n,)."""
        b = np.atleast_2d(np.asarray(bits, np.uint8))
        cptr, cidx, _, _ = self.csr()
        par = np.bitwise_xor.reduceat(b[:, cidx], cptr[:-1], axis=1)
        return par.sum(axis=1).astype(np.int64)


@functools.lru_cache(maxsize=None)
def get_code(rate, n):
    return LdpcCode(rate, int(n))
