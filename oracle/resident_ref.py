"""Numpy restatement of the resident FP16 min-sum (atsc3_minsum_resident_dev).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The decode is the plain normalized min-sum of lab/ldpc_kernel.c:65-146
(syndrome first; x = tot - m; min1/min2 of |x|; the first edge attaining
min1 takes alpha*min2, the others alpha*min1; negated iff parity(x < 0) ^
(x < 0); tot = lam + the messages summed in ascending check index), with
two storage roundings when ``storage="float16"``:

  * each message magnitude alpha*min is stored as FP16 (round to nearest
    even after saturating to +-65504) -- the FP16 message format;
  * each posterior tot = lam + s is stored as FP16 the same way (it lives
    in LDS as FP16), including the initial tot = lam.

Messages are kept per check as the compressed record (a1, a2, index of the
edge taking a2, per-edge signs).  With FP32 storage the variable sum visits,
per region, the entries whose check position wrapped past 359 first, then
the others, each in table order -- the ascending-check order of the
reference (lab/ldpc_kernel.c:137-142).  With FP16 or INT8 storage it visits
the entries in table order (every entry of the variable group's list, in
list order, regions included): these reduced-precision forms are not bound
to the reference's rounding, and the table order costs the kernel no
per-lane selects.

Pinning (tests/test_resident.py): with ``storage="float32"`` every rounding
is the identity and the function must reproduce the upstream kernel's own
outputs -- the golden fixtures tests/golden/minsum_*.npz written by
lab/ldpc_kernel.c compiled from its source, every code length, rate, alpha,
ties, +-inf/NaN LLRs -- bit for bit (bits, conv, its, bad), which checks the
record encoding, the edge order and the syndrome/stopping logic; the FP16
roundings are numpy's float16 casts (IEEE binary16, round to nearest even).
"""
from __future__ import annotations

import numpy as np

Z = 360


def _tables(code):
    q = code.qc()
    cg_ptr = q["cg_ptr"].astype(np.int64)
    be = q["be_word"].astype(np.int64)
    vg_ptr = q["vg_ptr"].astype(np.int64)
    ve = q["ve_word"].astype(np.int64)
    nbe = len(be)
    g_of_b = np.searchsorted(cg_ptr, np.arange(nbe), side="right") - 1
    return dict(ncg=q["ncg"], nvg=q["nvg"], cg_ptr=cg_ptr, s=be & 511, t=(be >> 9) & 511,
                miss=(be >> 18) & 1023, vg_ptr=vg_ptr, vb=ve & 2047, vu=(ve >> 11) & 511,
                vreg=(ve >> 20) & 1, vmiss=(ve >> 21) & 1023, g_of_b=g_of_b,
                k_of_b=np.arange(nbe) - cg_ptr[g_of_b])


def int8_step(llr_block):
    """The INT8 step the device derives for one codeword when the caller gives
    none: 6 * median|llr| / 127 in float32, with the lower median (the
    ((n-1)/2)-th smallest |llr|) kept to its exponent and top 8 mantissa bits
    and placed in the middle of that bucket.  (The device finds the bucket
    with two integer histograms; for non-negative floats the bit pattern
    orders as the value, so it is the bucket of the sorted rank-th value.)"""
    a = np.sort(np.abs(np.asarray(llr_block, np.float32)))
    u = np.array([a[(len(a) - 1) // 2]], np.float32).view(np.uint32)[0]
    med = np.array([(u & np.uint32(0xFFFF8000)) | np.uint32(1 << 14)], np.uint32).view(
        np.float32)[0]
    return np.float32(max(np.float32(np.float32(6.0) * med) / np.float32(127.0),
                          np.float32(1e-30)))


def resident_minsum(llr2d, code, iters=50, alpha=0.75, storage="float16", _plant=None,
                    steps=None, return_post=False):
    """(nblk, n) LLRs -> (bits u8 (nblk, n), conv bool, its i32, bad i32).

    storage "int8": posteriors as for "float16"; each message magnitude is an
    INT8 level of the codeword's step, min(rint(a / step), 127) * step (the
    INT8 message format's codec), with ``steps`` one per codeword or None
    (``int8_step`` of each codeword's LLRs).  return_post: also return each
    codeword's posteriors when it finished (nblk, n) float32 (the stored
    values: FP16-rounded in the FP16 / INT8 modes).

    ``_plant="list_order"`` is a negative control for the tests: with FP32
    storage each variable's messages are summed in table order (the FP16 /
    INT8 order), ignoring the wrap rotation -- a plausible mistake the golden
    fixtures must catch."""
    T = _tables(code)
    llr = np.asarray(llr2d, np.float32)
    B = llr.shape[0]
    ncg, nvg = T["ncg"], T["nvg"]
    llr = llr.reshape(B, nvg, Z)
    f32 = np.float32
    step = inv = None
    if storage == "int8":
        step = (np.array([int8_step(row.reshape(-1)) for row in llr], f32) if steps is None
                else np.asarray(steps, f32).reshape(-1))
        inv = (f32(1.0) / step).astype(f32)
        storage_tot = "float16"
    else:
        storage_tot = storage
    table_order = _plant == "list_order" or storage != "float32"
    if storage_tot == "float16":
        def rnd(x):
            return np.clip(x, -65504.0, 65504.0).astype(np.float16).astype(f32)
    elif storage_tot == "float32":
        def rnd(x):
            return np.array(x, f32)             # a copy: tot must not alias lam
    else:
        raise ValueError(storage)
    alpha = f32(alpha)
    if storage == "int8":
        def rnd_msg(a):
            # a: (B, Z) magnitudes; level per codeword's step
            q = np.minimum(np.rint((a * inv[:, None]).astype(f32)), f32(127.0))
            return (q * step[:, None]).astype(f32)
    else:
        rnd_msg = rnd
    cg = T["cg_ptr"]
    dcmax = int(np.diff(cg).max())
    tot = rnd(llr)
    a1 = np.zeros((B, ncg, Z), f32)
    a2 = np.zeros((B, ncg, Z), f32)
    idx = np.zeros((B, ncg, Z), np.int64)
    sg = np.zeros((B, ncg, dcmax, Z), bool)
    bits = np.zeros((B, nvg * Z), np.uint8)
    conv = np.zeros(B, bool)
    its = np.zeros(B, np.int32)
    bad = np.zeros(B, np.int32)
    post = np.zeros((B, nvg * Z), f32)
    done = np.zeros(B, bool)
    j = np.arange(Z)
    with np.errstate(invalid="ignore", over="ignore"):
        for it in range(iters + 1):
            unsat = np.zeros(B, np.int64)
            n1, n2 = np.zeros_like(a1), np.zeros_like(a2)
            nidx, nsg = np.zeros_like(idx), np.zeros_like(sg)
            for g in range(ncg):
                b0, b1 = cg[g], cg[g + 1]
                dc = b1 - b0
                ks = np.arange(dc)
                pos = (j[None, :] + T["s"][b0:b1, None]) % Z                    # (dc, Z)
                tv = tot[:, T["t"][b0:b1, None], pos]                            # (B, dc, Z)
                mb = T["miss"][b0:b1, None]
                valid = (mb == 0) | (j[None, :] != mb - 1)                       # (dc, Z)
                mo = np.where(ks[None, :, None] == idx[:, g][:, None, :],
                              a2[:, g][:, None, :], a1[:, g][:, None, :])
                mo = np.where(sg[:, g, :dc], -mo, mo)
                x = (tv - mo).astype(f32)
                syn = np.logical_xor.reduce((tv < 0) & valid[None], axis=1)
                neg = (x < 0) & valid[None]
                par = np.logical_xor.reduce(neg, axis=1)
                mag = np.where(valid[None], np.fmin(np.abs(x), f32(np.inf)), f32(np.inf))
                m1 = mag.min(axis=1)
                m2 = np.partition(mag, 1, axis=1)[:, 1] if dc > 1 else np.full_like(m1, np.inf)
                unsat += syn.sum(axis=1)
                n1[:, g] = rnd_msg(m1 * alpha)
                n2[:, g] = rnd_msg(m2 * alpha)
                nidx[:, g] = mag.argmin(axis=1)
                nsg[:, g, :dc] = par[:, None, :] ^ neg
            fin = ~done & ((unsat == 0) | (it == iters))
            if fin.any():
                bits[fin] = (tot[fin] < 0).reshape(int(fin.sum()), -1)
                post[fin] = tot[fin].reshape(int(fin.sum()), -1)
                conv[fin] = unsat[fin] == 0
                its[fin] = np.where(unsat[fin] == 0, it, iters)
                bad[fin] = unsat[fin]
                done |= fin
            if done.all():
                break
            a1, a2, idx, sg = n1, n2, nidx, nsg
            for tg in range(nvg):
                ents = np.arange(T["vg_ptr"][tg], T["vg_ptr"][tg + 1])
                s = np.zeros((B, Z), f32)
                for r in ((None,) if table_order else (0, 1)):
                    er = ents if r is None else ents[T["vreg"][ents] == r]
                    for wrapped in ((None,) if table_order else (True, False)):
                        for e in er:
                            b = T["vb"][e]
                            g, k, u, vm = T["g_of_b"][b], T["k_of_b"][b], T["vu"][e], T["vmiss"][e]
                            jj = (j + u) % Z
                            use = (vm == 0) | (j != vm - 1)
                            if wrapped is not None:
                                use = use & ((j + u >= Z) == wrapped)
                            m = np.where(idx[:, g, jj] == k, a2[:, g, jj], a1[:, g, jj])
                            m = np.where(sg[:, g, k, jj], -m, m)
                            s = np.where(use[None, :], (s + m).astype(f32), s)
                tot[:, tg] = rnd(llr[:, tg] + s)
    if return_post:
        return bits, conv, its, bad, post
    return bits, conv, its, bad
