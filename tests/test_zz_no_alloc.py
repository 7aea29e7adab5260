"""atsc3_minsum_resident_noalloc_dev (no allocation; its workspace is a
fixed device array inside the library) against atsc3_minsum_resident_dev,
both called through the C ABI: bit-identical bits, conv, its, bad on two
calls in a row (the second smaller, so it runs over the first call's
leftover scratch contents), in the FP16 and FP32 modes, and an oversized
launch refused with ATSC3_E_SIZE and nothing written."""
import numpy as np
import pytest

from atsc3_amd import synth
from atsc3_amd.codes import get_code

pytestmark = pytest.mark.gpu
E_SIZE = -4


def _llr(rate, n, nblk, seed):
    cw = synth.codewords(rate, n, nblk, seed)
    half = nblk // 2
    a = synth.llr_bpsk_like(cw[:half], 2.9 if n == 64800 else 1.6, seed + 1)
    b = synth.llr_bpsk_like(cw[half:], 2.5 if n == 64800 else 1.2, seed + 2)
    return np.concatenate([a, b]).astype(np.float32)


def _outs(plan, device):
    import torch
    return (torch.full((plan.total,), 0xAB, dtype=torch.uint8, device=device),
            torch.full((plan.nblk,), 0xAB, dtype=torch.uint8, device=device),
            torch.full((plan.nblk,), -7, dtype=torch.int32, device=device),
            torch.full((plan.nblk,), -7, dtype=torch.int32, device=device))


def _noalloc(plan, llr, outs):
    from atsc3_amd import _lib
    from atsc3_amd.ldpc import _stream_ptr
    b = plan.bank
    return _lib.lib().atsc3_minsum_resident_noalloc_dev(
        llr.data_ptr(), plan.nblk, plan.blk_off.data_ptr(), plan.blk_code.data_ptr(),
        b.desc.data_ptr(), len(b), b.max_checks, b.max_global_rows, b.max_n, b.max_dc,
        b.max_dv, plan.iters, plan.alpha, plan.fmt, None, plan.grid,
        *[o.data_ptr() for o in outs], _stream_ptr(plan.device))


def _compare(device, rate, n, nblk, seed, msg, grid=None):
    import torch
    from atsc3_amd.resident import ResidentPlan
    print(f"[gpu] no-alloc {msg} {n} {rate} nblk={nblk} grid={grid}", flush=True)
    code = get_code(rate, n)
    plan = ResidentPlan([code] * nblk, device, 50, 0.75, grid, msg)
    llr = torch.from_numpy(_llr(rate, n, nblk, seed).ravel()).to(device)
    want = _outs(plan, device)
    plan.run(llr, want)
    got = _outs(plan, device)
    rc = _noalloc(plan, llr, got)
    assert rc == 0, rc
    torch.cuda.synchronize()
    for g, w, nm in zip(got, want, ("bits", "conv", "its", "bad")):
        assert torch.equal(g, w), f"{msg} {nblk}: {nm} differs"
    its = got[2].cpu().numpy()
    assert (its >= 0).all() and got[1].cpu().numpy().any()
    return its


def test_noalloc_equals_resident_two_calls(device):
    from atsc3_amd import _lib
    assert _lib.lib().atsc3_minsum_resident_noalloc_max_bytes() == 128 << 20
    # FP16 at full grid (the headline configuration), then a smaller call
    _compare(device, "10/15", 64800, 300, 501, "float16")
    _compare(device, "10/15", 64800, 20, 502, "float16")
    # FP32 (bit-exact with the reference kernel) at a grid whose workspace
    # fits, then a smaller call of another code
    _compare(device, "10/15", 64800, 40, 503, "float32", grid=64)
    _compare(device, "6/15", 16200, 12, 504, "float32", grid=8)


def test_noalloc_refuses_oversize(device):
    """FP32 records plus global posterior rows for 256 workgroups exceed the
    scratch: refused before anything is enqueued, outputs untouched."""
    import torch
    from atsc3_amd import _lib
    from atsc3_amd.resident import ResidentPlan
    code = get_code("10/15", 64800)
    plan = ResidentPlan([code] * 300, device, 50, 0.75, 256, "float32")
    need = _lib.lib().atsc3_minsum_resident_ws_bytes(plan.grid, plan.bank.max_checks,
                                                     plan.bank.max_global_rows, plan.fmt)
    assert need > _lib.lib().atsc3_minsum_resident_noalloc_max_bytes()
    llr = torch.zeros(plan.total, dtype=torch.float32, device=device)
    outs = _outs(plan, device)
    assert _noalloc(plan, llr, outs) == E_SIZE
    torch.cuda.synchronize()
    assert (outs[0] == 0xAB).all() and (outs[2] == -7).all()
