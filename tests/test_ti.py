"""Time de-interleavers fused into the demap read.

CPU: the oracle against A/327 Figure 6.5's printed example and the
structural gates of the upstream (bijection, round trips, CTI delay
structure, the 9.3.9.1 identity); the product's closed-form HTI index table
(an analytic count of skipped virtual positions) against the oracle's
permutation inverse; the oracle CTI stream returns the transmitted blocks.

GPU: demapping through the gather equals demapping the oracle's gathered
cells with the plain entry point, bit for bit (the same arithmetic on the
same cells), for the HTI table (with and without virtual FEC Blocks, with
CSI and with the sigma2 estimate) and for the CTI ring (frames pushed one at
a time, a missing frame zero-filled, the ring wrapping)."""
import numpy as np
import pytest

from oracle import ti as OT
from atsc3_amd import ti

# ---------------------------------------------------------------------------
# oracle pins
# ---------------------------------------------------------------------------


def test_a327_figure_6_5_gold_vector():
    g = OT.GOLD
    nr, nc, nv = g["nrows"], g["ncols"], g["n_virtual"]
    mem = list("abcd") + list("efgh")          # the two data FEC Blocks
    got = [mem[t] for t in OT.tbi_read_order(nr, nc, nv)]
    assert got == g["expect"]


def test_tbi_read_order_is_a_bijection():
    for nr in (2700, 4050, 8100):
        for nc in (1, 2, 3, 5, 16, 37, 64):
            for nv in sorted({0, 1, nc // 3}):
                if nv >= nc:
                    continue
                o = OT.tbi_read_order(nr, nc, nv)
                assert len(o) == nr * (nc - nv)
                assert np.array_equal(np.sort(o), np.arange(nr * (nc - nv)))


@pytest.mark.parametrize("nr,nc,nv", [(2700, 37, 0), (2700, 37, 5), (8100, 1, 0),
                                      (8100, 39, 3), (2025, 16, 7)])
def test_tbi_round_trip_and_block_positions(nr, nc, nv):
    rng = np.random.default_rng(nr + nc + nv)
    x = rng.permutation(nr * (nc - nv))
    tx = OT.tbi_interleave(x, nr, nc, nv)
    mem = OT.tbi_deinterleave(tx, nr, nc, nv)
    assert np.array_equal(mem, x)
    for j in range(nc - nv):
        assert np.array_equal(tx[OT.hti_block_positions(nr, nc, nv, j)],
                              x[j * nr:(j + 1) * nr])


@pytest.mark.parametrize("plp,nti,nr,nc,nv", [
    (199800, 1, 2700, 74, 0),            # 16200-bit 64QAM, 74 blocks in one TI block
    (2 * 37 * 2700 + 50, 2, 2700, 37, 0),
    (3 * 39 * 2025, 3, 2025, 39, 3),
    (16 * 8100, 1, 8100, 16, 7),
    (4 * 4, 1, 4, 3, 1)])
def test_product_hti_index_equals_permutation_inverse(plp, nti, nr, nc, nv):
    got = ti.hti_index(plp, nti, nr, nc, nv)
    want = OT.hti_frame_index(plp, nti, nr, nc, nv)
    assert got.dtype == np.int32 and np.array_equal(got, want)


def test_hti_gather_table_over_several_frames():
    g = ti.HtiGather(3 * 37 * 2700, 3, 2700, 37, 0, device="cpu", frames=4, stride=400000)
    one = OT.hti_frame_index(3 * 37 * 2700, 3, 2700, 37, 0)
    want = np.concatenate([f * 400000 + one for f in range(4)])
    assert g.nblk == 4 * 111 and np.array_equal(g.idx.numpy().reshape(g.nblk, 2700), want)
    assert g.plp_size == 3 * 400000 + 3 * 37 * 2700


def test_cti_structure_and_round_trip():
    for nrows, start in ((8, 5), (512, 191), (1024, 344)):
        n = 3 * nrows * nrows
        i = np.arange(n)
        q = OT.cti_out_index(i, nrows, start)
        assert len(np.unique(q)) == n
        d = q - i
        assert (d % nrows == 0).all() and d.max() < nrows * nrows
        assert (((start + q) % nrows) == d // nrows).all()
        src = np.random.default_rng(nrows).integers(0, 1 << 30, n)
        rx, valid = OT.cti_deinterleave(OT.cti_interleave(src, nrows, start), nrows, start)
        assert valid.mean() > 0.5 and np.array_equal(rx[valid], src[valid])


def test_cti_signalled_start_inverts():
    """A/322 9.3.9.1: fec_block_start = C + Nrows ((start_row + C) mod Nrows)."""
    rng = np.random.default_rng(1)
    for nrows in (512, 724, 887, 1024, 1254, 1448):
        for _ in range(20):
            start, C = int(rng.integers(nrows)), int(rng.integers(32400))
            fbs = C + nrows * ((start + C) % nrows)
            assert OT.solve_C(fbs, start, nrows) == C


def _cti_stream(nrows, ncell, plp, nframes, start_row, C, seed):
    """Pre-CTI blocks starting at C, interleaved and cut into frames."""
    rng = np.random.default_rng(seed)
    total = nframes * plp
    pre = (rng.standard_normal(total) + 1j * rng.standard_normal(total)).astype(np.complex64)
    tx = OT.cti_interleave(pre, nrows, start_row)
    frames = [tx[f * plp:(f + 1) * plp] for f in range(nframes)]
    fbs = C + nrows * ((start_row + C) % nrows)
    return pre, frames, fbs


def test_oracle_cti_stream_returns_the_transmitted_blocks():
    nrows, ncell, plp, C, start = 32, 500, 1700, 123, 7
    pre, frames, fbs = _cti_stream(nrows, ncell, plp, 8, start, C, 3)
    st = OT.CtiStream(nrows, ncell, plp)
    st.reset(0, start, fbs)
    got = []
    for f, fr in enumerate(frames):
        st.push(f, fr)
        got += st.take_blocks()
    assert len(got) >= 20
    for bid, cells in got:
        lo = C + bid * ncell
        assert np.array_equal(cells, pre[lo:lo + ncell])


def test_cti_stream_bookkeeping_with_overrun_on_cpu():
    """The device stream's host bookkeeping (ring wrap, zero-filled gaps, the
    overrun rule) against the oracle, with a CPU ring: the consumer takes
    nothing for 4 frames (an overrun), then keeps up."""
    import torch
    nrows, ncell, plp, C, start = 16, 300, 1000, 77, 5
    pre, frames, fbs = _cti_stream(nrows, ncell, plp, 14, start, C, 11)
    ref = OT.CtiStream(nrows, ncell, plp, cap_frames=3)
    dev = ti.CtiStreamDev(nrows, ncell, plp, cap_frames=3, device="cpu")
    ref.reset(0, start, fbs)
    dev.reset(0, start, fbs)
    got_ref = []
    for f, fr in enumerate(frames):
        if f == 9:
            continue                                  # lost frame
        ref.push(f, fr)
        dev.push(f, torch.from_numpy(fr))
        assert (dev.out_ptr, dev.end) == (ref.out_ptr, len(ref.cells))
        if 2 <= f <= 5:
            continue                                  # consumer stalls
        want = ref.take_blocks()
        first, cells = dev.take_blocks()
        got_ref += want
        if want:
            assert first == want[0][0]
            assert np.array_equal(cells.numpy(), np.stack([c for _, c in want]))
        else:
            assert cells.shape[0] == 0
    assert ref.overruns >= 1 and dev.stats["buffer_overrun"] == ref.overruns
    assert dev.stats["zero_filled_frames"] == 1 and len(got_ref) >= 20


# ---------------------------------------------------------------------------
# device
# ---------------------------------------------------------------------------

def _demapper(device, n, mod, rate):
    from atsc3_amd.demap import Demapper
    return Demapper(n, mod, rate, device)


@pytest.mark.gpu
@pytest.mark.parametrize("nti,nv,with_csi,est", [(1, 0, False, False), (2, 0, True, False),
                                                 (1, 3, True, True)])
def test_hti_gather_demap_equals_plain_demap(device, nti, nv, with_csi, est):
    import torch
    dm = _demapper(device, 16200, "64QAM", "11/15")
    nr = dm.ncell
    nc = 74 // nti
    plp = nti * nc * nr
    rng = np.random.default_rng(nti * 10 + nv)
    pool_h = (rng.standard_normal(plp) + 1j * rng.standard_normal(plp)).astype(np.complex64)
    csi_h = rng.uniform(0.2, 2.0, plp).astype(np.float32) if with_csi else None
    g = ti.HtiGather(plp, nti, nr, nc, nv, device)
    idx = OT.hti_frame_index(plp, nti, nr, nc, nv)
    pool = torch.from_numpy(pool_h).to(device)
    csi = torch.from_numpy(csi_h).to(device) if with_csi else None
    s2 = None if est else 0.05
    print(f"[gpu] hti gather demap nti={nti} nv={nv} csi={with_csi} est={est}", flush=True)
    llr, sig = g.demap(dm, pool, csi, s2)
    cells = torch.from_numpy(pool_h[idx]).to(device)
    wcsi = torch.from_numpy(csi_h[idx]).to(device) if with_csi else None
    want, wsig = dm.demap_dev(cells, wcsi, s2)
    torch.cuda.synchronize()
    assert llr.shape == (g.nblk, 16200)
    assert torch.equal(llr, want) and torch.equal(sig, wsig)
    assert torch.equal(g.gather(pool), cells)


@pytest.mark.gpu
def test_cti_stream_gather_demap_equals_oracle_blocks(device):
    """QPSK 16200 blocks (8100 cells) through a CTI of 64 rows; a frame is
    lost (zero-filled in both streams) and the ring (3 frames) wraps."""
    import torch
    dm = _demapper(device, 16200, "QPSK", "5/15")
    nrows, ncell, plp, C, start = 64, dm.ncell, 20000, 4321, 17
    pre, frames, fbs = _cti_stream(nrows, ncell, plp, 12, start, C, 9)
    ref = OT.CtiStream(nrows, ncell, plp)
    ref.reset(0, start, fbs)
    st = ti.CtiStreamDev(nrows, ncell, plp, cap_frames=3, device=device)
    st.reset(0, start, fbs)
    st2 = ti.CtiStreamDev(nrows, ncell, plp, cap_frames=3, device=device)
    st2.reset(0, start, fbs)
    print("[gpu] cti stream gather demap", flush=True)
    nblocks = 0
    for f, fr in enumerate(frames):
        if f == 5:
            continue                               # lost frame
        ref.push(f, fr)
        st.push(f, torch.from_numpy(fr).to(device))
        st2.push(f, torch.from_numpy(fr).to(device))
        want = ref.take_blocks()
        first, llr, s2 = st.take_demap(dm, sigma2=0.3)
        first2, cells = st2.take_blocks()
        assert st.block_id == ref.block_id == st2.block_id
        if not want:
            assert llr is None and cells.shape[0] == 0
            continue
        assert first == want[0][0] == first2
        wc = np.stack([c for _, c in want])
        assert np.array_equal(cells.cpu().numpy(), wc)
        plain, _ = dm.demap_dev(torch.from_numpy(wc).to(device), None, 0.3)
        torch.cuda.synchronize()
        assert torch.equal(llr, plain)
        nblocks += len(want)
    assert nblocks >= 15 and st.stats["zero_filled_frames"] == 1
    assert st.end > st.cap                          # the ring wrapped
