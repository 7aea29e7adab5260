"""GPU min-sum parity: the HIP kernels against the oracle and the upstream
fixtures, bit-exact on every field (bits, conv, its, bad) in plain FP32."""
import numpy as np
import pytest

from tests import _fixtures as F

F.ensure_oracle_built()

from oracle import ldpc as OL          # noqa: E402
from atsc3_amd import synth            # noqa: E402
from atsc3_amd.codes import get_code   # noqa: E402

pytestmark = pytest.mark.gpu


def _gpu_decode(llr, code, iters, alpha, device, engine="qc", msg="float32"):
    import torch
    from atsc3_amd import ldpc
    t = torch.from_numpy(np.ascontiguousarray(llr, np.float32)).to(device)
    out = ldpc.decode_dev(t, code, iters, alpha, msg, engine)
    torch.cuda.synchronize()
    b, c, i, d = (x.cpu().numpy() for x in out)
    return b, c.astype(bool), i, d


@pytest.mark.parametrize("name", F.fixture_names())
@pytest.mark.parametrize("engine", ["qc", "phase", "csr"])
def test_fixture_parity(device, name, engine):
    code, llr, iters, alpha, want = F.load(name)
    print(f"[gpu] fixture {name} engine {engine} nblk {len(llr)}", flush=True)
    F.assert_same(_gpu_decode(llr, code, iters, alpha, device, engine), want, name)


def test_host_dropin_matches_fixture(device):
    """ldpc_minsum_f32 with host pointers, exactly as lab/m3_ldpc.py drives
    the upstream library."""
    from atsc3_amd import ldpc
    for rep in range(2):                  # the second pass hits the table cache
        for name in F.fixture_names():
            code, llr, iters, alpha, want = F.load(name)
            print(f"[gpu] drop-in {name} pass {rep}", flush=True)
            F.assert_same(ldpc.minsum_f32_host(llr, code.checks, code.n, iters, alpha), want,
                          name)


def test_host_dropin_generic_graphs(device):
    """Graphs that are not A/322 codes take the drop-in's CSR phase kernels:
    equal to the C oracle on random graphs (degrees 2..12, and one check of
    degree 128, the upstream limit), iters 0, 1 and 40."""
    from atsc3_amd import _lib, ldpc
    from atsc3_amd.ldpc import _GenericCode
    rng = np.random.default_rng(55)
    for n, nc, lo, hi in ((600, 300, 2, 12), (4000, 1800, 3, 7)):
        checks = [sorted(rng.choice(n, rng.integers(lo, hi + 1), replace=False).tolist())
                  for _ in range(nc)]
        if n == 4000:
            checks[7] = sorted(rng.choice(n, 128, replace=False).tolist())
        code = _GenericCode(checks, n)
        cptr, cidx, vptr, vedge = code.csr()
        p = lambda a: a.ctypes.data
        assert _lib.lib().atsc3_dropin_engine(n, nc, len(cidx), p(cptr), p(cidx), p(vptr),
                                              p(vedge)) == 0
        llr = (rng.standard_normal((5, n)) * 2 + 1.5).astype(np.float32)
        for iters in (0, 1, 40):
            print(f"[gpu] drop-in generic graph n={n} iters={iters}", flush=True)
            got = ldpc.minsum_f32_host(llr, checks, n, iters, 0.75)
            want = OL.minsum_f32(llr, code, iters, 0.75)
            for x, y in zip(got, want):
                assert np.array_equal(np.asarray(x), np.asarray(y))


def test_host_dropin_threaded(device):
    """The upstream's calling pattern (lab/m3_ldpc.py:485-506,
    lab/m9_fast.py:893-925): a thread pool of 8 workers, each calling
    ldpc_minsum_f32 on a sub-batch with its own buffers; the library decodes
    calls that arrive together in one launch, but only calls of the same
    code, iteration limit and alpha: here two codes, three iteration limits
    (0 among them), two alphas and sub-batches of 1 to 8 blocks interleave.
    Every result equals the C oracle."""
    from concurrent.futures import ThreadPoolExecutor
    from atsc3_amd import ldpc
    jobs = []
    for rate, n, snr, iters, alpha in (("10/15", 64800, 2.4, 50, 0.75),
                                       ("6/15", 16200, 1.3, 30, 0.75),
                                       ("10/15", 64800, 2.6, 0, 0.75),
                                       ("10/15", 64800, 2.4, 50, 0.85),
                                       ("10/15", 64800, 2.3, 20, 0.75)):
        code = get_code(rate, n)
        cw = synth.codewords(rate, n, 32, 700 + n % 7 + iters)
        llr = synth.llr_bpsk_like(cw, snr, 701 + iters)
        want = OL.minsum_f32(llr, code, iters, alpha)
        lo = 0
        for size in (8, 1, 5, 8, 3, 7):
            jobs.append((code, llr[lo:lo + size], iters, alpha,
                         tuple(np.asarray(w)[lo:lo + size] for w in want)))
            lo += size
    order = np.random.default_rng(5).permutation(len(jobs))
    jobs = [jobs[i] for i in order]
    print(f"[gpu] drop-in threaded: {len(jobs)} calls on 8 threads", flush=True)

    def one(j):
        code, llr, iters, alpha, _ = j
        return ldpc.minsum_f32_host(llr, code.checks, code.n, iters, alpha)
    with ThreadPoolExecutor(8) as ex:
        got = list(ex.map(one, jobs * 2))
    for (code, llr, iters, alpha, want), g in zip(jobs * 2, got):
        for x, y in zip(g, want):
            assert np.array_equal(np.asarray(x), np.asarray(y))


@pytest.mark.parametrize("rate,n,snr,alpha,iters,nblk", [
    ("10/15", 64800, 2.3, 0.75, 50, 24), ("4/15", 64800, -1.8, 1.0, 30, 8),
    ("6/15", 64800, 0.8, 0.85, 50, 8), ("8/15", 16200, 1.8, 0.75, 50, 40),
    ("13/15", 16200, 5.5, 1.0, 25, 16), ("2/15", 64800, -4.0, 0.75, 20, 6),
    ("7/15", 64800, 0.9, 0.75, 15, 6), ("12/15", 64800, 4.6, 0.75, 50, 6),
    ("9/15", 16200, 2.2, 0.75, 0, 5), ("5/15", 16200, -1.0, 0.75, 1, 5),
    ("5/15", 64800, -1.0, 0.75, 20, 4), ("13/15", 64800, 5.5, 0.85, 20, 4)])
def test_live_parity_vs_oracle(device, rate, n, snr, alpha, iters, nblk):
    code = get_code(rate, n)
    cw = synth.codewords(rate, n, nblk, 900 + nblk)
    llr = synth.llr_bpsk_like(cw, snr, 901 + nblk)
    print(f"[gpu] live {n} {rate} snr {snr} alpha {alpha} iters {iters} nblk {nblk}", flush=True)
    want = OL.minsum_f32(llr, code, iters, alpha)
    for engine in ("qc", "phase"):
        got = _gpu_decode(llr, code, iters, alpha, device, engine)
        for x, y, nm in zip(got, want, ("bits", "conv", "its", "bad")):
            assert np.array_equal(x, y), (engine, nm)


def test_empty_batch(device):
    import torch
    from atsc3_amd import ldpc
    code = get_code("10/15", 64800)
    out = ldpc.decode_dev(torch.empty((0, 64800), device=device), code)
    assert all(x.shape[0] == 0 for x in out)
    for dt in (np.float32, np.float64):
        b, c, i, d = ldpc.min_sum_decode_batch(np.zeros((0, 64800), dt), code.checks, 64800,
                                               dtype=dt)
        assert b.shape == (0, 64800) and c.shape == (0,)


def test_determinism_and_large_batch_invariants(device):
    """Full-size property check: 64800 r10/15, 320 blocks over five SNRs
    (mixed convergence) -- repeated phased runs and the fused kernel
    bit-identical; conv <=> zero syndrome of the returned bits; bad equals
    the syndrome weight of non-converged blocks; a 12-block sample equals the
    oracle."""
    code = get_code("10/15", 64800)
    cw = synth.codewords("10/15", 64800, 64, 4242)
    llr = np.concatenate([synth.llr_bpsk_like(cw, snr, 4243 + k)
                          for k, snr in enumerate((2.25, 2.5, 2.8, 3.2, 4.0))])
    a = _gpu_decode(llr, code, 50, 0.75, device, "phase")
    b = _gpu_decode(llr, code, 50, 0.75, device, "phase")
    c = _gpu_decode(llr, code, 50, 0.75, device, "qc")
    for x, y, z in zip(a, b, c):
        assert np.array_equal(x, y) and np.array_equal(x, z)
    bits, conv, its, bad = a
    syn = code.syndrome_count(bits)
    assert np.array_equal(syn == 0, conv)
    assert np.array_equal(bad[~conv], syn[~conv])
    assert 0 < conv.sum() < len(conv)
    idx = np.linspace(0, len(llr) - 1, 12).astype(int)
    want = OL.minsum_f32(llr[idx], code, 50, 0.75)
    for x, y in zip((bits[idx], conv[idx], its[idx], bad[idx]), want):
        assert np.array_equal(x, y)


def test_reference_signatures(device):
    """min_sum_decode_batch / GpuMinSum.decode accept the upstream list of
    checks and return the upstream tuple types."""
    from atsc3_amd import ldpc
    code = get_code("11/15", 16200)
    cw = synth.codewords("11/15", 16200, 6, 5)
    llr = synth.llr_bpsk_like(cw, 3.9, 6)
    checks = [list(c) for c in code.checks]
    bits, conv, its, bad = ldpc.min_sum_decode_batch(llr, checks, 16200, 50, 0.75, np.float32)
    assert bits.dtype == np.uint8 and conv.dtype == bool and its.dtype == np.int32
    want = OL.minsum_f32(llr, code, 50, 0.75)
    assert np.array_equal(bits, want[0]) and np.array_equal(its, want[2])
    # GpuMinSum: the upstream's update-then-check contract (float32 here) ...
    from oracle.utc_ref import utc_minsum
    g = ldpc.GpuMinSum(checks, 16200, iters=50, alpha=0.75, device=device, dtype="float32")
    out = g.decode(llr)
    want_u = utc_minsum(llr, checks, 16200, 50, 0.75, np.float32)
    for x, y in zip(out, want_u):
        assert np.array_equal(np.asarray(x), np.asarray(y))
    assert out[1].dtype == bool and out[2].dtype == np.int32
    # ... or, asked for, the compiled kernel's
    k = ldpc.GpuMinSum(checks, 16200, iters=50, alpha=0.75, device=device, dtype="float32",
                       contract="kernel")
    out = k.decode(llr)
    assert np.array_equal(out[0], want[0]) and np.array_equal(out[3], want[3])


def test_ladder_matches_oracle(device):
    from atsc3_amd import ldpc
    code = get_code("6/15", 64800)
    cw = synth.codewords("6/15", 64800, 6, 204)
    llr = synth.llr_bpsk_like(cw, 0.8, 205)
    got = ldpc.ladder_decode(llr, code, 64800)
    want = OL.ladder(llr, code)
    for x, y in zip(got, want):
        assert np.array_equal(np.asarray(x), np.asarray(y))


@pytest.mark.parametrize("msg", ["float16", "int8"])
def test_low_precision_engines_agree(device, msg):
    """The fused and phased engines run the same arithmetic in every message
    format."""
    code = get_code("4/15", 64800)
    cw = synth.codewords("4/15", 64800, 8, 91)
    llr = synth.llr_bpsk_like(cw, -1.6, 92)
    a = _gpu_decode(llr, code, 30, 0.75, device, "qc", msg)
    b = _gpu_decode(llr, code, 30, 0.75, device, "phase", msg)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("msg", ["float16", "int8"])
def test_low_precision_messages_decode(device, msg):
    """FP16 / INT8 message storage: every converged block is a codeword, and
    at an SNR where FP32 decodes all blocks the low-precision forms decode
    (nearly) all as well."""
    code = get_code("10/15", 64800)
    cw = synth.codewords("10/15", 64800, 48, 77)
    llr = synth.llr_bpsk_like(cw, 3.4, 78)
    ref = _gpu_decode(llr, code, 50, 0.75, device)
    got = _gpu_decode(llr, code, 50, 0.75, device, msg=msg)
    assert ref[1].all()
    bits, conv, its, bad = got
    assert (code.syndrome_count(bits)[conv] == 0).all()
    assert conv.mean() >= 0.95
    assert (bits[conv] == cw[conv]).all()
