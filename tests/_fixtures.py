"""Loading of the golden min-sum fixtures (tests/golden/minsum_*.npz)."""
import glob
import hashlib
import json
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLDEN = os.path.join(HERE, "golden")


def ensure_oracle_built():
    so = os.path.join(ROOT, "oracle", "liboracle_minsum.so")
    if not os.path.exists(so):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"),
                        os.path.join(ROOT, "oracle", "liboracle_minsum.so")],
                       check=True, capture_output=True)
    return so


def fixture_names():
    return sorted(os.path.basename(p)[len("minsum_"):-4]
                  for p in glob.glob(os.path.join(GOLDEN, "minsum_*.npz")))


def load(name):
    """-> (code, llr float32 (B, n), iters, alpha, expected dict)."""
    from atsc3_amd.codes import get_code
    d = np.load(os.path.join(GOLDEN, f"minsum_{name}.npz"), allow_pickle=False)
    meta = json.loads(str(d["meta"]))
    code = get_code(meta["rate"], meta["n"])
    if "llr" in d.files:
        llr = np.asarray(d["llr"], np.float32)
    else:
        llr = regenerate(meta["spec"])
    sha = hashlib.sha256(np.ascontiguousarray(llr, np.float32).tobytes()).hexdigest()
    assert sha == str(d["llr_sha256"]), f"{name}: regenerated LLRs drifted"
    bits = np.unpackbits(d["bits_packed"], axis=1)[:, :code.n]
    want = dict(bits=bits, conv=d["conv"].astype(bool), its=d["its"], bad=d["bad"])
    return code, llr, meta["iters"], meta["alpha"], want


def ref_lam_of_q(n, mod, rate):
    """The upstream's de-interleave map lam_of_q for (n, mod, rate): stored
    whole in tests/golden/ref_bicm.npz at the demap tests' operating points,
    else the label-simulating restatement (oracle/bicm.py) after checking
    its sha256 against the upstream's (same fixture file)."""
    from oracle import bicm as OB
    ref = np.load(os.path.join(GOLDEN, "ref_bicm.npz"), allow_pickle=False)
    key = f"map_{n}_{mod}_{rate.replace('/', '_')}"
    sha = dict(zip(map(str, ref["keys"]), map(str, ref["sha256"])))[f"{n}|{mod}|{rate}"]
    m = ref[key].astype(np.int64) if key in ref.files else OB.lam_of_q(n, mod, rate)
    assert OB.map_sha256(m) == sha, f"lam_of_q {n} {mod} {rate} is not the upstream's"
    return m


def regenerate(spec):
    from atsc3_amd import bicm, synth
    if spec["gen"] == "bpsk":
        cw = synth.codewords(spec["rate"], spec["n"], spec["nblk"], spec["seed"])
        return synth.llr_bpsk_like(cw, spec["snr"], spec["seed"] + 1)
    from oracle import demap as OD
    n, rate, mod = spec["n"], spec["rate"], spec["mod"]
    _, y, n0 = synth.awgn_blocks(rate, n, mod, spec["nblk"], spec["snr"], spec["seed"])
    lam, _ = OD.demap(y, bicm.points(mod, rate), ref_lam_of_q(n, mod, rate), sigma2=n0)
    return lam.astype(np.float32)


def assert_same(got, want, tag=""):
    bits, conv, its, bad = got
    assert np.array_equal(np.asarray(conv, bool), want["conv"]), f"{tag} conv"
    assert np.array_equal(np.asarray(its), want["its"]), f"{tag} its {its} vs {want['its']}"
    assert np.array_equal(np.asarray(bad), want["bad"]), f"{tag} bad {bad} vs {want['bad']}"
    assert np.array_equal(np.asarray(bits), want["bits"]), f"{tag} bits differ"
