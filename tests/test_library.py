"""CPU-side checks of the product library: it loads, exports every symbol the
headers declare, its tables match the reference fixture, and its encoder and
workload generator are deterministic.  No GPU compute here."""
import ctypes
import hashlib
import os
import re

import numpy as np
import pytest

import tweakseq_amd as ta
from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = []
    for h in ("tweakalign.h", "tweakalign_workload.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        names += re.findall(r"^\s*(?:const\s+)?\w[\w\s\*]*?\b(ta_\w+)\s*\(", text, re.M)
    return sorted(set(names))


def test_exports_every_declared_symbol():
    L = ta.lib()
    names = declared_symbols()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n


def test_protein_matrix_equals_reference_table():
    d = po.blosum62_fixture()
    assert (ta.matrix(ta.PROTEIN) == np.array(d["matrix"], dtype=np.int8)).all()


def test_dna_matrix():
    m = ta.matrix(ta.DNA, 2, -3)
    assert m.shape == (4, 4) and (np.diag(m) == 2).all() and m[0, 1] == -3


def test_encoder_matches_oracle():
    s = "ARNDCQEGHILKMFPSTWYVBZXJOUarndcqeghilkmfpstwyvbzxjou"
    assert (ta.encode(s) == po.encode(s)).all()
    assert (ta.encode("ACGTacgt", ta.DNA) == po.encode("ACGTacgt", 1)).all()
    for bad in ("A-C", "A.C", "A1", "A C"):
        with pytest.raises(RuntimeError):
            ta.encode(bad)
    with pytest.raises(RuntimeError):
        ta.encode("ACGN", ta.DNA)


def workload_digest(w):
    h = hashlib.sha256()
    for k in range(w.npairs):
        a, b = w.pair_bytes(k)
        h.update(a + b"\n" + b + b"\n")
    return h.hexdigest()


def test_workloads_are_deterministic_and_prefix_consistent():
    a = ta.Workload(3, size=300)
    b = ta.Workload(3, size=300)
    c = ta.Workload(3, size=100)
    assert workload_digest(a) == workload_digest(b)
    assert [a.pair_bytes(k) for k in range(100)] == [c.pair_bytes(k) for k in range(100)]
    assert ta.Workload(3, seed=4, size=100).pair_bytes(0) != c.pair_bytes(0)


def test_workload_shapes():
    w1 = ta.Workload(1)
    assert w1.npairs == 1 and [len(x) for x in w1.pair_bytes(0)] == [300, 320]
    w2 = ta.Workload(2, size=40)
    assert w2.npairs == 40 * 39 // 2 and w2.params.traceback == 0
    lens = [len(w2.pair_bytes(k)[1]) for k in range(w2.npairs)]
    assert min(lens) >= 50 and max(lens) <= 700
    w3 = ta.Workload(3, size=2000)
    la = np.array([len(w3.pair_bytes(k)[0]) for k in range(w3.npairs)])
    assert la.min() >= 30 and la.max() <= 2000 and 300 < np.median(la) < 400
    assert w3.params.mode == ta.LOCAL and w3.params.gap_open == -11
    w5 = ta.Workload(5, size=3, max_len=20000)
    assert w5.params.alphabet == ta.DNA and all(len(w5.pair_bytes(k)[0]) >= 2000 for k in range(3))
    w4 = ta.Workload(4, size=200, max_len=300)
    req = w4.requests()
    assert len(req) == 200 and all(1 <= r[2] <= 256 for r in req)
    assert all(req[i][0] <= req[i + 1][0] for i in range(199))


# SHA-256 of the full-size inputs: CPU and GPU runs use identical workloads.
FULL_DIGESTS_FILE = os.path.join(ROOT, "tests", "golden", "workload_sha256.txt")


def test_full_workload_digests():
    """Every configuration at its full size: 1 pair, 130,816 all-vs-all pairs,
    65,536 local pairs, 120,000 served requests (1.92 M pairs, ~9.7 GB of
    residues: ~25 s here) and 256 long DNA pairs."""
    want = dict(l.split() for l in open(FULL_DIGESTS_FILE) if l.strip())
    for cfg in (1, 2, 3, 4, 5):
        key = "config%d" % cfg
        w = ta.Workload(cfg)
        assert workload_digest(w) == want[key], key
        w.close()


def test_no_gpu_fails_loudly():
    """Without a visible GPU the engine refuses to start (no CPU fallback)."""
    import subprocess, sys
    code = ("import tweakseq_amd as ta\n"
            "try:\n    ta.Engine(0)\nexcept RuntimeError as e:\n    print('refused', e)\n")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=ROOT,
                         env={**os.environ, "HIP_VISIBLE_DEVICES": "-1", "ROCR_VISIBLE_DEVICES": "-1"})
    assert "refused" in out.stdout


def test_tune_bucket_by_median_length():
    from tweakseq_amd.engine import _Pairs
    L = ta.lib()
    for lens, want in (([10, 20, 300], 0), ([300, 300, 10], 1), ([2000, 5, 1500], 2), ([5000, 9000, 1], 3)):
        ps = _Pairs([("A" * n, "C" * max(1, n // 2)) for n in lens])
        assert L.ta_tune_bucket(ps.arr, ps.n) == want, lens
    assert L.ta_tune_bucket(None, 0) == 0
    # element n // 2 of the sorted max(la, lb), odd and even counts, lengths
    # at and around the bucket limits
    import random
    rng = random.Random(11)
    edges = [1, 255, 256, 257, 1023, 1024, 1025, 4095, 4096, 4097, 6000]
    for _ in range(200):
        n = rng.randint(1, 12)
        prs = [(rng.choice(edges), rng.choice(edges)) for _ in range(n)]
        med = sorted(max(a, b) for a, b in prs)[n // 2]
        want = 0 if med < 256 else (1 if med < 1024 else (2 if med < 4096 else 3))
        ps = _Pairs([("A" * a, "C" * b) for a, b in prs])
        assert L.ta_tune_bucket(ps.arr, ps.n) == want, prs
