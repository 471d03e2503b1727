"""Long-pair path (strips pipelined across CUs, checkpointed traceback) against
the CPU oracle: scores and CIGARs identical.  Shapes cover one strip, exact
strip multiples, several checkpoint blocks, a single column and tall/wide
pairs; larger pairs of the long-DNA workload are checked against the oracle
where it fits in memory and through size-independent properties above."""
import os
import random

import numpy as np
import pytest

import tweakseq_amd as ta
from oracle import pyoracle as po

from test_gpu_parity import compare, mutate, oracle_results, rand_prot

pytestmark = pytest.mark.gpu

DNA = "ACGT"
DNA_PARAMS = dict(mode=ta.GLOBAL, alphabet=ta.DNA, gap_open=-5, gap_extend=-2, match=2, mismatch=-3)


@pytest.fixture(scope="module")
def eng():
    e = ta.Engine(0)
    yield e
    e.close()


def dna_pairs(rng, shapes):
    pairs = []
    for k, (la, lb) in enumerate(shapes):
        a = rand_prot(rng, la, DNA)
        if k % 2 and la > 4:
            b = mutate(rng, a, 0.10, 0.02, DNA)[:lb] or "A"
        else:
            b = rand_prot(rng, lb, DNA)
        pairs.append((a, b))
    return pairs


def run_check(eng, pairs, label, **kw):
    params = ta.engine.make_params(**kw)
    got = eng.align(pairs, params)
    want = oracle_results(params, pairs)
    compare(got, want, params, pairs, label)
    return got


SHAPES = [(1, 1), (1, 700), (5, 3), (511, 600), (512, 512), (513, 800), (1024, 64), (1100, 1), (64, 1100),
          (3000, 2900), (4096, 4100), (4097, 4000), (4600, 4500), (9000, 8700), (600, 9000)]


def test_long_path_dna_shapes(eng):
    pairs = dna_pairs(random.Random(5), SHAPES)
    got = run_check(eng, pairs, "long dna", lanes=ta.LANES_LONG, **DNA_PARAMS)
    assert all(g.flags & ta.F_LONG for g in got)


def test_long_path_long_gaps(eng):
    """Long gaps move the path far off the diagonal: the traceback leaves
    its regions through the left side and continues in new ones."""
    rng = random.Random(13)
    a = rand_prot(rng, 8000, DNA)
    pairs = [(a, a[:3000] + rand_prot(rng, 2500, DNA) + a[3000:]),   # long D run
             (a, a[:2000] + a[5000:]),                              # long I run
             (a[:1500], a),                                         # b much longer
             (a, a[6000:7000] + a[:200])]
    got = run_check(eng, pairs, "long gaps", lanes=ta.LANES_LONG, **DNA_PARAMS)
    assert all(g.flags & ta.F_LONG for g in got)


def test_long_path_score_only(eng):
    pairs = dna_pairs(random.Random(6), SHAPES[:10])
    run_check(eng, pairs, "long dna score", lanes=ta.LANES_LONG, traceback=0, **DNA_PARAMS)


def test_long_path_protein_and_gap_params(eng):
    rng = random.Random(7)
    pairs = []
    for la in (700, 1500, 5000):
        a = rand_prot(rng, la)
        pairs.append((a, mutate(rng, a)))
        pairs.append((a, rand_prot(rng, la // 2 + 3)))
    for go, ge in ((-10, -1), (-11, -1), (-3, -3), (-1, -4)):
        run_check(eng, pairs, "long protein %d/%d" % (go, ge), mode=ta.GLOBAL, gap_open=go, gap_extend=ge,
                  lanes=ta.LANES_LONG)


def test_long_path_ties(eng):
    """Low-entropy sequences: many co-optimal paths, the tie order decides."""
    rng = random.Random(8)
    pairs = []
    for la, lb in ((2000, 1900), (5200, 5000), (700, 4300)):
        pairs.append((rand_prot(rng, la, "AC"), rand_prot(rng, lb, "AC")))
        pairs.append(("A" * la, "A" * lb))
    for go, ge in ((-5, -2), (-2, -2), (-1, -3)):
        run_check(eng, pairs, "long ties %d/%d" % (go, ge), mode=ta.GLOBAL, alphabet=ta.DNA, gap_open=go,
                  gap_extend=ge, match=1, mismatch=-1, lanes=ta.LANES_LONG)


def test_auto_routes_large_global_pairs(eng):
    """Pairs of >= 2^25 cells take the long path by themselves."""
    rng = random.Random(9)
    a = rand_prot(rng, 6000, DNA)
    pairs = [(a, mutate(rng, a, 0.1, 0.02, DNA)), ("ACGT" * 10, "ACG" * 12)]
    got = run_check(eng, pairs, "auto", **DNA_PARAMS)
    assert got[0].flags & ta.F_LONG and not got[1].flags & ta.F_LONG


def test_long_dna_workload_prefix(eng):
    """Pairs of the long-DNA workload (lengths 10,000-16,000) against the oracle."""
    w = ta.Workload(5, size=4, max_len=16000)
    pairs = [w.pair_bytes(k) for k in range(w.npairs)]
    params = w.params
    got = eng.align(pairs, params)
    p = po.params(mode=params.mode, alphabet=params.alphabet, gap_open=params.gap_open,
                  gap_extend=params.gap_extend, match=params.match, mismatch=params.mismatch, traceback=1)
    codes, a_off, la, b_off, lb = [], [], [], [], []
    pos = 0
    for a, b in pairs:
        ea, eb = po.encode(a, params.alphabet), po.encode(b, params.alphabet)
        a_off.append(pos); la.append(len(ea)); codes.append(ea); pos += len(ea)
        b_off.append(pos); lb.append(len(eb)); codes.append(eb); pos += len(eb)
    res, cigs = po.align_batch(p, np.concatenate(codes), a_off, la, b_off, lb)
    want = [(r.score, r.a_begin, r.a_end, r.b_begin, r.b_end, c) for r, c in zip(res, cigs)]
    compare(got, want, params, pairs, "long DNA pairs, prefix")
    assert all(g.flags & ta.F_LONG for g in got)


def test_long_dna_full_lengths_properties(eng):
    """Lengths up to 60,000 (beyond the oracle's full matrix): every CIGAR
    consumes both sequences and rescores to the reported score, and the score
    equals the int32 short-path kernel's score-only result."""
    w = ta.Workload(5, size=4, max_len=60000)
    pairs = [w.pair_bytes(k) for k in range(w.npairs)]
    params = w.params
    got = eng.align(pairs, params)
    short = ta.engine.make_params(mode=params.mode, alphabet=params.alphabet, gap_open=params.gap_open,
                                  gap_extend=params.gap_extend, match=params.match, mismatch=params.mismatch,
                                  traceback=0, lanes=ta.LANES_I32)
    ref = eng.align(pairs, short)
    p = po.params(mode=params.mode, alphabet=params.alphabet, gap_open=params.gap_open,
                  gap_extend=params.gap_extend, match=params.match, mismatch=params.mismatch, traceback=1)
    for k, (g, r) in enumerate(zip(got, ref)):
        a, b = pairs[k]
        assert g.flags & ta.F_LONG and not r.flags & ta.F_LONG
        assert g.score == r.score, k
        ea, eb = po.encode(a, params.alphabet), po.encode(b, params.alphabet)
        assert po.rescore(p, ea, eb, (g.score, g.a_begin, g.a_end, g.b_begin, g.b_end, g.cigar)) == g.score, k


def test_plain_equals_batched(eng):
    pairs = dna_pairs(random.Random(10), [(3000, 3100), (700, 5000), (5000, 4900)])
    params = ta.engine.make_params(lanes=ta.LANES_LONG, **DNA_PARAMS)
    b = eng.batch(pairs, params)
    try:
        b.run()
        b.wait()
        batched = b.results()
        b.run_plain()
        b.wait()
        plain = b.results()
    finally:
        b.close()
    assert [(r.score, r.cigar) for r in batched] == [(r.score, r.cigar) for r in plain]


def test_long_path_rejects_local(eng):
    params = ta.engine.make_params(mode=ta.LOCAL, lanes=ta.LANES_LONG)
    with pytest.raises(RuntimeError):
        eng.align([("ACDE", "ACDE")], params)


def test_wait_limit_is_reported(eng):
    """A wait that gives up (forced from the host with a spin limit of one
    poll) ends the kernels normally and the run reports an error."""
    pairs = dna_pairs(random.Random(12), [(6000, 6000)])
    params = ta.engine.make_params(lanes=ta.LANES_LONG, **DNA_PARAMS)
    os.environ["TWEAKALIGN_SPIN_LIMIT"] = "1"
    try:
        b = eng.batch(pairs, params)
    finally:
        del os.environ["TWEAKALIGN_SPIN_LIMIT"]
    try:
        b.run()
        with pytest.raises(RuntimeError):
            b.wait()
    finally:
        b.close()
    # the engine stays usable
    run_check(eng, pairs, "after limit", lanes=ta.LANES_LONG, **DNA_PARAMS)
