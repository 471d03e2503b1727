"""GPU vs CPU oracle: scores, coordinates and CIGARs must be identical."""
import random

import numpy as np
import pytest

import tweakseq_amd as ta
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

AA = "ARNDCQEGHILKMFPSTWYV"


@pytest.fixture(scope="module")
def eng():
    e = ta.Engine(0)
    yield e
    e.close()


def rand_prot(rng, n, alphabet=AA):
    return "".join(rng.choice(alphabet) for _ in range(n))


def mutate(rng, s, psub=0.2, pindel=0.05, alphabet=AA):
    out, i = [], 0
    while i < len(s):
        if rng.random() < pindel:
            k = 1 + int(rng.expovariate(1 / 2.0))
            if rng.random() < 0.5:
                out += [rng.choice(alphabet) for _ in range(k)]
            else:
                i += k
                continue
        c = s[i]
        if rng.random() < psub:
            c = rng.choice(alphabet)
        out.append(c)
        i += 1
    return "".join(out)


def oracle_results(params, pairs):
    p = po.params(mode=params.mode, alphabet=params.alphabet, gap_open=params.gap_open,
                  gap_extend=params.gap_extend, match=params.match, mismatch=params.mismatch,
                  traceback=1)
    out = []
    for a, b in pairs:
        ea, eb = po.encode(a, params.alphabet), po.encode(b, params.alphabet)
        out.append(po.align(p, ea, eb))
    return out


def compare(got, want, params, pairs, label):
    assert len(got) == len(want)
    walk_failed = [k for k, g in enumerate(got) if g.flags & ta.F_WALK_FAILED]
    assert not walk_failed, "%s: walks gave up on pairs %r" % (label, walk_failed[:5])
    bad = []
    for k, (g, w) in enumerate(zip(got, want)):
        gt = (g.score, g.a_begin, g.a_end, g.b_begin, g.b_end)
        if gt != w[:5] or (params.traceback and g.cigar != w[5]):
            bad.append((k, len(pairs[k][0]), len(pairs[k][1]), gt, w[:5], g.cigar[:60], w[5][:60], g.flags))
    assert not bad, "%s: %d/%d mismatches, first: %r" % (label, len(bad), len(got), bad[:3])


def check(eng, pairs, label, **kw):
    params = ta.engine.make_params(**kw)
    got = eng.align(pairs, params)
    want = oracle_results(params, pairs)
    compare(got, want, params, pairs, label)
    return got


EDGE_LENGTHS = [1, 2, 3, 7, 8, 9, 15, 16, 17, 31, 127, 128, 129, 255, 256, 257, 300]


@pytest.mark.parametrize("mode", [ta.GLOBAL, ta.LOCAL])
@pytest.mark.parametrize("lanes", [ta.LANES_AUTO, ta.LANES_I32])
def test_edge_lengths(eng, mode, lanes):
    rng = random.Random(11 + mode + lanes)
    pairs = []
    for la in EDGE_LENGTHS:
        for lb in (1, 5, 16, 17, 130, la, la + 3):
            a = rand_prot(rng, la)
            b = mutate(rng, a) if rng.random() < 0.7 and lb == la else rand_prot(rng, lb)
            pairs.append((a, b))
    check(eng, pairs, "edge", mode=mode, lanes=lanes, gap_open=-11 if mode else -10)


@pytest.mark.parametrize("mode", [ta.GLOBAL, ta.LOCAL])
def test_random_related(eng, mode):
    rng = random.Random(21 + mode)
    pairs = []
    for _ in range(300):
        a = rand_prot(rng, rng.randint(20, 700))
        b = mutate(rng, a, 0.25, 0.04)
        pairs.append((a, b))
    check(eng, pairs, "related", mode=mode, gap_open=-11 if mode else -10)


def test_small_alphabet_ties(eng):
    # Three letters make ties in every cell kind; tie-breaks must match exactly.
    rng = random.Random(5)
    pairs = [(rand_prot(rng, rng.randint(1, 90), "AGW"), rand_prot(rng, rng.randint(1, 90), "AGW"))
             for _ in range(200)]
    for mode in (ta.GLOBAL, ta.LOCAL):
        for go, ge in ((-10, -1), (-3, -3), (-4, 0), (-6, -2)):
            check(eng, pairs, "ties %d %d %d" % (mode, go, ge), mode=mode, gap_open=go, gap_extend=ge)


def test_dna(eng):
    rng = random.Random(8)
    pairs = []
    for _ in range(120):
        a = rand_prot(rng, rng.randint(1, 900), "ACGT")
        pairs.append((a, mutate(rng, a, 0.1, 0.02, "ACGT")))
    for mode in (ta.GLOBAL, ta.LOCAL):
        check(eng, pairs, "dna", mode=mode, alphabet=ta.DNA, gap_open=-5, gap_extend=-2, match=2,
              mismatch=-3)


def test_empty_and_letters(eng):
    pairs = [("", ""), ("", "ACD"), ("WWW", ""), ("a", "A"), ("xjou", "XXXX"), ("bz", "DE")]
    for mode in (ta.GLOBAL, ta.LOCAL):
        check(eng, pairs, "empty", mode=mode)
    with pytest.raises(RuntimeError):
        eng.align([("AC-D", "ACD")], ta.engine.make_params())


def test_score_only_global(eng):
    rng = random.Random(3)
    seqs = [rand_prot(rng, rng.randint(100, 600)) for _ in range(24)]
    pairs = [(seqs[i], seqs[j]) for i in range(24) for j in range(i + 1, 24)]
    params = ta.engine.make_params(mode=ta.GLOBAL, traceback=0)
    got = eng.align(pairs, params)
    want = oracle_results(params, pairs)
    assert [g.score for g in got] == [w[0] for w in want]
    assert all(g.cigar == "" for g in got)


def test_int16_saturation_reruns_in_int32(eng):
    # 3000 x 'W' against itself scores 33,000 > int16: the packed pass must
    # flag it and the int32 rerun must give the exact result.
    rng = random.Random(9)
    w = "W" * 3000
    pairs = [(w, w), (w[:2990] + "A" * 5, w), (rand_prot(rng, 400), rand_prot(rng, 380))]
    got = check(eng, pairs, "saturate", mode=ta.LOCAL, gap_open=-11)
    assert got[0].score == 33000 and got[0].flags & ta.F_RERUN
    assert not got[2].flags & ta.F_RERUN


def test_batched_equals_plain(eng):
    rng = random.Random(13)
    pairs = [(rand_prot(rng, rng.randint(30, 500)), rand_prot(rng, rng.randint(30, 500))) for _ in range(40)]
    for mode in (ta.GLOBAL, ta.LOCAL):
        params = ta.engine.make_params(mode=mode)
        b = eng.batch(pairs, params)
        b.run(); b.wait()
        batched = b.results()
        b.run_plain(); b.wait()
        plain = b.results()
        b.close()
        assert batched == plain


def test_plain_form_first_on_a_fresh_batch(eng):
    """The plain form run before any batched run, on a batch whose small
    device blocks come back from the library's cache holding other batches'
    data: its counters and error word start at 0 (a stale error word once
    failed the wait), and it equals the batched form."""
    rng = random.Random(15)
    pairs = [(rand_prot(rng, rng.randint(200, 700)), rand_prot(rng, rng.randint(200, 700))) for _ in range(24)]
    params = ta.engine.make_params(mode=ta.LOCAL)
    for _ in range(3):
        eng.align(pairs[:8], params)  # churn the cache's small blocks
        b = eng.batch(pairs, params)
        b.run_plain(); b.wait()
        plain = b.results()
        b.run(); b.wait()
        batched = b.results()
        b.close()
        assert plain == batched


def test_plain_reruns_saturated_pairs(eng):
    """The plain form (one launch per pair) reruns a saturating int16 pair in
    int32 as the batched form does: 3,000 x 'W' scores 33,000."""
    rng = random.Random(14)
    w = "W" * 3000
    pairs = [(w, w), (rand_prot(rng, 400), rand_prot(rng, 380)), (w[:2990] + "A" * 5, w)]
    params = ta.engine.make_params(mode=ta.LOCAL, gap_open=-11)
    b = eng.batch(pairs, params)
    b.run(); b.wait()
    batched = b.results()
    b.run_plain(); b.wait()
    plain = b.results()
    b.close()
    assert plain == batched
    assert plain[0].score == 33000 and plain[0].flags & ta.F_RERUN
    compare(plain, oracle_results(params, pairs), params, pairs, "plain saturating")


def test_workload_single_pair_300x320(eng):
    w = ta.Workload(1)
    a, b = w.pair_bytes(0)
    check(eng, [(a, b)], "300 x 320 global pair", mode=w.params.mode, gap_open=w.params.gap_open)


def test_workload_allvsall_with_derived_sequences(eng):
    """All-vs-all guide-tree scoring on 64 of the seeded sequences (2,016
    global pairs, score only; 30% of the sequences derived from a parent,
    so related high-scoring pairs are in the mix) against the oracle."""
    w = ta.Workload(2, size=64)
    pairs = [w.pair_bytes(k) for k in range(w.npairs)]
    assert w.params.traceback == 0 and len(pairs) == 64 * 63 // 2
    got = eng.align(w.pairs_struct(), w.params)
    p = po.params(mode=w.params.mode, gap_open=w.params.gap_open, gap_extend=w.params.gap_extend, traceback=0)
    want = po.align_pairs(p, pairs)
    compare(got, want, w.params, pairs, "all-vs-all 64")
    # the derived pairs score far above the unrelated ones: the mix is in the test
    assert max(g.score for g in got) > 300


def test_workload_local_batch_prefix(eng):
    w = ta.Workload(3, size=1500)
    pairs = [w.pair_bytes(k) for k in range(w.npairs)]
    check(eng, pairs, "65,536-pair local batch, first 1,500", mode=ta.LOCAL, gap_open=-11, gap_extend=-1)


@pytest.mark.parametrize("mode,traceback", [(ta.GLOBAL, 1), (ta.GLOBAL, 0), (ta.LOCAL, 1)])
def test_batch_in_eight_lane_units(eng, monkeypatch, mode, traceback):
    """A batch with more couples than resident fill waves (local, global with
    and without traceback), planned with
    8-lane units (64-row strips, 8 tasks per wave) beside 16-lane units and
    whole-wave tasks: by default (the smaller couples) and with every couple
    that is not whole-wave in 8-lane units, walked inside the fill and (the
    last) by the traceback kernel afterwards.  Both plans write fewer
    direction-plane bytes than the one without 8-lane units (fewer pad
    cells), and every pair equals the oracle."""
    rng = random.Random(71 + mode)
    pairs = []
    for _ in range(5000):
        a = rand_prot(rng, rng.randint(20, 300))
        pairs.append((a, mutate(rng, a, psub=0.25, pindel=0.04)))
    params = ta.engine.make_params(mode=mode, gap_open=-11, gap_extend=-1, traceback=traceback)
    want = oracle_results(params, pairs)
    planes = {}
    for label, limit, walk_in_fill in (("none", "0", True), ("default", None, True), ("all", "1000000", True),
                                       ("all, walks after the fill", "1000000", False)):
        if limit is None:
            monkeypatch.delenv("TWEAKALIGN_EIGHT_LANE", raising=False)
        else:
            monkeypatch.setenv("TWEAKALIGN_EIGHT_LANE", limit)
        if walk_in_fill:
            monkeypatch.delenv("TWEAKALIGN_WALK_IN_FILL", raising=False)
        else:
            monkeypatch.setenv("TWEAKALIGN_WALK_IN_FILL", "0")
        b = eng.batch(pairs, params)
        planes[label] = b.info()["tb_bytes"]
        b.run()
        b.wait()
        got = b.results()
        b.close()
        compare(got, want, params, pairs, "8-lane units: " + label)
    if params.mode == ta.LOCAL or traceback:
        assert planes["default"] < planes["none"] and planes["all"] < planes["default"], planes


@pytest.mark.parametrize("mode", [ta.GLOBAL, ta.LOCAL])
def test_shared_buffers_and_prefixes(eng, mode):
    """Pairs whose sequences point into the same buffers, at the same and at
    different lengths (prefixes): the batch encodes a (pointer, length) once
    and every prefix separately; results equal the oracle's on the strings."""
    import ctypes
    rng = random.Random(31 + mode)
    A = rand_prot(rng, 400)
    B = mutate(rng, A, 0.2, 0.04)
    ba = ctypes.create_string_buffer(A.encode(), len(A) + 1)
    bb = ctypes.create_string_buffer(B.encode(), len(B) + 1)
    lens = [(400, len(B)), (400, len(B)), (250, 300), (250, 300), (17, 400 if len(B) >= 400 else len(B)),
            (400, 60), (1, 1), (399, len(B) - 1), (250, 300)]
    ps = ta.engine._Pairs([])
    ps.arr = (ta.engine.Pair * len(lens))()
    ps.n = len(lens)
    ps.keep = [ba, bb]
    strs = []
    for k, (la, lb) in enumerate(lens):
        ps.arr[k].a = ctypes.cast(ba, ctypes.c_void_p)
        ps.arr[k].la = la
        ps.arr[k].b = ctypes.cast(bb, ctypes.c_void_p)
        ps.arr[k].lb = lb
        strs.append((A[:la], B[:lb]))
    params = ta.engine.make_params(mode=mode, gap_open=-11 if mode else -10)
    got = eng.align(ps, params)
    compare(got, oracle_results(params, strs), params, strs, "shared")


def test_lone_pairs(eng):
    """A batch of one fill task walks after its fill (the traceback kernel):
    lone pairs, global and local, one strip or several, each in a batch of
    its own; a lone identical-pointer couple too."""
    rng = random.Random(17)
    for la, lb in ((300, 320), (40, 700), (700, 40), (530, 515), (1200, 900), (1, 1)):
        a = rand_prot(rng, la)
        b = mutate(rng, a, 0.25, 0.04)[:lb] or "W"
        for mode in (ta.GLOBAL, ta.LOCAL):
            params = ta.engine.make_params(mode=mode)
            pairs = [(a, b)]
            compare(eng.align(pairs, params), oracle_results(params, pairs), params, pairs,
                    "lone %dx%d mode %d" % (la, len(b), mode))
    a = rand_prot(rng, 300)
    b = mutate(rng, a)
    params = ta.engine.make_params(mode=ta.GLOBAL)
    pairs = [(a, b), (a, b)]
    compare(eng.align(pairs, params), oracle_results(params, pairs), params, pairs, "lone couple")
