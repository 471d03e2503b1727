"""ta_align_pair -- the low-latency single-pair path (realign selection) --
against the CPU oracle and against ta_align_batch on the same pairs."""
import ctypes
import random

import pytest

import tweakseq_amd as ta
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

AA = "ARNDCQEGHILKMFPSTWYV"


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


@pytest.fixture(scope="module")
def eng():
    e = ta.Engine(0)
    yield e
    e.close()


def check(eng, pairs, label, **kw):
    params = ta.engine.make_params(**kw)
    # the oracle walks (and so gives local begin coordinates) only with a
    # traceback: ask it for one, compare the CIGAR only when it was asked for
    p = po.params(mode=params.mode, alphabet=params.alphabet, gap_open=params.gap_open,
                  gap_extend=params.gap_extend, match=params.match, mismatch=params.mismatch,
                  traceback=1)
    want = [w if params.traceback else w[:5] + ("",) for w in po.align_pairs(p, pairs)]
    bad = []
    for k, ((a, b), w) in enumerate(zip(pairs, want)):
        g = eng.align_pair(a, b, params)
        have = (g.score, g.a_begin, g.a_end, g.b_begin, g.b_end, g.cigar)
        if have != w or g.flags & ta.F_WALK_FAILED:
            bad.append((k, len(a), len(b), have[:5], w[:5], have[5][:40], w[5][:40]))
    assert not bad, "%s: %d/%d differ, first %r" % (label, len(bad), len(pairs), bad[:3])


EDGE = [1, 2, 3, 7, 8, 15, 16, 17, 63, 64, 65, 127, 128, 129, 130, 255, 256, 257, 300, 383, 384, 385]


@pytest.mark.parametrize("mode", [ta.GLOBAL, ta.LOCAL])
def test_edge_lengths(eng, mode):
    rng = random.Random(41 + mode)
    pairs = []
    for la in EDGE:
        for lb in (1, 2, 9, 64, 130, la):
            a = rand_prot(rng, la)
            b = mutate(rng, a)[:lb] if rng.random() < 0.6 else rand_prot(rng, lb)
            pairs.append((a, b or "W"))
    check(eng, pairs, "edge", mode=mode, gap_open=-11 if mode else -10)


@pytest.mark.parametrize("mode", [ta.GLOBAL, ta.LOCAL])
def test_related_and_random(eng, mode):
    rng = random.Random(51 + mode)
    pairs = []
    for _ in range(120):
        a = rand_prot(rng, rng.randint(10, 420))
        pairs.append((a, mutate(rng, a, 0.25, 0.04)) if rng.random() < 0.7 else (a, rand_prot(rng, rng.randint(5, 420))))
    check(eng, pairs, "related", mode=mode, gap_open=-11 if mode else -10)


def test_ties_and_gap_settings(eng):
    rng = random.Random(61)
    pairs = [(rand_prot(rng, rng.randint(1, 200), "AGW"), rand_prot(rng, rng.randint(1, 200), "AGW"))
             for _ in range(60)]
    for mode in (ta.GLOBAL, ta.LOCAL):
        for go, ge in ((-10, -1), (-3, -3), (-4, 0), (-6, -2), (0, 0)):
            check(eng, pairs, "ties %d %d %d" % (mode, go, ge), mode=mode, gap_open=go, gap_extend=ge)


def test_dna_and_score_only(eng):
    rng = random.Random(71)
    pairs = []
    for _ in range(40):
        a = rand_prot(rng, rng.randint(1, 400), "ACGT")
        pairs.append((a, mutate(rng, a, 0.1, 0.02, "ACGT")))
    for mode in (ta.GLOBAL, ta.LOCAL):
        check(eng, pairs, "dna", mode=mode, alphabet=ta.DNA, gap_open=-5, gap_extend=-2, match=2, mismatch=-3)
        check(eng, pairs[:10], "dna score only", mode=mode, alphabet=ta.DNA, gap_open=-5, gap_extend=-2,
              match=2, mismatch=-3, traceback=0)


def test_config1_pair_and_batch_agreement(eng):
    w = ta.Workload(1)
    a, b = w.pair_bytes(0)
    check(eng, [(a, b)], "300 x 320", mode=ta.GLOBAL)
    rng = random.Random(81)
    pairs = [(a, b)] + [(rand_prot(rng, rng.randint(1, 380)), rand_prot(rng, rng.randint(1, 380))) for _ in range(30)]
    for mode in (ta.GLOBAL, ta.LOCAL):
        params = ta.engine.make_params(mode=mode)
        batch = eng.align(pairs, params)
        for (x, y), g in zip(pairs, batch):
            r = eng.align_pair(x, y, params)
            assert r[:6] == g[:6]
            assert r.flags & ta.F_I32  # the pair kernel computes in int32


def test_large_pairs_and_empty_sides_take_the_batch_path(eng):
    rng = random.Random(91)
    a = rand_prot(rng, 1500)
    pairs = [(a, mutate(rng, a)), ("", "ACD"), ("WWW", "")]
    for mode in (ta.GLOBAL, ta.LOCAL):
        check(eng, pairs, "large/empty", mode=mode)


def test_bad_letters_and_small_buffer(eng):
    params = ta.engine.make_params()
    with pytest.raises(RuntimeError, match="alphabet"):
        eng.align_pair("AC-D", "ACD", params)
    with pytest.raises(RuntimeError, match="alphabet"):
        eng.align_pair("ACD", "AC1D", params)
    res = ta.engine.CResult()
    used = ctypes.c_int64()
    buf = ctypes.create_string_buffer(2)
    rc = ta.lib().ta_align_pair(eng.ctx, ctypes.byref(params), b"ACDEFGHIK", 9, b"ACDGHIK", 7, ctypes.byref(res),
                                buf, 2, ctypes.byref(used))
    assert rc == -3 and used.value == res.cigar_len > 2
