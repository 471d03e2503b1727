"""Resident sequences (ta_seq_add / ta_align_resident): sequences encoded once
into device memory and aligned by handle, with results identical to the
oracle's and to ta_align_batch on the same letters."""
import random

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


def mutate(rng, s, psub=0.25, pindel=0.04):
    out, i = [], 0
    while i < len(s):
        if rng.random() < pindel:
            if rng.random() < 0.5:
                out.append(rng.choice(AA))
            else:
                i += 1
            continue
        out.append(rng.choice(AA) if rng.random() < psub else s[i])
        i += 1
    return "".join(out)


def oracle(params, pairs):
    p = po.params(mode=params.mode, alphabet=params.alphabet, gap_open=params.gap_open,
                  gap_extend=params.gap_extend, match=params.match, mismatch=params.mismatch, traceback=1)
    return [po.align(p, po.encode(a, params.alphabet), po.encode(b, params.alphabet)) for a, b in pairs]


def same(got, want, traceback=True):
    assert len(got) == len(want)
    for k, (g, w) in enumerate(zip(got, want)):
        assert (g.score, g.a_begin, g.a_end, g.b_begin, g.b_end) == tuple(w[:5]), k
        if traceback:
            assert g.cigar == w[5], k
        assert not g.flags & ta.F_WALK_FAILED


@pytest.mark.parametrize("mode", [ta.GLOBAL, ta.LOCAL])
def test_resident_pairs_equal_oracle_and_batch(eng, mode):
    rng = random.Random(11 + mode)
    eng.seq_reset()
    seqs = []
    for _ in range(40):
        base = "".join(rng.choice(AA) for _ in range(rng.randint(0 if mode == ta.LOCAL else 1, 400)))
        seqs += [base, mutate(rng, base)]
    handles = [eng.seq_add(s) for s in seqs]
    idx = [(2 * k, 2 * k + 1) for k in range(40)] + [(rng.randrange(80), rng.randrange(80)) for _ in range(300)]
    params = ta.engine.make_params(mode=mode, gap_open=-11, gap_extend=-1)
    pairs = [(seqs[i], seqs[j]) for i, j in idx]
    got = eng.align_resident([(handles[i], handles[j]) for i, j in idx], params)
    same(got, oracle(params, pairs))
    # the same letters through ta_align_batch
    assert [tuple(g) for g in got] == [tuple(g) for g in eng.align(pairs, params)]
    # and the handles again, in another order
    again = eng.align_resident([(handles[j], handles[i]) for i, j in idx[:50]], params)
    same(again, oracle(params, [(seqs[j], seqs[i]) for i, j in idx[:50]]))


def test_resident_saturating_local_pair_reruns_in_int32(eng):
    """A local pair whose score leaves int16 (3,000 x W: 33,000) is flagged
    and rerun in int32; the resident store's residue counts give the bound."""
    eng.seq_reset()
    a = "W" * 3000
    h = eng.seq_add(a)
    g = eng.seq_add("AC" + a[:2900] + "G")
    params = ta.engine.make_params(mode=ta.LOCAL, gap_open=-10, gap_extend=-1)
    got = eng.align_resident([(h, h), (h, g)], params)
    same(got, oracle(params, [(a, a), (a, "AC" + a[:2900] + "G")]))
    assert got[0].score == 33000 and got[0].flags & ta.F_I32


def test_resident_store_grows_and_resets(eng):
    """Handles added before and after the arena grows stay valid; a reset
    forgets them all."""
    rng = random.Random(5)
    eng.seq_reset()
    small = ["".join(rng.choice(AA) for _ in range(rng.randint(50, 200))) for _ in range(6)]
    params = ta.engine.make_params(mode=ta.GLOBAL, gap_open=-10, gap_extend=-1)
    before = [eng.seq_add(s) for s in small[:3]]
    # the first call uploads the codes added so far
    same(eng.align_resident([(before[0], before[1])], params), oracle(params, [(small[0], small[1])]))
    big = "".join(rng.choice(AA) for _ in range(1000000))
    for _ in range(70):  # 70 MB: past the first arena (its uploaded part copied over) and
        eng.seq_add(big)  # past the 64 MB of codes held back before an upload
    after = [eng.seq_add(s) for s in small[3:]]
    hs = before + after
    idx = [(i, j) for i in range(6) for j in range(6)]
    got = eng.align_resident([(hs[i], hs[j]) for i, j in idx], params)
    same(got, oracle(params, [(small[i], small[j]) for i, j in idx]))
    eng.seq_reset()
    with pytest.raises(RuntimeError, match="ta_align_resident"):
        eng.align_resident([(hs[0], hs[1])], params)


def test_resident_errors(eng):
    eng.seq_reset()
    prot0 = ta.engine.make_params(mode=ta.LOCAL, gap_open=-11, gap_extend=-1)
    # no pairs: nothing to do, with or without sequences in the store (as
    # ta_align_batch with no pairs)
    assert eng.align([], prot0) == []
    assert eng.align_resident([], prot0) == []
    p = eng.seq_add("ACDEFGHIK")
    assert eng.align_resident([], prot0) == []
    d = eng.seq_add("ACGTACGT", alphabet=ta.DNA)
    with pytest.raises(RuntimeError, match="outside the alphabet|alphabet"):
        eng.seq_add("AC#DE")
    prot = ta.engine.make_params(mode=ta.GLOBAL, gap_open=-10, gap_extend=-1)
    dna = ta.engine.make_params(mode=ta.GLOBAL, alphabet=ta.DNA, gap_open=-5, gap_extend=-2, match=2, mismatch=-3)
    for params, pairs in ((prot, [(p, d)]), (dna, [(p, p)]), (prot, [(p, 99)]), (prot, [(-1, p)])):
        with pytest.raises(RuntimeError, match="ta_align_resident"):
            eng.align_resident(pairs, params)
    # empty sequences are allowed, as in ta_align_batch
    e = eng.seq_add("")
    got = eng.align_resident([(e, p), (p, e)], prot)
    same(got, oracle(prot, [("", "ACDEFGHIK"), ("ACDEFGHIK", "")]))
    got = eng.align_resident([(d, d)], dna)
    same(got, oracle(dna, [("ACGTACGT", "ACGTACGT")]))


def test_served_resident_requests_equal_oracle(eng):
    """Resident requests through ta_submit_resident / ta_poll, interleaved
    with letter requests of both modes; the store grows, and is then reset,
    while requests wait (their handles were resolved at submission, and they
    hold the arena they read)."""
    import time
    rng = random.Random(23)
    eng.seq_reset()
    seqs = []
    for _ in range(30):
        base = "".join(rng.choice(AA) for _ in range(rng.randint(30, 400)))
        seqs += [base, mutate(rng, base)]
    hs = [eng.seq_add(s) for s in seqs]
    local = ta.engine.make_params(mode=ta.LOCAL, gap_open=-11, gap_extend=-1)
    glob = ta.engine.make_params(mode=ta.GLOBAL, gap_open=-10, gap_extend=-1)
    tickets = []
    for q in range(30):
        params = local if q % 2 else glob
        idx = [(rng.randrange(60), rng.randrange(60)) for _ in range(rng.randint(1, 12))]
        if q % 3 == 0:
            t, n = eng.submit([(seqs[i], seqs[j]) for i, j in idx], params)
        else:
            t, n = eng.submit_resident([(hs[i], hs[j]) for i, j in idx], params)
        tickets.append((t, n, params, idx))
        if q == 10:  # past the first arena while requests wait
            big = "A" * 1000000
            for _ in range(18):
                eng.seq_add(big)
    eng.seq_reset()
    tickets.append(eng.submit_resident([], local) + (local, []))  # no pairs, empty store
    for t, n, params, idx in tickets:
        deadline = time.time() + 60
        got = eng.poll(t, n, params.traceback)
        while got is None:
            assert time.time() < deadline, "request %d still pending" % t
            time.sleep(0.001)
            got = eng.poll(t, n, params.traceback)
        same(got, oracle(params, [(seqs[i], seqs[j]) for i, j in idx]))
    with pytest.raises(RuntimeError, match="ta_submit_resident"):
        eng.submit_resident([(hs[0], hs[1])], local)  # the handles are gone
