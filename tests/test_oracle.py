"""Pins the CPU oracle: reference table fixture, brute force, closed forms,
symmetry, the linear-gap special case and CIGAR re-scoring."""
import random

import numpy as np
import pytest

from oracle import pyoracle as po
import bruteforce as bf

S_PROT, MAP = po.protein_matrix()
SP = S_PROT.astype(int).tolist()


def rand_codes(rng, n, k=20):
    return np.array([rng.randrange(k) for _ in range(n)], dtype=np.uint8)


def test_fixture_is_blosum62():
    d = po.blosum62_fixture()
    M = np.array(d["matrix"])
    assert d["order"] == "ARNDCQEGHILKMFPSTWYVBZX"
    assert (M == M.T).all()
    assert M.min() == -4 and M.max() == 11
    assert list(np.diag(M)) == [4, 5, 6, 6, 9, 5, 5, 6, 8, 4, 4, 5, 5, 6, 7, 4, 5, 11, 7, 4, 4, 4, -1]
    # A spot check against the published NCBI table (Henikoff & Henikoff 1992).
    idx = {c: i for i, c in enumerate(d["order"])}
    assert M[idx["W"], idx["C"]] == -2 and M[idx["H"], idx["Y"]] == 2 and M[idx["I"], idx["V"]] == 3
    # Letter map: J, O, U, X -> X row; B, Z rows; every letter maps to its own row.
    cmap = d["map"]
    for c in d["order"]:
        assert cmap[ord(c) - 65] == idx[c]
    for c in "JOUX":
        assert cmap[ord(c) - 65] == 22


def test_encode():
    assert list(po.encode("arNdX")) == [0, 1, 2, 3, 22]
    assert list(po.encode("ACGTacgt", alphabet=1)) == [0, 1, 2, 3] * 2
    with pytest.raises(ValueError):
        po.encode("AC-D")
    with pytest.raises(ValueError):
        po.encode("ACGN", alphabet=1)


@pytest.mark.parametrize("go,ge", [(-10, -1), (-5, -2), (-3, -3), (-11, -1), (-4, 0)])
def test_global_bruteforce(go, ge):
    rng = random.Random(1000 + go * 7 + ge)
    p = po.params(mode=0, gap_open=go, gap_extend=ge)
    for _ in range(60):
        m, n = rng.randint(0, 5), rng.randint(0, 5)
        k = rng.choice([3, 6, 23])  # small alphabets force ties
        a, b = rand_codes(rng, m, k), rand_codes(rng, n, k)
        got = po.align(p, a, b)
        want = bf.global_align(list(a), list(b), SP, go, ge)
        assert got[:5] == want[:5], (a, b, got, want)
        if m and n:
            assert got[5] == want[5], (a, b, got, want)


@pytest.mark.parametrize("go,ge", [(-10, -1), (-5, -2), (-3, -3), (-11, -1), (-2, -1)])
def test_local_bruteforce(go, ge):
    rng = random.Random(2000 - go * 3 + ge)
    p = po.params(mode=1, gap_open=go, gap_extend=ge)
    for _ in range(40):
        m, n = rng.randint(0, 5), rng.randint(0, 5)
        k = rng.choice([3, 6, 23])
        a, b = rand_codes(rng, m, k), rand_codes(rng, n, k)
        got = po.align(p, a, b)
        want = bf.local_align(list(a), list(b), SP, go, ge)
        assert got == want, (a, b, got, want)


@pytest.mark.parametrize("mode", [0, 1])
def test_bruteforce_to_seven(mode):
    """Lengths 6 and 7 (up to D(7,7) = 48,639 global alignments per pair),
    three-letter alphabet for ties, two gap settings."""
    import time
    rng = random.Random(77 + mode)
    t0 = time.time()
    for go, ge in ((-10, -1), (-3, -3)):
        p = po.params(mode=mode, gap_open=go, gap_extend=ge)
        for _ in range(6 if mode == 0 else 4):
            m, n = rng.randint(6, 7), rng.randint(6, 7)
            if mode == 1:
                m, n = 7, rng.randint(6, 7)  # local enumerates every start and end as well
            a, b = rand_codes(rng, m, 3), rand_codes(rng, n, 3)
            got = po.align(p, a, b)
            fn = bf.global_align if mode == 0 else bf.local_align
            want = fn(list(a), list(b), SP, go, ge)
            assert got == want, (a, b, got, want)
    assert time.time() - t0 < 120


def test_dna_bruteforce():
    rng = random.Random(7)
    p = po.params(mode=0, alphabet=1, gap_open=-5, gap_extend=-2, match=2, mismatch=-3)
    S = po.dna_matrix(2, -3).astype(int).tolist()
    for mode in (0, 1):
        p.mode = mode
        for _ in range(40):
            a, b = rand_codes(rng, rng.randint(1, 5), 4), rand_codes(rng, rng.randint(1, 5), 4)
            got = po.align(p, a, b)
            fn = bf.global_align if mode == 0 else bf.local_align
            assert got == fn(list(a), list(b), S, -5, -2)


def test_closed_forms():
    rng = random.Random(3)
    p = po.params(mode=0)
    assert po.align(p, [], []) == (0, 0, -1, 0, -1, "")
    for la in (1, 2, 7, 50):
        a = rand_codes(rng, la)
        assert po.align(p, a, [])[0] == -10 + (la - 1) * -1
        assert po.align(p, a, [])[5] == "%dI" % la
        assert po.align(p, [], a)[5] == "%dD" % la
    # Identical X-free sequences: gap-free alignment of the diagonal.
    for la in (1, 10, 300):
        a = rand_codes(rng, la)
        diag = int(sum(SP[c][c] for c in a))
        g = po.align(p, a, a)
        assert g[0] == diag and g[5] == "%dM" % la
        l = po.align(po.params(mode=1), a, a)
        assert l[0] == diag and l[1:5] == (0, la - 1, 0, la - 1) and l[5] == "%dM" % la
    # Local of unrelated sequences is never negative; empty when nothing scores.
    assert po.align(po.params(mode=1), po.encode("WWW"), po.encode("GGG"))[0] == 0
    assert po.align(po.params(mode=1), po.encode("WWW"), po.encode("GGG"))[1] == -1


def test_linear_gap_equals_textbook_nw():
    rng = random.Random(4)
    for g in (-1, -4, -8):
        p = po.params(mode=0, gap_open=g, gap_extend=g)
        for _ in range(30):
            a, b = rand_codes(rng, rng.randint(0, 40)), rand_codes(rng, rng.randint(0, 40))
            assert po.align(p, a, b)[0] == bf.linear_nw(list(a), list(b), SP, g)


def test_symmetry_and_rescoring():
    rng = random.Random(5)
    for mode in (0, 1):
        p = po.params(mode=mode, gap_open=-11 if mode else -10)
        for _ in range(40):
            a, b = rand_codes(rng, rng.randint(1, 120)), rand_codes(rng, rng.randint(1, 120))
            r = po.align(p, a, b)
            assert po.align(p, b, a)[0] == r[0]
            assert po.rescore(p, a, b, r) == r[0]


def test_batch_matches_single():
    rng = random.Random(6)
    seqs = [rand_codes(rng, rng.randint(0, 80)) for _ in range(20)]
    codes = np.concatenate([s for s in seqs] + [np.zeros(1, np.uint8)])
    offs = np.cumsum([0] + [len(s) for s in seqs])[:-1]
    pairs = [(i, j) for i in range(20) for j in range(i + 1, 20)]
    p = po.params(mode=1, gap_open=-11)
    res, cig = po.align_batch(p, codes, [offs[i] for i, _ in pairs], [len(seqs[i]) for i, _ in pairs],
                              [offs[j] for _, j in pairs], [len(seqs[j]) for _, j in pairs], nthreads=4)
    for k, (i, j) in enumerate(pairs):
        one = po.align(p, seqs[i], seqs[j])
        assert (res[k].score, res[k].a_begin, res[k].a_end, res[k].b_begin, res[k].b_end, cig[k]) == one


# ---- Consensus::calculate (Consensus.cpp:80-161)

CONS_CHARS = "ACDEFGHIKLMNPQRSTVWYBZXJOU-.acdefgw?*1"


def consensus_by_counts(rows, plurality=-1.0):
    """Per-column class counts, then one score per class present (an
    algorithm of its own: the oracle loops over all pairs of sequences)."""
    d = po.blosum62_fixture()
    S, lmap = d["matrix"], d["map"]
    n = len(rows)
    if plurality < 0:
        plurality = n / 2.0

    def cls(ch):
        x = ord(ch) - 65
        return lmap[x] if 0 <= x <= 25 else 99

    def pair(a, b):
        if a == 99 and b == 99:
            return 1, 1
        if a == 99 or b == 99:
            return -4, 0
        return S[a][b], int(S[a][b] > 0)

    out = []
    for c in range(len(rows[0])):
        col = [cls(r[c]) for r in rows]
        counts = {}
        for k in col:
            counts[k] = counts.get(k, 0) + 1
        per = {}
        for k in counts:
            sc = mt = 0
            for j, cj in counts.items():
                s, m = pair(k, j)
                sc += cj * s
                mt += cj * m
            s, m = pair(k, k)
            per[k] = (sc - s, mt - m)
        best = max(per[k][0] for k in per)
        ri = next(i for i, k in enumerate(col) if per[k][0] == best)
        out.append(rows[ri][c] if per[col[ri]][1] >= plurality else "?")
    return "".join(out)


def test_consensus_closed_forms():
    # identical rows: every pairing is the diagonal score, positive
    assert po.consensus(["ACDEW", "ACDEW", "ACDEW"]) == b"ACDEW"
    # one row: no pairings, no matches, below the default plurality 0.5
    assert po.consensus(["ACD"]) == b"???"
    # gaps pair +1 with gaps: the gap wins a gap column
    assert po.consensus(["-", "-", "-"]) == b"-"
    # W W A: W scores 11 - 3 = 8 (twice), A scores -6; W's matches: 1 < 1.5
    assert po.consensus(["W", "W", "A"]) == b"?"
    assert po.consensus(["W", "W", "A"], plurality=1.0) == b"W"
    # ties go to the first sequence: A/A/S/S, A and S both score 4 + 1 + 1 = 6
    assert po.consensus(["A", "S", "A", "S"]) == b"A"
    assert po.consensus(["S", "A", "A", "S"]) == b"S"
    # lowercase and other bytes are the non-letter class; the winner keeps its byte
    assert po.consensus(["a", "-", "."]) == b"a"


def test_consensus_matches_counting():
    rng = random.Random(31)
    for _ in range(60):
        n, L = rng.randint(1, 40), rng.randint(1, 60)
        alpha = CONS_CHARS[: rng.randint(3, len(CONS_CHARS))]
        rows = ["".join(rng.choice(alpha) for _ in range(L)) for _ in range(n)]
        pl = rng.choice([-1.0, 0.0, 1.0, n / 3.0, float(n)])
        assert po.consensus(rows, pl).decode("latin-1") == consensus_by_counts(rows, pl)
