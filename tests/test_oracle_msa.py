"""Pins the oracle's progressive multiple alignment (oracle/msa.c): two
sequences reduce to the pairwise global alignment, identical sequences stay
gap-free, columns are consistent, and a join of a profile with a sequence is
the brute-force optimum (with the pairwise tie-breaks) over all alignments of
their columns."""
import random

import numpy as np
import pytest

from oracle import pyoracle as po
import bruteforce as bf

S_PROT, _ = po.protein_matrix()
SP = S_PROT.astype(int).tolist()


def rand_codes(rng, n, k=20):
    return np.array([rng.randrange(k) for _ in range(n)], dtype=np.uint8)


def cigar_columns(cigar, la, lb):
    """Columns of a's and b's residues in the alignment a CIGAR describes."""
    ca, cb, q, i, j = [], [], 0, 0, 0
    num = ""
    for ch in cigar:
        if ch.isdigit():
            num += ch
            continue
        for _ in range(int(num)):
            if ch in "MI":
                ca.append(q)
            if ch in "MD":
                cb.append(q)
            q += 1
        num = ""
    assert len(ca) == la and len(cb) == lb
    return q, ca, cb


def ops_of(cigar):
    out, num = [], ""
    for ch in cigar:
        if ch.isdigit():
            num += ch
        else:
            out += [ch] * int(num)
            num = ""
    return out


@pytest.mark.parametrize("go,ge", [(-10, -1), (-4, -4), (-11, -1)])
def test_two_sequences_are_the_pairwise_alignment(go, ge):
    rng = random.Random(50 + go)
    p = po.params(mode=0, gap_open=go, gap_extend=ge)
    for _ in range(60):
        k = rng.choice([3, 6, 23])
        a, b = rand_codes(rng, rng.randint(1, 30), k), rand_codes(rng, rng.randint(1, 30), k)
        alen, cols, joins = po.msa([a, b], gap_open=go, gap_extend=ge)
        assert joins == [(0, 1)]
        _, _, _, _, _, cig = po.align(p, a, b)
        L, ca, cb = cigar_columns(cig, len(a), len(b))
        assert alen == L
        assert list(cols[0]) == ca and list(cols[1]) == cb


def test_identical_sequences_stay_gap_free():
    rng = random.Random(7)
    s = rand_codes(rng, 40)
    alen, cols, _ = po.msa([s, s.copy(), s.copy(), s.copy()])
    assert alen == 40
    for c in cols:
        assert list(c) == list(range(40))


def test_columns_are_consistent():
    rng = random.Random(11)
    for _ in range(20):
        n = rng.randint(3, 8)
        base = rand_codes(rng, rng.randint(5, 40))
        seqs = []
        for _ in range(n):
            s = [c for c in base if rng.random() > 0.15]
            s += list(rand_codes(rng, rng.randint(0, 4)))
            seqs.append(np.array(s or [1], dtype=np.uint8))
        alen, cols, joins = po.msa(seqs)
        assert len(joins) == n - 1
        used = np.zeros(alen, dtype=int)
        for s, c in zip(seqs, cols):
            assert len(c) == len(s)
            assert all(c[k] < c[k + 1] for k in range(len(c) - 1))
            assert len(c) == 0 or (c[0] >= 0 and c[-1] < alen)
            used[c] += 1
        assert (used > 0).all()  # no all-gap column


def test_join_of_a_profile_and_a_sequence_is_the_bruteforce_optimum():
    """Three tiny sequences: the second join aligns the lone sequence (the
    lower cluster id, DP rows) with the two-row profile of the first join
    (DP columns) under M(i,j) = sum of S over the column's residues and gap
    costs times 2.  Brute force over every alignment of the columns, with the
    pairwise tie-breaks, must give the oracle's final columns."""
    rng = random.Random(21)
    go, ge = -5, -2
    checked = 0
    for _ in range(120):
        k = rng.choice([3, 4, 23])
        seqs = [rand_codes(rng, rng.randint(1, 4), k) for _ in range(3)]
        alen, cols, joins = po.msa(seqs, gap_open=go, gap_extend=ge)
        a, b = joins[0]
        c = ({0, 1, 2} - {a, b}).pop()
        assert joins[1] == (c, 3)
        # the first join's profile
        plen, pcols, _ = po.msa([seqs[a], seqs[b]], gap_open=go, gap_extend=ge)
        prof = [[] for _ in range(plen)]
        for s, cc in ((seqs[a], pcols[0]), (seqs[b], pcols[1])):
            for r, q in enumerate(cc):
                prof[q].append(int(s[r]))
        M = [[sum(SP[int(x)][y] for y in prof[j]) for j in range(plen)] for x in seqs[c]]
        _, _, _, _, _, cig = bf.global_align(list(range(len(seqs[c]))), list(range(plen)), M, 2 * go, 2 * ge)
        ops = ops_of(cig)
        want_c, pmap, q, i, j = [], {}, 0, 0, 0
        for o in ops:
            if o in "MI":
                want_c.append(q)
            if o in "MD":
                pmap[j] = q
                j += 1
            q += 1
        assert alen == len(ops)
        assert list(cols[c]) == want_c
        assert list(cols[a]) == [pmap[x] for x in pcols[0]]
        assert list(cols[b]) == [pmap[x] for x in pcols[1]]
        checked += 1
    assert checked == 120


def test_upgma_joins_the_closest_pair_first():
    rng = random.Random(3)
    x = rand_codes(rng, 60)
    y = x.copy()
    y[::10] = (y[::10] + 1) % 20  # close to x
    z = rand_codes(rng, 60)        # unrelated
    _, _, joins = po.msa([z, x, y])
    assert joins[0] == (1, 2)
    assert joins[1] == (0, 3)
