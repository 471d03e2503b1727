"""Progressive multiple alignment on the GPU (ta_msa) against the oracle's
restatement (oracle/msa.c): same guide tree, same columns for every residue,
row for row."""
import random

import numpy as np
import pytest

import tweakseq_amd as ta
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

AA = "ARNDCQEGHILKMFPSTWYV"


def family(rng, n, length, psub=0.2, pindel=0.05, alphabet=AA):
    base = "".join(rng.choice(alphabet) for _ in range(length))
    out = []
    for _ in range(n):
        s = []
        for c in base:
            r = rng.random()
            if r < pindel / 2:
                continue
            if r < pindel:
                s.append(rng.choice(alphabet))
            s.append(rng.choice(alphabet) if rng.random() < psub else c)
        out.append("".join(s) or alphabet[0])
    return out


@pytest.fixture(scope="module")
def eng():
    e = ta.Engine(0)
    yield e
    e.close()


def oracle_rows(seqs, alphabet=0, go=-10, ge=-1):
    if alphabet == 0:
        alen, cols, joins = po.msa([po.encode(s) for s in seqs], gap_open=go, gap_extend=ge)
    else:
        alen, cols, joins = po.msa([po.encode(s, alphabet=1) for s in seqs], smat=po.dna_matrix(2, -3),
                                   gap_open=go, gap_extend=ge)
    return po.msa_rows(seqs, alen, cols), joins


@pytest.mark.parametrize("n,length,seed", [(3, 40, 1), (5, 120, 2), (9, 250, 3), (17, 90, 4), (40, 300, 5)])
def test_matches_oracle(eng, n, length, seed):
    rng = random.Random(seed)
    seqs = family(rng, n, length)
    rows, joins = eng.msa(seqs)
    want, wjoins = oracle_rows(seqs)
    assert joins == wjoins
    assert rows == want


def test_unrelated_and_ragged(eng):
    rng = random.Random(9)
    seqs = ["".join(rng.choice(AA) for _ in range(rng.randint(1, 200))) for _ in range(12)]
    rows, joins = eng.msa(seqs)
    want, wjoins = oracle_rows(seqs)
    assert joins == wjoins and rows == want
    assert [r.replace("-", "") for r in rows] == seqs


def test_dna_and_gap_costs(eng):
    rng = random.Random(10)
    seqs = family(rng, 8, 150, alphabet="ACGT")
    p = ta.engine.make_params(mode=ta.GLOBAL, alphabet=ta.DNA, gap_open=-5, gap_extend=-2)
    rows, joins = eng.msa(seqs, p)
    want, wjoins = oracle_rows(seqs, alphabet=1, go=-5, ge=-2)
    assert joins == wjoins and rows == want


@pytest.mark.parametrize("alphabet", [0, 1])
def test_long_profiles_in_strips(eng, alphabet):
    """Profiles longer than one strip of DP rows (2,048): the join sweeps the
    rows in strips, the next reading the last row of the one before."""
    rng = random.Random(12 + alphabet)
    letters = AA if alphabet == 0 else "ACGT"
    seqs = family(rng, 4, 2300, psub=0.15, pindel=0.04, alphabet=letters)
    seqs.append("".join(rng.choice(letters) for _ in range(1500)))
    if alphabet == 0:
        rows, joins = eng.msa(seqs)
        want, wjoins = oracle_rows(seqs)
    else:
        p = ta.engine.make_params(mode=ta.GLOBAL, alphabet=ta.DNA, gap_open=-5, gap_extend=-2)
        rows, joins = eng.msa(seqs, p)
        want, wjoins = oracle_rows(seqs, alphabet=1, go=-5, ge=-2)
    assert max(len(s) for s in seqs) > 2048
    assert joins == wjoins and rows == want


def test_two_sequences_are_the_pairwise_alignment(eng):
    rng = random.Random(11)
    a, b = family(rng, 2, 200)
    rows, joins = eng.msa([a, b])
    assert joins == [(0, 1)]
    p = ta.engine.make_params(mode=ta.GLOBAL)
    r = eng.align([(a, b)], p)[0]
    ra, rb, i, j = [], [], 0, 0
    num = ""
    for ch in r.cigar:
        if ch.isdigit():
            num += ch
            continue
        for _ in range(int(num)):
            ra.append(a[i] if ch in "MI" else "-")
            rb.append(b[j] if ch in "MD" else "-")
            i += ch in "MI"
            j += ch in "MD"
        num = ""
    assert rows == ["".join(ra), "".join(rb)]


def test_identical_and_single(eng):
    s = "MKTAYIAKQRQISFVKSHFSRQ"
    rows, _ = eng.msa([s, s, s])
    assert rows == [s, s, s]
    rows, joins = eng.msa([s])
    assert rows == [s] and joins == []


def test_rejects_bad_input(eng):
    with pytest.raises(RuntimeError):
        eng.msa(["ACD", ""])
    with pytest.raises(RuntimeError):
        eng.msa(["ACD", "AC-D"])
    with pytest.raises(RuntimeError):
        eng.msa(["ACD", "ACD"], ta.engine.make_params(mode=ta.LOCAL))
