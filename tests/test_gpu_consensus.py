"""GPU column consensus (ta_consensus) against the oracle's restatement of
Consensus::calculate: byte-equal."""
import random

import pytest

import tweakseq_amd as ta
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

CHARS = "ACDEFGHIKLMNPQRSTVWYBZXJOU-.acdefgw?*1"


@pytest.fixture(scope="module")
def eng():
    e = ta.Engine(0)
    yield e
    e.close()


def test_consensus_small_and_edge(eng):
    for rows, pl in ((["ACDEW"] * 3, -1.0), (["ACD"], -1.0), (["-", "-", "-"], -1.0), (["W", "W", "A"], 1.0),
                     (["A", "S", "A", "S"], -1.0), (["a", "-", "."], -1.0)):
        assert eng.consensus(rows, pl) == po.consensus(rows, pl), rows


@pytest.mark.parametrize("seed", range(4))
def test_consensus_random(eng, seed):
    rng = random.Random(100 + seed)
    for _ in range(6):
        n, L = rng.randint(1, 300), rng.randint(1, 1500)
        alpha = CHARS[: rng.randint(3, len(CHARS))]
        rows = ["".join(rng.choice(alpha) for _ in range(L)) for _ in range(n)]
        pl = rng.choice([-1.0, 0.0, 2.0, n / 3.0, float(n)])
        assert eng.consensus(rows, pl) == po.consensus(rows, pl), (n, L, pl)


def test_consensus_alignment_like(eng):
    # related rows with gap runs, 1,000 sequences x 600 columns
    rng = random.Random(7)
    base = [rng.choice(CHARS[:20]) for _ in range(600)]
    rows = []
    for _ in range(1000):
        r = [c if rng.random() > 0.2 else rng.choice(CHARS[:20]) for c in base]
        for _ in range(rng.randint(0, 4)):
            s = rng.randrange(600)
            for k in range(s, min(600, s + rng.randint(1, 30))):
                r[k] = "-"
        rows.append("".join(r))
    assert eng.consensus(rows) == po.consensus(rows)
