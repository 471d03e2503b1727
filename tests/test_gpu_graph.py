"""Captured (hipGraph) runs give the eager results on every replay, with other
GPU work and host synchronisation between the replays."""
import random

import pytest

import tweakseq_amd as ta

from test_gpu_parity import mutate, rand_prot

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = ta.Engine(0)
    yield e
    e.close()


def key(res):
    return [(r.score, r.a_begin, r.a_end, r.b_begin, r.b_end, r.cigar, r.flags) for r in res]


def replay_equals_eager(eng, pairs_struct, params, other):
    b = eng.batch(pairs_struct, params)
    try:
        b.run(); b.wait()
        want = key(b.results())
        b.graph(True)
        for rep in range(4):
            b.run(); b.wait()
            got = key(b.results())
            assert got == want, "replay %d differs" % rep
            t = b.timings()
            assert t[0] == -1 and t[3] > 0
            other()  # unrelated eager work between replays
        b.graph(False)
        b.run(); b.wait()
        assert key(b.results()) == want
        assert b.timings()[0] >= 0
    finally:
        b.close()


def test_graph_local_protein(eng):
    w = ta.Workload(3, size=3000)
    rng = random.Random(3)
    side = [(rand_prot(rng, 200), rand_prot(rng, 180)) for _ in range(50)]
    replay_equals_eager(eng, w.pairs_struct(), w.params,
                        lambda: eng.align(side, ta.engine.make_params(mode=ta.LOCAL)))


def test_graph_global_and_long(eng):
    rng = random.Random(4)
    a = rand_prot(rng, 3000, "ACGT")
    pairs = [(a, mutate(rng, a, 0.1, 0.02, "ACGT")), (rand_prot(rng, 700, "ACGT"), rand_prot(rng, 650, "ACGT"))]
    params = ta.engine.make_params(mode=ta.GLOBAL, alphabet=ta.DNA, gap_open=-5, gap_extend=-2, match=2, mismatch=-3,
                                   lanes=ta.LANES_LONG)
    from tweakseq_amd.engine import _Pairs
    replay_equals_eager(eng, _Pairs(pairs), params,
                        lambda: eng.align([("ACDEFG" * 30, "ACDFG" * 33)], ta.engine.make_params()))
    w = ta.Workload(2, size=96)
    replay_equals_eager(eng, w.pairs_struct(), w.params, lambda: None)
