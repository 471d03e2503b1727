"""Autotuner: every candidate launch shape gives the same alignments, and the
fastest one is stored on the context and used by later batches."""
import pytest

import tweakseq_amd as ta

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = ta.Engine(0)
    yield e
    e.close()


def key(res):
    return [(r.score, r.a_begin, r.a_end, r.b_begin, r.b_end, r.cigar) for r in res]


@pytest.mark.parametrize("config,size", [(3, 1500), (2, 48)])
def test_results_do_not_depend_on_tuning(eng, config, size):
    w = ta.Workload(config, size=size)
    ps = w.pairs_struct()
    bucket = ta.lib().ta_tune_bucket(ps.arr, ps.n)
    base = key(eng.align(ps, w.params))
    try:
        for waves in (4, 8, 12):
            for lanes in (ta.LANES_I16, ta.LANES_I32):
                for wt in (0, 1):
                    eng.tune_set(w.params, bucket, waves, lanes, wt)
                    got = eng.align(ps, w.params)
                    assert key(got) == base, (waves, lanes, wt)
                    if lanes == ta.LANES_I32:
                        assert all(r.flags & ta.F_I32 for r in got if r.score)
    finally:
        eng.tune_set(w.params, bucket, 0, ta.LANES_AUTO, 0)


def test_autotune_keeps_the_winner(eng):
    w = ta.Workload(3, size=2000)
    ps = w.pairs_struct()
    t, ms = eng.autotune(ps, w.params, repeats=1)
    assert t.waves_per_cu in (4, 8, 12) and t.lanes in (ta.LANES_I16, ta.LANES_I32) and ms > 0
    got = eng.tune_get(w.params, ta.lib().ta_tune_bucket(ps.arr, ps.n))
    assert (got.waves_per_cu, got.lanes, got.wave_tasks) == (t.waves_per_cu, t.lanes, t.wave_tasks)
    assert key(eng.align(ps, w.params)) == key(eng.align(ps, ta.engine.make_params(
        mode=w.params.mode, gap_open=w.params.gap_open, gap_extend=w.params.gap_extend, lanes=ta.LANES_I16)))
