"""Walks inside the fill kernel (a walker wave beside each fill wave) give the
same results as walks after the fill, and both match the oracle."""
import os
import random

import pytest

import tweakseq_amd as ta
from test_gpu_parity import compare, oracle_results, rand_prot, mutate

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = ta.Engine(0)
    yield e
    e.close()


def mixed_pairs(seed, n):
    # many short pairs (G-lane groups) and some long ones (whole-wave tasks)
    rng = random.Random(seed)
    out = []
    for k in range(n):
        la = rng.randint(600, 2400) if k % 37 == 0 else rng.randint(20, 300)
        a = rand_prot(rng, la)
        out.append((a, mutate(rng, a)))
    out.append(("", "ACD"))  # pairs without a fill task go through the traceback kernel
    out.append(("W", ""))
    return out


def run_batch(eng, pairs, params, walk, graph=False, runs=3):
    old = os.environ.get("TWEAKALIGN_WALK_IN_FILL")
    os.environ["TWEAKALIGN_WALK_IN_FILL"] = "1" if walk else "0"
    try:
        b = eng.batch(pairs, params)
    finally:
        if old is None:
            del os.environ["TWEAKALIGN_WALK_IN_FILL"]
        else:
            os.environ["TWEAKALIGN_WALK_IN_FILL"] = old
    if graph:
        b.graph(True)
    outs = []
    for _ in range(runs):
        b.run(); b.wait()
        outs.append(b.results())
    b.close()
    return outs


@pytest.mark.parametrize("mode", [ta.LOCAL, ta.GLOBAL])
@pytest.mark.parametrize("lanes", [ta.LANES_AUTO, ta.LANES_I32])
def test_walk_in_fill_matches_after_fill_and_oracle(eng, lanes, mode):
    pairs = mixed_pairs(71 + lanes + 5 * mode, 1200)
    params = ta.engine.make_params(mode=mode, lanes=lanes)
    after = run_batch(eng, pairs, params, walk=False)
    inside = run_batch(eng, pairs, params, walk=True)
    for r in after[1:] + inside:
        assert r == after[0]
    compare(inside[0], oracle_results(params, pairs), params, pairs, "walk in fill")


def test_walk_in_fill_few_pairs_and_graph(eng):
    # fewer pairs than waves: every task whole-wave, most workgroups idle
    pairs = mixed_pairs(74, 60)
    params = ta.engine.make_params(mode=ta.LOCAL)
    ref = oracle_results(params, pairs)
    for r in run_batch(eng, pairs, params, walk=True, graph=True, runs=4):
        compare(r, ref, params, pairs, "walk in fill, graph replay")


def test_walk_in_fill_local_batch_prefix(eng):
    w = ta.Workload(3, size=4000)
    pairs = [w.pair_bytes(k) for k in range(w.npairs)]
    params = w.params
    after = run_batch(eng, pairs, params, walk=False, runs=1)[0]
    inside = run_batch(eng, pairs, params, walk=True, runs=2)
    for r in inside:
        assert r == after


def run_env(eng, pairs, params, env, runs=2):
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        b = eng.batch(pairs, params)
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v
    outs = []
    for _ in range(runs):
        b.run(); b.wait()
        outs.append(b.results())
    b.close()
    return outs


@pytest.mark.parametrize("lanes", [ta.LANES_AUTO, ta.LANES_I32])
@pytest.mark.parametrize("mode", [ta.LOCAL, ta.GLOBAL])
def test_cooperative_long_tasks(eng, mode, lanes):
    # few pairs, several 512-row strips each: a workgroup's fill waves sweep
    # each pair together (strips pipelined); one short pair rides along
    rng = random.Random(81 + mode)
    pairs = []
    for la in (4100, 2600, 1800, 1030, 700, 513, 90):
        a = rand_prot(rng, la)
        pairs.append((a, mutate(rng, a)))
    pairs.append((rand_prot(rng, 3000), rand_prot(rng, 2500)))  # unrelated: scores near 0
    params = ta.engine.make_params(mode=mode, lanes=lanes)
    coop = run_env(eng, pairs, params, {"TWEAKALIGN_COOP": "1"})
    plain = run_env(eng, pairs, params, {"TWEAKALIGN_COOP": "0"}, runs=1)[0]
    for r in coop:
        assert r == plain
    compare(coop[0], oracle_results(params, pairs), params, pairs, "cooperative")
