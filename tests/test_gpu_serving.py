"""Serving: continuous batching returns, for every request, exactly what the
same request gives when aligned alone; cancellation and admission limits.

Each test runs through both admission paths: the device queue (requests
pushed into the running persistent kernel) and batches (requests grouped and
launched per batch, TWEAKALIGN_DEVICE_QUEUE=0).  The switch is read when a
context's serving loop first admits work, so it is set before the Engine
exists."""
import random
import time

import pytest

import tweakseq_amd as ta

pytestmark = pytest.mark.gpu

AA = "ARNDCQEGHILKMFPSTWYV"


def rand_prot(rng, n):
    return "".join(rng.choice(AA) for _ in range(n))


def make_requests(seed, nreq, maxlen=700):
    rng = random.Random(seed)
    reqs = []
    for _ in range(nreq):
        k = rng.randint(1, 12)
        mode = rng.choice([ta.GLOBAL, ta.LOCAL])
        pairs = []
        for _ in range(k):
            a = rand_prot(rng, rng.randint(1, maxlen))
            b = a[: len(a) // 2] + rand_prot(rng, rng.randint(0, maxlen // 2)) if rng.random() < 0.5 else \
                rand_prot(rng, rng.randint(1, maxlen))
            pairs.append((a, b))
        reqs.append((ta.engine.make_params(mode=mode, gap_open=-11 if mode else -10), pairs))
    return reqs


@pytest.fixture(params=["device_queue", "batches"])
def path(request, monkeypatch):
    monkeypatch.setenv("TWEAKALIGN_DEVICE_QUEUE", "1" if request.param == "device_queue" else "0")
    return request.param


def wait_all(eng, tickets, timeout=120):
    out = {}
    t0 = time.time()
    while len(out) < len(tickets):
        for key, (t, n) in tickets.items():
            if key in out:
                continue
            r = eng.poll(t, n)
            if r is not None:
                out[key] = r
        assert time.time() - t0 < timeout, "serving did not finish"
        time.sleep(0.001)
    return out


def test_results_identical_alone_and_in_the_mix(path):
    eng = ta.Engine(0)
    try:
        eng.serve_config(64 << 30, 4)
        reqs = make_requests(1, 60)
        tickets = {q: eng.submit(pairs, params) for q, (params, pairs) in enumerate(reqs)}
        mixed = wait_all(eng, tickets)
        for q, (params, pairs) in enumerate(reqs):
            alone = eng.align(pairs, params)
            assert mixed[q] == alone, "request %d differs in the mix" % q
    finally:
        eng.close()


def test_served_config4_mix_matches_oracle(path):
    """The served workload's mix -- 80 requests of 1-256 pairs (geometric,
    mean 16), lengths uniform in [50, 5,000], global and local interleaved
    per request, -10/-1 -- submitted at once through ta_submit / ta_poll; every
    pair's score, coordinates and CIGAR equal the CPU oracle's."""
    from oracle import pyoracle as po
    w = ta.Workload(4, size=80, max_len=5000)
    reqs = w.requests()
    assert {m for _, _, _, m in reqs} == {ta.GLOBAL, ta.LOCAL}
    eng = ta.Engine(0)
    try:
        eng.serve_config(64 << 30, 4)
        tickets = {}
        for q, (_, first, count, mode) in enumerate(reqs):
            params = ta.engine.make_params(mode=mode, gap_open=-10, gap_extend=-1)
            tickets[q] = eng.submit(w.pairs_struct(first, count), params)
        served = wait_all(eng, tickets, timeout=300)
    finally:
        eng.close()
    for mode in (ta.GLOBAL, ta.LOCAL):
        idx = [(q, k) for q, (_, first, count, m) in enumerate(reqs) if m == mode for k in range(count)]
        pairs = [w.pair_bytes(reqs[q][1] + k) for q, k in idx]
        want = po.align_pairs(po.params(mode=mode, gap_open=-10, gap_extend=-1), pairs)
        bad = []
        for (q, k), (a, b), wt in zip(idx, pairs, want):
            g = served[q][k]
            if (g.score, g.a_begin, g.a_end, g.b_begin, g.b_end, g.cigar) != wt or g.flags & ta.F_WALK_FAILED:
                bad.append((q, k, len(a), len(b), (g.score, g.a_begin, g.a_end, g.b_begin, g.b_end), wt[:5]))
        assert not bad, "mode %d: %d/%d served pairs differ from the oracle, first %r" % (
            mode, len(bad), len(idx), bad[:3])


def test_small_budget_serialises_but_completes(path):
    eng = ta.Engine(0)
    try:
        reqs = make_requests(2, 12, maxlen=400)
        # Budget of about one request at a time: admission must still drain the queue.
        eng.serve_config(6 << 20, 1)
        tickets = {q: eng.submit(pairs, params) for q, (params, pairs) in enumerate(reqs)}
        done = wait_all(eng, tickets)
        for q, (params, pairs) in enumerate(reqs):
            assert done[q] == eng.align(pairs, params)
    finally:
        eng.close()


def test_budget_rejects_oversized_request(path):
    eng = ta.Engine(0)
    try:
        eng.serve_config(1 << 20, 2)
        big = [(rand_prot(random.Random(3), 3000), rand_prot(random.Random(4), 3000))]
        with pytest.raises(RuntimeError, match="budget"):
            eng.submit(big, ta.engine.make_params(mode=ta.LOCAL))
    finally:
        eng.close()


def test_lowered_budget_fails_queued_request_and_serves_the_rest(path):
    eng = ta.Engine(0)
    try:
        eng.serve_config(64 << 30, 1)
        params = ta.engine.make_params(mode=ta.LOCAL)
        big = [(rand_prot(random.Random(3), 3000), rand_prot(random.Random(4), 3000))]
        small = [(rand_prot(random.Random(5), 120), rand_prot(random.Random(6), 130))]
        # One in-flight slot, taken by a request of other parameters (so that it
        # forms a batch of its own) submitted just before the others.
        rng = random.Random(7)
        first_pairs = [(rand_prot(rng, 1500), rand_prot(rng, 1500)) for _ in range(200)]
        first, n_first = eng.submit(first_pairs, ta.engine.make_params(mode=ta.GLOBAL))
        t_big, _ = eng.submit(big, params)
        t_small, _ = eng.submit(small, params)
        eng.serve_config(1 << 20, 0)  # now below the queued big request's planes
        t0 = time.time()
        r = None
        while r is None:
            r = eng.poll(t_small, 1)
            assert time.time() - t0 < 60
        assert r == eng.align(small, params)
        if path == "batches":
            with pytest.raises(RuntimeError, match="budget"):
                while eng.poll(t_big, 1) is None:
                    assert time.time() - t0 < 60
        else:
            # The device queue takes requests as they arrive (max_inflight
            # counts batches, which it does not form): the big request was
            # admitted before the change, and then runs exactly, or it was
            # still waiting and failed.
            try:
                rb = None
                while rb is None:
                    rb = eng.poll(t_big, 1)
                    assert time.time() - t0 < 60
                assert rb == eng.align(big, params)
            except RuntimeError as e:
                assert "budget" in str(e)
        # The first request ran, or (if the scheduler first saw it after the
        # change) failed the same way; either way it left the queue.
        try:
            while eng.poll(first, n_first) is None:
                assert time.time() - t0 < 60
        except RuntimeError as e:
            assert "budget" in str(e)
    finally:
        eng.close()


def test_cancel(path):
    eng = ta.Engine(0)
    try:
        eng.serve_config(64 << 30, 1)
        reqs = make_requests(5, 20, maxlen=900)
        tickets = [eng.submit(pairs, params) for params, pairs in reqs]
        # Cancel the later half; the rest must still complete exactly.
        for t, _ in tickets[10:]:
            assert eng.cancel(t) in (0, -8)
        for q, (t, n) in enumerate(tickets[:10]):
            r = None
            t0 = time.time()
            while r is None:
                r = eng.poll(t, n)
                assert time.time() - t0 < 60
            assert r == eng.align(reqs[q][1], reqs[q][0])
        for t, n in tickets[10:]:
            with pytest.raises(RuntimeError):
                eng.poll(t, n)
    finally:
        eng.close()
