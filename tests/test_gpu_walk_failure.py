"""A traceback walk that gives up is an error at the boundary, never a
silent short CIGAR.  The walk's step budget is lowered from the host
(TWEAKALIGN_WALK_GUARD, read when a batch is created or a single-pair call
starts), so every walk ends
normally, flagged TA_F_WALK_FAILED, and each entry point reports TA_E_WALK:
ta_align_batch, ta_batch_results, ta_poll, ta_align_pair, and the CLI, which then exits 1
and writes no output file (the reference shows "Alignment not completed",
tweakseq/UI/SeqEditMainWin.cpp:852-856)."""
import ctypes
import os
import random
import subprocess
import time

import pytest

import tweakseq_amd as ta

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
AA = "ARNDCQEGHILKMFPSTWYV"


def rand_prot(rng, n):
    return "".join(rng.choice(AA) for _ in range(n))


@pytest.fixture
def short_walks(monkeypatch):
    monkeypatch.setenv("TWEAKALIGN_WALK_GUARD", "5")
    yield


def pairs_of(seed, n=6, lo=60, hi=400):
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        a = rand_prot(rng, rng.randint(lo, hi))
        out.append((a, a[: len(a) * 3 // 4] + rand_prot(rng, 30)))
    return out


def align_raw(eng, pairs, params):
    ps = ta.engine._Pairs(pairs)
    res = (ta.engine.CResult * len(pairs))()
    cap = 4 * sum(len(a) + len(b) for a, b in pairs) + 16
    buf = ctypes.create_string_buffer(cap)
    used = ctypes.c_int64()
    rc = ta.lib().ta_align_batch(eng.ctx, ctypes.byref(params), ps.arr, ps.n, res, buf, cap, ctypes.byref(used))
    return rc, [res[k].flags for k in range(len(pairs))]


@pytest.mark.parametrize("mode", [ta.GLOBAL, ta.LOCAL])
def test_align_batch_reports_walk_failure(short_walks, mode):
    eng = ta.Engine(0)
    try:
        params = ta.engine.make_params(mode=mode)
        rc, flags = align_raw(eng, pairs_of(1 + mode), params)
        assert rc == ta.E_WALK, rc
        assert all(f & ta.F_WALK_FAILED for f in flags), flags
        assert "walk" in ta.lib().ta_strerror(rc).decode()
        # batch results: the same error after the results are written
        b = eng.batch(pairs_of(3 + mode), params)
        b.run(); b.wait()
        with pytest.raises(RuntimeError, match="walk"):
            b.results()
        b.close()
    finally:
        eng.close()


@pytest.mark.parametrize("mode", [ta.GLOBAL, ta.LOCAL])
def test_align_pair_reports_walk_failure(short_walks, mode):
    eng = ta.Engine(0)
    try:
        a, b = pairs_of(11 + mode, n=1)[0]
        with pytest.raises(RuntimeError, match="walk"):
            eng.align_pair(a, b, ta.engine.make_params(mode=mode))
    finally:
        eng.close()


def test_poll_reports_walk_failure(short_walks):
    eng = ta.Engine(0)
    try:
        eng.serve_config(64 << 30, 2)
        params = ta.engine.make_params(mode=ta.LOCAL)
        t, n = eng.submit(pairs_of(5), params)
        t0 = time.time()
        with pytest.raises(RuntimeError, match="walk"):
            while eng.poll(t, n) is None:
                assert time.time() - t0 < 60
        # delivered (results and CIGARs written): the ticket is retired
        with pytest.raises(RuntimeError, match="ticket"):
            eng.poll(t, n)
    finally:
        eng.close()


def test_walks_complete_without_the_cap():
    eng = ta.Engine(0)
    try:
        rc, flags = align_raw(eng, pairs_of(1), ta.engine.make_params(mode=ta.GLOBAL))
        assert rc == 0 and not any(f & ta.F_WALK_FAILED for f in flags)
    finally:
        eng.close()


def test_cli_fails_instead_of_writing_a_misalignment(tmp_path):
    cli = os.path.join(ROOT, "tweakseq_amd", "tweakalign")
    fin, fout = tmp_path / "in.fa", tmp_path / "out.fa"
    a, b = pairs_of(9, n=1)[0]
    fin.write_text(">a first\n%s\n>b second\n%s\n" % (a, b))
    env = dict(os.environ, TWEAKALIGN_WALK_GUARD="5")
    r = subprocess.run([cli, "-in", str(fin), "-out", str(fout)], capture_output=True, text=True, env=env,
                       timeout=120)
    assert r.returncode == 1 and "walk" in r.stderr and not fout.exists()
    env.pop("TWEAKALIGN_WALK_GUARD")
    r = subprocess.run([cli, "-in", str(fin), "-out", str(fout)], capture_output=True, text=True, env=env,
                       timeout=120)
    assert r.returncode == 0 and fout.exists()
