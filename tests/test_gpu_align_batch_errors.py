"""ta_align_batch end to end: letters are encoded on the GPU (a letter outside
the alphabet anywhere in a large batch is TA_E_ALPHABET, found on the device),
pairs are deduplicated by pointer on the host threads, and the call returns
what the batch API returns for the same pairs; error codes as documented
(TA_E_SPACE with the size needed, TA_E_INVALID before any work)."""
import ctypes

import numpy as np
import pytest

import tweakseq_amd as ta

pytestmark = pytest.mark.gpu

AA = np.frombuffer(b"ARNDCQEGHILKMFPSTWYV", dtype=np.uint8)
N = 20000  # enough pairs for the multi-threaded deduplication


def make_pairs(seed, n=N, maxlen=500):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        la = int(rng.integers(0, maxlen)) if k % 97 else 0   # a few empty sides
        a = AA[rng.integers(0, 20, la)].tobytes()
        if rng.random() < 0.5 and la > 4:
            b = a[: la * 3 // 4] + AA[rng.integers(0, 20, int(rng.integers(0, 40)))].tobytes()
        else:
            b = AA[rng.integers(0, 20, int(rng.integers(1, maxlen)))].tobytes()
        out.append((a, b))
    return out


def align_raw(eng, ps, n, params, cap):
    res = (ta.engine.CResult * n)()
    buf = ctypes.create_string_buffer(max(cap, 1))
    used = ctypes.c_int64(-1)
    rc = ta.lib().ta_align_batch(eng.ctx, ctypes.byref(params), ps.arr, n, res, buf, cap, ctypes.byref(used))
    rows = []
    if rc == 0:
        raw = buf.raw
        for k in range(n):
            r = res[k]
            cig = raw[r.cigar_off:r.cigar_off + r.cigar_len].decode() if r.cigar_len > 0 else ""
            rows.append((r.score, r.a_begin, r.a_end, r.b_begin, r.b_end, r.flags, cig))
    return rc, used.value, rows


@pytest.fixture
def engine():
    eng = ta.Engine(0)
    yield eng
    eng.close()


@pytest.mark.parametrize("mode", [ta.GLOBAL, ta.LOCAL])
def test_align_batch_equals_batch_api_with_shared_sequences(engine, mode):
    pairs = make_pairs(10 + mode)
    # every 5th pair reuses an earlier pair's buffers (deduplicated by pointer)
    ps = ta.engine._Pairs(pairs)
    for k in range(5, N, 5):
        ps.arr[k].a, ps.arr[k].la = ps.arr[k - 3].a, ps.arr[k - 3].la
        ps.arr[k].b, ps.arr[k].lb = ps.arr[k - 1].a, ps.arr[k - 1].la
    eff = [(ctypes.string_at(ps.arr[k].a, ps.arr[k].la), ctypes.string_at(ps.arr[k].b, ps.arr[k].lb))
           for k in range(N)]
    params = ta.engine.make_params(mode=mode)
    cap = 4 * sum(len(a) + len(b) for a, b in eff) + 16
    rc, used, rows = align_raw(engine, ps, N, params, cap)
    assert rc == 0
    want = engine.align(eff, params)
    bad = [k for k in range(N) if rows[k] != (want[k].score, want[k].a_begin, want[k].a_end, want[k].b_begin,
                                              want[k].b_end, want[k].flags, want[k].cigar)]
    assert not bad, "pair %d: %r" % (bad[0], rows[bad[0]])
    assert used == sum(len(r[6]) for r in rows)


def test_align_batch_errors(engine):
    pairs = make_pairs(20)
    params = ta.engine.make_params(mode=ta.LOCAL)
    ps = ta.engine._Pairs(pairs)
    cap = 4 * sum(len(a) + len(b) for a, b in pairs) + 16
    # a CIGAR buffer too small: TA_E_SPACE and the size needed
    rc, used, _ = align_raw(engine, ps, N, params, 1000)
    rc_full, used_full, _ = align_raw(engine, ps, N, params, cap)
    assert rc == ta.engine.E_SPACE and rc_full == 0 and used == used_full > 1000
    # a letter outside the alphabet near the end: found by the GPU encoding
    bad = list(pairs)
    bad[-5] = (b"ACD#" + bad[-5][0], bad[-5][1])
    assert align_raw(engine, ta.engine._Pairs(bad), N, params, cap + 16)[0] == ta.engine.E_ALPHABET
    # and the context still works afterwards
    assert align_raw(engine, ps, N, params, cap)[0] == 0
    # a pair past the length limit: TA_E_INVALID before any work
    bad = list(pairs)
    bad[-3] = (b"A" * 1000001, b"A")
    assert align_raw(engine, ta.engine._Pairs(bad), N, params, cap + 4000016)[0] == ta.engine.E_INVALID
