"""ta_align_pair_static -- the allocation-free variant of ta_align_pair --
through the C ABI: identical results to ta_align_pair on consecutive calls
(the second with a smaller pair, so a stale larger result in the shared
device array would show), pairs with an empty side, and TA_E_LIMIT beyond
the fixed scratch."""
import ctypes
import random

import pytest

import tweakseq_amd as ta

pytestmark = pytest.mark.gpu

AA = "ARNDCQEGHILKMFPSTWYV"
E_LIMIT = -11


def rand_prot(rng, n):
    return "".join(rng.choice(AA) for _ in range(n))


def call(fn, eng, a, b, params):
    a, b = a.encode(), b.encode()
    res = ta.engine.CResult()
    cap = 2 * (len(a) + len(b)) + 16
    buf = ctypes.create_string_buffer(cap)
    used = ctypes.c_int64()
    f = getattr(ta.lib(), fn)
    f.argtypes = [ctypes.c_void_p, ctypes.POINTER(ta.engine.Params), ctypes.c_char_p, ctypes.c_int32,
                  ctypes.c_char_p, ctypes.c_int32, ctypes.POINTER(ta.engine.CResult), ctypes.c_char_p,
                  ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
    f.restype = ctypes.c_int
    rc = f(eng.ctx, ctypes.byref(params), a, len(a), b, len(b), ctypes.byref(res), buf, cap, ctypes.byref(used))
    cig = buf.raw[res.cigar_off:res.cigar_off + res.cigar_len].decode() if rc == 0 else ""
    return rc, (res.score, res.a_begin, res.a_end, res.b_begin, res.b_end, cig, res.flags)


@pytest.fixture(scope="module")
def eng():
    e = ta.Engine(0)
    yield e
    e.close()


@pytest.mark.parametrize("mode", [ta.GLOBAL, ta.LOCAL])
def test_static_equals_pair_on_consecutive_calls(eng, mode):
    rng = random.Random(101 + mode)
    params = ta.engine.make_params(mode=mode, gap_open=-11 if mode else -10)
    big = rand_prot(rng, 380)
    pairs = [(big, big[:200] + rand_prot(rng, 150)),       # larger first
             (rand_prot(rng, 40), rand_prot(rng, 35)),       # then smaller
             (rand_prot(rng, 300), rand_prot(rng, 320)),
             (rand_prot(rng, 7), rand_prot(rng, 9))]
    for a, b in pairs:
        rc_s, got = call("ta_align_pair_static", eng, a, b, params)
        rc_p, want = call("ta_align_pair", eng, a, b, params)
        assert rc_s == 0 and rc_p == 0
        assert got == want, (len(a), len(b), got[:5], want[:5])


def test_empty_sides_and_limits(eng):
    for mode in (ta.GLOBAL, ta.LOCAL):
        params = ta.engine.make_params(mode=mode)
        for a, b in (("", "ACD"), ("WWW", ""), ("", "")):
            rc_s, got = call("ta_align_pair_static", eng, a, b, params)
            rc_p, want = call("ta_align_pair", eng, a, b, params)
            assert rc_s == rc_p == 0 and got == want, (a, b, got, want)
    params = ta.engine.make_params()
    rc, _ = call("ta_align_pair_static", eng, "A" * 3000, "C" * 1200, params)  # la + lb > 4096
    assert rc == E_LIMIT
    assert "fixed scratch" in ta.lib().ta_strerror(E_LIMIT).decode()
    rc, _ = call("ta_align_pair_static", eng, "A" * 2000, "C" * 2000, params)  # codes do not fit on chip
    assert rc == E_LIMIT
    rc, _ = call("ta_align_pair_static", eng, "AC-D", "ACD", params)
    assert rc == -2  # TA_E_ALPHABET, found by the kernel
