"""ctypes binding of libtweakalign.so (include/tweakalign.h, tweakalign_workload.h)."""
import ctypes
import os
from collections import namedtuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtweakalign.so")

GLOBAL, LOCAL = 0, 1
PROTEIN, DNA = 0, 1
LANES_AUTO, LANES_I16, LANES_I32, LANES_LONG = 0, 16, 32, 64
F_I32, F_RERUN, F_LONG, F_EMPTY = 0x1, 0x2, 0x4, 0x8
F_WALK_FAILED = 0x100
E_PENDING = 1
E_INVALID, E_ALPHABET, E_SPACE, E_WALK = -1, -2, -3, -10

Result = namedtuple("Result", "score a_begin a_end b_begin b_end cigar flags")


class Params(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in
                ("mode", "alphabet", "gap_open", "gap_extend", "match", "mismatch",
                 "traceback", "lanes")]


class Pair(ctypes.Structure):
    _fields_ = [("a", ctypes.c_void_p), ("la", ctypes.c_int32),
                ("b", ctypes.c_void_p), ("lb", ctypes.c_int32)]


class SeqPair(ctypes.Structure):
    _fields_ = [("a", ctypes.c_int64), ("b", ctypes.c_int64)]


class Tuning(ctypes.Structure):
    _fields_ = [("waves_per_cu", ctypes.c_int32), ("lanes", ctypes.c_int32), ("wave_tasks", ctypes.c_int32)]


class CResult(ctypes.Structure):
    _fields_ = [("score", ctypes.c_int32), ("a_begin", ctypes.c_int32),
                ("a_end", ctypes.c_int32), ("b_begin", ctypes.c_int32),
                ("b_end", ctypes.c_int32), ("cigar_off", ctypes.c_int64),
                ("cigar_len", ctypes.c_int32), ("flags", ctypes.c_int32)]


_LIB = None


def lib():
    """Load libtweakalign.so; raises when it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libtweakalign.so is not built (python -m tweakseq_amd.build)")
    L = ctypes.CDLL(LIB_PATH)
    P, vp, i32, i64 = ctypes.POINTER, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
    sig = {
        "ta_strerror": ([ctypes.c_int], ctypes.c_char_p),
        "ta_create": ([P(vp), ctypes.c_int], ctypes.c_int),
        "ta_destroy": ([vp], None),
        "ta_align_batch": ([vp, P(Params), P(Pair), i32, P(CResult), ctypes.c_char_p, i64, P(i64)], ctypes.c_int),
        "ta_align_pair": ([vp, P(Params), ctypes.c_char_p, i32, ctypes.c_char_p, i32, P(CResult), ctypes.c_char_p,
                           i64, P(i64)], ctypes.c_int),
        "ta_batch_create": ([vp, P(Params), P(Pair), i32, P(vp)], ctypes.c_int),
        "ta_batch_run": ([vp, vp], ctypes.c_int),
        "ta_batch_run_plain": ([vp, vp, i32, i32], ctypes.c_int),
        "ta_batch_wait": ([vp, vp], ctypes.c_int),
        "ta_batch_results": ([vp, vp, P(CResult), ctypes.c_char_p, i64, P(i64)], ctypes.c_int),
        "ta_batch_timings": ([vp, P(ctypes.c_double), i32], ctypes.c_int),
        "ta_batch_graph": ([vp, i32], ctypes.c_int),
        "ta_tune_bucket": ([P(Pair), i32], ctypes.c_int),
        "ta_autotune": ([vp, P(Params), P(Pair), i32, i32, P(Tuning), P(ctypes.c_double)], ctypes.c_int),
        "ta_tune_set": ([vp, P(Params), i32, P(Tuning)], ctypes.c_int),
        "ta_tune_get": ([vp, P(Params), i32, P(Tuning)], ctypes.c_int),
        "ta_batch_info": ([vp, P(i64), P(i64), P(i32)], ctypes.c_int),
        "ta_batch_destroy": ([vp], None),
        "ta_seq_add": ([vp, i32, ctypes.c_char_p, i32, P(i64)], ctypes.c_int),
        "ta_seq_reset": ([vp], ctypes.c_int),
        "ta_align_resident": ([vp, P(Params), P(SeqPair), i32, P(CResult), ctypes.c_char_p, i64, P(i64)],
                              ctypes.c_int),
        "ta_submit_resident": ([vp, P(Params), P(SeqPair), i32, P(ctypes.c_uint64)], ctypes.c_int),
        "ta_serve_config": ([vp, i64, i32], ctypes.c_int),
        "ta_submit": ([vp, P(Params), P(Pair), i32, P(ctypes.c_uint64)], ctypes.c_int),
        "ta_poll": ([vp, ctypes.c_uint64, P(CResult), ctypes.c_char_p, i64, P(i64)], ctypes.c_int),
        "ta_cancel": ([vp, ctypes.c_uint64], ctypes.c_int),
        "ta_matrix": ([i32, i32, i32, P(ctypes.c_int8), P(i32)], ctypes.c_int),
        "ta_encode": ([i32, ctypes.c_char_p, i32, P(ctypes.c_uint8)], ctypes.c_int),
        "ta_consensus": ([vp, P(ctypes.c_char_p), i32, i32, ctypes.c_double, ctypes.c_char_p], ctypes.c_int),
        "ta_msa": ([vp, P(Params), P(ctypes.c_char_p), P(i32), i32, P(i32), P(i32), P(i32)], ctypes.c_int),
        "ta_workload_create": ([i32, ctypes.c_uint64, i32, i32, P(vp)], ctypes.c_int),
        "ta_workload_destroy": ([vp], None),
        "ta_workload_npairs": ([vp], i32),
        "ta_workload_pairs": ([vp, P(Pair)], ctypes.c_int),
        "ta_workload_params": ([vp, P(Params)], ctypes.c_int),
        "ta_workload_nrequests": ([vp], i32),
        "ta_workload_request": ([vp, i32, P(ctypes.c_double), P(i32), P(i32), P(i32)], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _LIB = L
    return L


def _check(rc, what):
    if rc < 0:
        raise RuntimeError("%s: %s (%d)" % (what, lib().ta_strerror(rc).decode(), rc))
    return rc


def matrix(alphabet=PROTEIN, match=2, mismatch=-3):
    out = (ctypes.c_int8 * 1024)()
    k = ctypes.c_int32()
    _check(lib().ta_matrix(alphabet, match, mismatch, out, ctypes.byref(k)), "ta_matrix")
    K = k.value
    return np.frombuffer(bytes(out)[:K * K], dtype=np.int8).reshape(K, K).copy()


def encode(s, alphabet=PROTEIN):
    raw = s.encode() if isinstance(s, str) else bytes(s)
    out = (ctypes.c_uint8 * max(len(raw), 1))()
    _check(lib().ta_encode(alphabet, raw, len(raw), out), "ta_encode")
    return np.frombuffer(bytes(out)[:len(raw)], dtype=np.uint8).copy()


def make_params(mode=GLOBAL, alphabet=PROTEIN, gap_open=-10, gap_extend=-1, match=2, mismatch=-3,
                traceback=1, lanes=LANES_AUTO):
    return Params(mode, alphabet, gap_open, gap_extend, match, mismatch, traceback, lanes)


class _Pairs:
    """Keeps the byte strings alive while the C side holds pointers to them."""

    def __init__(self, pairs):
        self.keep = []
        self.arr = (Pair * max(len(pairs), 1))()
        for k, (a, b) in enumerate(pairs):
            a = a.encode() if isinstance(a, str) else bytes(a)
            b = b.encode() if isinstance(b, str) else bytes(b)
            ba, bb = ctypes.create_string_buffer(a, len(a) + 1), ctypes.create_string_buffer(b, len(b) + 1)
            self.keep += [ba, bb]
            self.arr[k].a = ctypes.cast(ba, ctypes.c_void_p)
            self.arr[k].la = len(a)
            self.arr[k].b = ctypes.cast(bb, ctypes.c_void_p)
            self.arr[k].lb = len(b)
        self.n = len(pairs)


def _collect(res, buf, n, traceback):
    out = []
    raw = buf.raw if buf is not None else b""
    for k in range(n):
        r = res[k]
        cig = raw[r.cigar_off:r.cigar_off + r.cigar_len].decode() if traceback else ""
        out.append(Result(r.score, r.a_begin, r.a_end, r.b_begin, r.b_end, cig, r.flags))
    return out


class Batch:
    def __init__(self, engine, pairs, params):
        self.engine = engine
        self.params = params
        self._pairs = pairs if isinstance(pairs, _Pairs) else _Pairs(pairs)
        self.n = self._pairs.n
        h = ctypes.c_void_p()
        _check(lib().ta_batch_create(engine.ctx, ctypes.byref(params), self._pairs.arr, self.n,
                                     ctypes.byref(h)), "ta_batch_create")
        self.h = h

    def info(self):
        c, t, k = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32()
        _check(lib().ta_batch_info(self.h, ctypes.byref(c), ctypes.byref(t), ctypes.byref(k)), "ta_batch_info")
        return {"cells": c.value, "tb_bytes": t.value, "tasks": k.value}

    def run(self):
        _check(lib().ta_batch_run(self.engine.ctx, self.h), "ta_batch_run")

    def run_plain(self, first=0, count=None):
        count = self.n - first if count is None else count
        _check(lib().ta_batch_run_plain(self.engine.ctx, self.h, first, count), "ta_batch_run_plain")

    def wait(self):
        _check(lib().ta_batch_wait(self.engine.ctx, self.h), "ta_batch_wait")

    def graph(self, on=True):
        """Replay later runs as one captured hipGraph."""
        _check(lib().ta_batch_graph(self.h, 1 if on else 0), "ta_batch_graph")

    def timings(self):
        ms = (ctypes.c_double * 4)()
        k = lib().ta_batch_timings(self.h, ms, 4)
        return list(ms)[:k]

    def results(self):
        res = (CResult * max(self.n, 1))()
        used = ctypes.c_int64()
        cap = 0
        buf = None
        if self.params.traceback:
            rc = lib().ta_batch_results(self.engine.ctx, self.h, res, None, 0, ctypes.byref(used))
            cap = used.value + 1
            buf = ctypes.create_string_buffer(cap)
        _check(lib().ta_batch_results(self.engine.ctx, self.h, res, buf, cap, ctypes.byref(used)),
               "ta_batch_results")
        return _collect(res, buf, self.n, self.params.traceback)

    def close(self):
        if getattr(self, "h", None):
            lib().ta_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Engine:
    """One HIP context (device + stream) of the engine."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        _check(lib().ta_create(ctypes.byref(h), device), "ta_create")
        self.ctx = h
        self._seq_len = {}  # resident handle -> length

    def batch(self, pairs, params):
        return Batch(self, pairs, params)

    def align(self, pairs, params):
        b = Batch(self, pairs, params)
        try:
            b.run()
            b.wait()
            return b.results()
        finally:
            b.close()

    def seq_add(self, letters, alphabet=PROTEIN):
        """A resident sequence (ta_seq_add): encoded once on the device; its handle."""
        s = letters.encode() if isinstance(letters, str) else bytes(letters)
        h = ctypes.c_int64()
        _check(lib().ta_seq_add(self.ctx, alphabet, s, len(s), ctypes.byref(h)), "ta_seq_add")
        self._seq_len[h.value] = len(s)
        return h.value

    def seq_reset(self):
        _check(lib().ta_seq_reset(self.ctx), "ta_seq_reset")
        self._seq_len = {}

    def align_resident(self, handle_pairs, params):
        """ta_align_resident on (handle a, handle b) pairs; a list of Results."""
        n = len(handle_pairs)
        arr = (SeqPair * max(n, 1))(*[SeqPair(a, b) for a, b in handle_pairs])
        res = (CResult * max(n, 1))()
        # a run of length k takes at most k + 1 <= 2k characters
        cap = (sum(2 * (self._seq_len.get(a, 0) + self._seq_len.get(b, 0)) for a, b in handle_pairs) + 16
               if params.traceback else 0)
        buf = ctypes.create_string_buffer(cap) if cap else None
        used = ctypes.c_int64()
        _check(lib().ta_align_resident(self.ctx, ctypes.byref(params), arr, n, res, buf, cap, ctypes.byref(used)),
               "ta_align_resident")
        return _collect(res, buf, n, params.traceback)

    def align_pair(self, a, b, params):
        """One pair through ta_align_pair (the low-latency path); a Result."""
        a = a.encode() if isinstance(a, str) else bytes(a)
        b = b.encode() if isinstance(b, str) else bytes(b)
        res = CResult()
        cap = 2 * (len(a) + len(b)) + 16
        buf = ctypes.create_string_buffer(cap)
        used = ctypes.c_int64()
        _check(lib().ta_align_pair(self.ctx, ctypes.byref(params), a, len(a), b, len(b), ctypes.byref(res), buf, cap,
                                   ctypes.byref(used)), "ta_align_pair")
        return _collect([res], buf, 1, params.traceback)[0]

    def consensus(self, rows, plurality=-1.0):
        """Column consensus of aligned rows (equal lengths; bytes or str) on
        the GPU, as Consensus::calculate computes it; returns bytes."""
        rows = [r.encode("latin-1") if isinstance(r, str) else bytes(r) for r in rows]
        n = len(rows)
        L = len(rows[0]) if rows else 0
        if any(len(r) != L for r in rows):
            raise ValueError("rows of unequal length")
        arr = (ctypes.c_char_p * max(n, 1))(*rows)
        out = ctypes.create_string_buffer(max(L, 1))
        _check(lib().ta_consensus(self.ctx, arr, n, L, float(plurality), out), "ta_consensus")
        return out.raw[:L]

    def msa(self, seqs, params=None):
        """Progressive multiple alignment on the GPU.  Returns (rows, joins):
        the aligned rows ('-' for gaps) in input order and the UPGMA joins."""
        raw = [s.encode("latin-1") if isinstance(s, str) else bytes(s) for s in seqs]
        n = len(raw)
        if params is None:
            params = make_params(mode=GLOBAL)
        lens = (ctypes.c_int32 * max(n, 1))(*[len(s) for s in raw])
        arr = (ctypes.c_char_p * max(n, 1))(*raw)
        total = sum(len(s) for s in raw)
        cols = (ctypes.c_int32 * max(total, 1))()
        joins = (ctypes.c_int32 * max(2 * (n - 1), 1))()
        alen = ctypes.c_int32()
        _check(lib().ta_msa(self.ctx, ctypes.byref(params), arr, lens, n, ctypes.byref(alen), cols, joins),
               "ta_msa")
        rows, k = [], 0
        for s in raw:
            r = bytearray(b"-" * alen.value)
            for ch in s:
                r[cols[k]] = ch
                k += 1
            rows.append(r.decode("latin-1"))
        return rows, [(joins[2 * t], joins[2 * t + 1]) for t in range(n - 1)]

    # autotuning
    def autotune(self, pairs, params, repeats=2):
        """Time candidate launch shapes on these pairs; the fastest is kept for
        this kind of batch.  Returns (Tuning, ms)."""
        ps = pairs if isinstance(pairs, _Pairs) else _Pairs(pairs)
        t, ms = Tuning(), ctypes.c_double()
        _check(lib().ta_autotune(self.ctx, ctypes.byref(params), ps.arr, ps.n, repeats, ctypes.byref(t),
                                 ctypes.byref(ms)), "ta_autotune")
        return t, ms.value

    def tune_set(self, params, bucket, waves_per_cu=0, lanes=LANES_AUTO, wave_tasks=0):
        _check(lib().ta_tune_set(self.ctx, ctypes.byref(params), bucket,
                                 ctypes.byref(Tuning(waves_per_cu, lanes, wave_tasks))), "ta_tune_set")

    def tune_get(self, params, bucket):
        t = Tuning()
        rc = lib().ta_tune_get(self.ctx, ctypes.byref(params), bucket, ctypes.byref(t))
        return t if rc == 0 else None

    # serving
    def serve_config(self, tb_budget_bytes, max_inflight=0):
        _check(lib().ta_serve_config(self.ctx, tb_budget_bytes, max_inflight), "ta_serve_config")

    def submit(self, pairs, params):
        p = pairs if isinstance(pairs, _Pairs) else _Pairs(pairs)
        t = ctypes.c_uint64()
        _check(lib().ta_submit(self.ctx, ctypes.byref(params), p.arr, p.n, ctypes.byref(t)), "ta_submit")
        return t.value, p.n

    def submit_resident(self, handle_pairs, params):
        """ta_submit_resident on (handle a, handle b) pairs; (ticket, n) as submit."""
        n = len(handle_pairs)
        arr = (SeqPair * max(n, 1))(*[SeqPair(a, b) for a, b in handle_pairs])
        t = ctypes.c_uint64()
        _check(lib().ta_submit_resident(self.ctx, ctypes.byref(params), arr, n, ctypes.byref(t)),
               "ta_submit_resident")
        return t.value, n

    def poll(self, ticket, n, traceback=True):
        """None while pending, else the list of Results."""
        res = (CResult * max(n, 1))()
        used = ctypes.c_int64()
        rc = lib().ta_poll(self.ctx, ticket, res, None, 0, ctypes.byref(used))
        if rc == E_PENDING:
            return None
        if rc == 0:  # delivered (no CIGAR text)
            return _collect(res, ctypes.create_string_buffer(1), n, traceback)
        if rc != -3:
            _check(rc, "ta_poll")
        cap = used.value + 1
        buf = ctypes.create_string_buffer(cap)
        _check(lib().ta_poll(self.ctx, ticket, res, buf, cap, ctypes.byref(used)), "ta_poll")
        return _collect(res, buf, n, traceback)

    def cancel(self, ticket):
        return lib().ta_cancel(self.ctx, ticket)

    def close(self):
        if getattr(self, "ctx", None):
            lib().ta_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Workload:
    """Seeded synthetic workload (include/tweakalign_workload.h)."""

    def __init__(self, config, seed=None, size=0, max_len=0):
        h = ctypes.c_void_p()
        _check(lib().ta_workload_create(config, config if seed is None else seed, size, max_len,
                                        ctypes.byref(h)), "ta_workload_create")
        self.h = h
        self.config = config
        n = lib().ta_workload_npairs(h)
        self.npairs = n
        self.arr = (Pair * max(n, 1))()
        _check(lib().ta_workload_pairs(h, self.arr), "ta_workload_pairs")
        self.params = Params()
        _check(lib().ta_workload_params(h, ctypes.byref(self.params)), "ta_workload_params")

    def pairs_struct(self, first=0, count=None):
        """A _Pairs-like view (no copies) usable by Batch / submit."""
        count = self.npairs - first if count is None else count
        v = _Pairs([])
        v.arr = (Pair * max(count, 1))()
        ctypes.memmove(v.arr, ctypes.byref(self.arr, first * ctypes.sizeof(Pair)), count * ctypes.sizeof(Pair))
        v.n = count
        v.keep = [self]
        return v

    def pair_bytes(self, k):
        p = self.arr[k]
        return ctypes.string_at(p.a, p.la), ctypes.string_at(p.b, p.lb)

    def cells(self):
        return sum(self.arr[k].la * self.arr[k].lb for k in range(self.npairs))

    def requests(self):
        out = []
        t, f, c, m = ctypes.c_double(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        for q in range(lib().ta_workload_nrequests(self.h)):
            lib().ta_workload_request(self.h, q, ctypes.byref(t), ctypes.byref(f), ctypes.byref(c), ctypes.byref(m))
            out.append((t.value, f.value, c.value, m.value))
        return out

    def close(self):
        if getattr(self, "h", None):
            lib().ta_workload_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
