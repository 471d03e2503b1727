"""ctypes binding of the CPU oracle (liboracle.so) -- test infrastructure only.

Imported by tests/, __graft_entry__.smoke() (as the checker) and bench.py's
cpu_baseline leg (as the timed CPU path).  Never imported by tweakseq_amd/.
"""
import ctypes
import json
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class Params(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in
                ("mode", "alphabet", "gap_open", "gap_extend", "match", "mismatch",
                 "traceback", "lanes")]


class Result(ctypes.Structure):
    _fields_ = [("score", ctypes.c_int32), ("a_begin", ctypes.c_int32),
                ("a_end", ctypes.c_int32), ("b_begin", ctypes.c_int32),
                ("b_end", ctypes.c_int32), ("cigar_off", ctypes.c_int64),
                ("cigar_len", ctypes.c_int32), ("flags", ctypes.c_int32)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = ctypes.CDLL(path)
        P = ctypes.POINTER
        L.or_encode.argtypes = [P(ctypes.c_int32), ctypes.c_int32, ctypes.c_char_p,
                                ctypes.c_int32, P(ctypes.c_uint8)]
        L.or_align.argtypes = [P(Params), P(ctypes.c_int8), ctypes.c_int32,
                               P(ctypes.c_uint8), ctypes.c_int32, P(ctypes.c_uint8),
                               ctypes.c_int32, P(Result), ctypes.c_char_p, ctypes.c_int64]
        L.or_align_batch.argtypes = [P(Params), P(ctypes.c_int8), ctypes.c_int32,
                                     P(ctypes.c_uint8), P(ctypes.c_int64), P(ctypes.c_int32),
                                     P(ctypes.c_int64), P(ctypes.c_int32), ctypes.c_int32,
                                     P(Result), ctypes.c_char_p, P(ctypes.c_int64),
                                     P(ctypes.c_int64), ctypes.c_int32]
        L.or_rescore.argtypes = [P(Params), P(ctypes.c_int8), ctypes.c_int32,
                                 P(ctypes.c_uint8), P(ctypes.c_uint8), P(Result),
                                 ctypes.c_char_p, P(ctypes.c_int32)]
        L.or_time_align.argtypes = [P(Params), P(ctypes.c_int8), ctypes.c_int32, P(ctypes.c_uint8), ctypes.c_int32,
                                    P(ctypes.c_uint8), ctypes.c_int32, ctypes.c_int32, ctypes.c_char_p,
                                    ctypes.c_int64, P(ctypes.c_double)]
        L.or_consensus.argtypes = [P(ctypes.c_int8), P(ctypes.c_int32), ctypes.c_char_p, ctypes.c_int32,
                                   ctypes.c_int32, ctypes.c_double, ctypes.c_char_p]
        L.or_msa.argtypes = [P(ctypes.c_int8), ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                             P(ctypes.c_uint8), P(ctypes.c_int64), P(ctypes.c_int32), ctypes.c_int32,
                             P(ctypes.c_int32), P(ctypes.c_int32), P(ctypes.c_int32)]
        _LIB = L
    return _LIB


def blosum62_fixture():
    """The reference's BLOSUM62 table + letter map, from tests/golden/blosum62.json."""
    path = os.path.join(_HERE, "..", "tests", "golden", "blosum62.json")
    with open(path) as f:
        return json.load(f)


_PROTEIN = None


def protein_matrix():
    global _PROTEIN
    if _PROTEIN is None:
        d = blosum62_fixture()
        _PROTEIN = (np.array(d["matrix"], dtype=np.int8), np.array(d["map"], dtype=np.int32))
    return _PROTEIN[0].copy(), _PROTEIN[1].copy()


def dna_matrix(match, mismatch):
    m = np.full((4, 4), mismatch, dtype=np.int8)
    np.fill_diagonal(m, match)
    return m


def params(mode=0, alphabet=0, gap_open=-10, gap_extend=-1, match=2, mismatch=-3,
           traceback=1):
    return Params(mode, alphabet, gap_open, gap_extend, match, mismatch, traceback, 0)


def _u8(a):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def encode(s, alphabet=0):
    _, cmap = protein_matrix()
    raw = s.encode() if isinstance(s, str) else bytes(s)
    out = np.zeros(max(len(raw), 1), dtype=np.uint8)
    rc = lib().or_encode(cmap.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), alphabet,
                         raw, len(raw), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    if rc:
        raise ValueError("residue outside the alphabet")
    return out[:len(raw)]


def matrix_for(p):
    if p.alphabet == 0:
        return protein_matrix()[0]
    return dna_matrix(p.match, p.mismatch)


def align(p, a, b, smat=None):
    """a, b: uint8 code arrays.  Returns (score, a_begin, a_end, b_begin, b_end, cigar)."""
    if smat is None:
        smat = matrix_for(p)
    smat = np.ascontiguousarray(smat, dtype=np.int8)
    a, pa = _u8(a)
    b, pb = _u8(b)
    res = Result()
    cap = 2 * (len(a) + len(b)) + 16
    buf = ctypes.create_string_buffer(cap)
    rc = lib().or_align(ctypes.byref(p), smat.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)),
                        smat.shape[0], pa, len(a), pb, len(b), ctypes.byref(res), buf, cap)
    if rc:
        raise RuntimeError("or_align failed: %d" % rc)
    cig = buf.raw[:res.cigar_len].decode() if p.traceback else ""
    return (res.score, res.a_begin, res.a_end, res.b_begin, res.b_end, cig)


def align_batch(p, codes, a_off, la, b_off, lb, nthreads=0, with_cigar=True, smat=None):
    """Many pairs over OpenMP.  Returns (results structured array, cigars list or None)."""
    if smat is None:
        smat = matrix_for(p)
    smat = np.ascontiguousarray(smat, dtype=np.int8)
    codes, pc = _u8(codes)
    a_off = np.ascontiguousarray(a_off, dtype=np.int64)
    b_off = np.ascontiguousarray(b_off, dtype=np.int64)
    la = np.ascontiguousarray(la, dtype=np.int32)
    lb = np.ascontiguousarray(lb, dtype=np.int32)
    n = len(la)
    res = (Result * n)()
    P = ctypes.POINTER
    if with_cigar and p.traceback:
        cap = (2 * (la.astype(np.int64) + lb) + 16)
        off = np.zeros(n, dtype=np.int64)
        off[1:] = np.cumsum(cap)[:-1]
        buf = ctypes.create_string_buffer(int(cap.sum()) + 1)
        cptr, offp, capp = buf, off.ctypes.data_as(P(ctypes.c_int64)), cap.ctypes.data_as(P(ctypes.c_int64))
    else:
        buf, cptr, offp, capp = None, None, None, None
    rc = lib().or_align_batch(ctypes.byref(p), smat.ctypes.data_as(P(ctypes.c_int8)), smat.shape[0],
                              pc, a_off.ctypes.data_as(P(ctypes.c_int64)), la.ctypes.data_as(P(ctypes.c_int32)),
                              b_off.ctypes.data_as(P(ctypes.c_int64)), lb.ctypes.data_as(P(ctypes.c_int32)),
                              n, res, cptr, offp, capp, nthreads)
    if rc:
        raise RuntimeError("or_align_batch failed: %d" % rc)
    cigars = None
    if buf is not None:
        raw = buf.raw
        cigars = [raw[res[k].cigar_off:res[k].cigar_off + res[k].cigar_len].decode() for k in range(n)]
    return res, cigars


def rescore(p, a, b, result_tuple, smat=None):
    if smat is None:
        smat = matrix_for(p)
    smat = np.ascontiguousarray(smat, dtype=np.int8)
    a, pa = _u8(a)
    b, pb = _u8(b)
    score, ab, ae, bb, be, cig = result_tuple
    r = Result(score, ab, ae, bb, be, 0, len(cig), 0)
    out = ctypes.c_int32()
    rc = lib().or_rescore(ctypes.byref(p), smat.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)),
                          smat.shape[0], pa, pb, ctypes.byref(r), cig.encode(), ctypes.byref(out))
    if rc:
        raise ValueError("CIGAR does not consume the reported ranges")
    return out.value


def consensus(rows, plurality=-1.0):
    """Consensus::calculate on equal-length rows (bytes or str); returns bytes."""
    rows = [r.encode() if isinstance(r, str) else bytes(r) for r in rows]
    n, L = len(rows), len(rows[0]) if rows else 0
    assert all(len(r) == L for r in rows)
    smat, lmap = protein_matrix()
    out = ctypes.create_string_buffer(max(L, 1))
    rc = lib().or_consensus(smat.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)),
                            lmap.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), b"".join(rows), n, L,
                            float(plurality), out)
    if rc:
        raise RuntimeError("or_consensus: %d" % rc)
    return out.raw[:L]


def msa(seqs, smat=None, gap_open=-10, gap_extend=-1):
    """Progressive MSA (or_msa) of code arrays; returns (alen, cols, joins):
    cols[s] = column of each residue of sequence s."""
    if smat is None:
        smat, _ = protein_matrix()
    smat = np.ascontiguousarray(smat, dtype=np.int8)
    K = smat.shape[0]
    seqs = [np.ascontiguousarray(x, dtype=np.uint8) for x in seqs]
    n = len(seqs)
    lens = np.array([len(x) for x in seqs], dtype=np.int32)
    off = np.zeros(n, dtype=np.int64)
    if n > 1:
        off[1:] = np.cumsum(lens[:-1])
    codes = np.concatenate(seqs + [np.zeros(1, dtype=np.uint8)]).astype(np.uint8)
    col = np.zeros(max(int(lens.sum()), 1), dtype=np.int32)
    joins = np.zeros(max(2 * (n - 1), 1), dtype=np.int32)
    alen = ctypes.c_int32()
    P = ctypes.POINTER
    rc = lib().or_msa(smat.ctypes.data_as(P(ctypes.c_int8)), K, gap_open, gap_extend,
                      codes.ctypes.data_as(P(ctypes.c_uint8)), off.ctypes.data_as(P(ctypes.c_int64)),
                      lens.ctypes.data_as(P(ctypes.c_int32)), n, joins.ctypes.data_as(P(ctypes.c_int32)),
                      ctypes.byref(alen), col.ctypes.data_as(P(ctypes.c_int32)))
    if rc:
        raise RuntimeError("or_msa: %d" % rc)
    cols = [col[off[s]:off[s] + lens[s]].copy() for s in range(n)]
    return alen.value, cols, [tuple(joins[2 * t:2 * t + 2]) for t in range(n - 1)]


def msa_rows(seqs_text, alen, cols):
    """Aligned rows ('-' for gaps) from residue columns."""
    rows = []
    for s, c in zip(seqs_text, cols):
        r = ["-"] * alen
        for k, q in enumerate(c):
            r[q] = s[k]
        rows.append("".join(r))
    return rows


def align_pairs(p, pairs, nthreads=0, alphabet=None):
    """(score, a_begin, a_end, b_begin, b_end, cigar) of every (a, b) letter
    pair (bytes or str), through or_align_batch on `nthreads` OpenMP threads
    (<= 0: all).  The CIGAR is "" when p.traceback is 0."""
    alphabet = p.alphabet if alphabet is None else alphabet
    seqs, a_off, la, b_off, lb = [], [], [], [], []
    pos = 0
    for a, b in pairs:
        ea, eb = encode(a, alphabet), encode(b, alphabet)
        seqs += [ea, eb]
        a_off.append(pos); la.append(len(ea)); pos += len(ea)
        b_off.append(pos); lb.append(len(eb)); pos += len(eb)
    codes = np.concatenate(seqs + [np.zeros(1, np.uint8)]) if seqs else np.zeros(1, np.uint8)
    res, cigars = align_batch(p, codes, a_off, la, b_off, lb, nthreads=nthreads)
    return [(res[k].score, res[k].a_begin, res[k].a_end, res[k].b_begin, res[k].b_end,
             cigars[k] if cigars is not None else "") for k in range(len(pairs))]


def time_align(p, a, b, repeats=1000, smat=None):
    """Median seconds of `repeats` or_align calls on codes a, b, timed in C."""
    if smat is None:
        smat = matrix_for(p)
    smat = np.ascontiguousarray(smat, dtype=np.int8)
    a, pa = _u8(a)
    b, pb = _u8(b)
    cap = 2 * (len(a) + len(b)) + 16
    buf = ctypes.create_string_buffer(cap)
    med = ctypes.c_double()
    rc = lib().or_time_align(ctypes.byref(p), smat.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)), smat.shape[0],
                             pa, len(a), pb, len(b), repeats, buf, cap, ctypes.byref(med))
    if rc:
        raise RuntimeError("or_time_align failed: %d" % rc)
    return med.value
