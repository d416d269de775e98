"""ctypes loader of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, and only as the checker. The functions restate the reference's scalar
path; see oracle.h for the reference lines and how each one is pinned.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
AVX2_LIB = os.path.join(HERE, "libcpu_avx2.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        u8p, i8p, i32p, i64p, u64p, u32p = (C.POINTER(C.c_uint8), C.POINTER(C.c_int8),
                                           C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                                           C.POINTER(C.c_uint64), C.POINTER(C.c_uint32))
        L.or_sw_reference.argtypes = [u8p, C.c_int64, u8p, C.c_int64, i32p, i8p, C.c_int64, C.c_int64,
                                      i64p, i64p, i64p]
        L.or_banded_score.argtypes = [u8p, C.c_int64, u8p, C.c_int64, i8p, i32p, C.c_int32, C.c_int32,
                                      C.c_int64, C.c_int64, i32p, i64p, i64p, u64p]
        L.or_banded_score_ij.argtypes = [u8p, C.c_int64, u8p, C.c_int64, i8p, i32p, C.c_int32,
                                         C.c_int32, C.c_int64, C.c_int64, i32p, i64p, i64p]
        L.or_banded_many.argtypes = [u8p, u8p, i8p, u64p, u32p, u64p, u32p, i32p, i32p, C.c_int64,
                                     i32p, C.c_int32, C.c_int32, i32p, u32p, u32p, u64p, C.c_int]
        L.or_xdrop.argtypes = [u8p, C.c_int64, i8p, u8p, C.c_int64, i32p, C.c_int64, C.c_int64,
                               C.c_int32, i32p, i64p, i64p, u64p]
        L.or_xdrop_scalar_twin.argtypes = [u8p, C.c_int64, i8p, u8p, C.c_int64, i32p, C.c_int64,
                                           C.c_int64, C.c_int32, i32p, i64p, i64p]
        L.or_best_ungapped_on_diagonal.argtypes = [u8p, C.c_int64, u8p, C.c_int64, C.c_int64, i32p, i8p]
        L.or_best_ungapped_on_diagonal.restype = C.c_int64
        L.or_composition_bias.argtypes = [u8p, C.c_int64, i32p, i8p]
        L.or_band_align.argtypes = [u8p, C.c_int64, u8p, C.c_int64, i8p, i32p, C.c_int32, C.c_int32,
                                    C.c_int64, C.c_int64, C.c_int32, C.c_uint64, i32p, i64p, i64p, i64p,
                                    i64p, u64p, u8p, u32p, i64p]
        L.or_band_align.restype = C.c_int
        _lib = L
    return _lib


def _u8(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint8))


def _p(a, t):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(t))


def _mat(m):
    return np.ascontiguousarray(np.asarray(m, dtype=np.int32).reshape(576))


def _bias(b):
    return None if b is None else np.ascontiguousarray(np.asarray(b, dtype=np.int8))


def sw_reference(q, s, matrix, bias, go, ge):
    q, s, mat, b = _u8(q), _u8(s), _mat(matrix), _bias(bias)
    sc, qe, se = C.c_int64(), C.c_int64(), C.c_int64()
    lib().or_sw_reference(_p(q, C.c_uint8), len(q), _p(s, C.c_uint8), len(s), _p(mat, C.c_int32),
                          _p(b, C.c_int8), go, ge, C.byref(sc), C.byref(qe), C.byref(se))
    return sc.value, qe.value, se.value


def banded_score(q, s, bias, matrix, go, ge, lo, hi):
    q, s, mat, b = _u8(q), _u8(s), _mat(matrix), _bias(bias)
    sc, qe, se, ce = C.c_int32(), C.c_int64(), C.c_int64(), C.c_uint64()
    lib().or_banded_score(_p(q, C.c_uint8), len(q), _p(s, C.c_uint8), len(s), _p(b, C.c_int8),
                          _p(mat, C.c_int32), go, ge, lo, hi, C.byref(sc), C.byref(qe), C.byref(se),
                          C.byref(ce))
    return sc.value, qe.value, se.value, ce.value


def banded_score_ij(q, s, bias, matrix, go, ge, lo, hi):
    q, s, mat, b = _u8(q), _u8(s), _mat(matrix), _bias(bias)
    sc, qe, se = C.c_int32(), C.c_int64(), C.c_int64()
    lib().or_banded_score_ij(_p(q, C.c_uint8), len(q), _p(s, C.c_uint8), len(s), _p(b, C.c_int8),
                             _p(mat, C.c_int32), go, ge, lo, hi, C.byref(sc), C.byref(qe), C.byref(se))
    return sc.value, qe.value, se.value


def banded_many(qres, qoff, sres, soff, jobs, matrix, go, ge, qbias=None, nthreads=8):
    """Oracle for a band batch given as flat arrays (same layout as the GPU
    batch). Returns (score, q_end, s_end, cells)."""
    qres, sres = _u8(qres), _u8(sres)
    qoff = np.asarray(qoff, np.uint64)
    soff = np.asarray(soff, np.uint64)
    jq = np.ascontiguousarray(qoff[jobs["query"]])
    jql = np.ascontiguousarray((qoff[jobs["query"] + 1] - qoff[jobs["query"]]).astype(np.uint32))
    js = np.ascontiguousarray(soff[jobs["subject"]] + jobs["win_lo"].astype(np.uint64))
    jsl = np.ascontiguousarray((jobs["win_hi"] - jobs["win_lo"]).astype(np.uint32))
    lo = np.ascontiguousarray(jobs["band_lo"].astype(np.int32))
    hi = np.ascontiguousarray(jobs["band_hi"].astype(np.int32))
    n = len(jobs)
    score = np.zeros(n, np.int32)
    qe = np.zeros(n, np.uint32)
    se = np.zeros(n, np.uint32)
    ce = np.zeros(n, np.uint64)
    b = _bias(qbias)
    mat = _mat(matrix)
    lib().or_banded_many(_p(qres, C.c_uint8), _p(sres, C.c_uint8), _p(b, C.c_int8), _p(jq, C.c_uint64),
                         _p(jql, C.c_uint32), _p(js, C.c_uint64), _p(jsl, C.c_uint32), _p(lo, C.c_int32),
                         _p(hi, C.c_int32), n, _p(mat, C.c_int32), go, ge, _p(score, C.c_int32),
                         _p(qe, C.c_uint32), _p(se, C.c_uint32), _p(ce, C.c_uint64), nthreads)
    return score, qe, se, ce


_avx2 = None


def avx2_lib():
    global _avx2
    if _avx2 is None:
        if not os.path.exists(AVX2_LIB):
            subprocess.run(["make", "-s", "-C", HERE, "avx2"], check=True)
        L = C.CDLL(AVX2_LIB)
        u8p, i8p, i32p, u64p, u32p = (C.POINTER(C.c_uint8), C.POINTER(C.c_int8), C.POINTER(C.c_int32),
                                      C.POINTER(C.c_uint64), C.POINTER(C.c_uint32))
        L.cpu_banded_many_avx2.argtypes = [u8p, u8p, i8p, u64p, u32p, u64p, u32p, i32p, i32p, C.c_int64,
                                           i32p, C.c_int32, C.c_int32, i32p, C.c_int]
        L.cpu_banded_many_avx2.restype = None
        L.cpu_score_many_avx2.argtypes = [u8p, u64p, C.c_uint32, i8p, u8p, u64p, C.c_uint32, i32p, C.c_int32,
                                          C.c_int32, i32p, C.c_int]
        L.cpu_score_many_avx2.restype = None
        L.cpu_xdrop_avx2.argtypes = [u8p, u64p, i8p, u8p, u64p, u32p, C.c_uint64, i32p, C.c_int32, i32p, u32p,
                                     u32p, u64p, C.c_int]
        L.cpu_xdrop_avx2.restype = None
        _avx2 = L
    return _avx2


def banded_many_avx2(qres, qoff, sres, soff, jobs, matrix, go, ge, qbias=None, nthreads=8):
    """Scores of a band batch by the AVX2 restatement of score_banded_many
    (jobs of one query must be consecutive to share u8 batches)."""
    qres, sres = _u8(qres), _u8(sres)
    qoff = np.asarray(qoff, np.uint64)
    soff = np.asarray(soff, np.uint64)
    jq = np.ascontiguousarray(qoff[jobs["query"]])
    jql = np.ascontiguousarray((qoff[jobs["query"] + 1] - qoff[jobs["query"]]).astype(np.uint32))
    js = np.ascontiguousarray(soff[jobs["subject"]] + jobs["win_lo"].astype(np.uint64))
    jsl = np.ascontiguousarray((jobs["win_hi"] - jobs["win_lo"]).astype(np.uint32))
    lo = np.ascontiguousarray(jobs["band_lo"].astype(np.int32))
    hi = np.ascontiguousarray(jobs["band_hi"].astype(np.int32))
    n = len(jobs)
    score = np.zeros(n, np.int32)
    b = _bias(qbias)
    mat = _mat(matrix)
    avx2_lib().cpu_banded_many_avx2(_p(qres, C.c_uint8), _p(sres, C.c_uint8), _p(b, C.c_int8),
                                    _p(jq, C.c_uint64), _p(jql, C.c_uint32), _p(js, C.c_uint64),
                                    _p(jsl, C.c_uint32), _p(lo, C.c_int32), _p(hi, C.c_int32), n,
                                    _p(mat, C.c_int32), go, ge, _p(score, C.c_int32), nthreads)
    return score


def score_many_avx2(qres, qoff, sres, soff, matrix, go, ge, qbias=None, nthreads=8):
    """Full-matrix SW scores [nq, ns] by the AVX2 restatement of score_many
    (sw_simd.rs:95-153: 8-bit pass, 16-bit re-score, int32 repair), queries
    spread over nthreads threads."""
    qres, sres = _u8(qres), _u8(sres)
    qoff = np.ascontiguousarray(qoff, np.uint64)
    soff = np.ascontiguousarray(soff, np.uint64)
    nq, ns = len(qoff) - 1, len(soff) - 1
    out = np.zeros((max(nq, 1), max(ns, 1)), np.int32)
    b = _bias(qbias)
    avx2_lib().cpu_score_many_avx2(_p(qres, C.c_uint8), _p(qoff, C.c_uint64), nq, _p(b, C.c_int8),
                                   _p(sres, C.c_uint8), _p(soff, C.c_uint64), ns, _p(_mat(matrix), C.c_int32), go, ge,
                                   _p(out, C.c_int32), nthreads)
    return out[:nq, :ns]


def xdrop_many_avx2(qres, qoff, sres, soff, hits, matrix, x, qbias=None, nthreads=8):
    """X-drop of every hit row (query, subject, q_pos, s_pos) by the AVX2
    restatement of xdrop_avx2 (scan.rs:934-1031), hits spread over nthreads
    threads: (score, q_lo, q_hi, steps) arrays, equal to or_xdrop's."""
    qres, sres = _u8(qres), _u8(sres)
    qoff = np.ascontiguousarray(qoff, np.uint64)
    soff = np.ascontiguousarray(soff, np.uint64)
    hits = np.ascontiguousarray(hits, np.uint32).reshape(-1, 4)
    n = len(hits)
    score = np.zeros(max(n, 1), np.int32)
    lo = np.zeros(max(n, 1), np.uint32)
    hi = np.zeros(max(n, 1), np.uint32)
    steps = np.zeros(max(n, 1), np.uint64)
    b = _bias(qbias)
    avx2_lib().cpu_xdrop_avx2(_p(qres, C.c_uint8), _p(qoff, C.c_uint64), _p(b, C.c_int8), _p(sres, C.c_uint8),
                              _p(soff, C.c_uint64), _p(hits, C.c_uint32), n, _p(_mat(matrix), C.c_int32), x,
                              _p(score, C.c_int32), _p(lo, C.c_uint32), _p(hi, C.c_uint32),
                              _p(steps, C.c_uint64), nthreads)
    return score[:n], lo[:n], hi[:n], steps[:n]


def xdrop(q, s, bias, matrix, qp, sp, x):
    q, s, mat, b = _u8(q), _u8(s), _mat(matrix), _bias(bias)
    sc, lo_, hi_, st = C.c_int32(), C.c_int64(), C.c_int64(), C.c_uint64()
    lib().or_xdrop(_p(q, C.c_uint8), len(q), _p(b, C.c_int8), _p(s, C.c_uint8), len(s),
                   _p(mat, C.c_int32), qp, sp, x, C.byref(sc), C.byref(lo_), C.byref(hi_), C.byref(st))
    return sc.value, lo_.value, hi_.value, st.value


def xdrop_scalar_twin(q, s, bias, matrix, qp, sp, x):
    q, s, mat, b = _u8(q), _u8(s), _mat(matrix), _bias(bias)
    sc, lo_, hi_ = C.c_int32(), C.c_int64(), C.c_int64()
    lib().or_xdrop_scalar_twin(_p(q, C.c_uint8), len(q), _p(b, C.c_int8), _p(s, C.c_uint8), len(s),
                               _p(mat, C.c_int32), qp, sp, x, C.byref(sc), C.byref(lo_), C.byref(hi_))
    return sc.value, lo_.value, hi_.value


def best_ungapped_on_diagonal(q, s, d, matrix, bias):
    q, s, mat, b = _u8(q), _u8(s), _mat(matrix), _bias(bias)
    return lib().or_best_ungapped_on_diagonal(_p(q, C.c_uint8), len(q), _p(s, C.c_uint8), len(s), d,
                                              _p(mat, C.c_int32), _p(b, C.c_int8))


def composition_bias(q, matrix):
    q, mat = _u8(q), _mat(matrix)
    out = np.zeros(len(q), np.int8)
    lib().or_composition_bias(_p(q, C.c_uint8), len(q), _p(mat, C.c_int32), _p(out, C.c_int8))
    return out


def band_align(q, s, matrix, bias, go, ge, band_lo, band_hi, stop_at=2**31 - 1, max_dir_bytes=2**63):
    """banded_align_until (banded.rs:472-521): None when there is no
    alignment, "cap" when the direction matrix exceeds max_dir_bytes, else
    (score, q_start, q_end, s_start, s_end, ops, cells) with ops a list of
    (op, length), op in b"MID"."""
    q, s, mat, b = _u8(q), _u8(s), _mat(matrix), _bias(bias)
    cap = len(q) + len(s) + 1
    codes = np.zeros(cap, np.uint8)
    lens = np.zeros(cap, np.uint32)
    sc = C.c_int32()
    qs, qe, ss, se, nops = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
    cells = C.c_uint64()
    r = lib().or_band_align(_p(q, C.c_uint8), len(q), _p(s, C.c_uint8), len(s), _p(b, C.c_int8),
                            _p(mat, C.c_int32), go, ge, band_lo, band_hi, int(stop_at), int(max_dir_bytes),
                            C.byref(sc), C.byref(qs), C.byref(qe), C.byref(ss), C.byref(se), C.byref(cells),
                            _p(codes, C.c_uint8), _p(lens, C.c_uint32), C.byref(nops))
    if r == 1:
        return None
    if r == 2:
        return "cap"
    ops = [(bytes([int(codes[k])]), int(lens[k])) for k in range(nops.value)]
    return (sc.value, qs.value, qe.value, ss.value, se.value, ops, cells.value)
