"""Synthetic workloads (host side, deterministic from a seed).

band_regions: the batched banded-SW workload -- n candidate regions, each a
(query window, subject window, centre diagonal) triple, per_query regions
sharing one query (as the engine's per-query band batches do); window lengths
lognormal (median 150, sigma 0.6, clamped 20..2000); band half-widths drawn
uniformly from {8, 16, 32, 64}; residues i.i.d. from the Robinson background;
a fraction of regions (0.4) are planted homologues (BLOSUM62-conditional
substitutions at 25-60 % identity, geometric indels at 0.02/residue with
mean length 3).
"""
import ctypes as C

import numpy as np

from . import _ffi
from .matrices import BLOSUM62

JOB_DTYPE = np.dtype([("query", "<u4"), ("subject", "<u4"), ("win_lo", "<u4"), ("win_hi", "<u4"),
                      ("band_lo", "<i4"), ("band_hi", "<i4")])


def _mat(matrix):
    return np.ascontiguousarray(np.asarray(matrix, dtype=np.int32).reshape(576))


def band_regions(seed, n, per_query=32, median=150.0, sigma=0.6, lmin=20, lmax=2000,
                 homolog_frac=0.4, half_widths=(8, 16, 32, 64), matrix=BLOSUM62):
    W = _ffi.wlib()
    hw = np.asarray(half_widths, dtype=np.int32)
    mat = _mat(matrix)
    h = W.wl_band_regions(seed, n, per_query, median, sigma, lmin, lmax, homolog_frac,
                          _ffi.ptr(hw, C.c_int32), len(hw), _ffi.ptr(mat, C.c_int32))
    try:
        a, b, c, d = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_uint64()
        W.wl_sizes(h, C.byref(a), C.byref(b), C.byref(c), C.byref(d))
        qres = np.zeros(max(a.value, 1), np.uint8)
        qoff = np.zeros(b.value + 1, np.uint64)
        sres = np.zeros(max(c.value, 1), np.uint8)
        soff = np.zeros(d.value + 1, np.uint64)
        lo = np.zeros(n, np.int32)
        hi = np.zeros(n, np.int32)
        hom = np.zeros(n, np.uint8)
        W.wl_copy(h, _ffi.ptr(qres, C.c_uint8), _ffi.ptr(qoff, C.c_uint64), _ffi.ptr(sres, C.c_uint8),
                  _ffi.ptr(soff, C.c_uint64), _ffi.ptr(lo, C.c_int32), _ffi.ptr(hi, C.c_int32),
                  _ffi.ptr(hom, C.c_uint8))
    finally:
        W.wl_free(h)
    jobs = np.zeros(n, JOB_DTYPE)
    jobs["query"] = (np.arange(n, dtype=np.uint64) // max(per_query, 1)).astype(np.uint32)
    jobs["subject"] = np.arange(n, dtype=np.uint32)
    jobs["win_lo"] = 0
    jobs["win_hi"] = (soff[1:] - soff[:-1]).astype(np.uint32)
    jobs["band_lo"] = lo
    jobs["band_hi"] = hi
    return dict(qres=qres, qoff=qoff, sres=sres, soff=soff, jobs=jobs, homolog=hom.astype(bool))


def sequences(seed, n, median, sigma, lmin, lmax, matrix=BLOSUM62):
    """n sequences with lognormal lengths: (residues, offsets)."""
    W = _ffi.wlib()
    mat = _mat(matrix)
    h = W.wl_sequences(seed, n, median, sigma, lmin, lmax, _ffi.ptr(mat, C.c_int32))
    try:
        a, b, c, d = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_uint64()
        W.wl_sizes(h, C.byref(a), C.byref(b), C.byref(c), C.byref(d))
        res = np.zeros(max(a.value, 1), np.uint8)
        off = np.zeros(b.value + 1, np.uint64)
        W.wl_copy(h, _ffi.ptr(res, C.c_uint8), _ffi.ptr(off, C.c_uint64), None, None, None, None, None)
    finally:
        W.wl_free(h)
    return res, off


def xdrop_hits(regions, seed, per_region=2):
    """Seed hits over a band_regions batch: for each region, per_region hits
    (query, subject, qp, sp); half on the region's centre diagonal (where
    planted homologues extend far), half uniformly inside the rectangle."""
    rng = np.random.default_rng(seed)
    jobs = regions["jobs"]
    n = len(jobs)
    q = np.repeat(jobs["query"].astype(np.int64), per_region)
    s = np.repeat(jobs["subject"].astype(np.int64), per_region)
    m = (regions["qoff"][q + 1] - regions["qoff"][q]).astype(np.int64)
    sl = (regions["soff"][s + 1] - regions["soff"][s]).astype(np.int64)
    c = np.repeat(((jobs["band_lo"].astype(np.int64) + jobs["band_hi"]) // 2), per_region)
    qp = (rng.random(n * per_region) * m).astype(np.int64)
    on_diag = (np.arange(n * per_region) % 2) == 0
    sp = np.where(on_diag, qp + c, (rng.random(n * per_region) * sl).astype(np.int64))
    sp = np.clip(sp, 0, sl - 1)
    return np.stack([q, s, qp, sp], axis=1).astype(np.uint32)


def _take(h, nsrc):
    W = _ffi.wlib()
    try:
        a, b, c, d = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_uint64()
        W.wl_sizes(h, C.byref(a), C.byref(b), C.byref(c), C.byref(d))
        qres = np.zeros(max(a.value, 1), np.uint8)
        qoff = np.zeros(b.value + 1, np.uint64)
        sres = np.zeros(max(c.value, 1), np.uint8)
        soff = np.zeros(d.value + 1, np.uint64)
        W.wl_copy(h, _ffi.ptr(qres, C.c_uint8), _ffi.ptr(qoff, C.c_uint64), _ffi.ptr(sres, C.c_uint8),
                  _ffi.ptr(soff, C.c_uint64), None, None, None)
        src = np.zeros(max(nsrc(b.value, d.value), 1), np.int32)
        W.wl_source(h, _ffi.ptr(src, C.c_int32))
    finally:
        W.wl_free(h)
    return qres[:a.value], qoff, sres[:c.value], soff, src[:nsrc(b.value, d.value)]


def search_set(seed=0, nq=100, nsub=5000, median=170.0, sigma=0.6, lmin=30, lmax=1000, homolog_frac=0.3,
               id_lo=0.25, id_hi=0.60, matrix=BLOSUM62):
    """The small blastp workload: nq queries, nsub subjects of which a fraction
    are planted homologues of a query (source = query index + 1, 0 = none)."""
    mat = _mat(matrix)
    h = _ffi.wlib().wl_search_set(seed, nq, nsub, median, sigma, lmin, lmax, homolog_frac, id_lo, id_hi,
                                  _ffi.ptr(mat, C.c_int32))
    qres, qoff, sres, soff, src = _take(h, lambda nq_, ns_: ns_)
    return dict(qres=qres, qoff=qoff, sres=sres, soff=soff, source=src)


def family_set(seed=0, nq=11206, median=170.0, sigma=0.6, lmin=30, lmax=1000, fam_extra=3.6, fam_max=30,
               id_lo=0.20, id_hi=0.50, ndecoy=320000, decoy_median=380.0, decoy_sigma=0.6, decoy_max=3000,
               matrix=BLOSUM62):
    """The self-search at scale: nq queries in planted families (family id per
    query), searched against themselves plus ndecoy reversed unrelated
    sequences (the database: queries first)."""
    mat = _mat(matrix)
    h = _ffi.wlib().wl_family_set(seed, nq, median, sigma, lmin, lmax, fam_extra, fam_max, id_lo, id_hi, ndecoy,
                                  decoy_median, decoy_sigma, decoy_max, _ffi.ptr(mat, C.c_int32))
    qres, qoff, sres, soff, src = _take(h, lambda nq_, ns_: nq_)
    return dict(qres=qres, qoff=qoff, sres=sres, soff=soff, family=src)
