"""MI355X (gfx950) candidate-scoring path of SABER, behind SABER's own scoring
interfaces.

Host mirror of the reference's band-scoring and backend API (names, argument
meaning and error behaviour follow the Rust originals):

  Band, BandJob, BandScore, Scheme, GapCosts   src/algorithm/banded.rs:21-99,
                                               band_simd8.rs:29-33, scoring/stats.rs:59-103
  banded_score(q, s, bias, scheme, band)       src/algorithm/banded.rs:408-449
  score_banded_many(query, bias, jobs, scheme) src/algorithm/band_simd8.rs:78-185
  score_many(query, bias, subjects, matrix, gaps)
                                               src/algorithm/sw_simd.rs:95-153
  HipBackend.score_windows(...)                trait ScoreBackend, src/gpu/mod.rs:45-59

Every call runs the HIP kernels of libsaber_hip.so on the GPU; there is no
CPU path here. Inputs are alignment codes 0..23 (codes >= 24 score as X).
"""
from dataclasses import dataclass
import atexit
import time
import ctypes as C
import os
import weakref

import numpy as np

from . import _ffi
from ._ffi import HipError
from .matrices import ALPHABET, BLOSUM45, BLOSUM62, PAM30

# One hit for Context.traceback (saber_hip_traceback_job).
TRACEBACK_JOB = np.dtype([("query", "<u4"), ("subject", "<u4"), ("win_lo", "<u4"), ("win_hi", "<u4"),
                          ("d_min", "<i8"), ("d_max", "<i8"), ("best", "<i4"), ("want", "<i4"),
                          ("has_limit", "<i4"), ("reserved", "<i4"), ("limit_lo", "<i8"), ("limit_hi", "<i8")])

# Live batches and contexts, released before interpreter exit: the HIP runtime
# tears itself down after atexit handlers, and a stream destroyed after that
# aborts the process.
_live_batches = weakref.WeakSet()
_live_contexts = weakref.WeakSet()


@atexit.register
def _release_all():
    for b in list(_live_batches):
        b.close()
    for c in list(_live_contexts):
        c.close()

__all__ = [
    "ALPHABET", "BLOSUM45", "BLOSUM62", "PAM30", "Band", "BandJob", "BandScore", "GapCosts",
    "Scheme", "Context", "HipBackend", "HipError", "banded_score", "score_banded_many",
    "score_many", "encode", "default_context",
]

_ENCODE = np.full(256, 22, dtype=np.uint8)
for _i, _c in enumerate(ALPHABET):
    _ENCODE[ord(_c)] = _i
    if "A" <= _c <= "Z":
        _ENCODE[ord(_c.lower())] = _i


def encode(ascii_seq):
    """ASCII protein -> alignment codes (src/alphabet.rs:48-78: unknown -> X)."""
    if isinstance(ascii_seq, str):
        ascii_seq = ascii_seq.encode()
    return _ENCODE[np.frombuffer(bytes(ascii_seq), dtype=np.uint8)]


@dataclass(frozen=True)
class GapCosts:
    """SABER gap convention: a gap of length k costs open + (k-1)*extend."""
    open: int
    extend: int


@dataclass(frozen=True)
class Band:
    """Inclusive diagonal band lo <= j - i <= hi (src/algorithm/banded.rs:30-54)."""
    lo: int
    hi: int

    @staticmethod
    def full(m, n):
        return Band(-int(m), int(n))

    def width(self):
        return max(self.hi - self.lo + 1, 0)

    def clip(self, m, n):
        return Band(max(self.lo, -(int(m) - 1)), min(self.hi, int(n) - 1))


@dataclass(frozen=True)
class BandScore:
    score: int
    q_end: int
    s_end: int
    cells: int


@dataclass(frozen=True)
class BandJob:
    subject: object
    band: Band


@dataclass(frozen=True)
class Scheme:
    matrix: tuple
    gap_open: int
    gap_extend: int

    def ffi(self):
        s = _ffi.Scheme()
        flat = np.asarray(self.matrix, dtype=np.int32).reshape(576)
        s.matrix[:] = [int(v) for v in flat]
        s.gap_open = int(self.gap_open)
        s.gap_extend = int(self.gap_extend)
        return s


def _codes(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint8))


def _seqset(seqs, bias=None):
    """Concatenate sequences (list of code arrays) into an ffi SeqSet."""
    lens = np.fromiter((len(s) for s in seqs), dtype=np.uint64, count=len(seqs))
    off = np.zeros(len(seqs) + 1, dtype=np.uint64)
    np.cumsum(lens, out=off[1:])
    res = np.concatenate([_codes(s) for s in seqs]) if len(seqs) and off[-1] else np.zeros(1, np.uint8)
    res = np.ascontiguousarray(res, dtype=np.uint8)
    b = None
    if bias is not None:
        b = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.int8) for x in bias]) if off[-1]
                                 else np.zeros(1, np.int8))
    ss = _ffi.SeqSet(_ffi.ptr(res, C.c_uint8), _ffi.ptr(off, C.c_uint64), len(seqs),
                     _ffi.ptr(b, C.c_int8))
    return ss, (res, off, b)


class BandBatch:
    """Device-resident band batch (saber_hip_band_batch_*)."""

    def __init__(self, ctx, handle, njobs, keep, ends=True):
        self.ctx, self.h, self.njobs, self._keep, self.ends = ctx, handle, njobs, keep, ends
        _live_batches.add(self)
        ctx._children.add(self)

    def run(self):
        _ffi.check(_ffi.lib().saber_hip_band_batch_run(self.h))

    def wait(self):
        _ffi.check(_ffi.lib().saber_hip_band_batch_wait(self.h))

    def read(self):
        """(score, q_end, s_end, cells); a scores-only batch returns
        (score, None, None, cells)."""
        n = self.njobs
        score = np.zeros(n, np.int32)
        qe = np.zeros(n, np.uint32) if self.ends else None
        se = np.zeros(n, np.uint32) if self.ends else None
        cells = np.zeros(n, np.uint64)
        _ffi.check(_ffi.lib().saber_hip_band_batch_read(
            self.h, _ffi.ptr(score, C.c_int32), _ffi.ptr(qe, C.c_uint32), _ffi.ptr(se, C.c_uint32),
            _ffi.ptr(cells, C.c_uint64)))
        return score, qe, se, cells

    def time(self, warmup=1, iters=5):
        """(ms per run, ms of the packed kernel's launch per run,
        {variant index: cells per run}, cells per run)."""
        ms = C.c_double()
        kms = C.c_double()
        cap = 256
        kc = (C.c_uint64 * cap)()
        nv = C.c_int(cap)
        cells = C.c_uint64()
        _ffi.check(_ffi.lib().saber_hip_band_batch_time(self.h, warmup, iters, C.byref(ms), C.byref(kms), kc,
                                                          C.byref(nv), C.byref(cells)))
        per = {v: kc[v] for v in range(min(nv.value, cap)) if kc[v] > 0}
        return ms.value, kms.value, per, cells.value

    def plan(self):
        """The batch's plan: (pairs [npairs, 2] job ids (0xFFFFFFFF = none),
        blocks [nchunks, 4] (shape, first pair, pairs, flags) in dispatch
        order, geometry [njobs, 5] (band lo, width, first row, rows,
        transposed))."""
        npairs, nchunks = C.c_uint64(), C.c_uint64()
        lib = _ffi.lib()
        _ffi.check(lib.saber_hip_band_batch_plan(self.h, C.byref(npairs), C.byref(nchunks), None, None, None))
        pairs = np.zeros((npairs.value, 2), np.uint32)
        blocks = np.zeros((nchunks.value, 4), np.uint32)
        geom = np.zeros((self.njobs, 5), np.int32)
        _ffi.check(lib.saber_hip_band_batch_plan(
            self.h, C.byref(npairs), C.byref(nchunks), pairs.ctypes.data_as(C.POINTER(C.c_uint32)),
            blocks.ctypes.data_as(C.POINTER(C.c_uint32)), geom.ctypes.data_as(C.POINTER(C.c_int32))))
        return pairs, blocks, geom

    def close(self):
        if self.h:
            _ffi.lib().saber_hip_band_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FullBatch:
    """Device-resident exhaustive batch (saber_hip_full_batch_*)."""

    def __init__(self, ctx, handle, nq, ns, keep, ends=True):
        self.ctx, self.h, self.nq, self.ns, self._keep, self.ends = ctx, handle, nq, ns, keep, ends
        _live_batches.add(self)
        ctx._children.add(self)

    def run(self):
        _ffi.check(_ffi.lib().saber_hip_full_batch_run(self.h))

    def read(self):
        """(score, q_end, s_end) of shape [queries, subjects]; a scores-only
        batch returns (score, None, None)."""
        shape = (self.nq, self.ns)
        score = np.zeros(shape, np.int32)
        qe = np.zeros(shape, np.uint32) if self.ends else None
        se = np.zeros(shape, np.uint32) if self.ends else None
        _ffi.check(_ffi.lib().saber_hip_full_batch_read(
            self.h, _ffi.ptr(score, C.c_int32), _ffi.ptr(qe, C.c_uint32), _ffi.ptr(se, C.c_uint32)))
        return score, qe, se

    def time(self, warmup=1, iters=3):
        """(ms per run, cells per run, pairs rescored in int32)."""
        ms = C.c_double()
        cells = C.c_uint64()
        resc = C.c_uint64()
        _ffi.check(_ffi.lib().saber_hip_full_batch_time(self.h, warmup, iters, C.byref(ms),
                                                          C.byref(cells), C.byref(resc)))
        return ms.value, cells.value, resc.value

    def close(self):
        if self.h:
            _ffi.lib().saber_hip_full_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class XdropBatch:
    """Device-resident X-drop batch (saber_hip_xdrop_batch_*)."""

    def __init__(self, ctx, handle, nhits, keep):
        self.ctx, self.h, self.nhits, self._keep = ctx, handle, nhits, keep
        _live_batches.add(self)
        ctx._children.add(self)

    def run(self):
        _ffi.check(_ffi.lib().saber_hip_xdrop_batch_run(self.h))

    def read(self):
        n = self.nhits
        score = np.zeros(n, np.int32)
        lo = np.zeros(n, np.uint32)
        hi = np.zeros(n, np.uint32)
        steps = np.zeros(n, np.uint32)
        _ffi.check(_ffi.lib().saber_hip_xdrop_batch_read(
            self.h, _ffi.ptr(score, C.c_int32), _ffi.ptr(lo, C.c_uint32), _ffi.ptr(hi, C.c_uint32),
            _ffi.ptr(steps, C.c_uint32)))
        return score, lo, hi, steps

    def time(self, warmup=1, iters=5):
        """(ms per run, steps per run)."""
        ms = C.c_double()
        st = C.c_uint64()
        _ffi.check(_ffi.lib().saber_hip_xdrop_batch_time(self.h, warmup, iters, C.byref(ms),
                                                           C.byref(st)))
        return ms.value, st.value

    def close(self):
        if self.h:
            _ffi.lib().saber_hip_xdrop_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class IndexFile:
    """A SABER index file, format v4, opened for its residues and subject
    offsets (saber_hip_index_open: the checks of SeedIndex::open,
    src/search/index.rs:947-1009; verify also checks every section CRC-32).
    residues / offsets are read-only views of the file's memory map."""

    def __init__(self, path, verify=False):
        h = C.POINTER(_ffi.IndexFile)()
        _ffi.check(_ffi.lib().saber_hip_index_open(os.fsencode(path), int(bool(verify)), C.byref(h)))
        self.h = h
        x = h.contents
        self.num_subjects, self.total_residues = int(x.num_subjects), int(x.total_residues)
        self.protein, self.translated = bool(x.protein), bool(x.translated)
        self.residues = np.ctypeslib.as_array(x.residues, shape=(max(self.total_residues, 1),))[:self.total_residues]
        self.offsets = np.ctypeslib.as_array(x.offsets, shape=(self.num_subjects + 1,))

    def close(self):
        if self.h:
            self.residues = self.offsets = None
            _ffi.lib().saber_hip_index_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def crc32(data):
    """CRC-32 of the index file format (saber_hip_crc32, src/search/crc32.rs)."""
    b = np.frombuffer(bytes(data), np.uint8) if not isinstance(data, np.ndarray) else np.ascontiguousarray(data, np.uint8)
    return int(_ffi.lib().saber_hip_crc32(b.ctypes.data_as(C.POINTER(C.c_uint8)), len(b)))


def band_variants():
    """[(L, C, K, bias)] for every compiled packed-kernel variant."""
    out = []
    v = 0
    L_, C_, K_, B_ = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    while _ffi.lib().saber_hip_band_variant(v, C.byref(L_), C.byref(C_), C.byref(K_), C.byref(B_)) == 0:
        out.append((L_.value, C_.value, K_.value, B_.value))
        v += 1
    return out


class Context:
    """One device context (saber_hip_create)."""

    def __init__(self, device=0, force_exact32=False, graphs=True, inject_faults=False, plan=None):
        """plan: None (band batches of 65,536 jobs and more planned on the
        device, smaller ones on the host), "host" or "device" (every batch)."""
        if plan not in (None, "host", "device"):
            raise ValueError(f"plan must be None, 'host' or 'device', not {plan!r}")
        L = _ffi.lib()
        opts = _ffi.Opts()
        opts.flags = ((_ffi.FORCE_EXACT32 if force_exact32 else 0) | (0 if graphs else _ffi.NO_GRAPH) |
                      (_ffi.INJECT_FAULTS if inject_faults else 0) |
                      {None: 0, "host": _ffi.HOST_PLAN, "device": _ffi.DEVICE_PLAN}[plan])
        h = C.c_void_p()
        self._children = weakref.WeakSet()   # batches and models on this context
        _ffi.check(L.saber_hip_create(device, C.byref(opts), C.byref(h)))
        self.h = h
        _live_contexts.add(self)

    def close(self):
        """Release the context, after every batch and model created on it."""
        for ch in list(getattr(self, "_children", ())):
            ch.close()
        if self.h:
            _ffi.lib().saber_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def copy_probe(self, nbytes, iters=5):
        """Mean ms of a plain device copy of nbytes (read nbytes + write nbytes)."""
        ms = C.c_double()
        _ffi.check(_ffi.lib().saber_hip_copy_probe(self.h, int(nbytes), int(iters), C.byref(ms)))
        return ms.value

    def choose(self, w_lo, w_hi, L, Cc, K):
        _ffi.check(_ffi.lib().saber_hip_band_choose(self.h, w_lo, w_hi, L, Cc, K))

    def tune(self, latency_rows):
        """Band jobs with more rows than this are planned for per-row latency
        (0: never)."""
        _ffi.check(_ffi.lib().saber_hip_band_tune(self.h, int(latency_rows)))

    def autotune(self, qres, qoff, sres, soff, jobs, scheme, target_cells=2_000_000_000):
        """Measured shape per class of band widths on the first jobs of this
        batch (saber_hip_band_autotune); returns one dict per visited class."""
        qs, ss, keep = self._seqsets(qres, qoff, None, sres, soff)
        jobs = np.ascontiguousarray(jobs)
        sch = scheme.ffi()
        rep = (_ffi.TuneEntry * 64)()
        n = C.c_int()
        _ffi.check(_ffi.lib().saber_hip_band_autotune(
            self.h, C.byref(qs), None if ss is None else C.byref(ss),
            jobs.ctypes.data_as(C.POINTER(_ffi.BandJob)), len(jobs), C.byref(sch), int(target_cells), rep, 64,
            C.byref(n)))
        return [{"w_lo": e.w_lo, "w_hi": e.w_hi, "shape": (e.L, e.C, e.K), "changed": bool(e.reserved),
                 "cells": int(e.cells), "gcups": round(e.gcups, 1)} for e in rep[:min(n.value, 64)]]

    def band_align(self, qres, qoff, sres, soff, jobs, scheme, stop_at=None, qbias=None,
                   max_dir_bytes=32 << 20, time_iters=0):
        """Banded traceback per job (saber_hip_band_align: banded_align_until,
        banded.rs:472-521; the default cap is the engine's
        traceback_max_bytes, engine.rs:191). One entry per job: None (no
        alignment), "cap" (direction matrix over max_dir_bytes) or
        (score, q_start, q_end, s_start, s_end, ops, cells) with window
        coordinates and ops a list of (op, length), op in b"MID". With
        time_iters > 0, returns (list, mean kernel ms)."""
        qs, ss, keep = self._seqsets(qres, qoff, qbias, sres, soff)
        jobs = np.ascontiguousarray(jobs)
        n = len(jobs)
        qo = np.asarray(qoff, np.int64)
        so = np.asarray(soff if soff is not None else [0], np.int64)
        m = (qo[jobs["query"].astype(np.int64) + 1] - qo[jobs["query"].astype(np.int64)]) if n else np.zeros(0)
        w = jobs["win_hi"].astype(np.int64) - jobs["win_lo"].astype(np.int64) if n else np.zeros(0)
        cap = int((m + w).sum()) if n else 0
        out = (_ffi.Alignment * max(n, 1))()
        codes = np.zeros(max(cap, 1), np.uint8)
        lens = np.zeros(max(cap, 1), np.uint32)
        st = None if stop_at is None else np.ascontiguousarray(stop_at, dtype=np.int32)
        sch = scheme.ffi()
        args = (self.h, C.byref(qs), None if ss is None else C.byref(ss),
                jobs.ctypes.data_as(C.POINTER(_ffi.BandJob)), _ffi.ptr(st, C.c_int32), n, C.byref(sch),
                int(max_dir_bytes), out, _ffi.ptr(codes, C.c_uint8), _ffi.ptr(lens, C.c_uint32), cap)
        kms = C.c_double()
        if time_iters > 0:
            _ffi.check(_ffi.lib().saber_hip_band_align_time(*args, int(time_iters), C.byref(kms)))
        else:
            _ffi.check(_ffi.lib().saber_hip_band_align(*args))
        res = []
        for k in range(n):
            o = out[k]
            if o.status == 1:
                res.append(None)
            elif o.status == 2:
                res.append("cap")
            else:
                ops = [(bytes([int(codes[o.op_off + x])]), int(lens[o.op_off + x])) for x in range(o.nops)]
                res.append((o.score, o.q_start, o.q_end, o.s_start, o.s_end, ops, o.cells))
        return (res, kms.value) if time_iters > 0 else res

    def traceback(self, qres, qoff, sres, soff, jobs, scheme, band_pad=16, exhaustive=False, qbias=None,
                  max_dir_bytes=32 << 20, timed=False):
        """Engine::traceback (engine.rs:1328-1393, with traceback_from_end
        :1403-1446) for a batch of reported hits (saber_hip_traceback). jobs
        is a TRACEBACK_JOB array (window, the region's diagonals d_min/d_max
        in subject coordinates, its best X-drop score, want = the window's
        exact score, optional scoring band limit_lo/limit_hi in window
        diagonals). Returns (results, counters): per job None or
        ((score, q_start, q_end, s_start, s_end, ops, cells), exact), and the
        engine's counters summed over the batch; with timed, also the wall
        milliseconds of the library call alone."""
        qs, ss, keep = self._seqsets(qres, qoff, qbias, sres, soff)
        jobs = np.ascontiguousarray(jobs, dtype=TRACEBACK_JOB)
        n = len(jobs)
        qo = np.asarray(qoff, np.int64)
        m = (qo[jobs["query"].astype(np.int64) + 1] - qo[jobs["query"].astype(np.int64)]) if n else np.zeros(0)
        w = jobs["win_hi"].astype(np.int64) - jobs["win_lo"].astype(np.int64) if n else np.zeros(0)
        cap = int((m + w).sum()) if n else 0
        out = (_ffi.Alignment * max(n, 1))()
        ex = np.zeros(max(n, 1), np.uint8)
        codes = np.zeros(max(cap, 1), np.uint8)
        lens = np.zeros(max(cap, 1), np.uint32)
        cnt = _ffi.TracebackCounters()
        sch = scheme.ffi()
        t0 = time.perf_counter()
        _ffi.check(_ffi.lib().saber_hip_traceback(
            self.h, C.byref(qs), None if ss is None else C.byref(ss),
            jobs.ctypes.data_as(C.POINTER(_ffi.TracebackJob)), n, C.byref(sch), int(band_pad), int(bool(exhaustive)),
            int(max_dir_bytes), out, _ffi.ptr(ex, C.c_uint8), _ffi.ptr(codes, C.c_uint8),
            _ffi.ptr(lens, C.c_uint32), cap, C.byref(cnt)))
        call_ms = (time.perf_counter() - t0) * 1e3
        res = []
        for k in range(n):
            o = out[k]
            if o.status != 0:
                res.append(None)
                continue
            ops = [(bytes([int(codes[o.op_off + x])]), int(lens[o.op_off + x])) for x in range(o.nops)]
            res.append(((o.score, o.q_start, o.q_end, o.s_start, o.s_end, ops, o.cells), bool(ex[k])))
        counters = {f: int(getattr(cnt, f)) for f, _ in _ffi.TracebackCounters._fields_}
        return (res, counters, call_ms) if timed else (res, counters)

    def full_autotune(self, qres, qoff, sres, soff, scheme, qbias=None, ends=False, target_cells=50_000_000_000):
        """Measured full-matrix shape per subject-length group on a sample of
        whole queries (saber_hip_full_autotune); applies to full batches of
        this flavour created afterwards. One dict per visited group."""
        qs, ss, keep = self._seqsets(qres, qoff, qbias, sres, soff)
        sch = scheme.ffi()
        rep = (_ffi.TuneEntry * 64)()
        n = C.c_int()
        _ffi.check(_ffi.lib().saber_hip_full_autotune(
            self.h, C.byref(qs), None if ss is None else C.byref(ss), C.byref(sch), int(bool(ends)),
            int(target_cells), rep, 64, C.byref(n)))
        return [{"len_lo": e.w_lo, "len_hi": e.w_hi, "shape": (e.L, e.C), "changed": bool(e.reserved),
                 "cells": int(e.cells), "gcups": round(e.gcups, 1)} for e in rep[:min(n.value, 64)]]

    def upload_index(self, index):
        """Make an opened v4 index file's subjects the device database
        (saber_hip_upload_index)."""
        _ffi.check(_ffi.lib().saber_hip_upload_index(self.h, index.h))

    def upload_db(self, qres, qoff):
        off = np.ascontiguousarray(qoff, dtype=np.uint64)
        res = np.ascontiguousarray(qres, dtype=np.uint8)
        ss = _ffi.SeqSet(_ffi.ptr(res, C.c_uint8), _ffi.ptr(off, C.c_uint64), len(off) - 1, None)
        _ffi.check(_ffi.lib().saber_hip_upload_db(self.h, C.byref(ss)))

    @staticmethod
    def _jobs(query_idx, subject_idx, win_lo, win_hi, band_lo, band_hi):
        n = len(query_idx)
        jobs = np.zeros(n, dtype=[("query", "<u4"), ("subject", "<u4"), ("win_lo", "<u4"),
                                  ("win_hi", "<u4"), ("band_lo", "<i4"), ("band_hi", "<i4")])
        jobs["query"] = query_idx
        jobs["subject"] = subject_idx
        jobs["win_lo"] = win_lo
        jobs["win_hi"] = win_hi
        jobs["band_lo"] = band_lo
        jobs["band_hi"] = band_hi
        return jobs

    def band_batch(self, qres, qoff, sres, soff, jobs, scheme, qbias=None, ends=True):
        """Create a device-resident batch from flat arrays (jobs: structured
        array with query, subject, win_lo, win_hi, band_lo, band_hi).
        ends=False: scores only (score_banded_many's output; faster)."""
        qres = np.ascontiguousarray(qres, np.uint8)
        qoff = np.ascontiguousarray(qoff, np.uint64)
        qb = None if qbias is None else np.ascontiguousarray(qbias, np.int8)
        qs = _ffi.SeqSet(_ffi.ptr(qres, C.c_uint8), _ffi.ptr(qoff, C.c_uint64), len(qoff) - 1,
                         _ffi.ptr(qb, C.c_int8))
        keep = [qres, qoff, qb]
        if sres is not None:
            sres = np.ascontiguousarray(sres, np.uint8)
            soff = np.ascontiguousarray(soff, np.uint64)
            ss = _ffi.SeqSet(_ffi.ptr(sres, C.c_uint8), _ffi.ptr(soff, C.c_uint64), len(soff) - 1, None)
            sp = C.byref(ss)
            keep += [sres, soff]
        else:
            sp = None
        jobs = np.ascontiguousarray(jobs)
        keep.append(jobs)
        sch = scheme.ffi()
        h = C.c_void_p()
        _ffi.check(_ffi.lib().saber_hip_band_batch_create(
            self.h, C.byref(qs), sp, jobs.ctypes.data_as(C.POINTER(_ffi.BandJob)), len(jobs),
            C.byref(sch), int(bool(ends)), C.byref(h)))
        return BandBatch(self, h, len(jobs), keep, bool(ends))

    def score_bands(self, qres, qoff, sres, soff, jobs, scheme, qbias=None, ends=True):
        """(score, q_end, s_end, cells) arrays, one entry per job (q_end and
        s_end None when ends=False: scores only)."""
        b = self.band_batch(qres, qoff, sres, soff, jobs, scheme, qbias, ends)
        try:
            b.run()
            return b.read()
        finally:
            b.close()

    @staticmethod
    def _seqsets(qres, qoff, qbias, sres, soff):
        qres = np.ascontiguousarray(qres, np.uint8)
        qoff = np.ascontiguousarray(qoff, np.uint64)
        qb = None if qbias is None else np.ascontiguousarray(qbias, np.int8)
        qs = _ffi.SeqSet(_ffi.ptr(qres, C.c_uint8), _ffi.ptr(qoff, C.c_uint64), len(qoff) - 1,
                         _ffi.ptr(qb, C.c_int8))
        keep = [qres, qoff, qb]
        ss = None
        if sres is not None:
            sres = np.ascontiguousarray(sres, np.uint8)
            soff = np.ascontiguousarray(soff, np.uint64)
            ss = _ffi.SeqSet(_ffi.ptr(sres, C.c_uint8), _ffi.ptr(soff, C.c_uint64), len(soff) - 1, None)
            keep += [sres, soff]
        return qs, ss, keep

    def xdrop_batch(self, qres, qoff, sres, soff, hits, scheme, x, qbias=None):
        """Device-resident X-drop batch; hits is an (n, 4) uint32 array of
        (query, subject, qp, sp)."""
        qs, ss, keep = self._seqsets(qres, qoff, qbias, sres, soff)
        hits = np.ascontiguousarray(hits, np.uint32).reshape(-1, 4)
        keep.append(hits)
        sch = scheme.ffi()
        h = C.c_void_p()
        _ffi.check(_ffi.lib().saber_hip_xdrop_batch_create(
            self.h, C.byref(qs), None if ss is None else C.byref(ss), _ffi.ptr(hits, C.c_uint32),
            len(hits), C.byref(sch), int(x), C.byref(h)))
        return XdropBatch(self, h, len(hits), keep)

    def full_batch(self, qres, qoff, sres, soff, scheme, qbias=None, ends=True):
        """Device-resident exhaustive batch: every query x every subject
        (ends=False: scores only, as score_many)."""
        qs, ss, keep = self._seqsets(qres, qoff, qbias, sres, soff)
        sch = scheme.ffi()
        h = C.c_void_p()
        _ffi.check(_ffi.lib().saber_hip_full_batch_create(
            self.h, C.byref(qs), None if ss is None else C.byref(ss), C.byref(sch), int(bool(ends)), C.byref(h)))
        nq = len(keep[1]) - 1
        ns = (len(keep[4]) - 1) if ss is not None else None
        return FullBatch(self, h, nq, ns, keep, bool(ends))

    def score_full(self, qres, qoff, sres, soff, scheme, qbias=None, ends=True):
        """(score, q_end, s_end) arrays of shape [queries, subjects] (q_end and
        s_end None when ends=False)."""
        b = self.full_batch(qres, qoff, sres, soff, scheme, qbias, ends)
        try:
            b.run()
            return b.read()
        finally:
            b.close()

    def xdrop(self, qres, qoff, sres, soff, hits, scheme, x, qbias=None):
        """(score, q_lo, q_hi, steps) per hit (xdrop_cat, scan.rs:859-904)."""
        b = self.xdrop_batch(qres, qoff, sres, soff, hits, scheme, x, qbias)
        try:
            b.run()
            return b.read()
        finally:
            b.close()

    def score_windows(self, query, bias, windows, scheme):
        q = _codes(query)
        n = len(windows)
        out = np.zeros(n, np.int32)
        if n == 0:
            return out
        wins = [_codes(w) for w in windows]
        ptrs = (C.POINTER(C.c_uint8) * n)(*[w.ctypes.data_as(C.POINTER(C.c_uint8)) for w in wins])
        lens = np.array([len(w) for w in wins], dtype=np.uint32)
        b = None if bias is None else np.ascontiguousarray(bias, np.int8)
        sch = scheme.ffi()
        _ffi.check(_ffi.lib().saber_hip_score_windows(
            self.h, _ffi.ptr(q, C.c_uint8), len(q), _ffi.ptr(b, C.c_int8), ptrs,
            _ffi.ptr(lens, C.c_uint32), n, C.byref(sch), _ffi.ptr(out, C.c_int32)))
        return out


_default = None


def default_context():
    global _default
    if _default is None:
        _default = Context(0)
    return _default


def score_banded_many(query, bias, jobs, scheme, ctx=None):
    """Exact banded SW score of `query` against every BandJob, caller order
    (src/algorithm/band_simd8.rs:78-185). bias must have the query's length."""
    ctx = ctx or default_context()
    q = _codes(query)
    if bias is not None and len(bias) != len(q):
        raise ValueError("bias length must equal query length")
    if len(jobs) == 0 or len(q) == 0:
        return [0] * len(jobs)
    subs = [_codes(j.subject) for j in jobs]
    soff = np.zeros(len(subs) + 1, np.uint64)
    np.cumsum([len(s) for s in subs], out=soff[1:])
    sres = np.concatenate(subs) if soff[-1] else np.zeros(1, np.uint8)
    n = len(jobs)
    jj = Context._jobs(np.zeros(n, np.uint32), np.arange(n, dtype=np.uint32), np.zeros(n, np.uint32),
                       np.array([len(s) for s in subs], np.uint32),
                       np.array([j.band.lo for j in jobs], np.int64).clip(-2**31, 2**31 - 1),
                       np.array([j.band.hi for j in jobs], np.int64).clip(-2**31, 2**31 - 1))
    score, _, _, _ = ctx.score_bands(q, np.array([0, len(q)], np.uint64), sres, soff, jj, scheme,
                                     None if bias is None else np.asarray(bias, np.int8))
    return [int(v) for v in score]


def banded_score(q, s, bias, scheme, band, ctx=None):
    """BandScore of one banded job (src/algorithm/banded.rs:408-449)."""
    ctx = ctx or default_context()
    q = _codes(q)
    s = _codes(s)
    if bias is not None and len(bias) != len(q):
        raise ValueError("bias length must equal query length")
    if len(q) == 0 or len(s) == 0:
        return BandScore(0, 0, 0, 0)
    jj = Context._jobs([0], [0], [0], [len(s)], [max(band.lo, -2**31)], [min(band.hi, 2**31 - 1)])
    sc, qe, se, ce = ctx.score_bands(q, np.array([0, len(q)], np.uint64), s,
                                     np.array([0, len(s)], np.uint64), jj, scheme,
                                     None if bias is None else np.asarray(bias, np.int8))
    return BandScore(int(sc[0]), int(qe[0]), int(se[0]), int(ce[0]))


def score_many(query, bias, subjects, matrix, gaps, ctx=None):
    """Exact full SW scores of `query` against every subject
    (src/algorithm/sw_simd.rs:95-153)."""
    ctx = ctx or default_context()
    q = _codes(query)
    if bias is not None and len(bias) != len(q):
        raise ValueError("bias length must equal query length")
    return [int(v) for v in ctx.score_windows(q, bias, subjects, Scheme(matrix, gaps.open, gaps.extend))]


class HipBackend:
    """ScoreBackend on the MI355X (src/gpu/mod.rs:45-59): exact i32 scores,
    identical to score_many; errors raise (the engine then falls back)."""

    def __init__(self, device=0, inject_faults=False):
        self.ctx = Context(device, inject_faults=inject_faults)

    def name(self):
        return "hip-gfx950"

    def score_windows(self, query, bias, windows, matrix, gaps):
        if bias is not None and len(bias) != len(query):
            raise ValueError("bias length mismatch")
        if len(windows) == 0:
            return []
        return [int(v) for v in self.ctx.score_windows(query, bias, windows,
                                                        Scheme(matrix, gaps.open, gaps.extend))]
