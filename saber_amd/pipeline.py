"""Host pipeline around the scoring kernels (the library's host code, C ABI
saber_hip_visit_hits / finish_subject / cap_regions / plan_windows /
score_regions), under the reference's names: visit_scan, finish_subject,
the region cap, plan_windows and the banded part of score_windows."""
import ctypes as C

import numpy as np

from . import _ffi

HSP_DTYPE = np.dtype([("diag", np.int64), ("q_lo", np.uint32), ("q_hi", np.uint32), ("score", np.int32),
                      ("reserved", np.uint32)])
REGION_DTYPE = np.dtype([("subject", np.uint64), ("d_min", np.int64), ("d_max", np.int64), ("best", np.int32),
                         ("n_hsp", np.uint32), ("hsp_sum", np.int32), ("reserved", np.uint32)])
PLAN_DTYPE = np.dtype([("lo", np.uint64), ("hi", np.uint64), ("band_lo", np.int64), ("band_hi", np.int64),
                       ("banded", np.int32), ("reserved", np.uint32)])


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def visit_scan(qlen, slen, qp, sp, score, q_lo, q_hi, hsp_threshold=40):
    """-> (HSP array, extensions, avoided)."""
    qp, sp = np.ascontiguousarray(qp, np.uint32), np.ascontiguousarray(sp, np.uint32)
    score = np.ascontiguousarray(score, np.int32)
    q_lo, q_hi = np.ascontiguousarray(q_lo, np.uint32), np.ascontiguousarray(q_hi, np.uint32)
    out = np.zeros(max(len(qp), 1), HSP_DTYPE)
    nh, ext, av = C.c_uint64(), C.c_uint64(), C.c_uint64()
    _ffi.check(_ffi.lib().saber_hip_visit_hits(int(qlen), int(slen), _p(qp, C.c_uint32), _p(sp, C.c_uint32), len(qp),
                                               _p(score, C.c_int32), _p(q_lo, C.c_uint32), _p(q_hi, C.c_uint32),
                                               int(hsp_threshold), _p(out, _ffi.Hsp), C.byref(nh), C.byref(ext),
                                               C.byref(av)))
    return out[:nh.value].copy(), ext.value, av.value


def finish_subject(subject, hsps, window_slack=64, hsp_threshold=40, region_thr=40, max_regions_per_subject=4):
    """hsps: an HSP array or (diag, q_lo, q_hi, score) tuples."""
    if isinstance(hsps, np.ndarray) and hsps.dtype == HSP_DTYPE:
        h = hsps.copy()
    else:
        h = np.array([tuple(x) + (0,) for x in hsps], HSP_DTYPE)
    out = np.zeros(max(len(h), 1), REGION_DTYPE)
    n = C.c_uint32()
    _ffi.check(_ffi.lib().saber_hip_finish_subject(int(subject), _p(h, _ffi.Hsp), len(h), int(window_slack),
                                                   int(hsp_threshold), int(region_thr), int(max_regions_per_subject),
                                                   _p(out, _ffi.Region), C.byref(n)))
    return out[:n.value].copy()


def cap_regions(regions, cap):
    r = np.array(regions, REGION_DTYPE)
    n = C.c_uint64()
    _ffi.check(_ffi.lib().saber_hip_cap_regions(_p(r, _ffi.Region), len(r), int(cap), C.byref(n)))
    return r[:n.value].copy()


def plan_windows(regions, qlen, subject_len, window_slack=64, score_band=0, hsp_threshold=40, exhaustive=False):
    r = np.ascontiguousarray(regions, REGION_DTYPE)
    sl = np.ascontiguousarray(subject_len, np.uint64)
    out = np.zeros(max(len(r), 1), PLAN_DTYPE)
    pp = _ffi.PlanParams(int(window_slack), int(score_band), int(hsp_threshold), 1 if exhaustive else 0)
    _ffi.check(_ffi.lib().saber_hip_plan_windows(_p(r, _ffi.Region), len(r), int(qlen), _p(sl, C.c_uint64),
                                                 C.byref(pp), _p(out, _ffi.Plan)))
    return out[:len(r)].copy()


def score_regions(ctx, qres, qoff, qidx, regions, plans, scheme, narrow_band=0, full_band=0, thr=None,
                  sres=None, soff=None, qbias=None):
    """Scores of the regions (banded part of score_windows) and the rescored
    flags; subjects default to the context's uploaded database."""
    qres = np.ascontiguousarray(qres, np.uint8)
    qoff = np.ascontiguousarray(qoff, np.uint64)
    qb = None if qbias is None else np.ascontiguousarray(qbias, np.int8)
    qs = _ffi.SeqSet(_ffi.ptr(qres, C.c_uint8), _ffi.ptr(qoff, C.c_uint64), len(qoff) - 1, _ffi.ptr(qb, C.c_int8))
    ss = None
    if sres is not None:
        sres = np.ascontiguousarray(sres, np.uint8)
        soff = np.ascontiguousarray(soff, np.uint64)
        ss = _ffi.SeqSet(_ffi.ptr(sres, C.c_uint8), _ffi.ptr(soff, C.c_uint64), len(soff) - 1, None)
    qidx = np.ascontiguousarray(qidx, np.uint32)
    r = np.ascontiguousarray(regions, REGION_DTYPE)
    p = np.ascontiguousarray(plans, PLAN_DTYPE)
    t = None if thr is None else np.ascontiguousarray(thr, np.float64)
    n = len(r)
    out = np.zeros(max(n, 1), np.int32)
    resc = np.zeros(max(n, 1), np.uint8)
    sch = scheme.ffi()
    _ffi.check(_ffi.lib().saber_hip_score_regions(ctx.h, C.byref(qs), C.byref(ss) if ss is not None else None,
                                                  _p(qidx, C.c_uint32), _p(r, _ffi.Region), _p(p, _ffi.Plan), n,
                                                  C.byref(sch), int(narrow_band), int(full_band),
                                                  _ffi.ptr(t, C.c_double), _p(out, C.c_int32), _p(resc, C.c_uint8)))
    return out[:n].copy(), resc[:n].astype(bool)
