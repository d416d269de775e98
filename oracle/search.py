"""Test infrastructure: the plain CPU search the GPU search is checked
against -- Engine::search_scan (src/search/engine/scan.rs:1261-1497) from the
queries to each query's reported hits, restated over the oracle's parts:

  preparation     composition_bias (or_composition_bias, bias.rs:80-122),
                  the query's floors and region threshold (engine.rs:718-786)
  scan + regions  or_scan (oracle/scan.c: seed tables, window filter, visit
                  order, coverage, X-drop, finish_subject_into, the cap)
  plans           plan_windows (oracle/host.py, engine.rs:1247-1324)
  scores          score_windows' two stages (engine.rs:1021-1099) with
                  or_banded_score (banded.rs:134-224, 408-449)
  finish          the E-value order, the reporting loop and make_hit
                  (engine.rs:1127-1221, 1447-1530) with Engine::traceback
                  restated over or_band_align (traceback() below)

Statistics (E-values, bit scores, score floors) come from the library's
host statistics (saber_amd.stats), which tests/test_stats.py pins against
the reference's own tests of src/scoring/stats.rs; everything else here is
independent of the product. Only tests/ and bench.py's cpu_baseline leg use
this module, and only as the checker.
"""
import ctypes as C
import math

import numpy as np

from . import host as H
from . import oracle as O

_I32 = C.c_int32


class ScanParams(C.Structure):
    _fields_ = [("alpha_map", C.POINTER(C.c_uint8)), ("asize", C.c_uint32), ("nshapes", C.c_uint32),
                ("offsets", C.POINTER(C.c_uint16)), ("weight", C.POINTER(C.c_uint32)),
                ("span", C.POINTER(C.c_uint32)), ("min_ids", C.c_uint32), ("kx", C.c_uint32), ("x", C.c_int32),
                ("hsp_threshold", C.c_int32), ("window_slack", C.c_int64), ("max_regions_per_subject", C.c_uint32),
                ("pad", C.c_uint32), ("cap", C.c_uint64), ("seed_max_occ", C.c_uint64),
                ("ref_db_residues", C.c_uint64)]


COUNTERS = ["seed_positions", "seeds_frequent_skipped", "scan_candidates", "hits", "hits_filtered",
            "ungapped_extensions", "extensions_avoided", "xdrop_steps", "hsps", "regions_weak_dropped",
            "regions_dropped_by_cap", "regions", "scan_groups"]


class ScanCounters(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in COUNTERS]


REGION = np.dtype([("subject", "<u8"), ("d_min", "<i8"), ("d_max", "<i8"), ("best", "<i4"), ("n_hsp", "<u4"),
                   ("hsp_sum", "<i4"), ("pad", "<u4")])

_bound = False


def _bind():
    global _bound
    L = O.lib()
    if not _bound:
        p = C.c_void_p
        L.or_scan.restype = p
        L.or_scan.argtypes = [C.POINTER(C.c_uint8), C.POINTER(C.c_uint64), C.c_uint32, C.POINTER(C.c_int8),
                              C.POINTER(C.c_uint8), C.POINTER(C.c_uint64), C.c_uint32, C.POINTER(_I32),
                              C.POINTER(_I32), C.POINTER(ScanParams), C.c_int, C.POINTER(ScanCounters)]
        L.or_scan_nregions.restype = C.c_uint64
        L.or_scan_nregions.argtypes = [p, C.c_uint32]
        L.or_scan_regions.restype = None
        L.or_scan_regions.argtypes = [p, C.c_uint32, C.c_void_p]
        L.or_scan_nsurvivors.restype = C.c_uint64
        L.or_scan_nsurvivors.argtypes = [p]
        L.or_scan_survivors.restype = None
        L.or_scan_survivors.argtypes = [p, C.POINTER(C.c_uint32)]
        L.or_scan_free.restype = None
        L.or_scan_free.argtypes = [p]
        L.or_auto_occ_cap.restype = C.c_uint64
        L.or_auto_occ_cap.argtypes = [C.c_uint64] * 4
        _bound = True
    return L


def parse_shape(p):
    """SeedShape::parse (alphabet.rs:364-384): care offsets and span."""
    if not p or p[0] != "1" or p[-1] != "1" or set(p) - {"0", "1"}:
        raise ValueError(f"bad shape {p}")
    offs = [i for i, ch in enumerate(p) if ch == "1"]
    if not 2 <= len(offs) <= 13:
        raise ValueError("weight must be in [2, 13]")
    return offs, len(p)


def scan(qres, qoff, qbias, sres, soff, matrix, alpha_map, asize, shapes, region_thr, *, min_ids=27, kx=2, x=20,
         hsp_threshold=40, window_slack=64, max_regions_per_subject=4, cap=5000, seed_max_occ=0,
         ref_db_residues=40_000_000, log_survivors=False):
    """or_scan: per query the capped region array (REGION), the counters and,
    with log_survivors, every survivor (query, subject, qp, sp, shape) in visit
    order."""
    L = _bind()
    qres = np.ascontiguousarray(qres, np.uint8)
    qoff = np.ascontiguousarray(qoff, np.uint64)
    sres = np.ascontiguousarray(sres, np.uint8)
    soff = np.ascontiguousarray(soff, np.uint64)
    nq, ns = len(qoff) - 1, len(soff) - 1
    qb = None if qbias is None else np.ascontiguousarray(qbias, np.int8)
    mat = np.ascontiguousarray(np.asarray(matrix, np.int32).reshape(576))
    am = np.ascontiguousarray(alpha_map, np.uint8)
    offs = np.zeros((len(shapes), 16), np.uint16)
    wt = np.zeros(len(shapes), np.uint32)
    sp = np.zeros(len(shapes), np.uint32)
    for i, s in enumerate(shapes):
        o, span = parse_shape(s)
        offs[i, :len(o)] = o
        wt[i] = len(o)
        sp[i] = span
    rt = np.ascontiguousarray(region_thr, np.int32)
    P = ScanParams(am.ctypes.data_as(C.POINTER(C.c_uint8)), asize, len(shapes),
                   offs.ctypes.data_as(C.POINTER(C.c_uint16)), wt.ctypes.data_as(C.POINTER(C.c_uint32)),
                   sp.ctypes.data_as(C.POINTER(C.c_uint32)), min_ids, kx, x, hsp_threshold, window_slack,
                   max_regions_per_subject, 0, cap, seed_max_occ, ref_db_residues)
    ctr = ScanCounters()
    h = L.or_scan(qres.ctypes.data_as(C.POINTER(C.c_uint8)), qoff.ctypes.data_as(C.POINTER(C.c_uint64)), nq,
                  None if qb is None else qb.ctypes.data_as(C.POINTER(C.c_int8)),
                  sres.ctypes.data_as(C.POINTER(C.c_uint8)), soff.ctypes.data_as(C.POINTER(C.c_uint64)), ns,
                  mat.ctypes.data_as(C.POINTER(_I32)), rt.ctypes.data_as(C.POINTER(_I32)), C.byref(P),
                  1 if log_survivors else 0, C.byref(ctr))
    try:
        regs = []
        for q in range(nq):
            n = L.or_scan_nregions(h, q)
            r = np.zeros(n, REGION)
            if n:
                L.or_scan_regions(h, q, r.ctypes.data)
            regs.append(r)
        surv = None
        if log_survivors:
            n = L.or_scan_nsurvivors(h)
            surv = np.zeros((n, 5), np.uint32)
            if n:
                L.or_scan_survivors(h, surv.ctypes.data_as(C.POINTER(C.c_uint32)))
    finally:
        L.or_scan_free(h)
    return regs, {k: int(getattr(ctr, k)) for k in COUNTERS}, surv


def _from_end(q, s, bias, matrix, go, ge, want, limit, band_pad, cap, cnt):
    """traceback_from_end (engine.rs:1403-1446)."""
    m, n = len(q), len(s)
    scope = limit if limit is not None else (-m, n)
    sc, qe, se, cells = O.banded_score(q, s, bias, matrix, go, ge, scope[0], scope[1])
    cnt["banded_cells"] += cells
    if sc <= 0:
        return None
    cnt["end_anchored_tracebacks"] += 1
    target = max(sc, want)
    d_end = se - qe
    pad = max(band_pad, 4)
    while True:
        lo, hi = max(d_end - pad, scope[0]), min(d_end + pad, scope[1])
        covers = lo <= scope[0] and hi >= scope[1]
        cnt["tracebacks"] += 1
        r = O.band_align(q, s, matrix, bias, go, ge, lo, hi, target, cap)
        if r == "cap":
            return None
        if r is None:
            if covers:
                return None
        else:
            cnt["banded_cells"] += r[6]
            if r[0] >= target or covers:
                return r, r[0] >= want
        cnt["band_widenings"] += 1
        pad *= 2


def traceback(q, s, bias, matrix, go, ge, d_min, d_max, win_lo, best, want, limit, band_pad, exhaustive, cap, cnt):
    """Engine::traceback (engine.rs:1328-1393) over or_band_align: None or
    ((score, q_start, q_end, s_start, s_end, ops, cells), exact)."""
    m, n = len(q), len(s)
    pad = max(band_pad, 4)
    full = (-m, n)
    base = full if exhaustive or best == 0 else (d_min - win_lo, d_max - win_lo)
    best_al = None
    while True:
        lo, hi = full if base == full else (base[0] - pad, base[1] + pad)
        covers = lo <= -(m - 1) and hi >= n - 1
        if limit is not None:
            lo, hi = max(lo, limit[0]), min(hi, limit[1])
            covers = covers or (lo <= limit[0] and hi >= limit[1])
        cnt["tracebacks"] += 1
        r = O.band_align(q, s, matrix, bias, go, ge, lo, hi, want, cap)
        if r == "cap":
            found = _from_end(q, s, bias, matrix, go, ge, want, limit, band_pad, cap, cnt)
            if found is not None:
                return found
            cnt["band_capped"] += 1
            return None if best_al is None else (best_al, False)
        if r is None:
            if covers:
                return None
        else:
            cnt["banded_cells"] += r[6]
            best_al = r
            if r[0] >= want or covers:
                return r, r[0] >= want
        cnt["band_widenings"] += 1
        pad *= 2


def make_hit(q, s_full, subject, s_off, al, exact, qs, matrix):
    """make_hit (engine.rs:1447-1530) without the alignment strings."""
    score, q_start, q_end, s_start, s_end, ops, _ = al
    s = s_full[s_off:]
    i, j = q_start, s_start
    matches = mism = pos = gopen = gcols = cols = 0
    for op, l in ops:
        cols += l
        if op == b"M":
            for _ in range(l):
                a, b = int(q[i]), int(s[j])
                ident = a == b and a < 20
                ca, cb = (a if a < 24 else 22), (b if b < 24 else 22)
                positive = ident or matrix[ca][cb] > 0
                matches += ident
                mism += not ident
                pos += positive
                i += 1
                j += 1
        elif op == b"I":
            gopen += 1
            gcols += l
            i += l
        else:
            gopen += 1
            gcols += l
            j += l
    return dict(subject=subject, score=score, bitscore=qs.bitscore(score), evalue=qs.evalue_for(score, len(s_full)),
                q_start=q_start + 1, q_end=q_end, s_start=s_start + s_off + 1, s_end=s_end + s_off, columns=cols,
                matches=matches, mismatches=mism, positives=pos, gap_opens=gopen, gap_columns=gcols,
                q_len=len(q), s_len=len(s_full), ops=[(op[0], l) for op, l in ops], exact=exact)


def hit_key(h):
    cig = "".join(f"{l}{chr(o)}" for o, l in h["ops"])
    return (h["subject"], h["score"], h["q_start"], h["q_end"], h["s_start"], h["s_end"], cig, h["exact"])


def _region_order(r):
    return (-int(r["best"]), int(r["subject"]), int(r["d_min"]), int(r["d_max"]))


def search(qres, qoff, sres, soff, matrix, go, ge, stats, seeds_map, seeds_size, shapes, *, evalue=10.0,
           max_target_seqs=10, max_hsps=1, hsp_threshold=40, x_drop=20, hsp_floor_fraction=0.65, min_ids=27, kx=2,
           band_pad=16, window_slack=64, score_band=48, prefilter_band=16, prefilter_frac=0.75, max_candidates=5000,
           max_regions_per_subject=4, comp_bias=True, seed_max_occ=0, traceback_max_bytes=32 << 20,
           ref_db_residues=40_000_000, per_subject=True):
    """The plain search: per query the reported hits (dicts, hit_key order
    fields) and the counters. stats: saber_amd.stats.SearchStatistics of the
    database (db_length = residues, num seqs = subjects)."""
    mat = np.asarray(matrix, np.int64).reshape(24, 24)
    nq = len(qoff) - 1
    db_len = int(soff[-1])
    re = max(evalue, 10.0)
    if ref_db_residues > 0 and db_len > ref_db_residues:
        re *= db_len / ref_db_residues
    bias_all = np.zeros(int(qoff[-1]), np.int8)
    preps = []
    for qi in range(nq):
        q = qres[int(qoff[qi]):int(qoff[qi + 1])]
        b = O.composition_bias(q, matrix) if comp_bias and len(q) else np.zeros(len(q), np.int8)
        bias_all[int(qoff[qi]):int(qoff[qi + 1])] = b
        qs = stats.for_query(max(len(q), 1))
        if len(q) == 0:
            floor = region_floor = 2**63 - 1
        else:
            floor = qs.min_score(evalue)
            region_floor = floor if re == evalue else qs.min_score(re)
        if hsp_floor_fraction > 0 and region_floor > 0:
            thr = max(hsp_threshold, int(min(math.ceil(region_floor * hsp_floor_fraction), 2**31 - 1)))
        else:
            thr = hsp_threshold
        preps.append(dict(bias=b, has_bias=bool((b != 0).any()), qs=qs, floor=floor, region_floor=region_floor,
                          region_thr=thr))
    regs, ctr, _ = scan(qres, qoff, bias_all, sres, soff, matrix, seeds_map, seeds_size, shapes,
                        [p["region_thr"] for p in preps], min_ids=min_ids, kx=kx, x=x_drop,
                        hsp_threshold=hsp_threshold, window_slack=window_slack,
                        max_regions_per_subject=max_regions_per_subject, cap=max_candidates,
                        seed_max_occ=seed_max_occ, ref_db_residues=ref_db_residues)
    ctr = dict(ctr)
    for k in ("regions_below_evalue", "band_scored_regions", "band_rescored", "tracebacks", "band_widenings",
              "band_capped", "end_anchored_tracebacks", "filtered_identity_coverage", "alignments"):
        ctr[k] = 0
    tcnt = dict(tracebacks=0, band_widenings=0, band_capped=0, end_anchored_tracebacks=0, banded_cells=0)
    ecut = evalue if evalue > 0 and math.isfinite(evalue) else math.inf
    out = []
    for qi in range(nq):
        q = qres[int(qoff[qi]):int(qoff[qi + 1])]
        P = preps[qi]
        qs = P["qs"]
        bias = P["bias"] if P["has_bias"] else None
        R = regs[qi]
        slens = [int(soff[int(r["subject"]) + 1] - soff[int(r["subject"])]) for r in R]
        rd = [dict(subject=int(r["subject"]), d_min=int(r["d_min"]), d_max=int(r["d_max"]), best=int(r["best"]))
              for r in R]
        plans = H.plan_windows(rd, len(q), slens, window_slack, score_band, hsp_threshold)
        scores = [0] * len(R)
        subj = lambda i: sres[int(soff[rd[i]["subject"]]):int(soff[rd[i]["subject"] + 1])]
        full = []
        for i, (lo, hi, band) in enumerate(plans):
            if band is None:
                full.append((i, (-(len(q) - 1), hi - lo - 1)))
                continue
            ctr["band_scored_regions"] += 1
            if 0 < prefilter_band < score_band:
                nb = H.clip(rd[i]["d_min"] - prefilter_band - lo, rd[i]["d_max"] + prefilter_band - lo, len(q),
                            hi - lo)
                sc = O.banded_score(q, subj(i)[lo:hi], bias, matrix, go, ge, nb[0], nb[1])[0]
                t = prefilter_frac * qs.floor_lower_bound(re, None, slens[i], P["region_floor"])
                if sc >= t:
                    full.append((i, band))
                    ctr["band_rescored"] += 1
                else:
                    scores[i] = sc
            else:
                full.append((i, band))
        for i, (blo, bhi) in full:
            lo, hi, _ = plans[i]
            scores[i] = O.banded_score(q, subj(i)[lo:hi], bias, matrix, go, ge, blo, bhi)[0]
        # finish (engine.rs:1127-1221)
        order = []
        for i in range(len(R)):
            s = scores[i]
            if not per_subject:
                if s >= P["floor"]:
                    order.append((qs.evalue(s), i))
                continue
            if s <= 0:
                continue
            n = slens[i]
            if qs.evalue_in(s, qs.space_lower_bound(n)) > ecut:
                continue
            e = qs.evalue_for(s, n)
            if e <= ecut:
                order.append((e, i))
        ctr["regions_below_evalue"] += len(R) - len(order)
        order.sort(key=lambda t: (t[0], -scores[t[1]], _region_order(R[t[1]])))
        hits, per_subject_cnt, reported = [], {}, 0
        for _, ri in order:
            sj = rd[ri]["subject"]
            if reported >= max_target_seqs and sj not in per_subject_cnt:
                break
            cnt = per_subject_cnt.get(sj, 0)
            if cnt >= max_hsps:
                continue
            lo, hi, band = plans[ri]
            tb = traceback(q, subj(ri)[lo:hi], bias, matrix, go, ge, rd[ri]["d_min"], rd[ri]["d_max"], lo,
                           rd[ri]["best"], scores[ri], band, band_pad, False, traceback_max_bytes, tcnt)
            if tb is None:
                continue
            al, exact = tb
            h = make_hit(q, subj(ri), sj, lo, al, exact, qs, mat)
            if cnt > 0 and any(o["subject"] == sj and o["q_start"] <= h["q_end"] and h["q_start"] <= o["q_end"]
                               and o["s_start"] <= h["s_end"] and h["s_start"] <= o["s_end"] for o in hits):
                continue
            passes = h["evalue"] <= ecut if per_subject else h["score"] >= P["floor"]
            if not passes:
                ctr["filtered_identity_coverage"] += 1
                continue
            if cnt == 0:
                reported += 1
            per_subject_cnt[sj] = cnt + 1
            hits.append(h)
        hits.sort(key=lambda h: (h["evalue"], -h["score"], h["subject"], h["s_start"]))
        ctr["alignments"] += len(hits)
        out.append(hits)
    for k in ("tracebacks", "band_widenings", "band_capped", "end_anchored_tracebacks"):
        ctr[k] = tcnt[k]
    return out, ctr
