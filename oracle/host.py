"""Test infrastructure: plain-Python restatement of the host pipeline around
the scoring kernels (hit visiting, region joining, region cap, window and band
planning, two-stage band selection). Only tests/ use it, as the checker of
the library's host code (saber_amd/csrc/host_pipeline.cpp, api.hip).

Parity pins beyond this restatement: tests/test_pipeline.py checks invariants
the reference's semantics fix (every HSP lands in exactly one region, regions
are separated by more than the join distance, plan windows and bands contain
the region's diagonals, a covering band collapses to an unbanded plan,
visit order decides coverage) and hand-worked cases.
"""

BAND_SCORE_MIN_CELLS = 1 << 20   # src/search/engine.rs:401
BAND_COST_RATIO = 12             # engine.rs:402
BAND_BRIDGE = 1024               # engine.rs:404
I32_MAX, I32_MIN = 2**31 - 1, -2**31


def visit_hits(qlen, slen, hits, xres, hsp_threshold):
    """visit_scan (src/search/engine/scan.rs:2384-2431) over precomputed
    X-drop results xres[k] = (score, q_lo, q_hi) of hits[k] = (qp, sp)."""
    cover = {}
    hsps, ext, avoided = [], 0, 0
    for (qp, sp), (score, lo, hi) in zip(hits, xres):
        d = sp - qp
        di = d + qlen
        if di < 0 or di >= qlen + slen + 2 or sp >= slen or qp >= qlen:
            continue
        if qp < cover.get(di, 0):
            avoided += 1
            continue
        ext += 1
        cover[di] = hi
        if score >= hsp_threshold:
            hsps.append((d, lo, hi, score))
    return hsps, ext, avoided


def finish_subject(subject, hsps, window_slack, hsp_threshold, region_thr, max_regions_per_subject):
    """finish_subject (src/search/engine.rs:2350-2412): regions as dicts."""
    join = 2 * window_slack
    regs = []
    for d, lo, hi, score in sorted(hsps, key=lambda h: h[0]):
        if regs and d - regs[-1]["d_max"] <= join:
            r = regs[-1]
            r["d_max"] = max(r["d_max"], d)
            r["best"] = max(r["best"], score)
            r["n"] += 1
            r["sum"] = max(I32_MIN, min(I32_MAX, r["sum"] + score))
        else:
            regs.append(dict(subject=subject, d_min=d, d_max=d, best=score, n=1, sum=score))
    if region_thr > hsp_threshold:
        regs = [r for r in regs if r["best"] >= region_thr or (r["n"] >= 2 and r["sum"] >= region_thr)]
    regs.sort(key=lambda r: (-r["best"], r["d_min"]))
    return regs[:max(max_regions_per_subject, 1)]


def region_key(r):
    """region_order (engine.rs:493-495)."""
    return (-r["best"], r["subject"], r["d_min"], r["d_max"])


def cap_regions(regs, cap):
    return sorted(regs, key=region_key)[:cap]


def window(d_min, d_max, qlen, sl, slack):
    """Engine::window (engine.rs:1223-1234)."""
    lo = min(max(d_min - slack, 0), sl)
    hi = max(min(qlen + d_max + slack, sl), lo)
    if (hi - lo) * 4 >= sl * 3:
        return 0, sl
    return lo, hi


def clip(lo, hi, m, n):
    return max(lo, -(m - 1)), min(hi, n - 1)


def covers(lo, hi, m, n):
    return lo <= -(m - 1) and hi >= n - 1


def plan_windows(regs, qlen, slens, window_slack, score_band, hsp_threshold, exhaustive=False):
    """plan_windows (engine.rs:1247-1324): (lo, hi, band or None) per region."""
    span = [(r["d_min"], r["d_max"]) for r in regs]
    by_subject = {}
    for i, r in enumerate(regs):
        by_subject.setdefault(r["subject"], []).append(i)
    for idx in by_subject.values():
        top = max(regs[i]["best"] for i in idx)
        strong = max(2 * hsp_threshold, int(top / 4) if top >= 0 else -((-top) // 4))
        idx = sorted((i for i in idx if regs[i]["best"] >= strong),
                     key=lambda i: (regs[i]["d_min"], regs[i]["d_max"], i))
        start = 0
        while start < len(idx):
            lo, hi = regs[idx[start]]["d_min"], regs[idx[start]]["d_max"]
            end = start + 1
            while end < len(idx) and regs[idx[end]]["d_min"] - hi <= BAND_BRIDGE:
                hi = max(hi, regs[idx[end]]["d_max"])
                end += 1
            for i in idx[start:end]:
                span[i] = (lo, hi)
            start = end
    out = []
    for ri, r in enumerate(regs):
        sl = slens[ri]
        lo, hi = window(r["d_min"], r["d_max"], qlen, sl, window_slack)
        own = (lo, hi, None)
        if exhaustive or r["best"] == 0 or window_slack >= (2**63 - 1) // 8:
            out.append(own)
            continue
        dmin, dmax = span[ri]
        if score_band > 0:
            w = score_band
            wlo = min(max(dmin - w, 0), sl)
            whi = max(min(qlen + dmax + w, sl), wlo)
            n = whi - wlo
            band = clip(dmin - w - wlo, dmax + w - wlo, qlen, n)
            if n == 0 or covers(*band, qlen, n):
                out.append((wlo, whi, None))
            else:
                out.append((wlo, whi, band))
            continue
        if qlen * (hi - lo) < BAND_SCORE_MIN_CELLS:
            out.append(own)
            continue
        ulo, uhi = window(dmin, dmax, qlen, sl, window_slack)
        n = uhi - ulo
        band = clip(dmin - ulo - window_slack, dmax - ulo + window_slack, qlen, n)
        if covers(*band, qlen, n):
            out.append(own)
            continue
        width = max(band[1] - band[0] + 1, 0)
        rows = min(qlen, n + width)
        if width * rows * BAND_COST_RATIO < qlen * n:
            out.append((ulo, uhi, band))
        else:
            out.append(own)
    return out


def two_stage_jobs(regs, plans, qlens, narrow_band, full_band):
    """The job lists of Engine::score_windows (engine.rs:1021-1099): narrow
    jobs (region index, band) and, without a narrow stage, the full jobs."""
    narrow, full = [], []
    for i, (r, (lo, hi, band)) in enumerate(zip(regs, plans)):
        m, n = qlens[i], hi - lo
        if band is None:
            full.append((i, (-(m - 1), n - 1)))
        elif 0 < narrow_band < full_band:
            narrow.append((i, clip(r["d_min"] - narrow_band - lo, r["d_max"] + narrow_band - lo, m, n)))
        else:
            full.append((i, band))
    return narrow, full
