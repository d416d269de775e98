"""Engine::traceback (src/search/engine.rs:1328-1393) with its fallback
traceback_from_end (:1403-1446), batched on the GPU (saber_hip_traceback),
against the same loop restated over the oracle's band_align
(banded_align_until, pinned in tests/test_oracle_align.py) and banded_score:
per hit the alignment, the exact flag, and the engine's counters summed over
the batch, identical. Cases: region bands too narrow for the alignment's
indels (widening), the scoring band as a limit, best == 0 and exhaustive
(full window), query bias, and direction caps that send hits to the
end-anchored fallback, which either succeeds or gives up (band_capped)."""
import numpy as np
import pytest

import saber_amd as S
from oracle import oracle as O
from saber_amd import BLOSUM62
from tests.gen import Lcg, flat, rand_seq

pytestmark = pytest.mark.gpu

KEYS = ("tracebacks", "band_widenings", "band_capped", "end_anchored_tracebacks", "banded_cells")


def oracle_from_end(q, s, bias, go, ge, want, limit, band_pad, cap, cnt):
    m, n = len(q), len(s)
    scope = limit if limit is not None else (-m, n)
    sc, qe, se, cells = O.banded_score(q, s, bias, BLOSUM62, go, ge, scope[0], scope[1])
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
        r = O.band_align(q, s, BLOSUM62, bias, go, ge, lo, hi, target, cap)
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


def oracle_traceback(q, s, bias, go, ge, d_min, d_max, win_lo, best, want, limit, band_pad, exhaustive, cap, cnt):
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
        r = O.band_align(q, s, BLOSUM62, bias, go, ge, lo, hi, want, cap)
        if r == "cap":
            found = oracle_from_end(q, s, bias, go, ge, want, limit, band_pad, cap, cnt)
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


def indel_copy(r, a):
    """a with substitutions and a few indels of 1-12 residues (the best
    alignment leaves the diagonal it starts on)."""
    out, i = [], 0
    while i < len(a):
        x = r.next(100)
        if x < 3:
            i += 1 + r.next(12)             # deletion
        elif x < 6:
            out.extend(rand_seq(r, 1 + r.next(12)))   # insertion
        elif x < 25:
            out.append(r.next(20))
            i += 1
        else:
            out.append(int(a[i]))
            i += 1
    return np.array(out if out else [0], np.uint8)


def make_case(seed, nhits, limit_every=3, zero_best_every=7):
    r = Lcg(seed)
    qs, ss, rows, biases = [], [], [], []
    for t in range(nhits):
        a = rand_seq(r, 40 + r.next(360))
        core = indel_copy(r, a)
        left, right = rand_seq(r, 10 + r.next(60)), rand_seq(r, 10 + r.next(60))
        subj = np.concatenate([rand_seq(r, 30), left, core, right, rand_seq(r, 30)]).astype(np.uint8)
        qs.append(a)
        ss.append(subj)
        biases.append(np.array([r.next(5) - 2 for _ in range(len(a))], np.int8))
        win_lo = r.next(30)
        win_hi = len(subj) - r.next(30)
        d0 = 30 + len(left)                 # subject - query diagonal where the copy starts
        spread = r.next(3)
        rows.append(dict(query=t, subject=t, win_lo=win_lo, win_hi=win_hi, d_min=d0 - spread, d_max=d0 + spread,
                         best=0 if t % zero_best_every == 5 else 50 + r.next(100),
                         has_limit=int(t % limit_every == 1),
                         limit_lo=d0 - win_lo - 20 - r.next(40), limit_hi=d0 - win_lo + 20 + r.next(40)))
    return qs, ss, rows, biases


def run_case(ctx, qs, ss, rows, biases, band_pad, exhaustive, cap, use_bias, go=11, ge=1, use_db=False):
    jobs = np.zeros(len(rows), S.TRACEBACK_JOB)
    cnt = dict.fromkeys(KEYS, 0)
    want_res = []
    for k, row in enumerate(rows):
        q = qs[row["query"]]
        s = ss[row["subject"]][row["win_lo"]:row["win_hi"]]
        b = biases[k] if use_bias else None
        limit = (row["limit_lo"], row["limit_hi"]) if row["has_limit"] else None
        scope = limit if limit is not None else (-len(q), len(s))
        want = O.banded_score(q, s, b, BLOSUM62, go, ge, scope[0], scope[1])[0]
        for f in row:
            jobs[k][f] = row[f]
        jobs[k]["want"] = want
        want_res.append(oracle_traceback(q, s, b, go, ge, row["d_min"], row["d_max"], row["win_lo"], row["best"],
                                         want, limit, band_pad, exhaustive, cap, cnt))
    qres, qoff = flat(qs)
    sres, soff = flat(ss)
    qb = np.concatenate(biases).astype(np.int8) if use_bias else None
    if use_db:
        ctx.upload_db(sres, soff)
    got, gcnt = ctx.traceback(qres, qoff, None if use_db else sres, None if use_db else soff, jobs,
                              S.Scheme(BLOSUM62, go, ge), band_pad=band_pad, exhaustive=exhaustive, qbias=qb,
                              max_dir_bytes=cap)
    for k, (g, w) in enumerate(zip(got, want_res)):
        assert g == w, (k, rows[k], g, w)
    assert gcnt == cnt, (gcnt, cnt)
    return cnt


@pytest.mark.parametrize("use_bias", [False, True])
def test_widening_limits_and_full_windows(ctx, use_bias):
    qs, ss, rows, biases = make_case(31 + use_bias, 60)
    cnt = run_case(ctx, qs, ss, rows, biases, band_pad=2, exhaustive=False, cap=32 << 20, use_bias=use_bias)
    assert cnt["band_widenings"] > 20          # narrow region bands had to widen
    assert cnt["band_capped"] == 0 and cnt["end_anchored_tracebacks"] == 0


def test_exhaustive_takes_the_full_window(ctx):
    qs, ss, rows, biases = make_case(41, 25)
    cnt = run_case(ctx, qs, ss, rows, biases, band_pad=8, exhaustive=True, cap=32 << 20, use_bias=False)
    assert cnt["band_widenings"] == 0 and cnt["tracebacks"] == 25


@pytest.mark.parametrize("cap", [6000, 20000, 60000])
def test_cap_sends_hits_to_the_end_cell(ctx, cap):
    """Direction caps of a few thousand bytes: wide attempts fail, the hit
    falls back to a score pass and a band around its end cell's diagonal,
    which succeeds or in turn reaches the cap (band_capped, the best
    alignment so far returned as not exact)."""
    qs, ss, rows, biases = make_case(51, 50, limit_every=2)
    cnt = run_case(ctx, qs, ss, rows, biases, band_pad=4, exhaustive=False, cap=cap, use_bias=True)
    assert cnt["end_anchored_tracebacks"] > 0
    if cap == 6000:
        assert cnt["band_capped"] > 0


def test_uploaded_database(ctx):
    """The engine's call: subjects are the uploaded database (subjects ==
    NULL), including the end-cell fallback's score passes."""
    qs, ss, rows, biases = make_case(61, 40, limit_every=2)
    cnt = run_case(ctx, qs, ss, rows, biases, band_pad=16, exhaustive=False, cap=20000, use_bias=False,
                   use_db=True)
    assert cnt["tracebacks"] >= 40 and cnt["end_anchored_tracebacks"] > 0
