"""Host pipeline (library host code) against the plain-Python restatement in
oracle/host.py, plus invariants the reference's semantics fix and
hand-worked cases. CPU only."""
import numpy as np
import pytest

from oracle import host as H
from saber_amd import pipeline as PL


def regions_as_dicts(arr):
    return [dict(subject=int(r["subject"]), d_min=int(r["d_min"]), d_max=int(r["d_max"]), best=int(r["best"]),
                 n=int(r["n_hsp"]), sum=int(r["hsp_sum"])) for r in arr]


def test_visit_coverage_hand_case():
    # three hits on diagonal 5 in visit order: the second lies inside the
    # first extension (qp 12 < q_hi 20) and is skipped; the third starts at
    # the first's end (qp 20 == q_hi) and is extended
    hits = [(10, 15), (12, 17), (20, 25), (3, 3)]
    xres = [(50, 8, 20), (99, 10, 30), (30, 20, 26), (41, 0, 9)]
    h, ext, av = PL.visit_scan(100, 100, *zip(*hits), *zip(*xres), hsp_threshold=40)
    assert (ext, av) == (3, 1)
    assert [(int(x["diag"]), int(x["score"])) for x in h] == [(5, 50), (0, 41)]
    want, e2, a2 = H.visit_hits(100, 100, hits, xres, 40)
    assert (e2, a2) == (ext, av) and [(d, s) for d, _, _, s in want] == [(5, 50), (0, 41)]


def test_visit_matches_restatement_random():
    rng = np.random.default_rng(1)
    for _ in range(200):
        qlen, slen = int(rng.integers(1, 300)), int(rng.integers(1, 300))
        n = int(rng.integers(0, 60))
        qp = rng.integers(0, qlen + 5, n)
        d = rng.integers(-5, 6, n)
        sp = np.clip(qp + d, 0, slen + 5)
        lo = np.maximum(qp - rng.integers(0, 10, n), 0)
        hi = qp + rng.integers(1, 30, n)
        sc = rng.integers(0, 90, n)
        got, ext, av = PL.visit_scan(qlen, slen, qp, sp, sc, lo, hi, 40)
        want, e2, a2 = H.visit_hits(qlen, slen, list(zip(qp.tolist(), sp.tolist())),
                                    list(zip(sc.tolist(), lo.tolist(), hi.tolist())), 40)
        assert (ext, av) == (e2, a2)
        assert [(int(x["diag"]), int(x["q_lo"]), int(x["q_hi"]), int(x["score"])) for x in got] == want


def test_finish_subject_hand_case():
    # diagonals 0, 128 join (128 <= 2*64); 257 starts a new region (257-128 > 128)
    hsps = [(128, 0, 5, 45), (0, 0, 5, 41), (257, 0, 5, 60), (300, 0, 5, 42)]
    got = regions_as_dicts(PL.finish_subject(7, hsps, 64, 40, 40, 4))
    assert got == [dict(subject=7, d_min=257, d_max=300, best=60, n=2, sum=102),
                   dict(subject=7, d_min=0, d_max=128, best=45, n=2, sum=86)]
    # region threshold 80 > hsp threshold: keep best >= 80 or (n >= 2, sum >= 80)
    got = regions_as_dicts(PL.finish_subject(7, hsps + [(1000, 0, 5, 79)], 64, 40, 80, 4))
    assert [r["d_min"] for r in got] == [257, 0]


def test_finish_subject_invariants_and_restatement():
    rng = np.random.default_rng(2)
    for _ in range(300):
        n = int(rng.integers(1, 40))
        diags = rng.integers(-400, 400, n)
        hs = [(int(d), 0, 1, int(s)) for d, s in zip(diags, rng.integers(40, 120, n))]
        slack = int(rng.integers(1, 80))
        rthr = int(rng.choice([40, 60, 100]))
        cap = int(rng.integers(1, 6))
        got = regions_as_dicts(PL.finish_subject(3, hs, slack, 40, rthr, cap))
        assert got == H.finish_subject(3, hs, slack, 40, rthr, cap)
        # without threshold and cap: a partition of the sorted diagonals
        allr = regions_as_dicts(PL.finish_subject(3, hs, slack, 40, 40, 10**6))
        allr.sort(key=lambda r: r["d_min"])
        assert sum(r["n"] for r in allr) == n
        for a, b in zip(allr, allr[1:]):
            assert b["d_min"] - a["d_max"] > 2 * slack
        for d, _, _, s in hs:
            owners = [r for r in allr if r["d_min"] <= d <= r["d_max"]]
            assert len(owners) == 1 and owners[0]["best"] >= s


def test_cap_regions_order():
    regs = [dict(subject=s, d_min=d, d_max=d + 1, best=b, n=1, sum=b)
            for s, d, b in [(2, 5, 50), (1, 9, 50), (1, 3, 50), (0, 0, 40), (3, 0, 90)]]
    arr = np.array([(r["subject"], r["d_min"], r["d_max"], r["best"], 1, r["best"], 0) for r in regs], PL.REGION_DTYPE)
    got = PL.cap_regions(arr, 4)
    assert [(int(r["subject"]), int(r["d_min"])) for r in got] == [(3, 0), (1, 3), (1, 9), (2, 5)]
    assert [(r["subject"], r["d_min"]) for r in H.cap_regions(regs, 4)] == [(3, 0), (1, 3), (1, 9), (2, 5)]


def plans_as_tuples(arr):
    return [(int(p["lo"]), int(p["hi"]), (int(p["band_lo"]), int(p["band_hi"])) if p["banded"] else None) for p in arr]


def test_plan_windows_hand_cases():
    # short pair: own window, clamped to the subject ([d_min-64, qlen+d_max+64))
    r = np.array([(0, 10, 20, 60, 1, 60, 0)], PL.REGION_DTYPE)
    assert plans_as_tuples(PL.plan_windows(r, 100, [1000])) == [(0, 184, None)]
    # window over 3/4 of the subject: the whole subject
    assert plans_as_tuples(PL.plan_windows(r, 100, [200])) == [(0, 200, None)]
    # long pair: query 3000, subject 20000, region on diagonal 5000 -> banded
    r = np.array([(0, 5000, 5000, 300, 1, 300, 0)], PL.REGION_DTYPE)
    got = plans_as_tuples(PL.plan_windows(r, 3000, [20000]))
    assert got == [(4936, 8064, (0, 128))]
    # score_band > 0: the band is [dmin - w - lo, dmax + w - lo], clipped
    got = plans_as_tuples(PL.plan_windows(r, 3000, [20000], score_band=48))
    assert got == [(4952, 8048, (0, 96))]


def test_plan_windows_restatement_and_invariants():
    rng = np.random.default_rng(4)
    for _ in range(200):
        nreg = int(rng.integers(1, 12))
        qlen = int(rng.integers(20, 4000))
        regs, slens = [], []
        for _ in range(nreg):
            subj = int(rng.integers(0, 4))
            sl = int(rng.integers(20, 30000))
            dmin = int(rng.integers(-qlen, sl))
            dmax = dmin + int(rng.integers(0, 2000))
            best = int(rng.choice([0, 40, 60, 85, 200, 400]))
            regs.append(dict(subject=subj, d_min=dmin, d_max=dmax, best=best, n=1, sum=best))
            slens.append(sl)
        arr = np.array([(r["subject"], r["d_min"], r["d_max"], r["best"], 1, r["best"], 0) for r in regs],
                       PL.REGION_DTYPE)
        for sb in (0, 48):
            got = plans_as_tuples(PL.plan_windows(arr, qlen, slens, 64, sb, 40))
            assert got == H.plan_windows(regs, qlen, slens, 64, sb, 40)
            for (lo, hi, band), sl in zip(got, slens):
                assert 0 <= lo <= hi <= sl
                if band is not None:
                    n = hi - lo
                    assert band[0] >= -(qlen - 1) and band[1] <= n - 1     # clipped
                    assert not H.covers(band[0], band[1], qlen, n)         # a covering band is unbanded
        ex = plans_as_tuples(PL.plan_windows(arr, qlen, slens, 64, 0, 40, exhaustive=True))
        assert all(b is None for _, _, b in ex)
