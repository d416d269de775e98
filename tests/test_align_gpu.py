"""Banded traceback on the GPU (saber_hip_band_align: banded_align_until,
banded.rs:472-521) against the oracle's restatement (or_band_align, pinned in
tests/test_oracle_align.py): score, coordinates, run-length operations and
cells identical, with and without the row stop at a known score; the memory
cap; the reference's own cases (banded.rs:565-767)."""
import numpy as np
import pytest

import saber_amd as S
from oracle import oracle as O
from saber_amd import BLOSUM62, PAM30, workload
from tests.gen import Lcg, flat, jobs_array, mutate, rand_seq

pytestmark = pytest.mark.gpu

W = 17


def oracle_align(qs, ss, jobs, mat, go, ge, qbias=None, stop=None, cap=32 << 20):
    out = []
    for k, J in enumerate(jobs):
        q = qs[J["query"]]
        s = ss[J["subject"]][J["win_lo"]:J["win_hi"]]
        b = None if qbias is None else qbias[J["query"]]
        st = 2**31 - 1 if stop is None else int(stop[k])
        out.append(O.band_align(q, s, mat, b, go, ge, int(J["band_lo"]), int(J["band_hi"]), st, cap))
    return out


def run(ctx, qs, ss, jobs, mat, go, ge, qbias=None, stop=None, cap=32 << 20):
    qres, qoff = flat(qs)
    sres, soff = flat(ss)
    qb = None if qbias is None else np.concatenate(qbias).astype(np.int8)
    return ctx.band_align(qres, qoff, sres, soff, jobs, S.Scheme(mat, go, ge), stop, qb, cap)


def check(got, want):
    assert len(got) == len(want)
    for k, (g, w) in enumerate(zip(got, want)):
        assert g == w, (k, g, w)


@pytest.mark.parametrize("gaps", [(11, 1), (10, 2), (6, 3), (8, 8)])
def test_random_pairs_and_bands(ctx, gaps):
    go, ge = gaps
    r = Lcg(2024 + go)
    qs, ss, bias = [], [], []
    for t in range(40):
        a = rand_seq(r, 1 + r.next(300 if t % 7 else 700))
        b = mutate(r, a) if t % 2 == 0 else rand_seq(r, 1 + r.next(260))
        qs.append(a)
        ss.append(b if len(b) else rand_seq(r, 5))
        bias.append(np.array([r.next(9) - 4 for _ in range(len(a))], np.int8))
    rows = []
    for t in range(40):
        m, n = len(qs[t]), len(ss[t])
        lo = -r.next(60)
        hi = r.next(60) + (300 if t % 5 == 0 else 0)
        for band in ((lo, hi), (-(m - 1), n - 1), (5, 40), (-2, 2)):
            rows.append((t, t, 0, n, band[0], band[1]))
    jobs = jobs_array(*zip(*rows))
    check(run(ctx, qs, ss, jobs, BLOSUM62, go, ge), oracle_align(qs, ss, jobs, BLOSUM62, go, ge))
    check(run(ctx, qs, ss, jobs, BLOSUM62, go, ge, qbias=bias),
          oracle_align(qs, ss, jobs, BLOSUM62, go, ge, qbias=bias))


def test_stop_at_known_score_and_windows(ctx):
    """Engine::traceback's call (engine.rs:1360): the window's exact score as
    stop_at; subject windows inside longer subjects."""
    r = Lcg(5)
    qs, ss, rows, stop = [], [], [], []
    for t in range(150):
        la = 20 + r.next(300)
        a = rand_seq(r, la)
        b = np.concatenate([rand_seq(r, 40), mutate(r, a[:la // 2]), rand_seq(r, 40)]).astype(np.uint8)
        qs.append(a)
        ss.append(b)
        lo_w = r.next(20)
        rows.append((t, t, lo_w, len(b), 0, 90))
    jobs = jobs_array(*zip(*rows))
    full = oracle_align(qs, ss, jobs, BLOSUM62, 11, 1)
    stop = np.array([f[0] if isinstance(f, tuple) else 2**31 - 1 for f in full], np.int32)
    got = run(ctx, qs, ss, jobs, BLOSUM62, 11, 1, stop=stop)
    check(got, oracle_align(qs, ss, jobs, BLOSUM62, 11, 1, stop=stop))
    saved = sum(f[6] - g[6] for f, g in zip(full, got) if isinstance(f, tuple))
    assert saved > 0
    for f, g in zip(full, got):
        if isinstance(f, tuple):
            assert g[:6] == f[:6]


def test_cap_known_answers_and_empty(ctx):
    a = np.full(1000, W, np.uint8)
    b = np.full(4000, W, np.uint8)
    wag = S.encode(b"WAGNER")
    qs = [a, b, wag, np.zeros(0, np.uint8), rand_seq(Lcg(3), 50)]
    ss = [a, b, wag, wag, rand_seq(Lcg(4), 60)]
    jobs = jobs_array([0, 1, 2, 3, 4, 4], [0, 1, 2, 3, 4, 4], [0, 0, 0, 0, 0, 0], [1000, 4000, 6, 6, 60, 60],
                      [-500, -20, -5, -5, 10, 100], [500, 20, 5, 5, 0, 120])
    got = run(ctx, qs, ss, jobs, BLOSUM62, 11, 1, cap=500_000)
    assert got[0] == "cap"                                   # 1000 x 1001 bytes; 4000 x 41 fits
    assert got[1][:6] == (44000, 0, 4000, 0, 4000, [(b"M", 4000)])
    assert got[2][:6] == (37, 0, 6, 0, 6, [(b"M", 6)])
    assert got[3] is None and got[4] is None and got[5] is None
    check(got, oracle_align(qs, ss, jobs, BLOSUM62, 11, 1, cap=500_000))


def test_workload_regions(ctx):
    d = workload.band_regions(seed=17, n=3000)
    qres, qoff, sres, soff, jobs = d["qres"], d["qoff"], d["sres"], d["soff"], d["jobs"]
    qs = [qres[qoff[i]:qoff[i + 1]] for i in range(len(qoff) - 1)]
    ss = [sres[soff[i]:soff[i + 1]] for i in range(len(soff) - 1)]
    got = ctx.band_align(qres, qoff, sres, soff, jobs, S.Scheme(BLOSUM62, 11, 1))
    check(got, oracle_align(qs, ss, jobs, BLOSUM62, 11, 1))
    assert sum(isinstance(g, tuple) for g in got) > 1000


@pytest.mark.parametrize("half,budget_mb", [(1800, 200), (400, 16)])
def test_jobs_spanning_several_launches(ctx, monkeypatch, half, budget_mb):
    """Direction matrices of ~10 MB (one thread per job, band width 3601) or
    ~3 MB (64 lanes per job, width 801) each, under a scratch budget of a few
    of them: the jobs run in several launches, results placed by job."""
    monkeypatch.setenv("SABER_HIP_ALIGN_SCRATCH_MB", str(budget_mb))
    r = Lcg(77)
    qs, ss, rows = [], [], []
    for t in range(52):
        a = rand_seq(r, 2800 + r.next(200))
        b = mutate(r, a)
        qs.append(a)
        ss.append(b if len(b) else a)
        rows.append((t, t, 0, len(ss[-1]), -half, half))
    jobs = jobs_array(*zip(*rows))
    check(run(ctx, qs, ss, jobs, PAM30, 9, 1), oracle_align(qs, ss, jobs, PAM30, 9, 1))


@pytest.mark.parametrize("stop_known", [False, True])
def test_long_jobs_four_cell_lanes(ctx, stop_known):
    """Jobs longer than the batch's mean lane load (and 64 row steps) with
    band widths up to 256 run 4 cells per lane on 4 to 64 lanes, the others
    16 cells per lane: widths 1 .. 257, jobs of 20 .. 1200 rows (both forms
    in one launch), homologous and random pairs, with query bias and the row
    stop."""
    r = Lcg(91)
    qs, ss, bias, rows = [], [], [], []
    widths = [1, 3, 16, 17, 32, 33, 64, 65, 128, 129, 200, 256, 257]
    t = 0
    for w in widths:
        for m in (20 + r.next(30), 383, 384, 700 + r.next(500)):
            a = rand_seq(r, m)
            b = mutate(r, a) if t % 3 else rand_seq(r, m + r.next(40))
            qs.append(a)
            ss.append(b if len(b) else a)
            bias.append(np.array([r.next(7) - 3 for _ in range(m)], np.int8))
            lo = -(w // 2) - r.next(3)
            rows.append((t, t, 0, len(ss[-1]), lo, lo + w - 1))
            t += 1
    jobs = jobs_array(*zip(*rows))
    want = oracle_align(qs, ss, jobs, BLOSUM62, 11, 1, qbias=bias)
    stop = None
    if stop_known:
        stop = np.array([f[0] if isinstance(f, tuple) else 2**31 - 1 for f in want], np.int32)
        want = oracle_align(qs, ss, jobs, BLOSUM62, 11, 1, qbias=bias, stop=stop)
    check(run(ctx, qs, ss, jobs, BLOSUM62, 11, 1, qbias=bias, stop=stop), want)
    assert sum(isinstance(g, tuple) for g in want) > len(widths) * 2
