"""The band planner on the device (band_plan.hip) against the host planner
and the oracle: the same scores, end cells and cell counts for every job, the
same work per kernel variant (the same buckets and pairs), and the same
errors for bad jobs. Contexts are forced to one planner (plan="host" /
plan="device"); by default batches of 65,536 jobs and more are planned on
the device, so the 10^6-region tests of test_band_gpu.py run it at scale."""
import numpy as np
import pytest

import saber_amd as S
from oracle import oracle as O
from saber_amd import BLOSUM62, workload
from tests.gen import Lcg, flat, jobs_array, mutate, rand_seq
from tests.test_band_gpu import random_cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctxs():
    h = S.Context(0, plan="host")
    d = S.Context(0, plan="device")
    yield h, d
    d.close()
    h.close()


def mixed_set(seed):
    """Random and related pairs with every band kind (wide bands for the
    int32 kernel, bands outside the rectangle, windows narrower than their
    band), queries shared by runs of jobs with the same first row (the
    pair-table flavour, odd runs included), transposable jobs, and a few
    many-row jobs (the latency buckets) and jobs beyond the packed kernel's
    32,767 rows."""
    qs, ss, lo, hi, bias = random_cases(seed, 400, 300, (11, 1), wide=True)
    r = Lcg(seed + 1)
    jq = list(range(len(qs)))
    # runs of 1-7 jobs on one query and one first row
    for t in range(120):
        q = int(r.next(len(qs)))
        for _ in range(1 + r.next(7)):
            b = mutate(r, qs[q])
            ss.append(b if len(b) > 4 else rand_seq(r, 40))
            jq.append(q)
            lo.append(-3 - int(r.next(30)))
            hi.append(12 + int(r.next(30)))
    # long rows: latency buckets, and the int32 route beyond 32,767 rows
    for m in (3000, 5000, 34000):
        a = rand_seq(r, m)
        qs.append(a)
        bias.append(np.array([r.next(9) - 4 for _ in range(m)], np.int8))
        ss.append(mutate(r, a))
        jq.append(len(qs) - 1)
        lo.append(-20)
        hi.append(20)
    qres, qoff = flat(qs)
    sres, soff = flat(ss)
    n = len(ss)
    jobs = jobs_array(np.asarray(jq), np.arange(n), np.zeros(n), [len(s) for s in ss], lo, hi)
    qb, _ = flat([np.asarray(b, np.int8).view(np.uint8) for b in bias])
    return qres, qoff, sres, soff, jobs, qb.view(np.int8)


def same(a, b, what):
    for name, x, y in zip(("score", "q_end", "s_end", "cells"), a, b):
        if x is None and y is None:
            continue
        bad = np.nonzero(np.asarray(x).astype(np.int64) != np.asarray(y).astype(np.int64))[0]
        assert len(bad) == 0, f"{what}: {name} differs at jobs {bad[:10].tolist()}"


@pytest.mark.parametrize("bias", [False, True])
@pytest.mark.parametrize("ends", [True, False])
def test_device_plan_matches_host_plan_and_oracle(ctxs, bias, ends):
    h, d = ctxs
    qres, qoff, sres, soff, jobs, qb = mixed_set(21)
    qb = qb if bias else None
    sch = S.Scheme(BLOSUM62, 11, 1)
    gh = h.score_bands(qres, qoff, sres, soff, jobs, sch, qb, ends=ends)
    gd = d.score_bands(qres, qoff, sres, soff, jobs, sch, qb, ends=ends)
    same(gd, gh, "device vs host plan")
    want = O.banded_many(qres, qoff, sres, soff, jobs, BLOSUM62, 11, 1, qb)
    same(gd, want if ends else (want[0], None, None, want[3]), "device plan vs oracle")


def test_device_plan_same_work_per_variant(ctxs):
    """Both planners put the same cells on every kernel variant (same
    buckets, shapes and transpositions), on the region workload (1000 jobs
    per query: shared pairs) and on the mixed set."""
    h, d = ctxs
    sch = S.Scheme(BLOSUM62, 11, 1)
    reg = workload.band_regions(seed=3, n=30000, per_query=1000)
    cases = [(reg["qres"], reg["qoff"], reg["sres"], reg["soff"], reg["jobs"], None)]
    cases.append(mixed_set(5))
    for qres, qoff, sres, soff, jobs, qb in cases:
        for ends in (True, False):
            per = []
            for c in (h, d):
                b = c.band_batch(qres, qoff, sres, soff, jobs, sch, qbias=qb, ends=ends)
                try:
                    _, _, variant_cells, cells = b.time(warmup=0, iters=1)
                    per.append((variant_cells, cells, b.read()))
                finally:
                    b.close()
            assert per[0][0] == per[1][0], f"ends={ends}: cells per variant differ: {per[0][0]} vs {per[1][0]}"
            assert per[0][1] == per[1][1]
            same(per[1][2], per[0][2], f"ends={ends}")


def test_device_plan_database_subjects_and_latency_rule(ctxs):
    """Subjects from the uploaded database, and a fixed many-row limit
    (saber_hip_band_tune) applied alike by both planners."""
    h, d = ctxs
    qres, qoff, sres, soff, jobs, qb = mixed_set(9)
    sch = S.Scheme(BLOSUM62, 11, 1)
    want = O.banded_many(qres, qoff, sres, soff, jobs, BLOSUM62, 11, 1, qb)
    for c in (h, d):
        c.upload_db(sres, soff)
    try:
        for rows in (-1, 120, 0):
            h.tune(rows)
            d.tune(rows)
            gh = h.score_bands(qres, qoff, None, None, jobs, sch, qb)
            gd = d.score_bands(qres, qoff, None, None, jobs, sch, qb)
            same(gd, gh, f"database subjects, latency rows {rows}")
            bh = h.band_batch(qres, qoff, None, None, jobs, sch, qbias=qb, ends=False)
            bd = d.band_batch(qres, qoff, None, None, jobs, sch, qbias=qb, ends=False)
            try:
                ph, pd = bh.time(warmup=0, iters=1), bd.time(warmup=0, iters=1)
                assert ph[2] == pd[2], f"latency rows {rows}: {ph[2]} vs {pd[2]}"
                rd, rh = bd.read(), bh.read()
                # each planner's batch against the oracle first, so that a
                # difference names the side that deviates
                for name, r in (("host", rh), ("device", rd)):
                    same((r[0], None, None, r[3]), (want[0], None, None, want[3]),
                         f"{name} plan vs oracle, scores only, latency rows {rows}")
                same(rd, rh, f"scores only, latency rows {rows}")
            finally:
                bh.close()
                bd.close()
    finally:
        h.tune(-1)
        d.tune(-1)
    same(gd, want, "database subjects vs oracle")


def test_device_plan_small_batches(ctxs):
    """0, 1, 2 and 3 jobs (odd pairs, a single chunk), and batches whose jobs
    are all empty or all for the int32 kernel."""
    h, d = ctxs
    sch = S.Scheme(BLOSUM62, 11, 1)
    r = Lcg(77)
    a = rand_seq(r, 90)
    qs = [a, rand_seq(r, 50), a]
    ss = [mutate(r, a), rand_seq(r, 70), mutate(r, a)]
    qres, qoff = flat(qs)
    sres, soff = flat(ss)
    for n in range(4):
        jobs = jobs_array(np.arange(n), np.arange(n), np.zeros(n), [len(s) for s in ss[:n]], [-5] * n, [9] * n)
        gd = d.score_bands(qres, qoff, sres, soff, jobs, sch)
        gh = h.score_bands(qres, qoff, sres, soff, jobs, sch)
        same(gd, gh, f"{n} jobs")
    empty = jobs_array(np.arange(3), np.arange(3), np.zeros(3), [len(s) for s in ss], [500] * 3, [600] * 3)
    wide = jobs_array(np.arange(3), np.arange(3), np.zeros(3), [len(s) for s in ss], [-400] * 3, [400] * 3)
    for jobs, what in ((empty, "empty bands"), (wide, "int32 kernel only")):
        gd = d.score_bands(qres, qoff, sres, soff, jobs, sch)
        same(gd, h.score_bands(qres, qoff, sres, soff, jobs, sch), what)
        want = O.banded_many(qres, qoff, sres, soff, jobs, BLOSUM62, 11, 1)
        same(gd, want, what + " vs oracle")


def test_device_plan_rejects_bad_jobs_like_the_host(ctxs):
    """The first bad job decides the error, with the host planner's message."""
    h, d = ctxs
    sch = S.Scheme(BLOSUM62, 11, 1)
    r = Lcg(3)
    qs = [rand_seq(r, 40) for _ in range(3)]
    ss = [rand_seq(r, 60) for _ in range(3)]
    qres, qoff = flat(qs)
    sres, soff = flat(ss)
    base = jobs_array(np.arange(3), np.arange(3), np.zeros(3), [60] * 3, [-5] * 3, [5] * 3)
    cases = []
    j = base.copy(); j["query"][2] = 3; cases.append((j, "query index"))
    j = base.copy(); j["subject"][1] = 7; cases.append((j, "subject index"))
    j = base.copy(); j["win_hi"][0] = 61; cases.append((j, "window"))
    j = base.copy(); j["win_lo"][1] = 30; j["win_hi"][1] = 20; cases.append((j, "window"))
    j = base.copy(); j["win_hi"][0] = 61; j["query"][1] = 9; cases.append((j, "window"))
    j = base.copy(); j["subject"][2] = 9; j["win_hi"][1] = 99; cases.append((j, "window"))
    for jobs, what in cases:
        msgs = []
        for c in (h, d):
            with pytest.raises(S.HipError) as e:
                c.score_bands(qres, qoff, sres, soff, jobs, sch)
            msgs.append(str(e.value))
        assert msgs[0] == msgs[1], msgs
        assert what in msgs[1], msgs
