"""Banded / full SW on the GPU vs the oracle: bit-exact scores, end cells and
cell counts. Cases follow the reference's tests (banded.rs:565-731,
band_simd8.rs:403-468, sw_simd.rs:488-550) and the batched-region workload."""
import numpy as np
import pytest

import saber_amd as S
from oracle import oracle as O
from saber_amd import BLOSUM45, BLOSUM62, PAM30, workload
from tests.gen import Lcg, flat, jobs_array, mutate, rand_seq

pytestmark = pytest.mark.gpu

W = 17


def run_both(ctx, qs, ss, lo, hi, mat, go, ge, bias=None):
    """One job per (q, s) pair; GPU and oracle results as tuples of arrays."""
    qres, qoff = flat(qs)
    sres, soff = flat(ss)
    n = len(qs)
    jobs = jobs_array(np.arange(n), np.arange(n), np.zeros(n), [len(s) for s in ss], lo, hi)
    qb = None
    if bias is not None:
        qb, _ = flat([np.asarray(b, np.int8).view(np.uint8) for b in bias])
        qb = qb.view(np.int8)
    sch = S.Scheme(mat, go, ge)
    got = ctx.score_bands(qres, qoff, sres, soff, jobs, sch, qb)
    want = O.banded_many(qres, qoff, sres, soff, jobs, mat, go, ge, qb)
    check_scores_only(ctx, qres, qoff, sres, soff, jobs, sch, qb, want)
    return got, want


def check_scores_only(ctx, qres, qoff, sres, soff, jobs, sch, qb, want):
    """The scores-only build (score_banded_many's output) gives the oracle's
    scores on the same jobs."""
    sc = ctx.score_bands(qres, qoff, sres, soff, jobs, sch, qb, ends=False)[0]
    bad = np.nonzero(sc.astype(np.int64) != np.asarray(want[0]).astype(np.int64))[0]
    assert len(bad) == 0, f"scores-only build differs at jobs {bad[:10].tolist()}"


def assert_same(got, want, what=""):
    names = ("score", "q_end", "s_end", "cells")
    for name, g, w in zip(names, got, want):
        bad = np.nonzero(np.asarray(g).astype(np.int64) != np.asarray(w).astype(np.int64))[0]
        assert len(bad) == 0, f"{what} {name} differs at jobs {bad[:10].tolist()}: got {np.asarray(g)[bad[:5]]} want {np.asarray(w)[bad[:5]]}"


def random_cases(seed, count, maxlen, gaps, bias_every=3, wide=False):
    r = Lcg(seed)
    qs, ss, lo, hi, bias = [], [], [], [], []
    for t in range(count):
        la = 1 + r.next(maxlen)
        a = rand_seq(r, la)
        lb = 1 + r.next(maxlen)
        b = mutate(r, a) if t % 2 == 0 else rand_seq(r, lb)
        if len(b) == 0:
            b = rand_seq(r, 3)
        qs.append(a)
        ss.append(b)
        bias.append(np.array([r.next(9) - 4 for _ in range(la)], np.int8))
        kind = t % 4
        if kind == 0:
            lo.append(-la)
            hi.append(len(b))
        elif kind == 1:
            lo.append(-int(r.next(60)))
            hi.append(int(r.next(60)) + (300 if wide and t % 5 == 0 else 0))
        elif kind == 2:
            lo.append(5)
            hi.append(40)
        else:
            c = int(r.next(len(b) + la)) - la
            h = [8, 16, 32, 64][r.next(4)]
            lo.append(c - h)
            hi.append(c + h)
    return qs, ss, lo, hi, bias


@pytest.mark.parametrize("gaps", [(11, 1), (10, 2), (6, 3), (15, 1), (8, 8)])
def test_random_and_related_pairs(ctx, gaps):
    # banded.rs:645-671 (simd_and_scalar_kernels_agree): 5 gap pairs, bias,
    # wide and full bands
    qs, ss, lo, hi, bias = random_cases(2024 + gaps[0], 300, 220, gaps, wide=True)
    got, want = run_both(ctx, qs, ss, lo, hi, BLOSUM62, *gaps)
    assert_same(got, want, f"no bias {gaps}")
    got, want = run_both(ctx, qs, ss, lo, hi, BLOSUM62, *gaps, bias=bias)
    assert_same(got, want, f"bias {gaps}")


@pytest.mark.parametrize("mat", [BLOSUM62, PAM30, BLOSUM45], ids=["B62", "PAM30", "B45"])
def test_u8_band_cases_all_matrices(ctx, mat):
    # band_simd8.rs:403-451: one query, many subjects with related stretches,
    # bands of width up to 200
    r = Lcg(7)
    for trial in range(6):
        ql = 1 + r.next(220)
        q = rand_seq(r, ql)
        bias = np.array([r.next(9) - 4 for _ in range(ql)], np.int8)
        subs = []
        for k in range(1 + r.next(70)):
            l = 1 + r.next(320)
            s = rand_seq(r, l)
            if k % 3 == 0:
                o = r.next(l)
                for t in range(min(max(l - o, 0), ql)):
                    if r.next(4) > 0:
                        s[o + t] = q[t]
            subs.append(s)
        lo, hi = [], []
        for s in subs:
            c = int(r.next(len(s) + ql)) - ql
            w = int(r.next(200 if trial % 5 == 0 else 90))
            lo.append(c - w // 2)
            hi.append(c + (w + 1) // 2)
        for go, ge in [(11, 1), (7, 2)]:
            b = [bias] * len(subs) if trial % 2 == 0 else None
            got, want = run_both(ctx, [q] * len(subs), subs, lo, hi, mat, go, ge, b)
            assert_same(got, want, f"trial {trial} gaps {go}/{ge}")


def test_exact32_kernel_matches_oracle(ctx32):
    qs, ss, lo, hi, bias = random_cases(99, 200, 300, (11, 1), wide=True)
    got, want = run_both(ctx32, qs, ss, lo, hi, BLOSUM62, 11, 1, bias=bias)
    assert_same(got, want, "exact32")


def test_saturation_rescored_exactly(ctx):
    # Scores above the int16 range: W x 4000 gives 44 000 (banded.rs:719-730),
    # W x 300 band +-5 gives 3300 next to an unrelated job (band_simd8.rs:454-468)
    a = np.full(4000, W, np.uint8)
    other = S.encode("MKVLAAGIVGKLLA")
    w300 = np.full(300, W, np.uint8)
    qs = [a, a, w300, w300, a[:3500]]
    ss = [a, a, w300, other, a[:3500]]
    lo = [-20, -4000, -5, -3, -3500]
    hi = [20, 4000, 5, 3, 3500]
    got, want = run_both(ctx, qs, ss, lo, hi, BLOSUM62, 11, 1)
    assert_same(got, want, "saturation")
    score, q_end, s_end, _ = got
    assert score.tolist() == [44000, 44000, 3300, want[0][3], 38500]
    assert (q_end[0], s_end[0]) == (4000, 4000)


def test_edge_cases(ctx):
    r = Lcg(11)
    qs, ss, lo, hi = [], [], [], []
    for t in range(400):
        m = 1 + r.next(40)
        n = 1 + r.next(40)
        qs.append(rand_seq(r, m))
        b = rand_seq(r, n) if t % 2 else mutate(r, qs[-1])
        ss.append(b if len(b) else rand_seq(r, 2))
        kind = t % 8
        if kind == 0:       # band entirely right of the rectangle
            lo.append(len(ss[-1]) + 3); hi.append(len(ss[-1]) + 10)
        elif kind == 1:     # entirely left
            lo.append(-m - 10); hi.append(-m - 1)
        elif kind == 2:     # single diagonal
            d = int(r.next(n + m)) - m
            lo.append(d); hi.append(d)
        elif kind == 3:     # hi < lo
            lo.append(4); hi.append(2)
        elif kind == 4:     # clipped on both sides
            lo.append(-1000); hi.append(1000)
        elif kind == 5:     # touching the corners
            lo.append(len(ss[-1]) - 1); hi.append(len(ss[-1]) + 50)
        elif kind == 6:
            lo.append(-m - 50); hi.append(-(m - 1))
        else:
            c = int(r.next(n + m)) - m
            h = int(r.next(70))
            lo.append(c - h); hi.append(c + h)
    got, want = run_both(ctx, qs, ss, lo, hi, BLOSUM62, 11, 1)
    assert_same(got, want, "edges")
    # out-of-alphabet codes score as X
    q = np.array([0, 1, 2, 200, 4], np.uint8)
    s1 = np.array([0, 1, 2, 22, 4], np.uint8)
    s2 = np.array([255, 255], np.uint8)
    got, want = run_both(ctx, [q, q], [s1, s2], [-5, -5], [5, 2], BLOSUM62, 11, 1)
    assert_same(got, want, "codes")
    # more rows than the packed kernel keeps (32767): the int32 route
    a = rand_seq(Lcg(5), 33000)
    b = mutate(Lcg(6), a)
    got, want = run_both(ctx, [a, a[:300]], [b, b[:300]], [-3, -3], [3, 3], BLOSUM62, 11, 1)
    assert_same(got, want, "long rows")


def test_every_variant_width(ctx):
    """Each width 1..340 (all packed variants and the int32 route above 328)
    on related pairs long enough to fill the band."""
    r = Lcg(123)
    qs, ss, lo, hi = [], [], [], []
    for w in range(1, 341):
        for _ in range(2):
            m = 60 + r.next(200)
            a = rand_seq(r, m)
            b = mutate(r, a)
            if len(b) < 5:
                b = rand_seq(r, 80)
            c = int(r.next(21)) - 10
            qs.append(a); ss.append(b)
            lo.append(c - w // 2); hi.append(c - w // 2 + w - 1)
    got, want = run_both(ctx, qs, ss, lo, hi, BLOSUM62, 11, 1)
    assert_same(got, want, "widths")


def run_shared(ctx, qs, ss, jq, lo, hi, mat, go, ge, bias=None):
    """Jobs (query jq[k], subject k, band [lo, hi]) where several jobs share a
    query: pairs of jobs with the same query and first row use the kernel's
    pair-table flavour, the rest the per-job flavour."""
    qres, qoff = flat(qs)
    sres, soff = flat(ss)
    n = len(ss)
    jobs = jobs_array(np.asarray(jq), np.arange(n), np.zeros(n), [len(s) for s in ss], lo, hi)
    qb = None
    if bias is not None:
        qb, _ = flat([np.asarray(b, np.int8).view(np.uint8) for b in bias])
        qb = qb.view(np.int8)
    sch = S.Scheme(mat, go, ge)
    got = ctx.score_bands(qres, qoff, sres, soff, jobs, sch, qb)
    want = O.banded_many(qres, qoff, sres, soff, jobs, mat, go, ge, qb)
    check_scores_only(ctx, qres, qoff, sres, soff, jobs, sch, qb, want)
    return got, want


def test_each_compiled_variant():
    """Every compiled packed variant, both kernel flavours (jobs sharing a
    query and first row, and not), with and without bias, forced over the
    widths it serves: the planner's cost model leaves some shapes unused by
    default."""
    r = Lcg(77)
    shapes = sorted({(L, C, K) for L, C, K, _ in S.band_variants()})
    for L, C, K in shapes:
        W = L * C
        lo_w = 1 if K == C else W - K + 1
        widths = sorted({lo_w, (lo_w + W) // 2, W})
        ctx = S.Context()
        ctx.choose(lo_w, W, L, C, K)
        qs, ss, jq, lo, hi, bias = [], [], [], [], [], []
        for w in widths:
            # long enough that the band is not clipped to a narrower width
            m = W // 2 + 40 + r.next(200)
            a = rand_seq(r, m)
            qs.append(a)
            bias.append(np.array([r.next(9) - 4 for _ in range(m)], np.int8))
            for t in range(7):
                if t == 6:          # an odd one out, alone on its own query
                    a = rand_seq(r, m)
                    qs.append(a)
                    bias.append(np.array([r.next(9) - 4 for _ in range(m)], np.int8))
                b = mutate(r, a) if t % 3 else rand_seq(r, m + r.next(40))
                if len(b) < W // 2 + 20:
                    b = rand_seq(r, m)
                c = int(r.next(31)) - 15
                ss.append(b)
                jq.append(len(qs) - 1)
                lo.append(c - w // 2); hi.append(c - w // 2 + w - 1)
        got, want = run_shared(ctx, qs, ss, jq, lo, hi, BLOSUM62, 11, 1)
        assert_same(got, want, f"variant {(L, C, K)}")
        got, want = run_shared(ctx, qs, ss, jq, lo, hi, BLOSUM62, 11, 1, bias=bias)
        assert_same(got, want, f"variant {(L, C, K)} bias")
        ctx.close()


def test_region_workload_sample(ctx):
    """20 000 regions of the batched banded-SW workload (lognormal windows,
    half-widths 8/16/32/64, 40 % homologues), with and without bias."""
    d = workload.band_regions(seed=5, n=20000)
    sch = S.Scheme(BLOSUM62, 11, 1)
    got = ctx.score_bands(d["qres"], d["qoff"], d["sres"], d["soff"], d["jobs"], sch)
    want = O.banded_many(d["qres"], d["qoff"], d["sres"], d["soff"], d["jobs"], BLOSUM62, 11, 1)
    assert_same(got, want, "regions")
    check_scores_only(ctx, d["qres"], d["qoff"], d["sres"], d["soff"], d["jobs"], sch, None, want)
    assert (got[0][d["homolog"]] > 60).mean() > 0.5
    qb = np.asarray(Lcg(1).next(1) + (np.arange(len(d["qres"])) % 9) - 4, np.int8)
    got = ctx.score_bands(d["qres"], d["qoff"], d["sres"], d["soff"], d["jobs"], sch, qb)
    want = O.banded_many(d["qres"], d["qoff"], d["sres"], d["soff"], d["jobs"], BLOSUM62, 11, 1, qb)
    assert_same(got, want, "regions+bias")
    check_scores_only(ctx, d["qres"], d["qoff"], d["sres"], d["soff"], d["jobs"], sch, qb, want)


@pytest.mark.parametrize("bias", [False, True])
def test_region_workload_at_scale(ctx, bias):
    """The benchmark's batch at its full size (10^6 regions, 1000 per query):
    the scores-only and end-cell builds agree on every job in two launches
    each, and a 4000-job sample equals the oracle. Small batches never fill
    the chip, so some timing-dependent faults only show at this size."""
    d = workload.band_regions(seed=1, n=1_000_000, per_query=1000)
    sch = S.Scheme(BLOSUM62, 11, 1)
    qb = np.asarray((np.arange(len(d["qres"])) % 11) - 5, np.int8) if bias else None
    runs = []
    for ends in (True, False, True, False):
        b = ctx.band_batch(d["qres"], d["qoff"], d["sres"], d["soff"], d["jobs"], sch, qbias=qb, ends=ends)
        try:
            b.run()
            b.wait()
            runs.append(tuple(np.asarray(x).copy() for x in b.read()))
        finally:
            b.close()
    for k in range(1, 4):
        bad = np.nonzero(runs[k][0] != runs[0][0])[0]
        assert len(bad) == 0, f"launch {k}: {len(bad)} scores differ, jobs {bad[:10].tolist()}"
    for f in range(1, 3):
        assert np.array_equal(runs[2][f], runs[0][f])
    idx = np.sort(np.random.default_rng(11).choice(len(d["jobs"]), 4000, replace=False))
    want = O.banded_many(d["qres"], d["qoff"], d["sres"], d["soff"], d["jobs"][idx], BLOSUM62, 11, 1, qb)
    for f, name in enumerate(("score", "q_end", "s_end", "cells")):
        g = runs[0][f][idx].astype(np.int64)
        bad = np.nonzero(g != np.asarray(want[f]).astype(np.int64))[0]
        assert len(bad) == 0, f"{name} differs at sampled jobs {idx[bad[:10]].tolist()}"


def test_score_windows_backend(ctx):
    # ScoreBackend contract (gpu/mod.rs:9-31) and doctor::check_gpu's shape
    # (doctor.rs:440-476): qlen 250, bias (i%5)-2, lengths 1+(13i)%700
    r = Lcg(4)
    q = rand_seq(r, 250)
    bias = np.array([(i % 5) - 2 for i in range(250)], np.int8)
    wins = [rand_seq(r, 1 + (13 * i) % 700) if i % 3 else mutate(r, q)[: 1 + (13 * i) % 700]
            for i in range(300)]
    wins = [w if len(w) else rand_seq(r, 1) for w in wins]
    be = S.HipBackend(0)
    got = be.score_windows(q, bias, wins, BLOSUM62, S.GapCosts(11, 1))
    want = [O.sw_reference(q, w, BLOSUM62, bias, 11, 1)[0] for w in wins]
    assert got == want
    assert be.score_windows(q, None, [], BLOSUM62, S.GapCosts(11, 1)) == []
    assert be.score_windows(np.zeros(0, np.uint8), None, wins[:3], BLOSUM62, S.GapCosts(11, 1)) == [0, 0, 0]
    with pytest.raises(ValueError):
        be.score_windows(q, bias[:10], wins[:3], BLOSUM62, S.GapCosts(11, 1))


def test_reference_named_interfaces(ctx):
    q = S.encode("WAGNER")
    sch = S.Scheme(BLOSUM62, 11, 1)
    assert S.banded_score(q, q, None, sch, S.Band.full(6, 6), ctx) == S.BandScore(37, 6, 6, 36)
    jobs = [S.BandJob(q, S.Band(-2, 2)), S.BandJob(S.encode("PPPP"), S.Band(-1, 1)),
            S.BandJob(np.zeros(0, np.uint8), S.Band(0, 0))]
    assert S.score_banded_many(q, None, jobs, sch, ctx) == [37, 0, 0]
    wag = S.encode("WAG")
    assert S.score_many(q, None, [q, wag], BLOSUM62, S.GapCosts(11, 1), ctx) == [
        37, O.sw_reference(q, wag, BLOSUM62, None, 11, 1)[0]]


def test_autotune_choice_is_exact():
    """The measured shape choice (saber_hip_band_autotune) installs covering
    shapes only, and batches planned with it stay bit-exact."""
    ctx = S.Context()
    d = workload.band_regions(seed=8, n=20000)
    sch = S.Scheme(BLOSUM62, 11, 1)
    rep = ctx.autotune(d["qres"], d["qoff"], d["sres"], d["soff"], d["jobs"], sch, target_cells=20_000_000)
    assert rep, "no width class was tuned"
    for e in rep:
        L, C, K = e["shape"]
        W = L * C
        lo_w = 1 if K == C else W - K + 1
        assert lo_w <= e["w_lo"] <= e["w_hi"] <= W, e
        assert e["gcups"] > 0
    got = ctx.score_bands(d["qres"], d["qoff"], d["sres"], d["soff"], d["jobs"], sch)
    want = O.banded_many(d["qres"], d["qoff"], d["sres"], d["soff"], d["jobs"], BLOSUM62, 11, 1)
    assert_same(got, want, "autotuned")
    ctx.close()


def test_graph_replays_match_direct_passes():
    """Band passes replayed from the captured graph equal passes enqueued
    directly, on every replay, with other work and synchronisations between
    the replays; saturated jobs (rescored by the int32 kernel, their count read
    on the device) and jobs too wide for the packed kernel included."""
    d = workload.band_regions(seed=13, n=6000)
    jobs = d["jobs"].copy()
    jobs["band_lo"][:20] = -600          # wider than every packed shape
    jobs["band_hi"][:20] = 600
    a = np.full(3000, W, np.uint8)        # W x 3000 saturates int16
    qres = np.concatenate([d["qres"], a])
    qoff = np.concatenate([d["qoff"], [d["qoff"][-1] + 3000]]).astype(np.uint64)
    sres = np.concatenate([d["sres"], a])
    soff = np.concatenate([d["soff"], [d["soff"][-1] + 3000]]).astype(np.uint64)
    extra = np.zeros(3, dtype=jobs.dtype)
    extra["query"] = len(qoff) - 2
    extra["subject"] = len(soff) - 2
    extra["win_lo"] = 0
    extra["win_hi"] = 3000
    extra["band_lo"] = [-8, -40, 0]
    extra["band_hi"] = [8, 40, 0]
    jobs = np.concatenate([jobs, extra])
    sch = S.Scheme(BLOSUM62, 11, 1)
    direct = S.Context(graphs=False)
    want = direct.score_bands(qres, qoff, sres, soff, jobs, sch)
    direct.close()
    assert (want[0][-3:] == 33000).sum() >= 2
    ctx = S.Context()
    for ends in (True, False):
        b = ctx.band_batch(qres, qoff, sres, soff, jobs, sch, ends=ends)
        other = ctx.band_batch(d["qres"], d["qoff"], d["sres"], d["soff"], d["jobs"][:500], sch)
        for rep in range(4):
            b.run()
            got = b.read()
            assert np.array_equal(got[0], want[0]), f"replay {rep} scores"
            if ends:
                assert all(np.array_equal(g, w) for g, w in zip(got[1:], want[1:])), f"replay {rep} ends/cells"
            other.run()          # other work between the replays
            other.wait()
        b.close()
        other.close()
    ctx.close()
