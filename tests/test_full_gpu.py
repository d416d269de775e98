"""Exhaustive full-matrix SW on the GPU vs the oracle (sw_reference,
reference.rs:15-52): every query x every subject, score and first row-major
end cell bit-exact. Cases follow simd_matches_independent_reference
(sw_simd.rs:488-519), saturation_is_detected_and_repaired (sw_simd.rs:521-533)
and the exhaustive-mode consistency test (tests/engine_consistency.rs:225-258)."""
import numpy as np
import pytest

import saber_amd as S
from oracle import oracle as O
from saber_amd import BLOSUM62, PAM30, workload
from tests.gen import Lcg, flat, mutate, rand_seq

pytestmark = pytest.mark.gpu

W = 17


def oracle_grid(qs, ss, mat, go, ge, qbias=None):
    sc = np.zeros((len(qs), len(ss)), np.int64)
    qe = np.zeros_like(sc)
    se = np.zeros_like(sc)
    for a, q in enumerate(qs):
        b = None if qbias is None else qbias[a]
        for c, s in enumerate(ss):
            sc[a, c], qe[a, c], se[a, c] = O.sw_reference(q, s, mat, b, go, ge)
    return sc, qe, se


def check(ctx, qs, ss, mat, go, ge, qbias=None):
    qres, qoff = flat(qs)
    sres, soff = flat(ss)
    qb = None if qbias is None else np.concatenate(qbias).astype(np.int8)
    got = ctx.score_full(qres, qoff, sres, soff, S.Scheme(mat, go, ge), qb)
    want = oracle_grid(qs, ss, mat, go, ge, qbias)
    for name, g, w in zip(("score", "q_end", "s_end"), got, want):
        bad = np.argwhere(g.astype(np.int64) != w)
        assert len(bad) == 0, f"{name} differs at {bad[:5].tolist()}"
    return got


@pytest.mark.parametrize("gaps", [(11, 1), (10, 2), (13, 2)])
def test_random_queries_and_subjects(ctx, gaps):
    r = Lcg(488 + gaps[0])
    qs = [rand_seq(r, 1 + r.next(300)) for _ in range(6)]
    ss = []
    for k in range(60):
        if k % 4 == 0:
            s = mutate(r, qs[k % len(qs)])
            ss.append(s if len(s) else rand_seq(r, 3))
        else:
            ss.append(rand_seq(r, 1 + r.next(600)))
    check(ctx, qs, ss, BLOSUM62, *gaps)
    bias = [np.array([r.next(9) - 4 for _ in range(len(q))], np.int8) for q in qs]
    check(ctx, qs, ss, BLOSUM62, *gaps, qbias=bias)


def test_subject_lengths_across_all_variants(ctx):
    """Subjects of every length class up to 3072 and beyond (int32 route)."""
    r = Lcg(7)
    q = rand_seq(r, 120)
    lens = [1, 2, 15, 16, 17, 33, 48, 49, 64, 65, 97, 128, 129, 200, 257, 384, 385, 512, 513, 700,
            768, 769, 1024, 1025, 1500, 1536, 1537, 2048, 2049, 2560, 3000, 3072, 3073, 3300]
    ss = []
    for n in lens:
        s = rand_seq(r, n)
        k = min(n, len(q))
        s[:k] = np.where(np.array([r.next(3) for _ in range(k)]) > 0, q[:k], s[:k])
        ss.append(s)
    check(ctx, [q, q[:50]], ss, PAM30, 10, 1)


def test_saturation_and_empty(ctx):
    # W x 3000 scores 33 000 (above int16): flagged and rescored exactly.
    a = np.full(3000, W, np.uint8)
    got = check(ctx, [a, np.zeros(0, np.uint8)], [a, a[:100], np.zeros(0, np.uint8)], BLOSUM62, 11, 1)
    assert got[0][0, 0] == 33000 and (got[1][0, 0], got[2][0, 0]) == (3000, 3000)
    assert (got[0][1] == 0).all()


def test_backend_routes_to_full_kernel(ctx):
    r = Lcg(3)
    q = rand_seq(r, 250)
    wins = [rand_seq(r, 1 + (13 * i) % 700) for i in range(100)]
    got = S.HipBackend(0).score_windows(q, None, wins, BLOSUM62, S.GapCosts(11, 1))
    assert got == [O.sw_reference(q, w, BLOSUM62, None, 11, 1)[0] for w in wins]


def test_workload_sample(ctx):
    qres, qoff = workload.sequences(11, 8, 250, 0.6, 30, 3000)
    sres, soff = workload.sequences(12, 300, 250, 0.6, 30, 3000)
    qs = [qres[qoff[i]:qoff[i + 1]] for i in range(len(qoff) - 1)]
    ss = [sres[soff[i]:soff[i + 1]] for i in range(len(soff) - 1)]
    check(ctx, qs, ss, BLOSUM62, 11, 1)


@pytest.mark.parametrize("with_bias", [False, True])
def test_scores_only_kernel(ctx, with_bias):
    """The scores-only kernels (score_many's output) agree with the oracle on
    every length class, including subjects past the widest end-cell shape
    (2048) that still fit the widest scores-only shape (3072); reading end
    cells from a scores-only batch is an error."""
    r = Lcg(95 + with_bias)
    qs = [rand_seq(r, 1 + r.next(400)) for _ in range(4)]
    lens = [1, 16, 17, 40, 48, 49, 96, 97, 257, 640, 641, 1025, 2048, 2049, 2600, 3072, 3073]
    ss = []
    for k, n in enumerate(lens):
        s = rand_seq(r, n)
        if k % 2 == 0:
            q = qs[k % len(qs)]
            kk = min(n, len(q))
            s[:kk] = np.where(np.array([r.next(3) for _ in range(kk)]) > 0, q[:kk], s[:kk])
        ss.append(s)
    qres, qoff = flat(qs)
    sres, soff = flat(ss)
    bias = [np.array([r.next(9) - 4 for _ in range(len(q))], np.int8) for q in qs] if with_bias else None
    qb = None if bias is None else np.concatenate(bias).astype(np.int8)
    sch = S.Scheme(BLOSUM62, 11, 1)
    b = ctx.full_batch(qres, qoff, sres, soff, sch, qb, ends=False)
    try:
        b.run()
        got = b.read()[0]
        with pytest.raises(S.HipError):
            import ctypes as C
            from saber_amd import _ffi
            qe = np.zeros(got.shape, np.uint32)
            _ffi.check(_ffi.lib().saber_hip_full_batch_read(b.h, None, _ffi.ptr(qe, C.c_uint32), None))
    finally:
        b.close()
    want = oracle_grid(qs, ss, BLOSUM62, 11, 1, bias)[0]
    bad = np.argwhere(got.astype(np.int64) != want)
    assert len(bad) == 0, f"score differs at {bad[:5].tolist()}"


@pytest.mark.parametrize("ends", [True, False])
def test_graph_replays_and_int32_paths(ctx, ends):
    """Full passes replayed from the captured graph equal a pass enqueued
    directly and the oracle, on every replay, with other work between the
    replays: subjects longer than the widest shape (the static int32 list) and
    saturating pairs (flagged by the packed kernels, rescored by the int32
    kernel from the count on the device) included; the forced-int32 context
    agrees. A matrix scoring 100 per identity saturates int16 within 330
    residues, which keeps the int32 work small."""
    r = Lcg(61 + ends)
    mat = np.full((24, 24), -30, np.int32)
    np.fill_diagonal(mat, 100)
    hot = rand_seq(r, 400)                     # self score 40 000: saturates
    qs = [rand_seq(r, 1 + r.next(350)) for _ in range(5)] + [hot]
    ss = [rand_seq(r, 1 + r.next(900)) for _ in range(40)] + [hot, rand_seq(r, 3300), hot[:300]]
    for k in range(0, 40, 5):
        q = qs[k % 5]
        kk = min(len(ss[k]), len(q))
        ss[k][:kk] = np.where(np.array([r.next(3) for _ in range(kk)]) > 0, q[:kk], ss[k][:kk])
    qres, qoff = flat(qs)
    sres, soff = flat(ss)
    sch = S.Scheme(mat, 11, 1)
    want = oracle_grid(qs, ss, mat, 11, 1)
    assert want[0][5, 40] == 40000 and want[0][5, 42] == 30000
    direct = S.Context(graphs=False)
    got = direct.score_full(qres, qoff, sres, soff, sch, ends=ends)
    direct.close()
    exact = S.Context(force_exact32=True)
    got32 = exact.score_full(qres, qoff, sres, soff, sch, ends=ends)
    exact.close()
    for g in (got, got32):
        assert np.array_equal(g[0], want[0])
        if ends:
            assert np.array_equal(g[1], want[1]) and np.array_equal(g[2], want[2])
    b = ctx.full_batch(qres, qoff, sres, soff, sch, ends=ends)
    other = ctx.full_batch(qres[:qoff[2]], qoff[:3], sres, soff, sch, ends=False)
    try:
        for rep in range(4):
            b.run()
            g = b.read()
            assert np.array_equal(g[0], want[0]), f"replay {rep} scores"
            if ends:
                assert np.array_equal(g[1], want[1]) and np.array_equal(g[2], want[2]), f"replay {rep} ends"
            other.run()
            other.read()
        ms, cells, rescued = b.time(warmup=0, iters=2)
        # every query with the 3300-long subject (column 41), plus the pairs
        # of the packed kernels reaching 32 767
        packed = np.delete(want[0], 41, axis=1)
        flagged = int((packed >= 32767).sum())
        assert flagged >= 1 and rescued == len(qs) + flagged, (rescued, flagged)
        assert cells == int(qoff[-1]) * int(soff[-1]) and ms > 0
    finally:
        b.close()
        other.close()


def test_autotune_choice_is_exact():
    """The measured shape choice (saber_hip_full_autotune) picks, per group of
    subject lengths, shapes that cover the group's longest subject, and full
    batches planned with it stay bit-exact."""
    c = S.Context()
    qres, qoff = workload.sequences(21, 4, 250, 0.6, 30, 3000)
    sres, soff = workload.sequences(22, 800, 250, 0.6, 30, 3000)
    sch = S.Scheme(BLOSUM62, 11, 1)
    try:
        rep = c.full_autotune(qres, qoff, sres, soff, sch, ends=False, target_cells=100_000_000)
        assert rep, "no length group was tuned"
        for e in rep:
            L, Cw = e["shape"]
            assert e["len_lo"] <= e["len_hi"] <= L * Cw, e
            assert e["gcups"] > 0 and e["cells"] > 0
        got = c.score_full(qres, qoff, sres, soff, sch, ends=False)[0]
    finally:
        c.close()
    qs = [qres[qoff[i]:qoff[i + 1]] for i in range(len(qoff) - 1)]
    ss = [sres[soff[i]:soff[i + 1]] for i in range(len(soff) - 1)]
    want = oracle_grid(qs, ss, BLOSUM62, 11, 1)[0]
    bad = np.argwhere(got.astype(np.int64) != want)
    assert len(bad) == 0, f"score differs at {bad[:5].tolist()}"
