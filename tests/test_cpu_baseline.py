"""The AVX2 CPU baseline (restatement of score_banded_many + banded_score's
int16 kernel) must give the oracle's scores; it is what bench.py times on the
host cores. CPU only."""
import numpy as np
import pytest

from oracle import oracle as O
from saber_amd import BLOSUM45, BLOSUM62, PAM30, workload
from tests.gen import Lcg, flat, jobs_array, mutate, rand_seq


def test_region_workload_scores_match_oracle():
    d = workload.band_regions(seed=21, n=6000, per_query=32)
    got = O.banded_many_avx2(d["qres"], d["qoff"], d["sres"], d["soff"], d["jobs"], BLOSUM62, 11, 1)
    want = O.banded_many(d["qres"], d["qoff"], d["sres"], d["soff"], d["jobs"], BLOSUM62, 11, 1)[0]
    assert np.array_equal(got, want)
    qb = ((np.arange(len(d["qres"])) * 7) % 9 - 4).astype(np.int8)
    got = O.banded_many_avx2(d["qres"], d["qoff"], d["sres"], d["soff"], d["jobs"], BLOSUM62, 11, 1, qb)
    want = O.banded_many(d["qres"], d["qoff"], d["sres"], d["soff"], d["jobs"], BLOSUM62, 11, 1, qb)[0]
    assert np.array_equal(got, want)


@pytest.mark.parametrize("mat,gaps", [(BLOSUM62, (11, 1)), (PAM30, (10, 1)), (BLOSUM45, (7, 2)),
                                      (BLOSUM62, (6, 3)), (BLOSUM62, (8, 8))])
def test_mixed_bands_one_query_per_group(mat, gaps):
    r = Lcg(17 + gaps[0])
    qs, ss, lo, hi = [], [], [], []
    for g in range(12):
        q = rand_seq(r, 1 + r.next(300))
        for k in range(1 + r.next(60)):
            s = mutate(r, q) if k % 3 == 0 else rand_seq(r, 1 + r.next(300))
            if len(s) == 0:
                s = rand_seq(r, 4)
            qs.append(q)
            ss.append(s)
            c = int(r.next(len(s) + len(q))) - len(q)
            w = int(r.next(300 if k % 7 == 0 else 100))
            lo.append(c - w // 2)
            hi.append(c + (w + 1) // 2)
    # one query entry per group: consecutive jobs share the query offset
    uq, qidx = [], []
    for q in qs:
        if not uq or uq[-1] is not q:
            uq.append(q)
        qidx.append(len(uq) - 1)
    qres, qoff = flat(uq)
    sres, soff = flat(ss)
    n = len(ss)
    jobs = jobs_array(qidx, np.arange(n), np.zeros(n), [len(s) for s in ss], lo, hi)
    got = O.banded_many_avx2(qres, qoff, sres, soff, jobs, mat, *gaps)
    want = O.banded_many(qres, qoff, sres, soff, jobs, mat, *gaps)[0]
    assert np.array_equal(got, want)


def test_overflow_and_int16_decline():
    W = 17
    a = np.full(4000, W, np.uint8)
    w300 = np.full(300, W, np.uint8)
    qres, qoff = flat([w300, a])
    sres, soff = flat([w300, np.array([12, 8, 17, 10, 0, 0], np.uint8), a, a])
    jobs = jobs_array([0, 0, 1, 1], [0, 1, 2, 3], [0] * 4, [300, 6, 4000, 4000],
                      [-5, -3, -20, -4000], [5, 3, 20, 4000])
    got = O.banded_many_avx2(qres, qoff, sres, soff, jobs, BLOSUM62, 11, 1)
    want = O.banded_many(qres, qoff, sres, soff, jobs, BLOSUM62, 11, 1)[0]
    assert got.tolist() == want.tolist()
    assert got[0] == 3300 and got[2] == 44000 and got[3] == 44000


def test_score_many_avx2_equals_sw_reference():
    """The AVX2 restatement of score_many (the exhaustive-mode CPU baseline)
    against the oracle's full Gotoh on random, related and saturating pairs,
    with and without a bias: the 8-bit pass, its 16-bit re-score and the
    int32 repair are all exercised."""
    from tests.gen import Lcg, flat, mutate, rand_seq
    from oracle import oracle as O
    from saber_amd import BLOSUM62
    r = Lcg(2024)
    qs = [rand_seq(r, 5 + r.next(300)) for _ in range(6)]
    ss = [rand_seq(r, 1 + r.next(400)) for _ in range(40)] + [mutate(r, qs[i % 6]) for i in range(20)]
    w = np.full(4000, 17, np.uint8)   # W x 4000 = 44 000 (banded.rs:719-728): the int32 repair
    qs.append(w)
    ss.append(w.copy())
    ss.append(np.full(120, 17, np.uint8))   # 1320 > 255: the 16-bit re-score
    qres, qoff = flat(qs)
    sres, soff = flat(ss)
    for bias in (None, ((np.arange(len(qres)) * 31) % 9 - 4).astype(np.int8)):
        got = O.score_many_avx2(qres, qoff, sres, soff, BLOSUM62, 11, 1, qbias=bias, nthreads=4)
        for i in range(len(qs)):
            q = qres[int(qoff[i]):int(qoff[i + 1])]
            b = None if bias is None else bias[int(qoff[i]):int(qoff[i + 1])]
            for j in range(len(ss)):
                s = sres[int(soff[j]):int(soff[j + 1])]
                if len(q) * len(s) > 300_000 and not (len(q) == 4000 and len(s) == 4000):
                    continue
                assert got[i, j] == O.sw_reference(q, s, BLOSUM62, b, 11, 1)[0], (i, j)
    assert got[len(qs) - 1, len(ss) - 2] == 44000 + int(bias[-4000:].sum())


@pytest.mark.parametrize("x", [1, 10, 20, 35, 300, 20000])
def test_xdrop_avx2_equals_oracle(x):
    """The AVX2 X-drop restatement (the X-drop leg's CPU baseline) against
    or_xdrop: related and random pairs, seeds at both ends of the sequences
    and in the middle, query bias of -128..127, blocks of 16 steps ending
    exactly at a sequence end; x = 20000 takes the scalar route."""
    r = Lcg(900 + x)
    qs, ss, hits = [], [], []
    for k in range(60):
        q = rand_seq(r, 1 + r.next(120))
        s = mutate(r, q) if k % 2 == 0 else rand_seq(r, 1 + r.next(120))
        if len(s) == 0:
            s = rand_seq(r, 1)
        if k % 5 == 0:
            s = np.concatenate([q, q])[:16 * (1 + r.next(4))] if len(q) else s
            q = s.copy()
        qs.append(q)
        ss.append(s)
        for _ in range(6):
            hits.append((k, k, r.next(len(q)), r.next(len(s))))
        hits.append((k, k, 0, 0))
        hits.append((k, k, len(q) - 1, len(s) - 1))
    qres, qoff = flat(qs)
    sres, soff = flat(ss)
    hits = np.array(hits, np.uint32)
    bias = np.array([r.next(256) - 128 for _ in range(len(qres))], np.int8)
    for mat, b in ((BLOSUM62, None), (PAM30, bias)):
        got = O.xdrop_many_avx2(qres, qoff, sres, soff, hits, mat, x, b, nthreads=3)
        for h, (qi, si, qp, sp) in enumerate(hits.tolist()):
            bb = None if b is None else b[int(qoff[qi]):int(qoff[qi + 1])]
            want = O.xdrop(qs[qi], ss[si], bb, mat, qp, sp, x)
            assert tuple(int(a[h]) for a in got) == want, (x, h)
