"""X-drop extension on the GPU vs the oracle (xdrop_cat, scan.rs:859-904):
score, query range and step count, bit-exact. Cases follow
profile_xdrop_matches_reference_xdrop (engine.rs:2550-2585: x in {10, 20, 35},
bias) and vector_xdrop_matches_scalar (scan.rs:2679-2728), plus the seeds at
sequence borders."""
import numpy as np
import pytest

import saber_amd as S
from oracle import oracle as O
from saber_amd import BLOSUM62, PAM30, workload
from tests.gen import Lcg, flat, mutate, rand_seq

pytestmark = pytest.mark.gpu


def oracle_all(qres, qoff, sres, soff, hits, mat, x, qbias=None):
    out = []
    for qi, si, qp, sp in hits.tolist():
        q = qres[qoff[qi]:qoff[qi + 1]]
        s = sres[soff[si]:soff[si + 1]]
        b = None if qbias is None else qbias[qoff[qi]:qoff[qi + 1]]
        out.append(O.xdrop(q, s, b, mat, qp, sp, x))
    return np.array(out, dtype=np.int64).reshape(-1, 4)


@pytest.mark.parametrize("x", [10, 20, 35])
def test_random_and_related(ctx, x):
    r = Lcg(2550 + x)
    qs, ss, hits = [], [], []
    for t in range(300):
        q = rand_seq(r, 5 + r.next(300))
        s = mutate(r, q) if t % 2 == 0 else rand_seq(r, 5 + r.next(300))
        if len(s) == 0:
            s = rand_seq(r, 5)
        qs.append(q)
        ss.append(s)
        for _ in range(5):
            hits.append((t, t, r.next(len(q)), r.next(len(s))))
        hits.append((t, t, 0, 0))
        hits.append((t, t, len(q) - 1, len(s) - 1))
        hits.append((t, t, 0, len(s) - 1))
        hits.append((t, t, len(q) - 1, 0))
    qres, qoff = flat(qs)
    sres, soff = flat(ss)
    hits = np.array(hits, np.uint32)
    sch = S.Scheme(BLOSUM62, 11, 1)
    bias = ((np.arange(len(qres)) * 5) % 9 - 4).astype(np.int8)
    for qb in (None, bias):
        got = np.stack(ctx.xdrop(qres, qoff, sres, soff, hits, sch, x, qb), axis=1).astype(np.int64)
        want = oracle_all(qres, qoff, sres, soff, hits, BLOSUM62, x, qb)
        bad = np.nonzero((got != want).any(axis=1))[0]
        assert len(bad) == 0, f"x={x} bias={qb is not None} hits {bad[:5]}: {got[bad[:3]]} vs {want[bad[:3]]}"


def test_region_workload_hits(ctx):
    d = workload.band_regions(seed=9, n=3000, per_query=100)
    hits = workload.xdrop_hits(d, seed=4, per_region=4)
    sch = S.Scheme(PAM30, 10, 1)
    got = np.stack(ctx.xdrop(d["qres"], d["qoff"], d["sres"], d["soff"], hits, sch, 20), axis=1)
    want = oracle_all(d["qres"], d["qoff"], d["sres"], d["soff"], hits, PAM30, 20)
    assert np.array_equal(got.astype(np.int64), want)


def test_bad_hit_is_rejected(ctx):
    qres, qoff = flat([rand_seq(Lcg(1), 10)])
    sres, soff = flat([rand_seq(Lcg(2), 10)])
    with pytest.raises(S.HipError):
        ctx.xdrop(qres, qoff, sres, soff, np.array([[0, 0, 10, 0]], np.uint32), S.Scheme(BLOSUM62, 11, 1), 20)


def test_wide_scores_and_uploaded_database(ctx):
    """Scores outside int8 (BLOSUM62 x 20) take the int16 table; subjects come
    from the uploaded database; the oracle is the same restatement."""
    d = workload.band_regions(seed=11, n=2000, per_query=50)
    hits = workload.xdrop_hits(d, seed=5, per_region=3)
    mat = (np.asarray(BLOSUM62, np.int64) * 20).tolist()
    sch = S.Scheme(mat, 220, 20)
    bias = ((np.arange(len(d["qres"])) * 7) % 11 - 5).astype(np.int8)
    ctx.upload_db(d["sres"], d["soff"])
    for qb in (None, bias):
        got = np.stack(ctx.xdrop(d["qres"], d["qoff"], None, None, hits, sch, 400, qb), axis=1).astype(np.int64)
        want = oracle_all(d["qres"], d["qoff"], d["sres"], d["soff"], hits, mat, 400, qb)
        bad = np.nonzero((got != want).any(axis=1))[0]
        assert len(bad) == 0, f"bias={qb is not None} hits {bad[:5]}: {got[bad[:3]]} vs {want[bad[:3]]}"
