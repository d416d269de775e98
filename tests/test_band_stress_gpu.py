"""Every body of the packed band kernel under load: for each compiled shape
(L, C, K), a batch of several hundred jobs whose clipped widths all fall in
the shape's range is forced onto that shape and run in every flavour --
query bias with per-job tables (one job per query), query bias with the
shared pair table (two jobs per query on the same first row), end cells, and
no bias -- on freshly created batches, three launches each, every score and
end cell equal to the oracle. A body that returns wrong values only when
many of its chunks run at once (as one compiled shape did, a different value
on each launch) fails here, where the small parity cases pass."""
import numpy as np
import pytest

import saber_amd as S
from oracle import oracle as O
from saber_amd import BLOSUM62
from tests.gen import Lcg, flat, jobs_array, mutate, rand_seq

pytestmark = pytest.mark.gpu


def shape_set(seed, wlo, whi, njobs, per_query):
    """njobs jobs whose bands lie inside their rectangles with widths in
    [wlo, whi] (so the clipped width is the band width), per_query jobs per
    query (each starting on query row 0), half of the windows related to their
    query; biases in [-7, 7]."""
    r = Lcg(seed)
    qs, ss, jq, lo, hi, bias = [], [], [], [], [], []
    while len(ss) < njobs:
        w = wlo + r.next(whi - wlo + 1)
        m = w + 20 + r.next(260)
        a = rand_seq(r, m)
        qs.append(a)
        bias.append(np.array([r.next(15) - 7 for _ in range(m)], np.int8))
        for _ in range(per_query):
            b = mutate(r, a) if r.next(2) else rand_seq(r, m)
            if len(b) < w + 2:
                b = rand_seq(r, w + 20)
            n = len(b)
            band_lo = -int(r.next(min(m - 1, w)))
            band_hi = band_lo + w - 1
            if band_hi > n - 1:
                band_lo, band_hi = n - w, n - 1
            ss.append(b)
            jq.append(len(qs) - 1)
            lo.append(band_lo)
            hi.append(band_hi)
    qres, qoff = flat(qs)
    sres, soff = flat(ss)
    nj = len(ss)
    jobs = jobs_array(np.asarray(jq), np.arange(nj), np.zeros(nj), [len(s) for s in ss], lo, hi)
    qb, _ = flat([np.asarray(b, np.int8).view(np.uint8) for b in bias])
    return qres, qoff, sres, soff, jobs, qb.view(np.int8)


# The compiled shapes, listed here so that collecting this module does not
# load the library (and its HIP runtime) ahead of torch's; the test checks the
# list against the library's variant table.
SHAPES = [(1, 9, 9), (1, 17, 8), (1, 17, 17), (2, 9, 9), (2, 13, 13), (2, 17, 8), (2, 17, 17),
          (4, 9, 9), (4, 13, 13), (4, 17, 8), (4, 17, 17), (8, 9, 9), (8, 13, 13), (8, 17, 8), (8, 17, 17),
          (16, 9, 9), (16, 13, 13), (16, 17, 16), (16, 17, 17), (16, 21, 21), (2, 33, 2), (4, 33, 4)]


def test_shape_list_is_the_variant_table():
    assert sorted({(v[0], v[1], v[2]) for v in S.band_variants()}) == sorted(SHAPES)


@pytest.mark.parametrize("shape", SHAPES, ids=[f"{L}x{C}-{K}" for L, C, K in SHAPES])
def test_every_shape_under_load(shape):
    L, C, K = shape
    W = L * C
    sch = S.Scheme(BLOSUM62, 11, 1)
    ctx = S.Context(0)
    try:
        ctx.choose(W - K + 1, W, L, C, K)
        for per_query in (1, 2):
            qres, qoff, sres, soff, jobs, qb = shape_set(1000 * L + 10 * C + per_query, W - K + 1, W, 600, per_query)
            ctx.upload_db(sres, soff)
            wb = O.banded_many(qres, qoff, sres, soff, jobs, BLOSUM62, 11, 1, qb)
            wn = O.banded_many(qres, qoff, sres, soff, jobs, BLOSUM62, 11, 1, None)
            for bias, ends, want in ((True, False, wb), (True, True, wb), (False, False, wn)):
                for rep in range(2):
                    b = ctx.band_batch(qres, qoff, None, None, jobs, sch, qbias=qb if bias else None, ends=ends)
                    try:
                        for launch in range(3):
                            b.run()
                            b.wait()
                            got = b.read()
                            fields = (0, 1, 2) if ends else (0,)
                            for f in fields:
                                g = np.asarray(got[f]).astype(np.int64)
                                bad = np.nonzero(g != np.asarray(want[f]).astype(np.int64))[0]
                                assert len(bad) == 0, (
                                    f"shape {shape} per_query {per_query} bias {bias} ends {ends} batch {rep} "
                                    f"launch {launch}: field {f} differs at jobs {bad[:8].tolist()} "
                                    f"({g[bad[:4]].tolist()} vs {np.asarray(want[f])[bad[:4]].tolist()})")
                    finally:
                        b.close()
    finally:
        ctx.close()


def test_choice_of_the_int32_kernel_and_of_wide_shapes_for_end_cells():
    """saber_hip_band_choose with L = 0 routes a width range to the exact
    int32 kernel for both builds; a 33-cell shape chosen for a range serves
    the scores-only build while end-cell batches keep a shape of at most 21
    cells. Scores and end cells equal the oracle either way."""
    sch = S.Scheme(BLOSUM62, 11, 1)
    ctx = S.Context(0)
    try:
        qres, qoff, sres, soff, jobs, qb = shape_set(4242, 129, 132, 200, 2)
        ctx.upload_db(sres, soff)
        want = O.banded_many(qres, qoff, sres, soff, jobs, BLOSUM62, 11, 1, None)
        for L, C, K in ((0, 0, 0), (4, 33, 4)):
            ctx.choose(129, 132, L, C, K)
            for ends in (False, True):
                b = ctx.band_batch(qres, qoff, None, None, jobs, sch, ends=ends)
                try:
                    b.run()
                    b.wait()
                    got = b.read()
                    for f in ((0, 1, 2) if ends else (0,)):
                        assert np.array_equal(np.asarray(got[f]).astype(np.int64),
                                              np.asarray(want[f]).astype(np.int64)), (L, C, K, ends, f)
                finally:
                    b.close()
    finally:
        ctx.close()
