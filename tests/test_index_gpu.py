"""A database uploaded from a v4 index file (saber_hip_upload_index) is the
database uploaded from the same residues and offsets (saber_hip_upload_db):
the same band scores over database subjects and the same search results
(hits and counters) on the small blastp workload."""
import numpy as np
import pytest

import saber_amd as S
from saber_amd import BLOSUM62, search, stats, workload
from tests import index_v4

pytestmark = pytest.mark.gpu


def test_index_upload_equals_direct_upload(ctx, tmp_path):
    d = workload.search_set(seed=5, nq=40, nsub=1500)
    subs = [d["sres"][int(d["soff"][i]):int(d["soff"][i + 1])] for i in range(len(d["soff"]) - 1)]
    data, _ = index_v4.build(subs)
    path = tmp_path / "db.sdx"
    path.write_bytes(data)
    k = stats.gapped_protein(62, 11, 1)
    sch = S.Scheme(BLOSUM62, 11, 1)
    ctx.upload_db(d["sres"], d["soff"])
    want = search.run(ctx, d["qres"], d["qoff"], sch, k)
    with S.IndexFile(path, verify=True) as idx:
        ctx.upload_index(idx)
    got = search.run(ctx, d["qres"], d["qoff"], sch, k)
    assert [[h.key() for h in q] for q in got.hits] == [[h.key() for h in q] for q in want.hits]
    assert got.counters == want.counters
    assert sum(len(q) for q in got.hits) > 0
    # band scores over the uploaded database's subjects
    r = np.random.default_rng(2)
    ns = len(subs)
    jobs = np.zeros(500, dtype=[("query", "<u4"), ("subject", "<u4"), ("win_lo", "<u4"), ("win_hi", "<u4"),
                                ("band_lo", "<i4"), ("band_hi", "<i4")])
    jobs["query"] = r.integers(0, len(d["qoff"]) - 1, 500)
    jobs["subject"] = r.integers(0, ns, 500)
    jobs["win_hi"] = [len(subs[s]) for s in jobs["subject"]]
    jobs["band_lo"] = -20
    jobs["band_hi"] = 20
    a = ctx.score_bands(d["qres"], d["qoff"], None, None, jobs, sch)
    ctx.upload_db(d["sres"], d["soff"])
    b = ctx.score_bands(d["qres"], d["qoff"], None, None, jobs, sch)
    for x, y in zip(a, b):
        assert np.array_equal(np.asarray(x), np.asarray(y))
