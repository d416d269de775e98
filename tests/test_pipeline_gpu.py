"""Two-stage region scoring on the GPU (saber_hip_score_regions) against the
oracle: the narrow-band scores, the selection against the thresholds and
the full-band scores of the selected regions."""
import numpy as np
import pytest

import saber_amd as S
from oracle import host as H
from oracle import oracle as O
from saber_amd import BLOSUM62, pipeline as PL
from tests.gen import Lcg, flat, jobs_array, mutate, rand_seq

pytestmark = pytest.mark.gpu


def test_two_stage_scores_match_oracle():
    r = Lcg(31)
    queries = [rand_seq(r, 2500 + r.next(800)) for _ in range(3)]
    subjects, regs, qidx = [], [], []
    for s in range(60):
        q = s % 3
        a = queries[q]
        if s % 2 == 0:
            # a homologous stretch planted at a random offset of a long subject
            pre = rand_seq(r, 3000 + r.next(3000))
            b = np.concatenate([pre, mutate(r, a[500:2000]), rand_seq(r, 2000)])
            d = len(pre) - 500
        else:
            b = rand_seq(r, 6000 + r.next(3000))
            d = int(r.next(4000))
        subjects.append(b)
        regs.append((s, d, d + int(r.next(6)), 45 + int(r.next(300)), 1, 60, 0))
        qidx.append(q)
    regs = np.array(regs, PL.REGION_DTYPE)
    qres, qoff = flat(queries)
    sres, soff = flat(subjects)
    plans = np.concatenate([PL.plan_windows(regs[i:i + 1], len(queries[qidx[i]]), [len(subjects[i])],
                                            64, 48, 40) for i in range(len(regs))])
    assert plans["banded"].sum() >= 50
    thr = np.full(len(regs), 60.0)
    ctx = S.Context()
    sch = S.Scheme(BLOSUM62, 11, 1)
    got, resc = PL.score_regions(ctx, qres, qoff, qidx, regs, plans, sch, narrow_band=16, full_band=48, thr=thr,
                                 sres=sres, soff=soff)
    # oracle: the same job lists (oracle/host.py) scored by the oracle band DP
    rd = [dict(subject=int(x["subject"]), d_min=int(x["d_min"]), d_max=int(x["d_max"]), best=int(x["best"]))
          for x in regs]
    pl = [(int(p["lo"]), int(p["hi"]), (int(p["band_lo"]), int(p["band_hi"])) if p["banded"] else None)
          for p in plans]
    qlens = [len(queries[q]) for q in qidx]
    narrow, full = H.two_stage_jobs(rd, pl, qlens, 16, 48)

    def oracle_scores(jobs):
        if not jobs:
            return []
        idx = [i for i, _ in jobs]
        ja = jobs_array(np.array([qidx[i] for i in idx]), np.array([rd[i]["subject"] for i in idx]),
                        np.array([pl[i][0] for i in idx]), np.array([pl[i][1] for i in idx]),
                        [b[0] for _, b in jobs], [b[1] for _, b in jobs])
        return O.banded_many(qres, qoff, sres, soff, ja, BLOSUM62, 11, 1)[0].tolist()

    want = np.zeros(len(regs), np.int64)
    want_resc = np.zeros(len(regs), bool)
    for (i, _), sc in zip(narrow, oracle_scores(narrow)):
        if sc >= thr[i]:
            full.append((i, pl[i][2]))
            want_resc[i] = True
        else:
            want[i] = sc
    for (i, _), sc in zip(full, oracle_scores(full)):
        want[i] = sc
    assert np.array_equal(resc, want_resc)
    assert np.array_equal(got.astype(np.int64), want)
    assert want_resc.sum() >= 10 and (~want_resc).sum() >= 10
    ctx.close()
