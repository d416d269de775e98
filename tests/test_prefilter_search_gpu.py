"""The learned region prefilter on the search's own regions (the scope's
config 4 on small blastp workloads): features of the capped regions against
their definition (numpy, from the region records the search returns), a
model trained on the GPU on regions labelled by the plain search, and the
filtered search against the plain one on a held-out workload -- what the
trained network gains (regions pruned, band cells saved) and what it costs
(reported hits lost <= 1 %), every hit of a kept region identical to the
plain search's."""
import math

import numpy as np
import pytest

import saber_amd as S
from saber_amd import BLOSUM62, prefilter as PF, search, stats, workload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(ctx):
    k = stats.gapped_protein(62, 11, 1)
    sch = S.Scheme(BLOSUM62, 11, 1)
    return k, sch


def test_features_match_definition(ctx, setup):
    k, sch = setup
    d = workload.search_set(seed=41, nq=40, nsub=2000)
    ctx.upload_db(d["sres"], d["soff"])
    r = search.run(ctx, d["qres"], d["qoff"], sch, k, keep_handle=True)
    try:
        X = PF.bf16_float(search.features(r))
    finally:
        search.close(r)
    ns = len(d["soff"]) - 1
    st = stats.SearchStatistics(k, int(d["soff"][-1]), ns, stats.PER_SUBJECT)
    rows = [(q, x) for q, reg in enumerate(r.regions) for x in reg]
    assert len(rows) == len(X) and len(rows) > 100
    for i in range(0, len(rows), max(1, len(rows) // 150)):
        q, x = rows[i]
        qq = d["qres"][int(d["qoff"][q]):int(d["qoff"][q + 1])]
        m = len(qq)
        rf = st.for_query(m).min_score(10.0)
        thr = max(40, math.ceil(rf * 0.65)) if rf > 0 else 40
        sub = d["sres"][int(d["soff"][x["subject"]]):int(d["soff"][x["subject"] + 1])]
        win = sub[int(x["win_lo"]):int(x["win_hi"])]
        n = len(win)
        blo, bhi = (int(x["band_lo"]), int(x["band_hi"])) if x["banded"] else (-(m - 1), n - 1)
        best = float(x["best"])
        f = np.zeros(64)
        f[0], f[1], f[2], f[3] = best / 64, math.log(1 + max(best, 0)), x["n_hsp"] / 4, x["hsp_sum"] / 128
        f[4], f[5] = (x["d_max"] - x["d_min"]) / 64, (bhi - blo + 1) / 128
        f[6], f[7], f[8] = math.log(1 + n), math.log(1 + m), best / math.log(2 + m * n)
        f[9], f[10], f[11] = x["hsp_sum"] / x["n_hsp"] / 64, float(x["banded"]), best / thr
        f[12:32] = np.bincount(np.minimum(qq, 23), minlength=24)[:20] / m
        f[32:52] = np.bincount(np.minimum(win, 23), minlength=24)[:20] / max(n, 1)
        f[52], f[53], f[54] = math.log(1 + len(sub)), (best - thr) / 64, x["hsp_sum"] / thr
        want = PF.bf16_float(PF.bf16_bits(f.astype(np.float32)))
        np.testing.assert_allclose(X[i], want, rtol=1e-2, atol=1e-2, err_msg=f"region {i}")


def test_trained_prefilter_gain_and_loss(ctx, setup):
    k, sch = setup
    train = [workload.search_set(seed=s) for s in range(51, 63)]
    valid = [workload.search_set(seed=s) for s in range(63, 69)]
    test = workload.search_set(seed=70)
    pf = search.train_prefilter(ctx, train, valid, sch, k, loss_target=0.004, epochs=12, batch=8192, seed=1,
                                log=lambda s: print(s, flush=True))
    print(f"# trained on {pf['train_regions']} regions ({pf['positives']} positive), threshold "
          f"{pf['threshold']:.3f}, validation kept {pf['valid_kept']:.3f}", flush=True)
    ctx.upload_db(test["sres"], test["soff"])
    plain = search.run(ctx, test["qres"], test["qoff"], sch, k, regions=True)
    filt = search.run(ctx, test["qres"], test["qoff"], sch, k, prefilter=search.prefilter_arg(pf), regions=True)
    pp, fp = search.hit_pairs(plain), search.hit_pairs(filt)
    loss = len(pp - fp) / max(1, len(pp))
    kept = 1 - filt.counters["regions_pruned"] / max(1, filt.counters["regions"])
    print(f"# test: plain hits {len(pp)}, lost {len(pp - fp)} ({loss:.4f}), kept {kept:.3f}, band cells "
          f"{plain.counters['banded_cells']} -> {filt.counters['banded_cells']}", flush=True)
    assert loss <= 0.01
    assert filt.counters["regions_pruned"] > 0 and kept < 1
    assert filt.counters["banded_cells"] < plain.counters["banded_cells"]
    # a plain hit whose region the filter kept is reported unchanged (pruning
    # only removes candidates ahead of it in the reporting loop)
    fk = {(q, h.subject): h.key() for q, hl in enumerate(filt.hits) for h in hl}
    checked = 0
    for q, hl in enumerate(plain.hits):
        rp, rf = plain.regions[q], filt.regions[q]
        assert np.array_equal(rp["subject"], rf["subject"]) and np.array_equal(rp["d_min"], rf["d_min"])
        for h in hl:
            i = [j for j in range(len(rp)) if rp["subject"][j] == h.subject and rp["flags"][j] & 4]
            assert len(i) == 1
            if not rf["flags"][i[0]] & 8:
                assert fk.get((q, h.subject)) == h.key()
                checked += 1
    assert checked > 0.9 * len(pp)
    pf["model"].close()
