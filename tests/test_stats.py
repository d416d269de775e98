"""Host statistics (saber_hip_stats_*) against the reference's own tests of
src/scoring/stats.rs (:760-890) and the defining property of BLAST's length
adjustment. CPU only: the host code needs no device."""
import math

import numpy as np
import pytest

from saber_amd import stats as ST


def b62(open_=12, ext=1):
    return ST.gapped_protein(62, open_, ext)


def test_parameter_rows():
    p = ST.gapped_protein(62, 11, 1)          # SABER 11/1 = BLAST 10/1
    assert abs(p.lambda_ - 0.243) < 1e-12 and abs(p.k - 0.024) < 1e-12 and p.calibrated
    p = ST.gapped_protein(62, 12, 1)          # BLAST 11/1
    assert abs(p.lambda_ - 0.267) < 1e-12 and abs(p.k - 0.041) < 1e-12
    assert abs(p.alpha_d_lambda - 1.9 / 0.267) < 1e-12 and p.beta == -30.0
    assert abs(ST.gapped_protein(45, 17, 1).lambda_ - 0.186) < 1e-12   # BLAST 16/1
    with pytest.raises(ValueError):
        ST.gapped_protein(62, 5, 1)
    u = ST.gapped_protein(62, 5, 1, allow_uncalibrated=True)
    assert not u.calibrated and abs(u.lambda_ - 0.3176) < 1e-12
    assert abs(u.alpha_d_lambda - 1 / 0.4012) < 1e-12 and u.beta == 0.0


def test_database_model_is_unchanged_by_the_subject_length():
    q = ST.SearchStatistics(b62(), 38_000_000, 91_000, ST.DATABASE).for_query(170)
    for n in (30, 170, 5000):
        for raw in (20, 60, 200):
            assert q.evalue_for(raw, n) == q.evalue(raw)
            assert q.space_lower_bound(n) == q.search_space


def test_per_subject_lower_bound_never_exceeds_the_exact_space():
    st = ST.SearchStatistics(b62(), 38_000_000, 91_000, ST.PER_SUBJECT)
    for m in (1, 5, 24, 25, 60, 170, 1000, 35_000):
        q = st.for_query(m)
        for n in (1, 3, 24, 25, 26, 40, 100, 420, 3000, 35_000):
            lb, ex = q.space_lower_bound(n), q.space_for(n)
            assert lb <= ex * (1 + 1e-12)
            assert ex >= st.db_num_seqs


def test_models_agree_on_equal_lengths_and_evalues_grow_with_length():
    z, n = 10_000, 300
    db = ST.SearchStatistics(b62(), z * n, z, ST.DATABASE).for_query(250)
    ps = ST.SearchStatistics(b62(), z * n, z, ST.PER_SUBJECT).for_query(250)
    assert 0.5 <= ps.evalue_for(80, n) / db.evalue(80) < 2.0
    last = 0.0
    for length in (50, 100, 300, 1000, 10_000):
        e = ps.evalue_for(80, length)
        assert e > last
        last = e
    assert ps.bitscore(80) == db.bitscore(80)


def test_length_adjustment_blast_scale_and_fixed_point():
    p = b62()
    adj, conv = ST.compute_length_adjustment(p.k, p.log_k, p.alpha_d_lambda, p.beta, 300, 100_000_000, 300_000)
    assert conv and 80 < adj < 160
    adj2, _ = ST.compute_length_adjustment(p.k, p.log_k, p.alpha_d_lambda, p.beta, 300, 1_000_000_000, 3_000_000)
    assert adj2 > adj
    # BLAST's adjustment is the largest integer l with f(l) >= l, where
    # f(l) = alpha/lambda (ln K + ln((m - l)(n - N l))) + beta
    rng = np.random.default_rng(1)
    for _ in range(300):
        m = int(rng.integers(30, 5000))
        nn = int(rng.integers(100, 500_000))
        n = nn * int(rng.integers(50, 600))
        a, conv = ST.compute_length_adjustment(p.k, p.log_k, p.alpha_d_lambda, p.beta, m, n, nn)
        if not conv:
            continue

        def f(ell):
            ss = (m - ell) * (n - nn * ell)
            return p.alpha_d_lambda * (p.log_k + math.log(ss)) + p.beta if ss > 0 else -math.inf
        assert f(a) >= a - 1e-9 and f(a + 1) < a + 1, (m, n, nn, a)


def test_min_score_is_exact_inverse():
    q = ST.SearchStatistics(ST.gapped_protein(62, 11, 1), 2_000_000, 11_000).for_query(170)
    for cut in (10.0, 1.0, 1e-3, 1e-10):
        s = q.min_score(cut)
        assert q.evalue(s) <= cut and q.evalue(s - 1) > cut
    s = q.min_score(10.0, 50.0)
    assert (q.bitscore(s) >= 50.0 and q.bitscore(s - 1) < 50.0) or q.evalue(s - 1) > 10.0


def test_evalue_batch_equals_single_calls():
    p = b62()
    st = ST.SearchStatistics(p, 147_000_000, 331_000, ST.PER_SUBJECT)
    rng = np.random.default_rng(9)
    m = np.sort(rng.integers(20, 2000, 300))
    n = rng.integers(20, 5000, 300)
    raw = rng.integers(-5, 300, 300)
    got = ST.evalues(p, 147_000_000, 331_000, ST.PER_SUBJECT, m, n, raw)
    want = [st.for_query(int(a)).evalue_for(int(r), int(b)) for a, b, r in zip(m, n, raw)]
    assert np.array_equal(got, np.array(want))


def test_floor_lower_bound_below_exact_floor():
    # the per-subject floor bound never exceeds the smallest passing score
    q = ST.SearchStatistics(b62(), 147_000_000, 331_000, ST.PER_SUBJECT).for_query(170)
    for n in (30, 100, 170, 600, 3000):
        lb = q.floor_lower_bound(10.0, None, n, 0)
        s = next(s for s in range(1, 1000) if q.evalue_for(s, n) <= 10.0)
        assert lb <= s
    qd = ST.SearchStatistics(b62(), 147_000_000, 331_000, ST.DATABASE).for_query(170)
    assert qd.floor_lower_bound(10.0, None, 100, 57) == 57.0
