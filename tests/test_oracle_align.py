"""The oracle's banded traceback (or_band_align: banded_align_scalar +
walk, banded.rs:227-403, with the AVX2 kernels' row stop) pinned by what fixes
it independently: the score, end cell and cells of the score-only oracle
(itself pinned by sw_reference), a CIGAR rescored column by column
(score_alignment, reference.rs:58-105) that must reproduce the score and end
at the end cell, every column inside the band, the reference's own cases
(banded.rs:565-767) and a closed-form known answer."""
import numpy as np
import pytest

from oracle import oracle as O
from saber_amd import BLOSUM62, PAM30
from tests.gen import Lcg, mutate, rand_seq

W = 17


def rescore(q, s, q_start, s_start, ops, mat, bias, go, ge):
    """score_alignment (reference.rs:58-105): M columns score S(q_i, s_j) +
    bias_i, a gap run of length n costs go + (n - 1) ge."""
    i, j, score = q_start, s_start, 0
    for op, n in ops:
        assert n > 0
        if op == b"M":
            for _ in range(n):
                assert i < len(q) and j < len(s)
                score += int(mat[q[i] if q[i] < 24 else 22][s[j] if s[j] < 24 else 22])   # codes >= 24 score as X
                score += 0 if bias is None else int(bias[i])
                i += 1
                j += 1
        elif op == b"I":
            assert i + n <= len(q)
            score -= go + (n - 1) * ge
            i += n
        else:
            assert op == b"D" and j + n <= len(s)
            score -= go + (n - 1) * ge
            j += n
    return score, i, j


def columns_inside(al, lo, hi):
    i, j = al[1], al[3]
    for op, n in al[5]:
        for _ in range(n):
            if not (lo <= j - i <= hi):
                return False
            if op == b"M":
                i, j = i + 1, j + 1
            elif op == b"I":
                i += 1
            else:
                j += 1
    return True


def test_full_band_equals_reference_and_cigar_rescores():
    r = Lcg(42)
    for t in range(300):
        a = rand_seq(r, 1 + r.next(160))
        b = mutate(r, a) if t % 2 == 0 else rand_seq(r, 1 + r.next(160))
        if len(b) == 0:
            continue
        bias = np.array([r.next(7) - 3 for _ in range(len(a))], np.int8) if t % 3 == 0 else None
        rs, ri, rj = O.sw_reference(a, b, BLOSUM62, bias, 11, 1)
        al = O.band_align(a, b, BLOSUM62, bias, 11, 1, -(len(a) - 1), len(b) - 1)
        if rs == 0:
            assert al is None
            continue
        sc, qs, qe, ss, se, ops, cells = al
        assert (sc, qe, se) == (rs, ri, rj), t
        assert rescore(a, b, qs, ss, ops, BLOSUM62, bias, 11, 1) == (rs, qe, se), (t, ops)
        assert cells == len(a) * len(b)


@pytest.mark.parametrize("gaps", [(11, 1), (10, 2), (6, 3), (8, 8)])
def test_band_alignment_matches_banded_score_and_stays_inside(gaps):
    go, ge = gaps
    r = Lcg(7 + go)
    for t in range(120):
        a = rand_seq(r, 50 + r.next(200))
        b = mutate(r, a) if t % 2 == 0 else rand_seq(r, 20 + r.next(200))
        if len(b) == 0:
            continue
        bias = np.array([r.next(9) - 4 for _ in range(len(a))], np.int8) if t % 3 == 0 else None
        for lo, hi in ((-2, 2), (-8, 8), (-32, 32), (5, 40), (-(r.next(60)), r.next(60) + 300)):
            want = O.banded_score(a, b, bias, BLOSUM62, go, ge, lo, hi)
            al = O.band_align(a, b, BLOSUM62, bias, go, ge, lo, hi)
            if want[0] == 0:
                assert al is None
                continue
            sc, qs, qe, ss, se, ops, cells = al
            assert (sc, qe, se, cells) == tuple(int(x) for x in want), (t, lo, hi)
            assert rescore(a, b, qs, ss, ops, BLOSUM62, bias, go, ge) == (sc, qe, se)
            assert columns_inside(al, lo, hi)


def test_stopping_at_the_known_score_gives_the_same_alignment():
    """banded.rs:734-758"""
    r = Lcg(5)
    saved = 0
    for _ in range(200):
        la = 20 + r.next(300)
        a = rand_seq(r, la)
        b = np.concatenate([rand_seq(r, 40), mutate(r, a[:la // 2]), rand_seq(r, 40)]).astype(np.uint8)
        full = O.band_align(a, b, BLOSUM62, None, 11, 1, 0, 90)
        if full is None:
            continue
        early = O.band_align(a, b, BLOSUM62, None, 11, 1, 0, 90, stop_at=full[0])
        assert early[:6] == full[:6]
        assert early[6] <= full[6]
        saved += full[6] - early[6]
    assert saved > 0


def test_memory_cap_and_known_answers():
    a = np.full(1000, W, np.uint8)
    assert O.band_align(a, a, BLOSUM62, None, 11, 1, -500, 500, max_dir_bytes=1000) == "cap"
    a = np.full(4000, W, np.uint8)
    al = O.band_align(a, a, BLOSUM62, None, 11, 1, -20, 20)
    assert al[:6] == (44000, 0, 4000, 0, 4000, [(b"M", 4000)])
    # WAGNER self-alignment (reference.rs:127-134): 37 at (6, 6) as 6M
    from saber_amd import encode
    wag = encode(b"WAGNER")
    al = O.band_align(wag, wag, BLOSUM62, None, 11, 1, -5, 5)
    assert al[:6] == (37, 0, 6, 0, 6, [(b"M", 6)])


def test_gapped_alignment_is_found_in_a_wide_band():
    """A deletion in the subject: the CIGAR carries it and rescores exactly."""
    r = Lcg(11)
    a = rand_seq(r, 200)
    b = np.concatenate([a[:90], a[100:]]).astype(np.uint8)   # 10 residues deleted from the subject
    al = O.band_align(a, b, PAM30, None, 9, 1, -30, 30)
    ops = al[5]
    assert any(op == b"I" and n == 10 for op, n in ops), ops
    assert rescore(a, b, al[1], al[3], ops, PAM30, None, 9, 1) == (al[0], al[2], al[4])
