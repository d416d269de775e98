"""Pins the CPU oracle (CPU only): known answers from the reference's tests,
brute force on tiny inputs, an independent (i, j)-coordinate banded DP, closed
forms for X-drop, the reference's own CUDA kernel source compiled as C++ (when
/root/reference is mounted), and the composition-bias invariants."""
import itertools
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from saber_amd import BLOSUM45, BLOSUM62, PAM30, encode
from tests.gen import Lcg, mutate, rand_seq

W = 17  # code of W


def nucleotide(match, mismatch):
    m = [[mismatch] * 24 for _ in range(24)]
    for i in range(4):
        m[i][i] = match
    return m


# ---- known answers (reference tests) ----

def test_wagner_self_score():
    # reference.rs:127-134
    q = encode("WAGNER")
    assert O.sw_reference(q, q, BLOSUM62, None, 11, 1) == (37, 6, 6)


def test_long_identical_runs():
    # banded.rs:719-728, sw_simd.rs:524-532, doctor.rs:284-296
    a = np.full(4000, W, np.uint8)
    assert O.sw_reference(a, a, BLOSUM62, None, 11, 1)[0] == 44000
    assert O.sw_reference(a[:3500], a[:3500], BLOSUM62, None, 11, 1)[0] == 38500
    assert O.banded_score(a, a, None, BLOSUM62, 11, 1, -20, 20)[0] == 44000


def test_w300_band_5():
    # band_simd8.rs:454-468
    q = np.full(300, W, np.uint8)
    assert O.banded_score(q, q, None, BLOSUM62, 11, 1, -5, 5)[0] == 3300


def test_nucleotide_scores():
    # gpu/mod.rs:178-182 (ACGTN -> 0..4)
    nt = nucleotide(2, -3)
    enc = {c: i for i, c in enumerate("ACGTN")}
    a = np.array([enc[c] for c in "ACGTACGTAC"], np.uint8)
    b = np.array([enc[c] for c in "ACGTNNACGT"], np.uint8)
    assert O.sw_reference(a, a, nt, None, 7, 2)[0] == 20
    assert O.sw_reference(a, b, nt, None, 7, 2)[0] == 8


# ---- brute force on tiny inputs ----

def brute_force(q, s, mat, bias, go, ge):
    """Max over every local alignment (M/I/D op strings) of its score, and
    the first row-major end cell (exclusive) reached by a best alignment
    ending in an aligned pair."""
    m, n = len(q), len(s)
    best_at = {}

    def S(i, j):
        a = q[i] if q[i] < 24 else 22
        b = s[j] if s[j] < 24 else 22
        return mat[a][b] + (int(bias[i]) if bias is not None else 0)

    def walk(i, j, sc, last):
        # an aligned pair may follow anything; a query-gap run (I) may follow
        # a pair or a subject-gap run, and vice versa (Gotoh opens a gap from
        # H, which may itself come from the other gap state)
        if i < m and j < n:
            v = sc + S(i, j)
            key = (i + 1, j + 1)
            best_at[key] = max(best_at.get(key, -10**9), v)
            walk(i + 1, j + 1, v, "M")
        if last in ("M", "D"):
            for k in range(1, m - i + 1):
                walk(i + k, j, sc - go - (k - 1) * ge, "I")
        if last in ("M", "I"):
            for k in range(1, n - j + 1):
                walk(i, j + k, sc - go - (k - 1) * ge, "D")

    for i in range(m):
        for j in range(n):
            walk(i, j, 0, None)
    best = max([0] + list(best_at.values()))
    if best <= 0:
        return 0, 0, 0
    end = min(k for k, v in best_at.items() if v == best)
    return best, end[0], end[1]


@pytest.mark.parametrize("seed", range(40))
def test_sw_reference_matches_brute_force(seed):
    r = Lcg(1000 + seed)
    mats = [BLOSUM62, PAM30, BLOSUM45]
    mat = mats[seed % 3]
    m = 1 + r.next(5)
    n = 1 + r.next(5)
    q = rand_seq(r, m, 20)
    s = q.copy() if seed % 4 == 0 and m == n else rand_seq(r, n, 20)
    bias = np.array([r.next(7) - 3 for _ in range(m)], np.int8) if seed % 2 else None
    go, ge = [(11, 1), (4, 2), (2, 1), (6, 6)][seed % 4]
    got = O.sw_reference(q, s, mat, bias, go, ge)
    want = brute_force(q, s, mat, bias, go, ge)
    assert got == want


# ---- banded invariants ----

def test_full_band_equals_reference_and_ij_form():
    # banded.rs:565-596 plus the independent (i, j)-coordinate DP
    r = Lcg(42)
    for t in range(120):
        la = 1 + r.next(120)
        a = rand_seq(r, la)
        lb = 1 + r.next(120)
        b = mutate(r, a) if t % 2 == 0 else rand_seq(r, lb)
        if len(b) == 0:
            b = rand_seq(r, 5)
        bias = np.array([r.next(7) - 3 for _ in range(la)], np.int8) if t % 3 == 0 else None
        rs, ri, rj = O.sw_reference(a, b, BLOSUM62, bias, 11, 1)
        bs = O.banded_score(a, b, bias, BLOSUM62, 11, 1, -la, len(b))
        assert bs[0] == rs
        if rs > 0:
            assert (bs[1], bs[2]) == (ri, rj)
        assert bs[3] == la * len(b)
        lo = -int(r.next(40))
        hi = int(r.next(40))
        x = O.banded_score(a, b, bias, BLOSUM62, 11, 1, lo, hi)
        y = O.banded_score_ij(a, b, bias, BLOSUM62, 11, 1, lo, hi)
        assert x[:3] == y
        assert x[0] <= rs


def test_band_cells_count():
    r = Lcg(3)
    for _ in range(200):
        m = 1 + r.next(50)
        n = 1 + r.next(50)
        lo = int(r.next(120)) - 60
        hi = lo + int(r.next(40))
        a = rand_seq(r, m)
        b = rand_seq(r, n)
        cells = O.banded_score(a, b, None, BLOSUM62, 11, 1, lo, hi)[3]
        want = sum(1 for i in range(m) for j in range(n) if lo <= j - i <= hi)
        assert cells == want


def test_shifted_diagonal_needs_band_on_that_diagonal():
    # banded.rs:630-643
    core = encode("MKVLAAGIVGHHWWPQRSTVYCCMKDEFGHIKLMNPQRSTVWY")
    subj = np.concatenate([np.zeros(300, np.uint8), core])
    rs = O.sw_reference(core, subj, BLOSUM62, None, 11, 1)[0]
    near = O.banded_score(core, subj, None, BLOSUM62, 11, 1, -10, 10)[0]
    on = O.banded_score(core, subj, None, BLOSUM62, 11, 1, 290, 310)[0]
    assert near < rs // 2
    assert on == rs


# ---- X-drop ----

def test_xdrop_matches_scalar_twin_and_closed_forms():
    r = Lcg(77)
    for t in range(400):
        m = 5 + r.next(120)
        q = rand_seq(r, m)
        s = mutate(r, q) if t % 2 == 0 else rand_seq(r, 5 + r.next(120))
        if len(s) == 0:
            continue
        bias = np.array([r.next(9) - 4 for _ in range(m)], np.int8) if t % 3 == 0 else None
        qp = int(r.next(m))
        sp = int(r.next(len(s)))
        x = [10, 20, 35][t % 3]
        sc, lo, hi, steps = O.xdrop(q, s, bias, BLOSUM62, qp, sp, x)
        assert (sc, lo, hi) == O.xdrop_scalar_twin(q, s, bias, BLOSUM62, qp, sp, x)
        # x -> infinity: best prefix sum to the right (incl. the seed pair) plus
        # best suffix sum to the left, first strict maxima.
        big = 10**6
        sc2, lo2, hi2, _ = O.xdrop(q, s, bias, BLOSUM62, qp, sp, big)

        def col(i, j):
            a = int(q[i]) if q[i] < 24 else 22
            b = int(s[j]) if s[j] < 24 else 22
            return BLOSUM62[a][b] + (int(bias[i]) if bias is not None else 0)

        cur = best = 0
        end = 0
        for t2 in range(min(m - qp, len(s) - sp)):
            cur += col(qp + t2, sp + t2)
            if cur > best:
                best, end = cur, t2 + 1
        curl = bestl = 0
        st = 0
        for t2 in range(1, min(qp, sp) + 1):
            curl += col(qp - t2, sp - t2)
            if curl > bestl:
                bestl, st = curl, t2
        assert (sc2, lo2, hi2) == (best + bestl, qp - st, qp + end)
        assert steps >= 1 or min(m - qp, len(s) - sp) == 0


def test_xdrop_bounded_by_best_ungapped_segment():
    r = Lcg(5)
    for _ in range(200):
        m = 10 + r.next(80)
        q = rand_seq(r, m, 20)
        s = mutate(r, q)
        if len(s) < 2:
            continue
        qp = int(r.next(m))
        sp = int(r.next(len(s)))
        sc = O.xdrop(q, s, None, BLOSUM62, qp, sp, 20)[0]
        assert sc <= O.best_ungapped_on_diagonal(q, s, sp - qp, BLOSUM62, None)


# ---- composition bias ----

def test_composition_bias_invariants():
    # bias.rs:186-206
    s = b"MKTAYIAKQRQISFVKSHFSRQ" + b"Q" * 40 + (
        b"LEDNWETLNDNLKVIEKADNAAQVKDALTKMRAAALDAQKATPPKLEDKSPDSPEMKDFRHGFDILVGQIDDALKLANEGKVKEAQAAAEQLKTTRNAYHQKYR")
    c = O.composition_bias(encode(s), BLOSUM62)
    assert c[22 + 20] <= -3
    assert abs(c[80:120].astype(float).mean()) < 1.5
    assert (O.composition_bias(encode("XXXXBZ*XXXX"), BLOSUM62) == 0).all()


def test_composition_bias_direct_formula():
    robinson = [78.05, 51.29, 44.87, 53.64, 19.25, 42.64, 62.95, 73.77, 21.99, 51.42, 90.19, 57.44,
                22.43, 38.56, 52.03, 71.20, 58.41, 13.30, 32.16, 64.41]
    tot = sum(robinson)
    p = [x / tot for x in robinson]
    r = Lcg(9)
    for _ in range(20):
        n = 2 + r.next(120)
        q = rand_seq(r, n, 24)
        got = O.composition_bias(q, BLOSUM62)
        for i in range(n):
            a = int(q[i])
            if a >= 20:
                assert got[i] == 0
                continue
            lo, hi = max(0, i - 20), min(n, i + 21)
            neigh = sum(BLOSUM62[a][min(int(q[j]), 23)] for j in range(lo, hi) if j != i) / (hi - lo - 1)
            bg = sum(p[x] * BLOSUM62[a][x] for x in range(20))
            f = bg - neigh
            v = int(f - 0.5) if f < 0 else int(f + 0.5)
            assert got[i] == max(-127, min(127, v))


# ---- the reference's own CUDA kernel compiled as C++ (oracle/_ref) ----

REF_EMU = os.path.join(os.path.dirname(O.HERE), "oracle", "_ref", "emulate")


@pytest.mark.skipif(not os.path.exists(REF_EMU), reason="reference emulator not built (needs /root/reference)")
def test_reference_kernel_emulation_agrees_with_oracle(tmp_path):
    """Vectors in the harness format of tools/gpu_emulate/emulate.cpp with
    expected scores from the oracle; the reference kernel must agree."""
    r = Lcg(0xC0DA)
    nt = nucleotide(2, -3)
    lines = []
    for case in range(24):
        mat, alpha, go, ge = [(BLOSUM62, 24, 11, 1), (PAM30, 20, 10, 1), (nt, 5, 7, 2),
                              (BLOSUM62, 24, 13, 2)][case % 4]
        ql = 1 + r.next(300)
        q = rand_seq(r, ql, alpha)
        bias = np.array([r.next(5) - 2 for _ in range(ql)], np.int8) if case % 3 == 0 else None
        wins = []
        for k in range(1 + r.next(40)):
            if k % 5 == 0:
                wins.append(mutate(r, q[: 1 + r.next(ql)]))
            else:
                wins.append(rand_seq(r, r.next(400), alpha))
        exp = [O.sw_reference(q, w, mat, bias, go, ge)[0] for w in wins]
        b = bias if bias is not None else np.zeros(ql, np.int8)
        lines.append(f"CASE {ql} {len(wins)} {go} {ge}")
        lines.append(" ".join(map(str, q.tolist())))
        lines.append(" ".join(map(str, b.tolist())))
        lines.append(" ".join(str(v) for row in mat for v in row))
        for w in wins:
            lines.append(f"{len(w)} " + " ".join(map(str, w.tolist())))
        lines.append(" ".join(map(str, exp)))
    vec = tmp_path / "vectors.txt"
    vec.write_text("\n".join(lines) + "\n")
    out = subprocess.run([REF_EMU, str(vec)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert '"mismatches": 0' in out.stdout
