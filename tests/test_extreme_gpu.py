"""Extreme scoring parameters through every GPU kernel against the oracle:
query biases of +127, -128 and mixed extremes at every position; a gap
extension dearer than the opening (ge > go, which the reference's int16
kernel declines, banded_simd.rs:412-415, so its scalar kernel answers) and a
free opening (go = 0); substitution scores of -128 and 127 in the matrix.

Kernels: the packed band kernel in both builds (scores only, end cells) and
forced through the int32 kernel; the full-matrix kernel (scores, end cells);
X-drop; the banded traceback. Results must equal the oracle's restatement
of the reference's scalar path exactly."""
import numpy as np
import pytest

import saber_amd as S
from oracle import oracle as O
from saber_amd import BLOSUM62
from tests.gen import Lcg, flat, jobs_array, mutate, rand_seq

pytestmark = pytest.mark.gpu

GAPS = [(11, 1), (1, 5), (0, 1), (0, 0), (4, 4)]


def extreme_matrix():
    m = np.array(BLOSUM62, np.int32).reshape(24, 24).copy()
    for a in range(0, 20, 3):
        m[a, a] = 127                   # a few residues match at the int8 top
    for a, b in [(1, 2), (2, 1), (5, 7), (7, 5), (10, 3), (3, 10), (22, 0)]:
        m[a, b] = -128
    return m


def bias_cases(n, r):
    return {
        "plus127": np.full(n, 127, np.int8),
        "minus128": np.full(n, -128, np.int8),
        "mixed": np.array([(127, -128, 0, 127)[r.next(4)] for _ in range(n)], np.int8),
        "random": np.array([r.next(256) - 128 for _ in range(n)], np.int8),
    }


def dataset(seed, nq=4, per_q=6):
    r = Lcg(seed)
    qs, ss, rows = [], [], []
    for qi in range(nq):
        q = rand_seq(r, 10 + r.next(70))
        qs.append(q)
        for k in range(per_q):
            s = mutate(r, q) if k % 2 == 0 else rand_seq(r, 5 + r.next(90))
            if len(s) == 0:
                s = rand_seq(r, 3)
            si = len(ss)
            ss.append(s)
            c = r.next(21) - 10
            h = (2, 8, 20, 70)[k % 4]
            rows.append((qi, si, 0, len(s), c - h, c + h))
    qres, qoff = flat(qs)
    sres, soff = flat(ss)
    jobs = jobs_array(*[np.array([x[i] for x in rows]) for i in range(6)])
    return qs, ss, qres, qoff, sres, soff, jobs, r


@pytest.mark.parametrize("go,ge", GAPS)
@pytest.mark.parametrize("mat_kind", ["blosum62", "extreme"])
def test_band_kernels(ctx, ctx32, go, ge, mat_kind):
    qs, ss, qres, qoff, sres, soff, jobs, r = dataset(100 + go * 7 + ge)
    mat = extreme_matrix() if mat_kind == "extreme" else np.array(BLOSUM62, np.int32).reshape(24, 24)
    sch = S.Scheme(mat.tolist(), go, ge)
    for name, bias in bias_cases(len(qres), r).items():
        want = [O.banded_score(qs[j["query"]], ss[j["subject"]], bias[int(qoff[j["query"]]):int(qoff[j["query"] + 1])],
                               mat, go, ge, int(j["band_lo"]), int(j["band_hi"])) for j in jobs]
        for c, label in ((ctx, "packed"), (ctx32, "int32")):
            print(f"# band {label} {mat_kind} go={go} ge={ge} bias={name}", flush=True)
            sc, qe, se, cells = c.score_bands(qres, qoff, sres, soff, jobs, sch, qbias=bias, ends=True)
            so = c.score_bands(qres, qoff, sres, soff, jobs, sch, qbias=bias, ends=False)[0]
            for k, w in enumerate(want):
                assert (int(sc[k]), int(qe[k]), int(se[k]), int(cells[k])) == tuple(w), (label, name, k)
                assert int(so[k]) == w[0], (label, name, k)


@pytest.mark.parametrize("go,ge", GAPS)
def test_full_kernel(ctx, go, ge):
    qs, ss, qres, qoff, sres, soff, jobs, r = dataset(200 + go * 7 + ge, nq=3, per_q=10)
    mat = extreme_matrix()
    sch = S.Scheme(mat.tolist(), go, ge)
    for name, bias in bias_cases(len(qres), r).items():
        print(f"# full go={go} ge={ge} bias={name}", flush=True)
        sc, qe, se = ctx.score_full(qres, qoff, sres, soff, sch, qbias=bias, ends=True)
        so = ctx.score_full(qres, qoff, sres, soff, sch, qbias=bias, ends=False)[0]
        for i, q in enumerate(qs):
            b = bias[int(qoff[i]):int(qoff[i + 1])]
            for j, s in enumerate(ss):
                w = O.sw_reference(q, s, mat, b, go, ge)
                assert (int(sc[i, j]), int(qe[i, j]), int(se[i, j])) == w, (name, i, j)
                assert int(so[i, j]) == w[0], (name, i, j)


@pytest.mark.parametrize("mat_kind", ["blosum62", "extreme"])
def test_xdrop_kernel(ctx, mat_kind):
    qs, ss, qres, qoff, sres, soff, jobs, r = dataset(300, nq=5, per_q=8)
    mat = extreme_matrix() if mat_kind == "extreme" else np.array(BLOSUM62, np.int32).reshape(24, 24)
    sch = S.Scheme(mat.tolist(), 11, 1)
    hits = []
    for j in jobs:
        q, s = qs[j["query"]], ss[j["subject"]]
        for _ in range(4):
            hits.append((int(j["query"]), int(j["subject"]), r.next(len(q)), r.next(len(s))))
    hits = np.array(hits, np.uint32)
    for name, bias in bias_cases(len(qres), r).items():
        for x in (1, 20, 300):
            print(f"# xdrop {mat_kind} x={x} bias={name}", flush=True)
            sc, lo, hi, st = ctx.xdrop(qres, qoff, sres, soff, hits, sch, x, qbias=bias)
            for k, (qi, si, qp, sp) in enumerate(hits.tolist()):
                b = bias[int(qoff[qi]):int(qoff[qi + 1])]
                w = O.xdrop(qs[qi], ss[si], b, mat, qp, sp, x)
                assert (int(sc[k]), int(lo[k]), int(hi[k]), int(st[k])) == w, (name, x, k)


@pytest.mark.parametrize("go,ge", GAPS)
def test_traceback_kernel(ctx, go, ge):
    qs, ss, qres, qoff, sres, soff, jobs, r = dataset(400 + go * 7 + ge)
    mat = extreme_matrix()
    sch = S.Scheme(mat.tolist(), go, ge)
    for name, bias in bias_cases(len(qres), r).items():
        print(f"# traceback go={go} ge={ge} bias={name}", flush=True)
        got = ctx.band_align(qres, qoff, sres, soff, jobs, sch, qbias=bias)
        for k, j in enumerate(jobs):
            b = bias[int(qoff[j["query"]]):int(qoff[j["query"] + 1])]
            w = O.band_align(qs[j["query"]], ss[j["subject"]], mat, b, go, ge, int(j["band_lo"]), int(j["band_hi"]))
            assert got[k] == w, (name, k)
