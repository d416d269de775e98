"""Pins the oracle's scan schedule (oracle/scan.c, or_scan) -- the checker of
the GPU search -- against what the definitions fix, on small inputs:

* the window-filter survivors equal a brute-force enumeration: every (shape,
  query position, subject position) whose residues at every care position
  are seed letters with equal letters (SeedShape / seed_code, alphabet.rs,
  scan.rs:811-823), whose code is not over-frequent in the database
  (auto_occ_cap, index.rs:329-336, counted directly), and whose 32-residue
  window value (seed-letter identities + kx x exact identities, scan.rs:
  1812-1856), evaluated position by position, reaches min_ids; in visit order
  (per subject, per query, by (query position, shape, subject position));
* the regions equal the host restatement (oracle/host.py visit_hits and
  finish_subject, pinned in tests/test_pipeline.py) replayed over those
  survivors with or_xdrop results, then capped under region_order;
* auto_occ_cap's closed form; the library's composition_bias equals the
  oracle's (bias.rs:80-122) on random and edge-case queries.
"""
import math

import numpy as np
import pytest

from oracle import host as H
from oracle import oracle as O
from oracle import search as OS
from saber_amd import BLOSUM62
from saber_amd.search import DS16, MURPHY10_GROUPS, alphabet_map, composition_bias, full_alphabet
from tests.gen import Lcg, flat, mutate, rand_seq


def brute_survivors(qs, ss, amap, asize, shapes, min_ids, kx, seed_max_occ):
    """Every window-filter survivor by direct enumeration, in visit order."""
    c_half = min(len(s) for s in shapes) // 2
    red = lambda x: int(amap[min(int(x), 23)])   # 255: not a seed letter
    surv = []
    # database counts of every code, per shape (masking off)
    counts = []
    for sh in shapes:
        care = [i for i, ch in enumerate(sh) if ch == "1"]
        cnt, tot = {}, 0
        for s in ss:
            for p in range(len(s) - len(sh) + 1):
                w = tuple(red(s[p + o]) for o in care)
                if 255 in w:
                    continue
                cnt[w] = cnt.get(w, 0) + 1
                tot += 1
        ncodes = asize ** len(care)
        cap = seed_max_occ if seed_max_occ > 0 else max(math.ceil(tot / ncodes * 64), 256)
        counts.append((cnt, cap))
    for si_, s in enumerate(ss):
        groups = {}
        for shi, sh in enumerate(shapes):
            care = [i for i, ch in enumerate(sh) if ch == "1"]
            cnt, cap = counts[shi]
            for p in range(len(s) - len(sh) + 1):
                w = tuple(red(s[p + o]) for o in care)
                if 255 in w:
                    continue
                for qi, q in enumerate(qs):
                    for qp in range(len(q) - len(sh) + 1):
                        if tuple(red(q[qp + o]) for o in care) != w or cnt.get(w, 0) > cap:
                            continue
                        if min_ids > 0:
                            r = e = 0
                            for t in range(32):
                                a, b = qp + c_half - 16 + t, p + c_half - 16 + t
                                if 0 <= a < len(q) and 0 <= b < len(s) and red(q[a]) != 255:
                                    if red(q[a]) == red(s[b]):
                                        r += 1
                                        if int(q[a]) == int(s[b]):
                                            e += 1
                            if r + kx * e < min_ids:
                                continue
                        groups.setdefault(qi, []).append((qp, shi, p))
        for qi in sorted(groups):
            for qp, shi, p in sorted(groups[qi]):
                surv.append((qi, si_, qp, p, shi))
    return surv


def point_mutate(r, s, pct):
    """s with each residue replaced at random with probability pct/100."""
    return np.array([r.next(20) if r.next(100) < pct else int(x) for x in s], np.uint8)


def related_set(seed, nq, ns, qlen=(25, 60), slen=(20, 70), alpha=20):
    """Queries; subjects cycling through an indel-mutated copy of a query
    (banded.rs-style mutate), a copy with 15 % point substitutions, and an
    unrelated sequence."""
    r = Lcg(seed)
    qs = [rand_seq(r, qlen[0] + r.next(qlen[1] - qlen[0]), alpha) for _ in range(nq)]
    ss = []
    for k in range(ns):
        flank = lambda: rand_seq(r, r.next(8), alpha)
        if k % 3 == 0:
            ss.append(np.concatenate([flank(), mutate(r, qs[r.next(nq)]), flank()]).astype(np.uint8))
        elif k % 3 == 1:
            ss.append(np.concatenate([flank(), point_mutate(r, qs[r.next(nq)], 15), flank()]).astype(np.uint8))
        else:
            ss.append(rand_seq(r, slen[0] + r.next(slen[1] - slen[0]), alpha))
    return qs, ss


@pytest.mark.parametrize("alpha,shapes,min_ids,kx,occ", [
    ("murphy10", ["1101", "111", "11011"], 8, 2, 0),
    ("murphy10", ["11110111", "110010101011"], 12, 2, 0),
    ("full", ["111", "1011"], 4, 1, 0),
    ("murphy10", ["1101", "111"], 0, 0, 0),      # no window filter
    ("murphy10", ["111", "1101"], 6, 2, 3),      # over-frequent codes skipped
])
def test_survivors_match_brute_force(alpha, shapes, min_ids, kx, occ):
    amap, asize = alphabet_map(MURPHY10_GROUPS) if alpha == "murphy10" else full_alphabet()
    qs, ss = related_set(7 + len(shapes) + min_ids, 3, 6, alpha=24 if alpha == "full" else 20)
    qres, qoff = flat(qs)
    sres, soff = flat(ss)
    regs, ctr, surv = OS.scan(qres, qoff, None, sres, soff, BLOSUM62, amap, asize, shapes, [40] * len(qs),
                              min_ids=min_ids, kx=kx, seed_max_occ=occ, hsp_threshold=10, log_survivors=True)
    want = brute_survivors(qs, ss, amap, asize, shapes, min_ids, kx, occ)
    got = [tuple(int(x) for x in row) for row in surv]
    assert len(want) > 20
    assert got == want
    assert ctr["hits"] - ctr["hits_filtered"] == len(want)
    if occ:
        assert ctr["seeds_frequent_skipped"] > 0


@pytest.mark.parametrize("seed", [1, 2])
def test_regions_match_host_restatement(seed):
    amap, asize = alphabet_map(MURPHY10_GROUPS)
    shapes = ["1101", "111", "11011"]
    qs, ss = related_set(seed, 4, 10)
    qres, qoff = flat(qs)
    sres, soff = flat(ss)
    bias = np.concatenate([O.composition_bias(q, BLOSUM62) for q in qs]).astype(np.int8)
    thr = [40, 55, 40, 70]
    regs, ctr, surv = OS.scan(qres, qoff, bias, sres, soff, BLOSUM62, amap, asize, shapes, thr, min_ids=6, kx=2,
                              hsp_threshold=30, window_slack=8, max_regions_per_subject=2, cap=6,
                              log_survivors=True)
    # replay per (subject, query) group
    want = [[] for _ in qs]
    ext = av = 0
    rows = [tuple(int(x) for x in r) for r in surv]
    k = 0
    while k < len(rows):
        qi, si = rows[k][0], rows[k][1]
        j = k
        while j < len(rows) and rows[j][:2] == (qi, si):
            j += 1
        q, s = qs[qi], ss[si]
        b = bias[int(qoff[qi]):int(qoff[qi + 1])]
        hits = [(r[2], r[3]) for r in rows[k:j]]
        xres = [O.xdrop(q, s, b, BLOSUM62, qp, sp, 20)[:3] for qp, sp in hits]
        hsps, e, a = H.visit_hits(len(q), len(s), hits, xres, 30)
        ext += e
        av += a
        for r in H.finish_subject(si, hsps, 8, 30, thr[qi], 2):
            want[qi].append(r)
        k = j
    assert ext == ctr["ungapped_extensions"] and av == ctr["extensions_avoided"]
    assert ctr["ungapped_extensions"] > 10
    for qi in range(len(qs)):
        w = H.cap_regions(want[qi], 6)
        g = regs[qi]
        assert [(r["subject"], r["d_min"], r["d_max"], r["best"], r["n"], r["sum"]) for r in w] == \
               [(int(r["subject"]), int(r["d_min"]), int(r["d_max"]), int(r["best"]), int(r["n_hsp"]),
                 int(r["hsp_sum"])) for r in g]


def test_auto_occ_cap_closed_form():
    L = OS._bind()
    for total, ncodes, db, ref in [(0, 10**7, 10**6, 40 * 10**6), (10**9, 10**7, 10**8, 4 * 10**7),
                                   (147 * 10**6, 10**7, 147 * 10**6, 40 * 10**6), (5, 3, 10, 0)]:
        floor = 256 if not (ref > 0 and db > ref) else math.ceil(256 * db / ref)
        assert L.or_auto_occ_cap(total, ncodes, db, ref) == max(math.ceil(total / ncodes * 64), floor)
    # the reference's at-scale run reports caps of 940 (k16 stats: 146 837 350 residues)
    assert L.or_auto_occ_cap(0, 10**7, 146_837_350, 40_000_000) == 940


def test_composition_bias_library_matches_oracle():
    r = Lcg(99)
    cases = [np.zeros(0, np.uint8), np.array([3], np.uint8), np.array([5, 5], np.uint8),
             np.full(80, 5, np.uint8), np.array([22, 23, 24, 200, 0, 1], np.uint8)]
    cases += [rand_seq(r, 1 + r.next(400), 24 if k % 3 == 0 else 20) for k in range(40)]
    for q in cases:
        assert np.array_equal(composition_bias(q, BLOSUM62), O.composition_bias(q, BLOSUM62)), q


def test_default_seed_scheme_is_the_default_preset():
    amap, asize = alphabet_map(MURPHY10_GROUPS)
    assert asize == 10 and len(DS16) == 16
    assert all(len([c for c in s if c == "1"]) == 7 for s in DS16)
