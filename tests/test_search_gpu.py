"""The GPU search (saber_hip_search_run: seed tables, window filter, survivor
order, X-drop, coverage and regions on the device; cap, plans, thresholds,
E-values and the reporting loop on the host; scoring and traceback passes on
the device) against the plain CPU search of oracle/search.py on the same
inputs: per query the reported hits -- (subject, score, q/s coordinates,
CIGAR, exact) as in tests/engine_consistency.rs:113-117 and 295-333, plus
E-value and bit score -- identical, and the engine's counters identical.

Cases: the small blastp workload (100 queries x 5000 subjects, 30 % planted
homologues at 25-60 % identity, Murphy-10 with the 16 weight-7 seeds, window
filter 27, masking off) and variants: composition bias off, no window
filter, a small candidate cap, forced over-frequent codes, several HSPs per
subject, a stricter region threshold, the full alphabet with contiguous and
spaced weight-4 seeds, single-pass scoring and window scoring.
"""
import math

import numpy as np
import pytest

import saber_amd as S
from oracle import search as OS
from saber_amd import BLOSUM62, search, stats, workload
from saber_amd.search import ScanSeeds, full_alphabet

pytestmark = pytest.mark.gpu

COMPARED = ["seed_positions", "seeds_frequent_skipped", "scan_candidates", "hits", "hits_filtered", "scan_groups",
            "ungapped_extensions", "extensions_avoided", "xdrop_steps", "hsps", "regions_weak_dropped",
            "regions_dropped_by_cap", "regions", "regions_below_evalue", "band_scored_regions", "band_rescored",
            "tracebacks", "band_widenings", "band_capped", "end_anchored_tracebacks", "alignments"]


@pytest.fixture(scope="module")
def sets():
    return {"c0": workload.search_set(seed=0), "small": workload.search_set(seed=5, nq=40, nsub=1500)}


def run_both(ctx, d, seeds=None, **kw):
    seeds = seeds or ScanSeeds.default()
    k = stats.gapped_protein(62, 11, 1)
    ns = len(d["soff"]) - 1
    st = stats.SearchStatistics(k, int(d["soff"][-1]), ns, stats.PER_SUBJECT)
    ctx.upload_db(d["sres"], d["soff"])
    names = dict(max_target_seqs="max_target_seqs", max_hsps="max_hsps", hsp_floor_fraction="hsp_floor_fraction",
                 min_ids="hit_filter_min_ids", kx="hit_filter_exact_weight", max_candidates="max_candidates",
                 comp_bias="comp_bias", seed_max_occ="seed_max_occ", score_band="score_band",
                 prefilter_band="score_prefilter_band")
    params = search.default_params(**{names[a]: (int(v) if not isinstance(v, float) else v) for a, v in kw.items()})
    print(f"# search {len(d['qoff']) - 1} x {ns} {kw}", flush=True)
    got = search.run(ctx, d["qres"], d["qoff"], S.Scheme(BLOSUM62, 11, 1), k, seeds=seeds, params=params, regions=True)
    print("#   gpu stages (ms):", {a: round(b, 2) for a, b in got.timing.items()}, flush=True)
    want, wctr = OS.search(d["qres"], d["qoff"], d["sres"], d["soff"], BLOSUM62, 11, 1, st, seeds.alpha_map,
                           seeds.alpha_size, seeds.shapes, **kw)
    return got, want, wctr


def check(got, want, wctr):
    assert len(got.hits) == len(want)
    nh = 0
    for q, (g, w) in enumerate(zip(got.hits, want)):
        gk = [h.key() for h in g]
        wk = [OS.hit_key(h) for h in w]
        assert gk == wk, f"query {q}"
        for a, b in zip(g, w):
            assert a.evalue == b["evalue"] and a.bitscore == b["bitscore"], f"query {q}"
            assert (a.matches, a.mismatches, a.positives, a.gap_opens, a.gap_columns, a.columns) == \
                   (b["matches"], b["mismatches"], b["positives"], b["gap_opens"], b["gap_columns"], b["columns"])
        nh += len(g)
    for k in COMPARED:
        assert got.counters[k] == wctr[k], (k, got.counters[k], wctr[k])
    return nh


def test_small_blastp_workload(ctx, sets):
    got, want, wctr = run_both(ctx, sets["c0"])
    nh = check(got, want, wctr)
    assert nh > 900 and got.counters["hits_filtered"] > 0 and got.counters["extensions_avoided"] > 0
    # regions carry their plan and exact score
    r = np.concatenate(got.regions)
    assert len(r) == got.counters["regions"] and ((r["flags"] & 4) != 0).sum() == nh


@pytest.mark.parametrize("kw", [
    dict(comp_bias=False),
    dict(min_ids=0, kx=0),
    dict(max_candidates=7),
    dict(seed_max_occ=3),
    dict(max_target_seqs=4, max_hsps=3),
    dict(hsp_floor_fraction=0.95),
    dict(prefilter_band=0),
    dict(score_band=0, prefilter_band=0),
])
def test_variants(ctx, sets, kw):
    got, want, wctr = run_both(ctx, sets["small"], **kw)
    check(got, want, wctr)
    if "seed_max_occ" in kw:
        assert got.counters["seeds_frequent_skipped"] > 0
    if "max_candidates" in kw:
        assert got.counters["regions_dropped_by_cap"] > 0


def test_full_alphabet_contiguous_and_spaced(ctx, sets):
    m, n = full_alphabet()
    got, want, wctr = run_both(ctx, sets["small"], seeds=ScanSeeds(m, n, ["1111", "11011"]), min_ids=12, kx=1)
    assert check(got, want, wctr) > 0


def test_empty_and_short_queries(ctx, sets):
    d = dict(sets["small"])
    qs = [d["qres"][int(d["qoff"][i]):int(d["qoff"][i + 1])] for i in range(5)]
    qs = [np.zeros(0, np.uint8), np.array([3, 4], np.uint8)] + qs
    off = np.zeros(len(qs) + 1, np.uint64)
    np.cumsum([len(x) for x in qs], out=off[1:])
    d["qres"] = np.concatenate(qs).astype(np.uint8)
    d["qoff"] = off
    got, want, wctr = run_both(ctx, d)
    check(got, want, wctr)
    assert got.hits[0] == [] and got.hits[1] == []


def test_many_survivors_regions_and_counters(ctx):
    """More survivors than one pass of the device grids (no window filter,
    30 000 subjects): the capped regions of every query and the scan-stage
    counters against the oracle's scan (or_scan) directly."""
    d = workload.search_set(seed=9, nq=60, nsub=30000)
    seeds = ScanSeeds.default()
    k = stats.gapped_protein(62, 11, 1)
    ctx.upload_db(d["sres"], d["soff"])
    params = search.default_params(hit_filter_min_ids=0, hit_filter_exact_weight=0)
    got = search.run(ctx, d["qres"], d["qoff"], S.Scheme(BLOSUM62, 11, 1), k, seeds=seeds, params=params, regions=True)
    nq, ns = len(d["qoff"]) - 1, len(d["soff"]) - 1
    st = stats.SearchStatistics(k, int(d["soff"][-1]), ns, stats.PER_SUBJECT)
    re = 10.0
    thr, bias = [], []
    for q in range(nq):
        qq = d["qres"][int(d["qoff"][q]):int(d["qoff"][q + 1])]
        bias.append(search.composition_bias(qq, BLOSUM62))
        rf = st.for_query(len(qq)).min_score(re)
        thr.append(max(40, int(math.ceil(rf * 0.65))) if rf > 0 else 40)
    regs, ctr, _ = OS.scan(d["qres"], d["qoff"], np.concatenate(bias), d["sres"], d["soff"], BLOSUM62,
                           seeds.alpha_map, seeds.alpha_size, seeds.shapes, thr, min_ids=0, kx=0)
    assert ctr["hits"] - ctr["hits_filtered"] > (1 << 20)
    for key in ("seed_positions", "scan_candidates", "hits", "hits_filtered", "scan_groups", "ungapped_extensions",
                "extensions_avoided", "xdrop_steps", "hsps", "regions_weak_dropped", "regions"):
        assert got.counters[key] == ctr[key], (key, got.counters[key], ctr[key])
    for q in range(nq):
        g = got.regions[q]
        w = regs[q]
        assert [(int(r["subject"]), int(r["d_min"]), int(r["d_max"]), int(r["best"]), int(r["n_hsp"]),
                 int(r["hsp_sum"])) for r in g] == \
               [(int(r["subject"]), int(r["d_min"]), int(r["d_max"]), int(r["best"]), int(r["n_hsp"]),
                 int(r["hsp_sum"])) for r in w], f"query {q}"
